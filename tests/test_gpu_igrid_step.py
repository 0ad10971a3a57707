"""GPU parity of the igrid time advancement (pdo_igrid_*) against the oracle's
restatement (oracle/igrid.py) on triply periodic grids: TVD-RK3 and SSP-RK45
steps of the rotational nonlinear term, the viscous term, HIT shell forcing
and the Smagorinsky SGS stress, with the CFL time step.

FP64: after two steps the spectral state, the physical velocity and the
time step agree with the oracle to relative L2 <= 1e-11 (two steps of five
stages each, every stage a projection and some 60 transforms and compact
solves).  FP32 + refinement: against the FP64 GPU run <= 1e-5.  The velocity
stays divergence-free (PadePoisson DivergenceCheck) to 1e-10 relative."""
import numpy as np
import pytest

import padeops_amd as P
from fields import isotropic_velocity, rel_l2
from oracle.igrid import Igrid

pytestmark = pytest.mark.gpu

FORCING = dict(kmin=1.0, kmax=3.0, nwaves=5, eps=0.1, seed=42)
SGS = dict(model="smagorinsky", explicit_edge=False, cglobal=0.01)
CASES = [(16, 16, 16, 2, True, None), (32, 16, 32, 2, True, SGS), (16, 16, 16, 1, False, None),
         (32, 16, 32, 1, True, SGS)]


def _initial(nx, ny, nz):
    u, v, w = isotropic_velocity(max(nx, ny, nz), k0=3.0, seed=3)
    u, v, w = (a[:nx, :ny, :nz] for a in (u, v, w))
    wE = np.concatenate([w, w[:, :, :1]], axis=2)
    return np.asfortranarray(u), np.asfortranarray(v), np.asfortranarray(wE)


def _dev(a, cuda, dt):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a.ravel(order="F"))).to(device=cuda, dtype=dt)


def _host(t, shape):
    import torch
    torch.cuda.synchronize()
    x = t.detach().cpu()
    x = x.to(torch.complex128) if x.is_complex() else x.double()
    return x.numpy().reshape(shape, order="F")


def _run_gpu(nx, ny, nz, scheme, forcing, sgs, prec, cuda, nsteps):
    import torch
    L = 2 * np.pi
    g = P.igrid()
    assert g.init(nx, ny, nz, L, L, L, Re=100.0, scheme=scheme, CFL=0.5, sgs=sgs,
                  forcing=FORCING if forcing else None, prec=prec) == 0
    dt = torch.float64 if prec == "f64" else torch.float32
    u, v, w = _initial(nx, ny, nz)
    assert g.set_velocity(_dev(u, cuda, dt), _dev(v, cuda, dt), _dev(w, cuda, dt)) == 0
    assert g.step(nsteps) == 0
    nxh = nx // 2 + 1
    out = {k: _host(g.field(k), s) for k, s in (("uhat", (nxh, ny, nz)), ("vhat", (nxh, ny, nz)),
                                                 ("what", (nxh, ny, nz + 1)), ("u", (nx, ny, nz)),
                                                 ("v", (nx, ny, nz)), ("w", (nx, ny, nz + 1)))}
    return g, out


@pytest.mark.parametrize("case", CASES, ids=["%dx%dx%d-rk%d-f%d-%s" % (c[0], c[1], c[2], c[3], c[4], bool(c[5]))
                                             for c in CASES])
def test_igrid_steps_fp64(cuda, case):
    nx, ny, nz, scheme, forcing, sgs = case
    L = 2 * np.pi
    ref = Igrid(nx, ny, nz, L, L, L, Re=100.0, scheme=scheme, CFL=0.5,
                sgs=None if sgs is None else dict(model=0, explicit_edge=False, cglobal=sgs["cglobal"]),
                forcing=FORCING if forcing else None)
    ref.set_velocity(*_initial(nx, ny, nz))
    for _ in range(2):
        ref.step()
    print("case", case, flush=True)
    g, out = _run_gpu(nx, ny, nz, scheme, forcing, sgs, "f64", cuda, 2)
    t, s, dt = g.time()
    assert s == 2
    assert abs(dt - ref.dt) <= 1e-13 * ref.dt
    assert abs(t - ref.tsim) <= 1e-13 * ref.tsim
    refs = dict(uhat=ref.state[0], vhat=ref.state[1], what=ref.state[2], u=ref.f["u"], v=ref.f["v"], w=ref.f["w"])
    for k, r in refs.items():
        e = rel_l2(out[k], r)
        print(k, e)
        assert e <= 1e-11, k
    # a stable run: the velocity stays of the initial field's size
    assert np.abs(out["u"]).max() <= 2 * np.abs(_initial(nx, ny, nz)[0]).max()
    div = ref.pp.divergence(out["uhat"], out["vhat"], out["what"])
    print("max |div| / max |u|", np.abs(div).max() / np.abs(out["u"]).max())
    assert np.abs(div).max() <= 1e-10 * np.abs(out["u"]).max()


def test_igrid_fp32_refine(cuda):
    nx, ny, nz = 32, 16, 32
    _, o64 = _run_gpu(nx, ny, nz, 2, True, SGS, "f64", cuda, 2)
    _, o32 = _run_gpu(nx, ny, nz, 2, True, SGS, "f32_refine", cuda, 2)
    for k in ("uhat", "vhat", "what", "u", "v", "w"):
        e = rel_l2(o32[k], o64[k])
        print(k, e)
        assert e <= 1e-5, k


def test_restart_round_trip(cuda, tmp_path):
    """dumpRestartFile then readRestartFile into a new object: the same
    physical fields (raw FP64 files of the reference's names and sizes), the
    time from the info record (5 significant digits, G15.5), and a next step
    identical to that of the object that never stopped."""
    import os
    nx, ny, nz = 32, 16, 32
    g, _ = _run_gpu(nx, ny, nz, 2, True, None, "f64", cuda, 3)
    t0, s0, _ = g.time()
    assert g.dumpRestartFile(tmp_path, 7) == 0
    for var, n in (("u", nx * ny * nz), ("v", nx * ny * nz), ("w", nx * ny * (nz + 1))):
        name = tmp_path / ("RESTART_Run07_%s.%06d" % (var, s0))
        assert os.path.getsize(name) == 8 * n
    info = (tmp_path / ("RESTART_Run07_info.%06d" % s0)).read_text()
    assert len(info.rstrip("\n")) == 15
    assert abs(float(info) - t0) <= 5e-5 * t0
    h = P.igrid()
    L = 2 * np.pi
    assert h.init(nx, ny, nz, L, L, L, Re=100.0, scheme=2, CFL=0.5, forcing=FORCING, prec="f64") == 0
    assert h.readRestartFile(tmp_path, s0, 7) == 0
    t1, s1, _ = h.time()
    assert s1 == s0 and t1 == float(info)
    u_file = np.fromfile(tmp_path / ("RESTART_Run07_u.%06d" % s0), dtype=np.float64)
    assert np.array_equal(_host(g.field("u"), (nx * ny * nz,)), u_file)
    for k in ("u", "v", "w", "uhat"):
        a = _host(g.field(k), (-1,))
        b = _host(h.field(k), (-1,))
        assert rel_l2(b, a) <= 1e-14, k
    assert g.step(1) == 0 and h.step(1) == 0
    for k in ("u", "v", "w"):
        assert rel_l2(_host(h.field(k), (-1,)), _host(g.field(k), (-1,))) <= 1e-13, k
