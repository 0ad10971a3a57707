"""GPU parity of the cgrid compressible right-hand side (pdo_cgrid_*) against
the oracle's restatement (oracle/cgrid.py Cgrid): getRHS with getLAD, get_tau,
get_q and the fluxes, and the cgrid filter, one species of ideal gas.

  FP64: every RHS component, the primitives, the transport properties after
        the artificial diffusivities, tau and q to 1e-12 relative L2 (per
        field, against its own norm);
  FP32 storage modes against the FP64 GPU result on the same FP32-rounded
        state: the RHS to 5e-6 (about 1.3e-6 measured);
  cgrid%filter: cf90 once and Gaussian twice per direction, 1e-12 (FP64).
Cases: triply periodic inviscid (the reference's default) and viscous
(power law), and x reflecting (bc 1, 1) with z one-sided (bc 0, 0)."""
import numpy as np
import pytest

import padeops_amd as P
from fields import rel_l2
from oracle.cgrid import Cgrid

pytestmark = pytest.mark.gpu

SHAPE = (32, 24, 16)
CASES = {
    "periodic-inviscid": dict(periodic=(True, True, True), bc=((0, 0), (0, 0), (0, 0)), inviscid=True),
    "periodic-viscous": dict(periodic=(True, True, True), bc=((0, 0), (0, 0), (0, 0)), inviscid=False,
                             mu_ref=0.02, T_ref=0.9, visc_n=0.7, Pr=0.72),
    "walls-viscous": dict(periodic=(False, True, False), bc=((1, 1), (0, 0), (0, 0)), inviscid=False,
                          mu_ref=0.02, T_ref=0.9, visc_n=0.7, Pr=0.72),
}


def _state(seed):
    nx, ny, nz = SHAPE
    L = (2 * np.pi, 1.6 * np.pi, 1.2 * np.pi)
    X = np.meshgrid(*[np.arange(SHAPE[i]) * (L[i] / SHAPE[i]) for i in range(3)], indexing="ij")
    rng = np.random.default_rng(seed)

    def smooth():
        a = rng.uniform(-1, 1, 6)
        return (a[0] * np.sin(X[0] + a[1]) * np.cos(2 * X[1] * 1.25) + a[2] * np.cos(X[2] * 5 / 3 + X[1] * 1.25)
                + a[3] * np.sin(2 * X[0] - X[2] * 5 / 3) + a[4] * np.cos(X[1] * 1.25 + a[5]))
    rho = 1.0 + 0.15 * smooth()
    u, v, w = 0.3 * smooth(), 0.3 * smooth(), 0.3 * smooth()
    p = 1.0 + 0.1 * smooth()
    e = p / (0.4 * rho)
    W = np.stack([rho, rho * u, rho * v, rho * w, rho * (e + 0.5 * (u * u + v * v + w * w))], axis=-1)
    return W, tuple(L[i] / SHAPE[i] for i in range(3))


def _dev(a, cuda, dt):
    import torch
    return torch.from_numpy(np.ascontiguousarray(np.asarray(a).ravel(order="F"))).to(device=cuda, dtype=dt)


def _field(cg, name, ncomp, dt):
    import ctypes
    import torch
    n = int(np.prod(SHAPE)) * ncomp
    out = torch.empty(n, dtype=dt)
    W = 8 if dt == torch.float64 else 4
    assert P._lib.lib().pdo_memcpy_d2h(ctypes.c_void_p(out.data_ptr()), ctypes.c_void_p(cg.field(name)),
                                       n * W) == 0
    return out.double().numpy().reshape(SHAPE + ((ncomp,) if ncomp > 1 else ()), order="F")


def _gpu(case, W, d, prec, cuda):
    import torch
    kw = dict(CASES[case])
    cg = P.cgrid()
    assert cg.init(SHAPE, d, prec=prec, **kw) == 0
    dt = torch.float64 if prec == "f64" else torch.float32
    Wd = _dev(W, cuda, dt)
    rhs = torch.empty_like(Wd)
    assert cg.getRHS(Wd, rhs) == 0
    torch.cuda.synchronize()
    return cg, rhs.double().cpu().numpy().reshape(SHAPE + (5,), order="F"), dt


@pytest.mark.parametrize("case", list(CASES))
def test_cgrid_rhs_fp64(cuda, case):
    W, d = _state(3)
    ref_rhs, parts = Cgrid(SHAPE, d, **CASES[case]).rhs(W, parts=True)
    cg, rhs, dt = _gpu(case, W, d, "f64", cuda)
    for c in range(5):
        e = rel_l2(rhs[..., c], ref_rhs[..., c])
        print(case, "rhs", c, e)
        assert e <= 1e-12
    for name in ("rho", "u", "v", "w", "e", "p", "T", "mu", "bulk", "kap"):
        ref = parts[name]
        got = _field(cg, name, 1, dt)
        if np.abs(ref).max() == 0:
            assert np.abs(got).max() == 0, name
            continue
        e = rel_l2(got, ref)
        print(case, name, e)
        assert e <= 1e-12, name
    tau = _field(cg, "tau", 6, dt)
    for k in range(6):
        assert rel_l2(tau[..., k], parts["tau"][k]) <= 1e-12
    g = _field(cg, "duidxj", 9, dt)
    for k in range(9):
        assert rel_l2(g[..., k], parts["grad"][k]) <= 1e-12
    q = _field(cg, "q", 3, dt)
    for k in range(3):
        assert rel_l2(q[..., k], parts["q"][k]) <= 1e-12


@pytest.mark.parametrize("prec", ["f32_refine", "f32_store"])
def test_cgrid_rhs_fp32(cuda, prec):
    W, d = _state(4)
    W32 = W.astype(np.float32).astype(np.float64)
    _, r64, _ = _gpu("periodic-viscous", W32, d, "f64", cuda)
    _, r32, _ = _gpu("periodic-viscous", W32, d, prec, cuda)
    for c in range(5):
        e = rel_l2(r32[..., c], r64[..., c])
        print(prec, "rhs", c, e)
        assert e <= 5e-6


@pytest.mark.parametrize("case", ["periodic-inviscid", "walls-viscous"])
def test_cgrid_filter(cuda, case):
    import torch
    W, d = _state(5)
    f = W[..., 4]
    ref = Cgrid(SHAPE, d, **CASES[case])
    cg = P.cgrid()
    assert cg.init(SHAPE, d, prec="f64", **CASES[case]) == 0
    for gaussian, times in ((False, 1), (True, 2), (False, 2)):
        fd = _dev(f, cuda, torch.float64)
        assert cg.filter(fd, numtimes=times, gaussian=gaussian) == 0
        torch.cuda.synchronize()
        got = fd.cpu().numpy().reshape(SHAPE, order="F")
        e = rel_l2(got, ref.filter(f, "gaussian" if gaussian else "cf90", times))
        print(case, gaussian, times, e)
        assert e <= 1e-12


def _state_at(shape, seed):
    """the smooth state of _state on another grid of the same box"""
    L = (2 * np.pi, 1.6 * np.pi, 1.2 * np.pi)
    X = np.meshgrid(*[np.arange(shape[i]) * (L[i] / shape[i]) for i in range(3)], indexing="ij")
    rng = np.random.default_rng(seed)

    def smooth():
        a = rng.uniform(-1, 1, 6)
        return (a[0] * np.sin(X[0] + a[1]) * np.cos(2 * X[1] * 1.25) + a[2] * np.cos(X[2] * 5 / 3 + X[1] * 1.25)
                + a[3] * np.sin(2 * X[0] - X[2] * 5 / 3) + a[4] * np.cos(X[1] * 1.25 + a[5]))
    rho = 1.0 + 0.15 * smooth()
    u, v, w = 0.3 * smooth(), 0.3 * smooth(), 0.3 * smooth()
    p = 1.0 + 0.1 * smooth()
    e = p / (0.4 * rho)
    W = np.stack([rho, rho * u, rho * v, rho * w, rho * (e + 0.5 * (u * u + v * v + w * w))], axis=-1)
    return W, tuple(L[i] / shape[i] for i in range(3))


def test_cgrid_fp32_error_is_not_the_artificial_terms(cuda):
    """Where the FP32+refine RHS error comes from, at 128^3: it is the same
    with the artificial coefficients off (Cmu = Cbeta = Ckap = 0) as with them
    on (measured 9.95e-6 and 9.94e-6), so it is not the LAD terms' fourth
    derivatives.  It comes from differentiating FP32-stored intermediates
    (the fluxes, and the viscous stresses and heat flux, which are
    derivatives of stored first derivatives): their storage rounding is
    lifted by the grid-scale wavenumber, so the error grows with the grid
    (1.3e-6 at 32 x 24 x 16, 1e-5 at 128^3)."""
    import torch
    shape = (128, 128, 128)
    W, d = _state_at(shape, 4)
    W32 = W.astype(np.float32).astype(np.float64)
    errs = {}
    for lad in (True, False):
        kw = dict(CASES["periodic-viscous"])
        if not lad:
            kw.update(Cmu=0.0, Cbeta=0.0, Ckap=0.0)
        out = {}
        for p in ("f64", "f32_refine"):
            cg = P.cgrid()
            assert cg.init(shape, d, prec=p, **kw) == 0
            dt = torch.float64 if p == "f64" else torch.float32
            Wd = _dev(W32, cuda, dt)
            rhs = torch.empty_like(Wd)
            assert cg.getRHS(Wd, rhs) == 0
            torch.cuda.synchronize()
            out[p] = rhs.double().cpu().numpy().reshape(shape + (5,), order="F")
            del cg, Wd, rhs
        errs[lad] = max(rel_l2(out["f32_refine"][..., c], out["f64"][..., c]) for c in range(5))
    print("128^3 FP32+refine vs FP64, LAD on", errs[True], "off", errs[False])
    assert errs[True] <= 2e-5
    assert abs(errs[True] - errs[False]) <= 0.1 * errs[False]
