"""GPU tests of the igrid time advancement with walls in z (pdo_igrid_* with
walls): PeriodicInZ = .false. with the boundary kinds of botWall / topWall,
the PadePoisson of the 2 nz extensions, the SGS wall model and the mean
pressure gradient.

  * parity: two steps against the oracle's restatement (oracle/igrid.py with
    walls), FP64, relative L2 <= 1e-11 for the spectra and the physical
    velocity; w is zero on both walls;
  * FP32 storage with refined solves against FP64 after two steps, 1e-5;
  * laminar Poiseuille flow between no-slip walls, from rest: the profile
    converges to the discrete steady state of the oracle's second-derivative
    operator (to 2e-6) and so to the parabola Re dpFdx z (Lz - z) / 2 within
    that operator's second-order error (3.3e-4 at nz = 32);
  * the C5 setting in small: a wall-modelled channel LES (wall model below,
    slip above, dpFdx forcing) with the learned closure in the loop at every
    stage and the dynamic model where the network's dissipation is negative,
    using the network trained on the solver's own 256^3 DNS
    (results/closure_c4_dns256_params.npy): it stays finite, w stays zero on
    both walls, the mean profile stays positive, and the fallback acts."""
import os

import numpy as np
import pytest

import padeops_amd as P
from fields import isotropic_velocity, rel_l2
from oracle.igrid import Igrid
import oracle as O
from oracle.spectral import pade6_apply

pytestmark = pytest.mark.gpu

L = 2 * np.pi
# Smagorinsky's multiplier (Cs Delta)^2, Cs = 0.12, Delta = (dx dy dz)^(1/3) of
# the 32 x 16 x 32 box of height 1: the time step follows the advective CFL
# only (igrid.F90:1374-1400), so a larger constant makes the explicit
# wall-normal diffusion of the eddy viscosity unstable
CG_WALLS = (0.12 * (L / 32 * L / 16 * 1.0 / 32) ** (1.0 / 3.0)) ** 2
CASES = {
    "noslip-noslip-rk45": dict(nx=16, ny=16, nz=32, scheme=2, Re=100.0, walls=dict(botWall=1, topWall=1, dpFdx=1.0),
                               sgs=None),
    "wallmodel-slip-smag-rk45": dict(nx=32, ny=16, nz=32, scheme=2, Re=0.0,
                                     walls=dict(botWall=3, topWall=2, dpFdx=1.0),
                                     sgs=dict(model="smagorinsky", cglobal=CG_WALLS, wall_model="moeng", z0=1e-4)),
    "noslip-noslip-nostokes-rk45": dict(nx=16, ny=16, nz=32, scheme=2, Re=100.0,
                                        walls=dict(botWall=1, topWall=1, dpFdx=1.0, stokes_pressure=False),
                                        sgs=None),
    "slip-slip-stokes-rk3": dict(nx=16, ny=16, nz=16, scheme=1, Re=200.0,
                                 walls=dict(botWall=2, topWall=2, dpFdx=0.5, stokes_pressure=True), sgs=None),
}


def _initial(nx, ny, nz, Lz, mean):
    u, v, w = isotropic_velocity(max(nx, ny, nz), k0=3.0, seed=5)
    u, v, w = (0.2 * a[:nx, :ny, :nz] for a in (u, v, w))
    zc = (np.arange(nz) + 0.5) * Lz / nz
    ze = np.arange(nz + 1) * Lz / nz
    env_c = np.sin(np.pi * zc / Lz)[None, None, :]
    u = mean(zc)[None, None, :] + u * env_c
    v = v * env_c
    wE = np.zeros((nx, ny, nz + 1))
    wE[:, :, :nz] = w
    wE *= np.sin(np.pi * ze / Lz)[None, None, :]
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


def _gpu_igrid(c, Lz, prec="f64", CFL=0.5, dt=0.0):
    g = P.igrid()
    assert g.init(c["nx"], c["ny"], c["nz"], L, L, Lz, Re=c["Re"], scheme=c["scheme"], CFL=CFL, dt=dt,
                  sgs=c["sgs"], walls=c["walls"], prec=prec) == 0
    return g


@pytest.mark.parametrize("case", list(CASES))
def test_walls_steps_fp64(cuda, case):
    import torch
    c = CASES[case]
    nx, ny, nz, Lz = c["nx"], c["ny"], c["nz"], 1.0
    mean = lambda z: 2.0 + np.log1p(20 * z)  # noqa: E731
    u, v, w = _initial(nx, ny, nz, Lz, mean)
    osgs = None
    if c["sgs"] is not None:
        osgs = dict(model=0, explicit_edge=False, cglobal=c["sgs"]["cglobal"], z0=c["sgs"]["z0"])
    ref = Igrid(nx, ny, nz, L, L, Lz, Re=c["Re"], scheme=c["scheme"], CFL=0.5, sgs=osgs, walls=c["walls"])
    ref.set_velocity(u, v, w)
    for _ in range(2):
        ref.step()
    print("case", case, flush=True)
    g = _gpu_igrid(c, Lz)
    assert g.set_velocity(_dev(u, cuda, torch.float64), _dev(v, cuda, torch.float64),
                          _dev(w, cuda, torch.float64)) == 0
    assert g.step(2) == 0
    t, s, dt = g.time()
    assert s == 2 and abs(dt - ref.dt) <= 1e-13 * ref.dt
    nxh = nx // 2 + 1
    shapes = dict(uhat=(nxh, ny, nz), vhat=(nxh, ny, nz), what=(nxh, ny, nz + 1), u=(nx, ny, nz), v=(nx, ny, nz),
                  w=(nx, ny, nz + 1))
    refs = dict(uhat=ref.state[0], vhat=ref.state[1], what=ref.state[2], u=ref.f["u"], v=ref.f["v"], w=ref.f["w"])
    for k, r in refs.items():
        e = rel_l2(_host(g.field(k), shapes[k]), r)
        print(k, e)
        assert e <= 1e-11, k
    wg = _host(g.field("w"), shapes["w"])
    assert np.abs(wg[:, :, 0]).max() <= 1e-13 and np.abs(wg[:, :, nz]).max() <= 1e-13
    # a stable run: the velocity stays of the initial field's size
    assert np.abs(_host(g.field("u"), shapes["u"])).max() <= 2 * np.abs(u).max()


@pytest.mark.parametrize("case", ["noslip-noslip-rk45", "wallmodel-slip-smag-rk45"])
def test_walls_fp32_refine(cuda, case):
    """FP32 storage with refined solves against FP64, two steps from the same
    initial field (the FP32 run takes it rounded): every spectrum and
    physical field to 1e-5 relative L2, w zero on both walls."""
    import torch
    c = CASES[case]
    nx, ny, nz, Lz = c["nx"], c["ny"], c["nz"], 1.0
    mean = lambda z: 2.0 + np.log1p(20 * z)  # noqa: E731
    u, v, w = _initial(nx, ny, nz, Lz, mean)
    nxh = nx // 2 + 1
    shapes = dict(uhat=(nxh, ny, nz), vhat=(nxh, ny, nz), what=(nxh, ny, nz + 1), u=(nx, ny, nz), v=(nx, ny, nz),
                  w=(nx, ny, nz + 1))
    out = {}
    for prec, dt in (("f64", torch.float64), ("f32_refine", torch.float32)):
        g = _gpu_igrid(c, Lz, prec=prec)
        assert g.set_velocity(_dev(u, cuda, dt), _dev(v, cuda, dt), _dev(w, cuda, dt)) == 0
        assert g.step(2) == 0
        out[prec] = {k: _host(g.field(k), sh) for k, sh in shapes.items()}
    for k in shapes:
        e = rel_l2(out["f32_refine"][k], out["f64"][k])
        print(case, k, e)
        assert e <= 1e-5, k
    w32 = out["f32_refine"]["w"]
    assert np.abs(w32[:, :, 0]).max() == 0 and np.abs(w32[:, :, nz]).max() == 0


def test_poiseuille_converges_to_the_discrete_steady_state(cuda):
    import torch
    nx, ny, nz, Re, dp, Lz = 16, 8, 32, 20.0, 1.0, 1.0
    c = dict(nx=nx, ny=ny, nz=nz, scheme=2, Re=Re, walls=dict(botWall=1, topWall=1, dpFdx=dp), sgs=None)
    g = _gpu_igrid(c, Lz, CFL=0.0, dt=0.005)
    z = torch.zeros(nx * ny * nz, dtype=torch.float64, device=cuda)
    assert g.set_velocity(z, z, torch.zeros(nx * ny * (nz + 1), dtype=torch.float64, device=cuda)) == 0
    for _ in range(6):
        assert g.step(1000) == 0
        torch.cuda.synchronize()
        print("t", g.time()[0], flush=True)
    u = _host(g.field("u"), (nx, ny, nz))
    # the steady state of the oracle's operator: Re^-1 d2dz2_C2C(u; odd, odd) = -dpFdx
    dz = Lz / nz
    I = np.eye(nz)
    D = np.stack([pade6_apply(O.STG_D2_C2C, I[:, k].reshape(1, 1, nz).astype(complex), dz, nz, -1, -1,
                              False).real.reshape(nz) for k in range(nz)], axis=1)
    steady = np.linalg.solve(D / Re, -dp * np.ones(nz))
    zc = (np.arange(nz) + 0.5) * dz
    par = 0.5 * Re * dp * zc * (Lz - zc)
    prof = u.mean(axis=(0, 1))
    e_steady = np.abs(prof - steady).max() / par.max()
    e_par = np.abs(prof - par).max() / par.max()
    print("to the discrete steady state", e_steady, "to the parabola", e_par)
    assert np.abs(u - prof[None, None, :]).max() <= 1e-12 * par.max()
    assert e_steady <= 2e-6
    assert e_par <= 4e-4
    assert np.abs(_host(g.field("w"), (nx, ny, nz + 1))).max() <= 1e-12 * par.max()


def test_wall_modelled_les_with_the_learned_closure(cuda):
    import torch
    pfile = os.path.join(os.path.dirname(__file__), "..", "results", "closure_c4_dns256_params.npy")
    params = np.load(pfile)
    nx, ny, nz, Lz, z0, dp = 32, 32, 16, 1.0, 1e-4, 1.0
    c = dict(nx=nx, ny=ny, nz=nz, scheme=2, Re=0.0, walls=dict(botWall=3, topWall=2, dpFdx=dp),
             sgs=dict(model="smagorinsky", cglobal=0.01, wall_model="moeng", z0=z0))
    g = _gpu_igrid(c, Lz, prec="f32_refine", CFL=0.5)
    kappa = 0.4
    ustar0 = np.sqrt(dp * Lz)
    mean = lambda zz: ustar0 / kappa * np.log(zz / z0)  # noqa: E731
    u, v, w = _initial(nx, ny, nz, Lz, mean)
    f32 = torch.float32
    assert g.set_velocity(_dev(u, cuda, f32), _dev(v, cuda, f32), _dev(w, cuda, f32)) == 0
    net = P.SGSNet(64)
    net.set_params(params.astype(np.float32))
    cl = P.Closure(net, nx * ny * (nz + 1))
    # the grid's filter width, against the 64^3 training grid's pi / 21
    delta = (L / nx * L / ny * Lz / nz) ** (1.0 / 3.0)
    assert g.set_closure(2, cl, delta=delta, dratio=delta / (np.pi / 21)) == 0
    assert g.step(40) == 0
    torch.cuda.synchronize()
    uu = _host(g.field("u"), (nx, ny, nz))
    ww = _host(g.field("w"), (nx, ny, nz + 1))
    assert np.isfinite(uu).all() and np.isfinite(ww).all()
    assert np.abs(ww[:, :, 0]).max() == 0 and np.abs(ww[:, :, nz]).max() == 0
    prof = uu.mean(axis=(0, 1))
    print("mean profile", np.round(prof, 3), "t", g.time())
    assert (prof > 0).all()
    nf = cl.fallback_count()
    print("fallback points at the last stage", nf, "of", nx * ny * (nz + 1))
    # the network was trained on isotropic turbulence: in the channel the
    # near-wall shear sets the rms gradient that normalises its inputs, and
    # its dissipation comes out negative at nearly every point, so the
    # dynamic model carries the run
    assert nf > 0


def test_wall_modelled_step_full_size(cuda):
    """The wall-modelled box at its full size (512 x 512 x 256 cells, dz = dx,
    Moeng wall model below, slip wall above, dpFdx = 1, Smagorinsky with
    Cs = 0.12, FP32 + refinement) from the log law plus the C2-spectrum
    turbulence, two SSP-RK45 steps: the fields stay finite and of the initial
    size, w is zero on both walls, and the divergence the step's projection
    leaves is at the FP32 rounding level (1e-7 of max |u| / dx).  This needs
    the Stokes pressure (on by default, as igrid's ComputeStokesPressure):
    without it the projection zeroes the wall w each stage's update leaves,
    and the wall cells keep a divergence of about 1e-4 of that scale."""
    import torch
    from padeops_amd.synthetic import wall_bounded
    nx, nz = 512, 256
    dx = 2 * np.pi / nx
    Lz = nz * dx
    cs = (0.12 * dx) ** 2
    c = dict(nx=nx, ny=nx, nz=nz, scheme=2, Re=0.0, walls=dict(botWall=3, topWall=2, dpFdx=1.0),
             sgs=dict(model="smagorinsky", cglobal=cs, wall_model="moeng", z0=1e-4))
    g = _gpu_igrid(c, Lz, prec="f32_refine")
    uC, _, wE = wall_bounded(nx, nx, nz, 4.0, 1234, torch.float32, cuda)
    wE.view(nz + 1, -1)[0].zero_()
    wE.view(nz + 1, -1)[nz].zero_()
    u0 = float(uC.abs().max())
    assert g.set_velocity(uC, torch.zeros_like(uC), wE) == 0
    del uC, wE
    print("case wall-modelled 512x512x256 f32_refine", flush=True)
    assert g.step(2) == 0
    torch.cuda.synchronize()
    u = g.field("u")
    w = g.field("w").view(nz + 1, -1)
    assert bool(torch.isfinite(u).all()) and bool(torch.isfinite(w).all())
    umax = float(u.abs().max())
    assert umax <= 2 * u0
    assert float(w[0].abs().max()) == 0 and float(w[nz].abs().max()) == 0
    sp = P.spectral()
    assert sp.init(nx, nx, nz, dx, dx, dx, init_periodicInZ=False) == 0
    pp = P.padepoisson()
    assert pp.init(sp, dx, PeriodicInZ=False) == 0
    div = torch.empty(nx * nx * nz, dtype=torch.float32, device=cuda)
    assert pp.DivergenceCheck(g.field("uhat"), g.field("vhat"), g.field("what"), div, prec="f32_refine") == 0
    torch.cuda.synchronize()
    dmax = float(div.abs().max())
    print("t", g.time(), "max |u|", umax, "max |div|", dmax, "scale", umax / dx)
    # measured 2e-8 of max |u| / dx (Stokes pressure on, the default)
    assert dmax <= 1e-7 * umax / dx
