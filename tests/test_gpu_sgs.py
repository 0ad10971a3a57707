"""GPU parity of the SGS kernels (Smagorinsky, sigma, AMD, model constants)
and of the standard dynamic procedure against oracle/sgs.py.  Pointwise
kernels: FP64 relative L2 <= 1e-14 (sigma <= 1e-10: its closed-form cubic
roots pass through acos and cos, whose last bits differ between libraries and
are amplified where two singular values nearly coincide); the dynamic
coefficient (planar sums): <= 1e-10; FP32 storage: <= 1e-6."""
import numpy as np
import pytest

import oracle.sgs as G
import padeops_amd as P
from fields import from_dev, rel_l2, to_dev
from oracle.spectral import Spectral

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("model", ["smagorinsky", "sigma", "amd"])
@pytest.mark.parametrize("prec", ["f64", "f32_store"])
def test_sgs_kernel(cuda, model, prec):
    import torch
    shape = (32, 24, 10)
    npts = int(np.prod(shape))
    g = np.random.default_rng(5).standard_normal(shape + (9,))
    dt = torch.float64 if prec == "f64" else torch.float32
    if dt == torch.float32:
        g = g.astype(np.float32).astype(np.float64)
    camd = (0.3, 0.2, 0.5)
    gd = to_dev(np.asfortranarray(g), dt, cuda)
    S = torch.empty(6 * npts, dtype=dt, device=cuda)
    nu = torch.empty(npts, dtype=dt, device=cuda)
    assert P.sgs_kernel(gd, S, nu, npts, model, camd, prec=prec) == 0
    Sref = G.sij(g)
    nref = {"smagorinsky": lambda: G.smagorinsky(Sref), "sigma": lambda: G.sigma(g),
            "amd": lambda: G.amd(g, Sref, *camd)}[model]()
    bar = 1e-14 if prec == "f64" else 1e-6
    assert rel_l2(from_dev(S, shape + (6,)), Sref) <= bar
    nbar = 1e-10 if (model == "sigma" and prec == "f64") else bar * 10
    e = rel_l2(from_dev(nu, shape), nref)
    print(model, prec, "nu", e)
    assert e <= nbar
    cpl = torch.tensor(np.linspace(0.5, 1.5, shape[2]), dtype=torch.float64, device=cuda)
    assert P.sgs_scale(nu, shape[0] * shape[1], shape[2], cpl, prec=prec) == 0
    assert rel_l2(from_dev(nu, shape), nref * np.linspace(0.5, 1.5, shape[2])[None, None, :]) <= nbar
    assert P.sgs_scale(nu, shape[0] * shape[1], shape[2], None, 2.0, prec=prec) == 0


@pytest.mark.parametrize("prec", ["f64", "f32_store"])
def test_dynamic_procedure(cuda, prec):
    import torch
    n, nz = 32, 8
    dx = 2 * np.pi / n
    rng = np.random.default_rng(11)
    sp = Spectral(n, n, nz, dx, dx, 0.1)
    u, v, w = (np.asfortranarray(rng.standard_normal((n, n, nz))) for _ in range(3))
    g = rng.standard_normal((n, n, nz, 9))
    single = prec != "f64"
    if single:
        u, v, w = (a.astype(np.float32).astype(np.float64) for a in (u, v, w))
        g = g.astype(np.float32).astype(np.float64)
    S = G.sij(g)
    D = G.smagorinsky(S)
    if single:
        S = S.astype(np.float32).astype(np.float64)
        D = D.astype(np.float32).astype(np.float64)
    uh, vh, wh = sp.fft(u), sp.fft(v), sp.fft(w)
    if single:
        r = lambda a: a.real.astype(np.float32).astype(np.float64) + 1j * a.imag.astype(np.float32)  # noqa: E731
        uh, vh, wh = r(uh), r(vh), r(wh)
    ref = G.dynamic_procedure(sp, dx, dx, 2.0, u, v, w, uh, vh, wh, S, D)
    dt = torch.float32 if single else torch.float64
    ct = torch.complex64 if single else torch.complex128

    def cdev(a):
        return torch.from_numpy(np.ascontiguousarray(a.ravel(order="F")).astype(
            np.complex64 if single else np.complex128)).to(cuda)
    s = P.spectral()
    assert s.init(n, n, nz, dx, dx, 0.1) == 0
    dp = P.DynamicProcedure()
    assert dp.init(s, 2.0) == 0
    lam = torch.empty(nz, dtype=torch.float64, device=cuda)
    assert dp.compute(to_dev(u, dt, cuda), to_dev(v, dt, cuda), to_dev(w, dt, cuda), cdev(uh), cdev(vh), cdev(wh),
                      to_dev(np.asfortranarray(S), dt, cuda), to_dev(D, dt, cuda), lam, prec=prec) == 0
    torch.cuda.synchronize()
    got = lam.cpu().numpy()
    print(prec, got, ref)
    assert rel_l2(got, ref) <= (1e-10 if not single else 1e-5)
    assert ct is not None
