"""GPU parity of the igrid spectral operators (2D fft/ifft, mTimes_ik,
dealias) and of the periodic PadePoisson projection / divergence check
against the CPU oracle (oracle/spectral.py), through the C ABI.

Bars: FP64 relative L2 <= 1e-12.  FP32 storage (FFT in FP32; the staggered
solves with one FP64 residual refinement): <= 2e-6 against the oracle on the
same FP32-rounded input."""
import numpy as np
import pytest

import padeops_amd as P
from fields import rel_l2
from oracle.spectral import PadePoissonPeriodic, Spectral

pytestmark = pytest.mark.gpu

L = 2 * np.pi
F32_BAR = 2e-6


def _dev(a, cuda, single=False):
    import torch
    a = np.asarray(a)
    flat = np.ascontiguousarray(a.ravel(order="F"))
    if np.iscomplexobj(a):
        return torch.from_numpy(flat.astype(np.complex64 if single else np.complex128)).to(cuda)
    return torch.from_numpy(flat.astype(np.float32 if single else np.float64)).to(cuda)


def _host(t, shape):
    import torch
    torch.cuda.synchronize()
    x = t.detach().cpu()
    x = x.to(torch.complex128) if x.is_complex() else x.double()
    return x.numpy().reshape(shape, order="F")


def _r32(a):
    a = np.asarray(a)
    if np.iscomplexobj(a):
        return a.real.astype(np.float32).astype(np.float64) + 1j * a.imag.astype(np.float32).astype(np.float64)
    return a.astype(np.float32).astype(np.float64)


def _pair(nx, ny, nz, periodic_z=False, fix=False):
    dx, dy, dz = L / nx, L / ny, L / nz
    s = P.spectral()
    assert s.init(nx, ny, nz, dx, dy, dz, fixOddball=fix, init_periodicInZ=periodic_z) == 0
    return s, Spectral(nx, ny, nz, dx, dy, dz, fix_oddball=fix, periodic_z=periodic_z)


@pytest.mark.parametrize("shape,pz,fix", [((32, 16, 6), False, False), ((64, 32, 8), True, False),
                                          ((16, 8, 8), True, True)])
def test_wavenumbers_and_masks(cuda, shape, pz, fix):
    s, o = _pair(*shape, periodic_z=pz, fix=fix)
    k1, k2, k3, g1, g2, g3 = s.wavenumbers()
    for a, b in ((k1, o.k1), (k2, o.k2), (k3, o.k3), (g1, o.g1), (g2, o.g2), (g3, o.g3)):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("shape", [(32, 16, 6), (64, 32, 3), (16, 8, 1)])
@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_fft_ifft(cuda, shape, prec):
    import torch
    nx, ny, nz = shape
    s, o = _pair(nx, ny, nz)
    single = prec != "f64"
    u = np.random.default_rng(nx + nz).standard_normal(shape)
    if single:
        u = _r32(u)
    fh = torch.empty(s.nxh * ny * nz, dtype=torch.complex64 if single else torch.complex128, device=cuda)
    assert s.fft(_dev(u, cuda, single), fh, prec=prec) == 0
    ref = o.fft(u)
    err = rel_l2(_host(fh, ref.shape), ref)
    bar = 1e-12 if not single else F32_BAR
    print("fft", shape, prec, err)
    assert err <= bar
    # inverse of a given spectrum (its own input is left unchanged), with and without the oddball
    g = ref * (1 + 0.1j)
    gd = _dev(g, cuda, single)
    g0 = gd.clone()
    out = torch.empty(nx * ny * nz, dtype=torch.float32 if single else torch.float64, device=cuda)
    for odd in (False, True):
        assert s.ifft(gd, out, setOddball=odd, prec=prec) == 0
        r = o.ifft(_r32(g) if single else g, set_oddball=odd)
        e = rel_l2(_host(out, shape), r)
        print("ifft", shape, prec, odd, e)
        assert e <= bar
    assert torch.equal(gd, g0)


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_mtimes_ik_and_dealias(cuda, prec):
    import torch
    nx, ny, nz = 32, 16, 5
    s, o = _pair(nx, ny, nz)
    single = prec != "f64"
    rng = np.random.default_rng(9)
    f = rng.standard_normal((s.nxh, ny, nz)) + 1j * rng.standard_normal((s.nxh, ny, nz))
    if single:
        f = _r32(f)
    bar = 1e-12 if not single else F32_BAR
    for dir in (1, 2):
        fd = _dev(f, cuda, single)
        out = torch.empty_like(fd)
        assert s.mTimes_ik(dir, fd, out, prec=prec) == 0
        assert rel_l2(_host(out, f.shape), o.mtimes_ik(f, dir)) <= bar
        assert s.mTimes_ik(dir, fd, prec=prec) == 0          # in place
        assert rel_l2(_host(fd, f.shape), o.mtimes_ik(f, dir)) <= bar
    fd = _dev(f, cuda, single)
    assert s.dealias(fd, prec=prec) == 0
    assert rel_l2(_host(fd, f.shape), o.dealias(f)) <= bar
    # the 2D case leaves edge fields alone
    fe = _dev(np.concatenate([f, f[:, :, :1]], axis=2), cuda, single)
    fe0 = fe.clone()
    assert s.dealias_edgeField(fe, prec=prec) == 0
    assert torch.equal(fe, fe0)


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_dealias_periodic_z(cuda, prec):
    nx, ny, nz = 32, 16, 16
    s, o = _pair(nx, ny, nz, periodic_z=True)
    single = prec != "f64"
    rng = np.random.default_rng(4)
    f = rng.standard_normal((s.nxh, ny, nz)) + 1j * rng.standard_normal((s.nxh, ny, nz))
    if single:
        f = _r32(f)
    bar = 1e-12 if not single else F32_BAR
    fd = _dev(f, cuda, single)
    assert s.dealias(fd, prec=prec) == 0
    assert rel_l2(_host(fd, f.shape), o.dealias(f)) <= bar
    fe = np.concatenate([f, f[:, :, :1] * 0.5], axis=2)
    fed = _dev(fe, cuda, single)
    assert s.dealias_edgeField(fed, prec=prec) == 0
    assert rel_l2(_host(fed, fe.shape), o.dealias_edge(fe)) <= bar


def _fields(nx, ny, nz, seed):
    rng = np.random.default_rng(seed)
    u, v = rng.standard_normal((2, nx, ny, nz))
    w = rng.standard_normal((nx, ny, nz + 1))
    w[:, :, nz] = w[:, :, 0]
    return u, v, w


@pytest.mark.parametrize("shape", [(32, 16, 32), (16, 32, 64), (64, 64, 16)])
@pytest.mark.parametrize("prec", ["f64", "f32_refine"])
def test_padepoisson_projection(cuda, shape, prec):
    import torch
    nx, ny, nz = shape
    s, o = _pair(nx, ny, nz)
    op = PadePoissonPeriodic(o, L / nx, L / ny, L / nz)
    pp = P.padepoisson()
    assert pp.init(s, L / nz) == 0
    single = prec != "f64"
    u, v, w = _fields(nx, ny, nz, sum(shape))
    uh, vh = o.fft(u), o.fft(v)
    wh = np.fft.fft(np.fft.rfft(w, axis=0), axis=1)
    if single:
        uh, vh, wh = _r32(uh), _r32(vh), _r32(wh)
    ud, vd, wd = _dev(uh, cuda, single), _dev(vh, cuda, single), _dev(wh, cuda, single)
    div = torch.empty(nx * ny * nz, dtype=torch.float32 if single else torch.float64, device=cuda)
    # divergence check before the projection
    assert pp.DivergenceCheck(ud, vd, wd, div, prec=prec) == 0
    dref = op.divergence(uh, vh, wh)
    e0 = rel_l2(_host(div, (nx, ny, nz)), dref)
    assert pp.PressureProjection(ud, vd, wd, prec=prec) == 0
    ru, rv, rw = op.projection(uh, vh, wh)
    errs = [rel_l2(_host(ud, uh.shape), ru), rel_l2(_host(vd, vh.shape), rv), rel_l2(_host(wd, wh.shape), rw)]
    assert pp.DivergenceCheck(ud, vd, wd, div, prec=prec) == 0
    maxdiv = float(div.abs().max())
    print("projection", shape, prec, "div-check err", e0, "u,v,w err", errs, "max div after", maxdiv)
    bar = 1e-12 if not single else F32_BAR
    assert e0 <= bar and max(errs) <= bar
    assert maxdiv < (1e-11 if not single else 1e-4) * np.abs(dref).max()
