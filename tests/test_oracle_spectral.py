"""Pins of the spectral / PadePoisson oracle (oracle/spectral.py) by what the
mathematics fixes: a brute-force DFT, round trips, analytic derivatives of
single Fourier modes, closed-form dealias mask counts, and the invariants of
a projection (zero divergence, idempotence, gradients removed, solenoidal
fields kept)."""
import numpy as np
import pytest

import oracle as O
from oracle.spectral import PadePoissonPeriodic, Spectral, getwavenums

L = 2 * np.pi


def _grid(n):
    return np.arange(n) * (L / n)


def test_wavenums_match_c_oracle_and_layout():
    for n in (8, 16, 64):
        dx = L / n
        k = getwavenums(n, dx)
        assert np.allclose(k, O.modk(0, n, dx), rtol=0, atol=1e-13)
        m = np.round(k * dx * n / (2 * np.pi)).astype(int)
        assert list(m) == list(range(n // 2)) + list(range(-n // 2, 0))


def test_fft2_brute_force_dft():
    rng = np.random.default_rng(0)
    nx, ny, nz = 8, 4, 3
    u = rng.standard_normal((nx, ny, nz))
    sp = Spectral(nx, ny, nz, 1.0, 1.0, 1.0)
    fh = sp.fft(u)
    x = np.arange(nx)[:, None]
    y = np.arange(ny)[:, None]
    Ex = np.exp(-2j * np.pi * np.arange(nx // 2 + 1)[None, :] * x / nx)  # (x, kx)
    Ey = np.exp(-2j * np.pi * np.arange(ny)[None, :] * y / ny)           # (y, ky)
    ref = np.einsum("xyz,xa,yb->abz", u, Ex, Ey)
    assert np.abs(fh - ref).max() < 1e-12 * np.abs(ref).max()


def test_ifft2_round_trip_and_oddball():
    rng = np.random.default_rng(1)
    nx, ny, nz = 16, 8, 2
    u = rng.standard_normal((nx, ny, nz))
    sp = Spectral(nx, ny, nz, 1.0, 1.0, 1.0)
    assert np.abs(sp.ifft(sp.fft(u)) - u).max() < 1e-13
    # dropping the kx = nx/2 plane removes exactly the (-1)^i component along x
    v = sp.ifft(sp.fft(u), set_oddball=True)
    alt = (-1.0) ** np.arange(nx)
    comp = np.einsum("xyz,x->yz", u, alt) / nx
    assert np.abs(v - (u - alt[:, None, None] * comp[None])).max() < 1e-13
    # after the y inverse, only the real parts of the kx = 0 and kx = nx/2
    # rows enter the c2r (FFTW's Hermitian c2r): imaginary parts added there
    # leave the result unchanged
    t = np.fft.ifft(sp.fft(u), axis=1)
    t[0] += 1j * rng.standard_normal(t[0].shape)
    t[nx // 2] += 1j * rng.standard_normal(t[nx // 2].shape)
    assert np.abs(sp.ifft(np.fft.fft(t, axis=1)) - u).max() < 1e-13


@pytest.mark.parametrize("dir", [1, 2])
def test_mtimes_ik_is_spectral_derivative(dir):
    nx, ny, nz = 32, 16, 3
    x, y = _grid(nx), _grid(ny)
    X, Y = np.meshgrid(x, y, indexing="ij")
    u = np.repeat((np.sin(3 * X + 0.3) * np.cos(5 * Y - 0.7))[:, :, None], nz, axis=2)
    du = (3 * np.cos(3 * X + 0.3) * np.cos(5 * Y - 0.7)) if dir == 1 else (-5 * np.sin(3 * X + 0.3) * np.sin(5 * Y - 0.7))
    sp = Spectral(nx, ny, nz, L / nx, L / ny, 1.0)
    d = sp.ifft(sp.mtimes_ik(sp.fft(u), dir))
    assert np.abs(d - du[:, :, None]).max() < 1e-12


def test_oddball_sign_and_fix():
    sp = Spectral(16, 8, 4, L / 16, L / 8, 1.0)
    assert sp.k1[8] == pytest.approx(8.0) and sp.k2[4] == pytest.approx(4.0)
    spf = Spectral(16, 8, 4, L / 16, L / 8, 1.0, fix_oddball=True)
    assert spf.k1[8] == 0.0 and spf.k2[4] == 0.0


def test_dealias_mask_counts():
    nx, ny = 64, 32
    sp = Spectral(nx, ny, 4, L / nx, L / ny, 1.0)
    # |m| < n/3: kx in 0..21 (22 of 33), ky in -10..10 (21 of 32)
    assert sp.g1.sum() == 22 and sp.g2.sum() == 21
    spz = Spectral(nx, ny, 16, L / nx, L / ny, L / 16, periodic_z=True)
    assert spz.g1.sum() == 22 and spz.g2.sum() == 21 and spz.g3.sum() == 11
    # with fix_oddball the zeroed oddball passes the periodic-z mask
    spf = Spectral(nx, ny, 16, L / nx, L / ny, L / 16, fix_oddball=True, periodic_z=True)
    assert spf.g1[nx // 2] == 1.0 and spf.g2[ny // 2] == 1.0


def test_dealias_periodic_z_single_modes():
    nx, ny, nz = 16, 16, 16
    sp = Spectral(nx, ny, nz, L / nx, L / ny, L / nz, periodic_z=True)
    z = _grid(nz)
    for kz, keep in ((3, True), (6, False)):
        fh = np.zeros((sp.nxh, ny, nz), complex)
        fh[1, 2] = np.exp(1j * kz * z)
        out = sp.dealias(fh)
        assert np.abs(out - (fh if keep else 0)).max() < 1e-13
    fe = np.zeros((sp.nxh, ny, nz + 1), complex)
    fe[1, 1, :nz] = np.cos(2 * z)
    oe = sp.dealias_edge(fe)
    assert np.abs(oe[:, :, nz] - oe[:, :, 0]).max() == 0


def _project_setup(nx=16, ny=16, nz=16, seed=3):
    dx, dy, dz = L / nx, L / ny, L / nz
    sp = Spectral(nx, ny, nz, dx, dy, dz)
    pp = PadePoissonPeriodic(sp, dx, dy, dz)
    rng = np.random.default_rng(seed)
    u, v = rng.standard_normal((2, nx, ny, nz))
    w = rng.standard_normal((nx, ny, nz + 1))
    w[:, :, nz] = w[:, :, 0]
    return sp, pp, sp.fft(u), sp.fft(v), np.fft.fft(np.fft.rfft(w, axis=0), axis=1)


def test_projection_removes_divergence_and_is_idempotent():
    sp, pp, uh, vh, wh = _project_setup()
    assert np.abs(pp.divergence(uh, vh, wh)).max() > 1.0
    u1, v1, w1 = pp.projection(uh, vh, wh)
    assert np.abs(pp.divergence(u1, v1, w1)).max() < 1e-12
    u2, v2, w2 = pp.projection(u1, v1, w1)
    for a, b in ((u1, u2), (v1, v2), (w1, w2)):
        assert np.abs(a - b).max() < 1e-12 * np.abs(a).max()


def test_projection_of_gradient_and_solenoidal_fields():
    sp, pp, uh, vh, wh = _project_setup()
    nz = sp.shape[2]
    phi = uh.copy()
    phi[0, 0, :] = 0.0      # the zero-wavenumber column carries no pressure
    from oracle.spectral import _stagg
    gu, gv = sp.mtimes_ik(phi, 1), sp.mtimes_ik(phi, 2)
    gw = _stagg(O.STG_D1_C2E, phi, pp.dz, nz)
    ru, rv, rw = pp.projection(gu, gv, gw)
    scale = max(np.abs(gu).max(), np.abs(gw).max())
    assert max(np.abs(ru).max(), np.abs(rv).max(), np.abs(rw).max()) < 1e-12 * scale
    # solenoidal: u = d psi/dy, v = -d psi/dx, w = 0
    su, sv = sp.mtimes_ik(phi, 2), -sp.mtimes_ik(phi, 1)
    sw = np.zeros_like(gw)
    pu, pv, pw = pp.projection(su, sv, sw)
    assert np.abs(pu - su).max() < 1e-12 * np.abs(su).max()
    assert np.abs(pv - sv).max() < 1e-12 * np.abs(sv).max()
    assert np.abs(pw).max() < 1e-12 * np.abs(su).max()


def _duidxj_case(nz, periodic):
    from oracle.spectral import duidxj
    nx, ny, Lz = 16, 16, 1.0
    dx, dy, dz = L / nx, L / ny, Lz / nz
    a = (2 * np.pi if periodic else np.pi) / Lz
    x, y = _grid(nx), _grid(ny)
    zc, ze = (np.arange(nz) + 0.5) * dz, np.arange(nz + 1) * dz
    X, Y, ZC = np.meshgrid(x, y, zc, indexing="ij")
    XE, YE, ZE = np.meshgrid(x, y, ze, indexing="ij")
    u = lambda X, Y, Z: np.cos(X) * np.sin(2 * Y) * np.cos(a * Z)          # noqa: E731  even at the walls
    w = lambda X, Y, Z: np.sin(2 * X + 1.0) * np.cos(Y) * np.sin(a * Z)    # noqa: E731  odd at the walls
    spC, spE = Spectral(nx, ny, nz, dx, dy, dz), Spectral(nx, ny, nz + 1, dx, dy, dz)
    bc = (1, 1, 1, 1, -1, -1, -1, -1, -1, -1, 1, 1)
    o = duidxj(spC, spE, dz, periodic, bc, False, spC.fft(u(X, Y, ZC)), spC.fft(u(X, Y, ZC)), spC.fft(w(X, Y, ZC)),
               spE.fft(w(XE, YE, ZE)), spE.fft(u(XE, YE, ZE)), spE.fft(u(XE, YE, ZE)))
    errs = [
        np.abs(o["duidxjE"][..., 2] + a * np.cos(XE) * np.sin(2 * YE) * np.sin(a * ZE)).max(),      # dudz on edges
        np.abs(o["duidxjC"][..., 2] + a * np.cos(X) * np.sin(2 * Y) * np.sin(a * ZC)).max(),        # dudz on cells
        np.abs(o["duidxjC"][..., 8] - a * np.sin(2 * X + 1) * np.cos(Y) * np.cos(a * ZC)).max(),   # dwdz
        np.abs(spC.ifft(o["d2udz2hatC"]) + a * a * np.cos(X) * np.sin(2 * Y) * np.cos(a * ZC)).max(),
    ]
    exact_xy = max(np.abs(o["duidxjC"][..., 0] + np.sin(X) * np.sin(2 * Y) * np.cos(a * ZC)).max(),
                   np.abs(o["duidxjE"][..., 7] + np.sin(2 * XE + 1) * np.sin(YE) * np.sin(a * ZE)).max())
    return np.array(errs), exact_xy


@pytest.mark.parametrize("periodic", [True, False])
def test_duidxj_manufactured_solution(periodic):
    """x/y derivatives are spectral (exact); the z derivatives, interpolations
    and d2/dz2 converge at sixth order with the boundary kinds of the fields
    (even u, odd w at the walls)."""
    e16, x16 = _duidxj_case(16, periodic)
    e32, x32 = _duidxj_case(32, periodic)
    assert max(x16, x32) < 1e-13
    ratio = e16 / e32
    assert np.all(ratio > 50) and np.all(ratio < 80), ratio
