"""CPU pins of the oracle's PadePoisson with walls (oracle/spectral.py
PadePoissonWalls, the reference's spectral extension algorithm restated).

  * The projected field is discretely divergence-free with the z derivative
    of the reference's DivergenceCheck, the non-periodic cd06stagg ddz_E2C
    with odd ends (a compact operator of the C oracle, independent of the
    extension algebra): a wrong sign of the half-cell phase, a wrong
    extension or a wrong k3' fails it.
  * It is a projection: applying it twice changes nothing.
  * It removes a discrete gradient (i k1 phi, i k2 phi, ddz_C2E(phi) with
    even ends) entirely.
  * w vanishes on both walls.

With the Stokes pressure (Lz given):
  * For u = v = 0 and w given on the walls only, the Stokes step is the
    potential flow of each (k1, k2) line: phi'' = lam^2 phi with
    phi'(0) = w(bottom), phi'(Lz) = w(top), solved here in closed form
    (phi = (w_top cosh(lam z) - w_bot cosh(lam (Lz - z))) / (lam sinh(lam Lz)));
    u = -i k1 phi, v = -i k2 phi on the cells, w = -phi' on the inner edges,
    and the stored pressure i phi.  A wrong sign of a sinh/cosh table, of
    denFact or the order of the two walls fails it.
  * A field with nonzero wall velocity (no mean and no Nyquist modes there)
    comes out divergence-free; without the Stokes pressure it does not.
  * getPressure: the velocity change of the projection is i k p of the
    returned pressure (its sign and scale), with the Stokes part added as
    i phi; getPressureAndUpdateRHS returns the projection and the same
    pressure; the periodic ones likewise."""
import numpy as np
import pytest

import oracle as O
from oracle.spectral import PadePoissonPeriodic, PadePoissonWalls, Spectral


def _setup(nx, ny, nz, seed):
    L = 2 * np.pi
    dx, dy, dz = L / nx, L / ny, 1.0 / nz
    sp = Spectral(nx, ny, nz, dx, dy, dz)
    rng = np.random.default_rng(seed)
    u, v = rng.standard_normal((2, nx, ny, nz))
    w = rng.standard_normal((nx, ny, nz + 1))
    w[:, :, 0] = 0
    w[:, :, nz] = 0
    return sp, PadePoissonWalls(sp, dx, dy, dz), sp.fft(u), sp.fft(v), Spectral(nx, ny, nz + 1, dx, dy, dz).fft(w), dz


def _div(sp, uh, vh, wh, dz, nz):
    k1 = sp.k1[:, None, None]
    k2 = sp.k2[None, :, None]
    dwdz = (O.stagg_np(O.STG_D1_E2C, np.asfortranarray(wh.real), dz, 0, nz) +
            1j * O.stagg_np(O.STG_D1_E2C, np.asfortranarray(wh.imag), dz, 0, nz))
    return dwdz + 1j * (k1 * uh + k2 * vh)


@pytest.mark.parametrize("shape", [(16, 8, 16), (8, 16, 32)])
def test_projection_is_divergence_free_and_idempotent(shape):
    nx, ny, nz = shape
    sp, pp, uh, vh, wh, dz = _setup(nx, ny, nz, sum(shape))
    d0 = np.abs(_div(sp, uh, vh, wh, dz, nz)).max()
    u1, v1, w1 = pp.projection(uh, vh, wh)
    d1 = np.abs(_div(sp, u1, v1, w1, dz, nz)).max()
    assert d1 <= 1e-12 * d0, (d0, d1)
    assert np.abs(w1[:, :, 0]).max() == 0 and np.abs(w1[:, :, nz]).max() == 0
    u2, v2, w2 = pp.projection(u1, v1, w1)
    scale = max(np.abs(u1).max(), np.abs(w1).max())
    assert max(np.abs(u2 - u1).max(), np.abs(v2 - v1).max(), np.abs(w2 - w1).max()) <= 1e-12 * scale


def test_projection_removes_a_discrete_gradient():
    nx, ny, nz = 16, 8, 16
    sp, pp, _, _, _, dz = _setup(nx, ny, nz, 1)
    rng = np.random.default_rng(2)
    phi = sp.fft(rng.standard_normal((nx, ny, nz)))
    k1 = sp.k1[:, None, None]
    k2 = sp.k2[None, :, None]
    g = lambda f: O.stagg_np(O.STG_D1_C2E, np.asfortranarray(f), dz, O.TOP_EVEN | O.BOT_EVEN, nz)  # noqa: E731
    uh, vh = 1j * k1 * phi, 1j * k2 * phi
    wh = g(phi.real) + 1j * g(phi.imag)
    u1, v1, w1 = pp.projection(uh, vh, wh)
    scale = max(np.abs(uh).max(), np.abs(wh).max())
    assert max(np.abs(u1).max(), np.abs(v1).max(), np.abs(w1).max()) <= 1e-12 * scale


def _stokes_setup(nx, ny, nz, Lz, seed):
    L = 2 * np.pi
    dx, dy, dz = L / nx, L / ny, Lz / nz
    sp = Spectral(nx, ny, nz, dx, dy, dz)
    return sp, PadePoissonWalls(sp, dx, dy, dz, Lz=Lz), np.random.default_rng(seed), dz


def test_stokes_step_is_the_potential_flow():
    nx, ny, nz, Lz = 16, 8, 16, 1.0
    sp, pp, rng, dz = _stokes_setup(nx, ny, nz, Lz, 3)
    nxh = sp.nxh
    W0 = rng.standard_normal((nxh, ny)) + 1j * rng.standard_normal((nxh, ny))
    W1 = rng.standard_normal((nxh, ny)) + 1j * rng.standard_normal((nxh, ny))
    w = np.zeros((nxh, ny, nz + 1), complex)
    w[:, :, 0], w[:, :, nz] = W0, W1
    zero = np.zeros((nxh, ny, nz), complex)
    u, v, wn, p1, p2 = pp.stokes_step(zero, zero, w)
    k1 = np.fft.fftfreq(nx, 1.0 / nx)[:nxh]
    k1[nx // 2] = -nx // 2
    k2 = np.fft.fftfreq(ny, 1.0 / ny)
    lam = np.sqrt(k1[:, None] ** 2 + k2[None, :] ** 2)
    zc = (np.arange(nz) + 0.5) * dz
    ze = np.arange(nz + 1) * dz
    m = lam > 0
    L3, S = lam[:, :, None], np.where(m, np.sinh(lam * Lz) * lam, 1.0)[:, :, None]
    phi = (W1[:, :, None] * np.cosh(L3 * zc) - W0[:, :, None] * np.cosh(L3 * (Lz - zc))) / S
    dphi = (W1[:, :, None] * L3 * np.sinh(L3 * ze) + W0[:, :, None] * L3 * np.sinh(L3 * (Lz - ze))) / S
    m3 = m[:, :, None]
    scale = np.abs(phi[np.broadcast_to(m3, phi.shape)]).max()
    assert np.abs(np.where(m3, u - (-1j * k1[:, None, None] * phi), 0)).max() <= 1e-11 * scale * nx
    assert np.abs(np.where(m3, v - (-1j * k2[None, :, None] * phi), 0)).max() <= 1e-11 * scale * ny
    assert np.abs(np.where(m3, (p1 + p2) - 1j * phi, 0)).max() <= 1e-11 * scale
    assert np.abs(np.where(m3, wn[:, :, 1:nz] + dphi[:, :, 1:nz], 0)).max() <= 1e-11 * np.abs(dphi).max()
    assert (wn[:, :, 0] == 0).all() and (wn[:, :, nz] == 0).all()
    # the mean line gets no correction
    assert np.abs(u[0, 0]).max() == 0 and np.abs(p1[0, 0]).max() == 0


def _wall_fields(sp, rng, nz):
    nxh, ny = sp.nxh, sp.shape[1]
    c = lambda *s: rng.standard_normal(s) + 1j * rng.standard_normal(s)  # noqa: E731
    u, v, w = c(nxh, ny, nz), c(nxh, ny, nz), c(nxh, ny, nz + 1)
    for a in (u, v, w):
        a[sp.shape[0] // 2] = 0      # no Nyquist modes: the Stokes step uses the plain wavenumbers
        a[:, ny // 2] = 0
    w[0, 0, 0] = w[0, 0, nz] = 0     # the mean line has no Stokes correction
    return u, v, w


@pytest.mark.parametrize("Lz", [1.0, 2.0])
def test_stokes_projection_removes_wall_velocity(Lz):
    nx, ny, nz = 16, 8, 16
    sp, pp, rng, dz = _stokes_setup(nx, ny, nz, Lz, 4)
    u, v, w = _wall_fields(sp, rng, nz)
    assert np.abs(w[:, :, 0]).max() > 0.5
    d0 = np.abs(_div(sp, u, v, w, dz, nz)).max()
    u1, v1, w1 = pp.projection(u, v, w)
    d1 = np.abs(_div(sp, u1, v1, w1, dz, nz)).max()
    assert d1 <= 1e-11 * d0, (d0, d1)
    assert (w1[:, :, 0] == 0).all() and (w1[:, :, nz] == 0).all()
    # without the Stokes pressure the wall velocity leaves divergence behind
    plain = PadePoissonWalls(sp, 2 * np.pi / nx, 2 * np.pi / ny, dz)
    d2 = np.abs(_div(sp, *plain.projection(u, v, w), dz, nz)).max()
    assert d2 >= 1e-3 * d0


def _inner(a, nx):
    return a[1:nx // 2]


@pytest.mark.parametrize("kind", ["periodic", "walls", "stokes"])
def test_pressure_is_the_projection_gradient(kind):
    nx, ny, nz = 16, 8, 16
    L = 2 * np.pi
    dx, dy = L / nx, L / ny
    rng = np.random.default_rng(5)
    if kind == "periodic":
        dz = L / nz
        sp = Spectral(nx, ny, nz, dx, dy, dz, periodic_z=True)
        pp = PadePoissonPeriodic(sp, dx, dy, dz)
        c = lambda *s: rng.standard_normal(s) + 1j * rng.standard_normal(s)  # noqa: E731
        u, v, w = c(sp.nxh, ny, nz), c(sp.nxh, ny, nz), c(sp.nxh, ny, nz + 1)
        w[:, :, nz] = w[:, :, 0]
    else:
        sp, pp, rng, dz = _stokes_setup(nx, ny, nz, 1.0, 6)
        if kind == "walls":
            pp = PadePoissonWalls(sp, dx, dy, dz)
        u, v, w = _wall_fields(sp, rng, nz)
    pr = pp.get_pressure(u, v, w)
    u1, v1, w1, pr2 = pp.get_pressure_and_update(u, v, w)
    pu, pv, pw = pp.projection(u, v, w)
    assert np.array_equal(pr, pr2)
    assert np.array_equal(u1, pu) and np.array_equal(v1, pv) and np.array_equal(w1, pw)
    ph = _inner(sp.fft(pr), nx)
    k1 = _inner(sp.k1, nx)[:, None, None]
    if kind == "stokes":
        us, _, _, p1, p2 = pp.stokes_step(u, v, w)
        pmain = _inner(us - u1, nx) / (1j * k1)
        ref = pmain + _inner(p1 + p2, nx)
    else:
        ref = _inner(u - u1, nx) / (1j * k1)
    assert np.abs(ph - ref).max() <= 1e-12 * np.abs(ref).max()
