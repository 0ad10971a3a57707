"""CPU pins of the oracle's igrid time advancement (oracle/igrid.py), the
reference for the GPU time-stepping tests.

  * A shear mode u = sin(y) is an exact solution of the forced-free
    Navier-Stokes equations (its nonlinear term is a gradient, which the
    projection removes): it decays as exp(-t/Re), and v, w stay zero.
  * A vertical shear mode u = sin(z) decays with the cd06 staggered second
    derivative's symbol in place of -k^2 (the viscous term's z part).
  * The Runge-Kutta schemes converge at their orders: the error of the decay
    falls by 2^4 (SSP-RK45) and 2^3 (TVD-RK3) when dt halves.
  * The HIT forcing at each drawn mode is fac conj(u_k) with fac = eps
    (nx ny nz)^2 / (|u_k|^2 + |v_k|^2 + |w_k|^2 + 1e-14) / nwaves, so the sum
    over all modes of u_k f_k (3D Fourier coefficients, w shifted to the
    cells) is real and equals eps (nx ny nz)^2 up to the 1e-14 guard: a
    dropped conjugate, a wrong normalisation or a lost shift breaks it.
"""
import numpy as np
import pytest

from oracle.igrid import Igrid
from test_oracle_pins import stagg_symbol
import oracle as O

L = 2 * np.pi


def _run(n, u, Re, scheme, dt, nsteps, nz=None):
    nz = nz or n
    g = Igrid(n, n, nz, L, L, L, Re=Re, scheme=scheme, dt=dt)
    z = np.zeros((n, n, nz))
    g.set_velocity(u, z, np.zeros((n, n, nz + 1)))
    for _ in range(nsteps):
        g.step()
    return g


def test_shear_mode_decays_exactly():
    n = 16
    y = np.arange(n) * L / n
    u = np.broadcast_to(np.sin(y)[None, :, None], (n, n, n)).copy()
    Re, dt, ns = 10.0, 0.01, 10
    g = _run(n, u, Re, 2, dt, ns)
    exact = np.exp(-ns * dt / Re) * u
    assert np.abs(g.f["u"] - exact).max() <= 1e-11
    assert np.abs(g.f["v"]).max() <= 1e-13 and np.abs(g.f["w"]).max() <= 1e-13


def test_vertical_shear_decays_with_the_compact_symbol():
    n = 16
    dz = L / n
    zc = (np.arange(n) + 0.5) * dz
    u = np.broadcast_to(np.sin(zc)[None, None, :], (n, n, n)).copy()
    Re, dt, ns = 10.0, 0.01, 10
    g = _run(n, u, Re, 2, dt, ns)
    lam = stagg_symbol(O.STG_D2_C2C, dz, dz).real / Re
    exact = np.exp(lam * ns * dt) * u
    assert abs(lam * Re + 1) < 1e-3          # close to -k^2 = -1, but not equal
    assert np.abs(g.f["u"] - exact).max() <= 1e-11
    # the exact -k^2 decay would miss by far more than that
    assert np.abs(g.f["u"] - np.exp(-ns * dt / Re) * u).max() > 1e-8


@pytest.mark.parametrize("scheme,order", [(2, 4), (1, 3)])
def test_runge_kutta_orders(scheme, order):
    n = 8
    y = np.arange(n) * L / n
    u = np.broadcast_to(np.sin(y)[None, :, None], (n, n, n)).copy()
    Re, T = 1.0, 0.8
    errs = []
    for ns in (4, 8):
        g = _run(n, u, Re, scheme, T / ns, ns)
        errs.append(np.abs(g.f["u"] - np.exp(-T / Re) * u).max())
    ratio = errs[0] / errs[1]
    assert 2 ** order * 0.8 < ratio < 2 ** order * 1.25, (errs, ratio)


def test_forcing_power_input():
    n = 16
    rng = np.random.default_rng(5)
    u, v = rng.standard_normal((2, n, n, n))
    w = rng.standard_normal((n, n, n + 1))
    w[:, :, n] = w[:, :, 0]
    F = dict(kmin=1.0, kmax=3.0, nwaves=7, eps=0.1, seed=9)
    g = Igrid(n, n, n, L, L, L, Re=0.0, scheme=2, dt=0.01, forcing=F)
    g.set_velocity(u, v, w)
    zC = np.zeros((n // 2 + 1, n, n), complex)
    fu, fv, fw = g.forcing_rhs(zC, zC.copy(), np.zeros((n // 2 + 1, n, n + 1), complex))
    k3 = 2 * np.pi * np.fft.fftfreq(n, d=1.0 / n) / L
    sh = np.exp(1j * k3 * (L / n) / 2)[None, None, :]
    U = np.fft.fft(g.f["uhat"], axis=2)
    V = np.fft.fft(g.f["vhat"], axis=2)
    W = np.fft.fft(g.f["what"][:, :, :n], axis=2) * sh
    Fu, Fv = np.fft.fft(fu, axis=2), np.fft.fft(fv, axis=2)
    Fw = np.fft.fft(fw[:, :, :n], axis=2) * sh
    assert np.abs(fw[:, :, n] - fw[:, :, 0]).max() == 0
    power = (U * Fu + V * Fv + W * Fw).sum()
    target = F["eps"] * float(n ** 3) ** 2
    assert abs(power.imag) <= 1e-10 * target
    assert abs(power.real - target) <= 1e-10 * target
    # nonzero only at the drawn wave vectors (kx, ky, kz >= 0, within kmax)
    mag = np.abs(Fu) + np.abs(Fv) + np.abs(Fw)
    nz_modes = np.argwhere(mag > 1e-10 * mag.max())
    assert 1 <= len(nz_modes) <= F["nwaves"]
    assert (nz_modes <= np.ceil(F["kmax"])).all()


def test_walls_uniform_acceleration_between_slip_walls():
    """walls in z, slip top and bottom, fluid at rest, mean pressure gradient
    dpFdx: a uniform u is even about both walls, so its viscous term is zero
    and u = dpFdx t exactly (v = w = 0) under both Runge-Kutta schemes."""
    nx, ny, nz, dp = 8, 8, 16, 0.7
    for scheme in (1, 2):
        g = Igrid(nx, ny, nz, L, L, 1.0, Re=50.0, scheme=scheme, dt=0.01,
                  walls=dict(botWall=2, topWall=2, dpFdx=dp))
        z = np.zeros((nx, ny, nz))
        g.set_velocity(z, z, np.zeros((nx, ny, nz + 1)))
        for _ in range(3):
            g.step()
        f = g.f
        assert np.abs(f["u"] - dp * g.tsim).max() <= 1e-13
        assert np.abs(f["v"]).max() <= 1e-14 and np.abs(f["w"]).max() <= 1e-14


def test_walls_poiseuille_profile_is_steady_away_from_the_walls():
    """no-slip walls: the Poiseuille parabola u = (Re dpFdx / 2) z (Lz - z)
    balances the pressure gradient wherever the compact second derivative is
    exact (a quadratic), i.e. away from the walls, where the odd reflection of
    the no-slip kind acts; one step leaves the centre cells unchanged.  A
    wrong sign or scale of the forcing or of the viscous term moves them."""
    nx, ny, nz, Re, dp, Lz = 8, 8, 32, 20.0, 1.0, 1.0
    g = Igrid(nx, ny, nz, L, L, Lz, Re=Re, scheme=2, dt=0.002, walls=dict(botWall=1, topWall=1, dpFdx=dp))
    zc = (np.arange(nz) + 0.5) * Lz / nz
    prof = 0.5 * Re * dp * zc * (Lz - zc)
    u0 = np.broadcast_to(prof[None, None, :], (nx, ny, nz)).copy()
    z = np.zeros((nx, ny, nz))
    g.set_velocity(u0, z, np.zeros((nx, ny, nz + 1)))
    g.step()
    du = g.f["u"] - u0
    centre = slice(nz // 2 - 3, nz // 2 + 3)
    assert np.abs(du[:, :, centre]).max() <= 1e-9 * prof.max()
    assert np.abs(du).max() > 1e-6 * prof.max()          # the walls do act
    assert np.abs(g.f["v"]).max() <= 1e-13 and np.abs(g.f["w"]).max() <= 1e-12 * prof.max()
