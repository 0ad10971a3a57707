"""CPU pins of the cgrid oracle (oracle/cgrid.py), exact discrete identities
that a transcription error in a flux, a stress, a sign or an exponent breaks:

  * conservation: on a periodic grid every RHS component sums to zero (the
    compact first derivative annihilates the mean);
  * an entropy wave (uniform velocity and pressure, varying density): the
    momentum RHS is u_i times the mass RHS; with the artificial conductivity
    off the energy RHS is |u|^2 / 2 times it;
  * viscous shears u_i(x_j) (all six) at uniform density and pressure:
    rhs(rho u_i) = mu Dj(Dj u_i), rhs(TE) = mu Dj(u_i Dj u_i), the rest zero;
  * a static gas: rhs(rho u_i) = -D_i p, rhs(TE) = sum_i D_i(kappa D_i T)
    with the power-law conductivity;
  * the artificial bulk viscosity acts in compression only;
  * reflecting ends (bc 1, 1) along each direction equal the periodic grid
    mirrored about both ends: every symmetric / antisymmetric choice;
  * the stress trace is 3 bulk div u;
  * dimensions: stretching the box by s with the same grid values scales each
    artificial coefficient (shear, bulk, conductivity) by exactly s, as a
    viscosity or conductivity must (dx^6 against four derivatives and a
    gradient, dx^4 dx^2 against four and one, dx^4 dx against four);
  * a discretely solenoidal (Taylor-Green) field has no artificial bulk
    viscosity (its switch is off where the dilatation is not negative)."""
import numpy as np
import pytest

from oracle.cgrid import Cgrid

N = (16, 12, 8)


def _d(L=(2 * np.pi, 2 * np.pi, 2 * np.pi), n=N):
    return tuple(L[i] / n[i] for i in range(3))


def _grid_xyz(L, n=N):
    return np.meshgrid(*[np.arange(n[i]) * (L[i] / n[i]) for i in range(3)], indexing="ij")


def _conserved(rho, u, v, w, p, gam=1.4):
    e = p / ((gam - 1.0) * rho)
    TE = rho * (e + 0.5 * (u * u + v * v + w * w))
    return np.stack([rho, rho * u, rho * v, rho * w, TE], axis=-1)


def _smooth(x, y, z, seed):
    rng = np.random.default_rng(seed)
    a = rng.uniform(0.1, 0.3, 4)
    return a[0] * np.sin(x + 0.3) * np.cos(2 * y) + a[1] * np.cos(z + y) + a[2] * np.sin(2 * x - z) + a[3] * np.cos(y)


@pytest.mark.parametrize("inviscid", [True, False])
def test_periodic_conservation(inviscid):
    L = (2 * np.pi,) * 3
    x, y, z = _grid_xyz(L)
    rho = 1.0 + 0.2 * _smooth(x, y, z, 1)
    u, v, w = 0.3 * _smooth(x, y, z, 2), 0.3 * _smooth(x, y, z, 3), 0.3 * _smooth(x, y, z, 4)
    p = 1.0 + 0.1 * _smooth(x, y, z, 5)
    cg = Cgrid(N, _d(L), inviscid=inviscid, mu_ref=0.01)
    r = cg.rhs(_conserved(rho, u, v, w, p))
    for c in range(5):
        assert abs(r[..., c].sum()) <= 1e-12 * np.abs(r[..., c]).sum(), c


@pytest.mark.parametrize("Ckap", [0.0, 0.01])
def test_entropy_wave(Ckap):
    L = (2 * np.pi,) * 3
    x, y, z = _grid_xyz(L)
    rho = 1.0 + 0.3 * _smooth(x, y, z, 6)
    U = np.array([0.7, -0.4, 0.25])
    one = np.ones(N)
    cg = Cgrid(N, _d(L), Ckap=Ckap)
    r = cg.rhs(_conserved(rho, U[0] * one, U[1] * one, U[2] * one, 1.3 * one))
    m = r[..., 0]
    scale = np.abs(m).max()
    assert scale > 1e-2
    for i in range(3):
        assert np.abs(r[..., 1 + i] - U[i] * m).max() <= 1e-12 * scale
    if Ckap == 0.0:
        assert np.abs(r[..., 4] - 0.5 * (U @ U) * m).max() <= 1e-12 * scale
    else:
        # the artificial conductivity of the varying internal energy adds heat flux
        assert np.abs(r[..., 4] - 0.5 * (U @ U) * m).max() > 1e-6 * scale


@pytest.mark.parametrize("i,j", [(i, j) for i in range(3) for j in range(3) if i != j])
def test_viscous_shear(i, j):
    """velocity component i varying along direction j only"""
    L = (2 * np.pi,) * 3
    X = _grid_xyz(L)
    ui = 0.5 * np.sin(X[j]) + 0.2 * np.cos(2 * X[j])
    vel = [np.zeros(N) for _ in range(3)]
    vel[i] = ui
    one = np.ones(N)
    mu_ref, T_ref, n = 0.02, 0.8, 0.75
    cg = Cgrid(N, _d(L), inviscid=False, mu_ref=mu_ref, T_ref=T_ref, visc_n=n, Cmu=0.0)
    p, rho = 1.1, 1.0
    r = cg.rhs(_conserved(rho * one, *vel, p * one))
    T = p / rho  # Rgas = 1
    mu = mu_ref * (T / T_ref) ** n
    Dj = lambda f: cg.dd(f, j, (0, 0))  # noqa: E731
    scale = np.abs(r[..., 1 + i]).max()
    assert scale > 1e-3
    assert np.abs(r[..., 1 + i] - mu * Dj(Dj(ui))).max() <= 1e-12 * scale
    assert np.abs(r[..., 4] - mu * Dj(ui * Dj(ui))).max() <= 1e-12 * scale
    for c in range(4):
        if c != 1 + i:
            assert np.abs(r[..., c]).max() <= 1e-12 * scale


def test_static_gas_pressure_and_heat_conduction():
    """no flow, varying pressure and density: rhs(rho u_i) = -D_i p and
    rhs(TE) = sum_i D_i(kappa D_i T) with kappa = Cp mu(T) / Pr (power law)"""
    L = (2 * np.pi,) * 3
    x, y, z = _grid_xyz(L)
    rho = 1.0 + 0.2 * _smooth(x, y, z, 13)
    p = 1.0 + 0.15 * _smooth(x, y, z, 14)
    zero = np.zeros(N)
    gam, Pr, mu_ref, T_ref, n = 1.4, 0.7, 0.03, 0.9, 0.7
    cg = Cgrid(N, _d(L), inviscid=False, mu_ref=mu_ref, T_ref=T_ref, visc_n=n, Pr=Pr, Ckap=0.0, gam=gam)
    r = cg.rhs(_conserved(rho, zero, zero, zero, p, gam=gam))
    T = p / rho
    kap = gam / (gam - 1.0) * mu_ref * (T / T_ref) ** n / Pr
    scale = np.abs(r[..., 1:4]).max()
    assert np.abs(r[..., 0]).max() <= 1e-13 * scale
    for i in range(3):
        assert np.abs(r[..., 1 + i] + cg.dd(p, i, (0, 0))).max() <= 1e-12 * scale
    e_ref = sum(cg.dd(kap * cg.dd(T, i, (0, 0)), i, (0, 0)) for i in range(3))
    assert np.abs(r[..., 4] - e_ref).max() <= 1e-12 * np.abs(e_ref).max()


def test_artificial_bulk_viscosity_acts_in_compression_only():
    """1D wave u = a sin x: raw artificial bulk viscosity only where the
    dilatation is negative; two Gaussian passes reach 8 points, so it is
    exactly zero 9 or more points inside the expanding half"""
    n = (64, 8, 8)
    L = (2 * np.pi,) * 3
    x, y, z = _grid_xyz(L, n)
    cg = Cgrid(n, _d(L, n))
    rho = 1.0 + 0.1 * np.sin(x + 0.5)
    zero = np.zeros(n)
    _, q = cg.rhs(_conserved(rho, 0.3 * np.sin(x), zero, zero, np.ones(n)), parts=True)
    b = q["bulk"][:, 0, 0]
    expanding = [i % 64 for i in range(-7, 8)]
    assert (b[expanding] == 0).all()
    assert b[32] > 0 and (b >= 0).all()


def _mirror(f, axis, parity):
    """period 2 (n - 1) data: f, then parity times f reflected about its last point"""
    g = np.moveaxis(f, axis, 0)
    ext = np.concatenate([g, parity * g[-2:0:-1]], axis=0)
    return np.moveaxis(ext, 0, axis)


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_reflecting_ends_equal_the_mirrored_periodic_grid(axis):
    """bc (1, 1) along one direction (even scalars, odd normal velocity): the
    RHS equals that of the periodic grid twice as long (mirrored about both
    ends), which fixes every symmetric / antisymmetric choice of the fluxes,
    gradients, LAD derivatives and filters along that direction"""
    n = [12, 10, 9]
    L = [1.1 * n[0], 0.9 * n[1], 1.3 * n[2]]
    X = _grid_xyz(L, n)
    rho = 1.0 + 0.2 * _smooth(*X, 15)
    p = 1.0 + 0.15 * _smooth(*X, 16)
    vel = [0.3 * _smooth(*X, 17 + c) for c in range(3)]
    # even about both ends along `axis`, except the normal velocity (odd)
    def even(f):
        g = np.moveaxis(f, axis, 0).copy()
        c = np.cos(np.pi * np.arange(n[axis]) / (n[axis] - 1))
        return np.moveaxis(g + 0.3 * c.reshape((-1,) + (1,) * 2), 0, axis)
    rho, p = even(rho), even(p)
    vel = [even(v) for v in vel]
    s = np.sin(np.pi * np.arange(n[axis]) / (n[axis] - 1)).reshape([-1 if k == axis else 1 for k in range(3)])
    vel[axis] = vel[axis] * s
    W = _conserved(rho, *vel, p)
    d = tuple(L[k] / n[k] for k in range(3))
    per = [True, True, True]
    per[axis] = False
    bc = [(0, 0), (0, 0), (0, 0)]
    bc[axis] = (1, 1)
    kw = dict(inviscid=False, mu_ref=0.02)
    r = Cgrid(tuple(n), d, periodic=tuple(per), bc=tuple(bc), **kw).rhs(W)
    Wm = np.stack([_mirror(W[..., c], axis, -1.0 if c == 1 + axis else 1.0) for c in range(5)], axis=-1)
    nm = list(n)
    nm[axis] = 2 * (n[axis] - 1)
    rm = Cgrid(tuple(nm), d, **kw).rhs(Wm)
    rm = np.moveaxis(np.moveaxis(rm, axis, 0)[: n[axis]], 0, axis)
    assert np.abs(r - rm).max() <= 1e-11 * np.abs(rm).max()


def test_stress_trace():
    L = (2 * np.pi,) * 3
    x, y, z = _grid_xyz(L)
    rho = 1.0 + 0.2 * _smooth(x, y, z, 7)
    u, v, w = 0.4 * _smooth(x, y, z, 8), 0.4 * _smooth(x, y, z, 9), 0.4 * _smooth(x, y, z, 10)
    cg = Cgrid(N, _d(L), inviscid=False, mu_ref=0.01)
    _, q = cg.rhs(_conserved(rho, u, v, w, 1.0 + 0.1 * _smooth(x, y, z, 11)), parts=True)
    g = q["grad"]
    div = g[0] + g[4] + g[8]
    txx, _, _, tyy, _, tzz = q["tau"]
    assert np.abs(q["bulk"]).max() > 0
    assert np.abs(txx + tyy + tzz - 3 * q["bulk"] * div).max() <= 1e-12 * np.abs(txx).max()


def test_artificial_coefficients_scale_as_a_length():
    rng = np.random.default_rng(12)
    vals = [rng.standard_normal(N) * 0.2 for _ in range(4)]
    rho = 1.0 + 0.3 * np.tanh(vals[0])
    P = 1.0 + 0.2 * np.tanh(vals[3])
    out = []
    for s in (1.0, 3.0):
        L = (2 * np.pi * s, 1.5 * np.pi * s, 2.5 * np.pi * s)
        cg = Cgrid(N, _d(L))
        W = _conserved(rho, vals[1], vals[2], 0.5 * vals[1] - vals[2], P)
        Pm = cg.primitive(W)
        neg = lambda b: (-b[0], -b[1])  # noqa: E731
        bc = cg.bc
        g = (cg.gradient(Pm["u"], [neg(bc[0]), bc[1], bc[2]]) + cg.gradient(Pm["v"], [bc[0], neg(bc[1]), bc[2]])
             + cg.gradient(Pm["w"], [bc[0], bc[1], neg(bc[2])]))
        out.append(cg.lad(Pm, g))
    for a, b in zip(out[0], out[1]):
        assert np.abs(a).max() > 0
        assert np.abs(b - 3.0 * a).max() <= 1e-11 * np.abs(b).max()


def test_solenoidal_field_has_no_artificial_bulk_viscosity():
    L, n = (2 * np.pi,) * 3, (16, 16, 8)
    x, y, z = _grid_xyz(L, n)
    u = np.sin(x) * np.cos(y)
    v = -np.cos(x) * np.sin(y)
    w = np.zeros(n)
    cg = Cgrid(n, _d(L, n))
    _, q = cg.rhs(_conserved(np.ones(n), u, v, w, 1.0 + 0.1 * np.cos(2 * x)), parts=True)
    g = q["grad"]
    assert np.abs(g[0] + g[4] + g[8]).max() <= 1e-13
    assert np.abs(q["bulk"]).max() == 0.0
    assert np.abs(q["mu"]).max() > 0 and np.abs(q["kap"]).max() > 0


def test_pressure_scaling_of_the_artificial_coefficients():
    """p -> t p at fixed density and velocity: the artificial shear and bulk
    viscosities do not change; the conductivity (rho c |D4 e| dx^5 / T, c the
    speed of sound) scales by sqrt(t)"""
    rng = np.random.default_rng(18)
    vals = [rng.standard_normal(N) * 0.2 for _ in range(4)]
    rho = 1.0 + 0.3 * np.tanh(vals[0])
    p = 1.0 + 0.2 * np.tanh(vals[3])
    L = (2 * np.pi, 1.5 * np.pi, 2.5 * np.pi)
    cg = Cgrid(N, _d(L))
    out = []
    for t in (1.0, 4.0):
        _, q = cg.rhs(_conserved(rho, vals[1], vals[2], vals[1] - vals[2], t * p), parts=True)
        out.append((q["mu"], q["bulk"], q["kap"]))
    (m1, b1, k1), (m4, b4, k4) = out
    assert np.abs(m4 - m1).max() <= 1e-12 * np.abs(m1).max()
    assert np.abs(b4 - b1).max() <= 1e-12 * np.abs(b1).max()
    assert np.abs(k4 - 2.0 * k1).max() <= 1e-12 * np.abs(k4).max()


@pytest.mark.parametrize("perm", [(2, 1, 0), (1, 0, 2), (1, 2, 0)])
@pytest.mark.parametrize("inviscid", [True, False])
def test_axis_permutation_equivariance(perm, inviscid):
    """relabelling the axes (with their spacings, sizes and velocity
    components) relabels the RHS: no direction-specific spacing or component
    is mixed up"""
    n = (12, 10, 9)
    L = (2 * np.pi, 1.7 * np.pi, 2.4 * np.pi)
    X = _grid_xyz(L, n)
    rho = 1.0 + 0.2 * _smooth(*X, 19)
    p = 1.0 + 0.15 * _smooth(*X, 20)
    vel = [0.3 * _smooth(*X, 21 + c) for c in range(3)]
    kw = dict(inviscid=inviscid, mu_ref=0.02)
    r = Cgrid(n, _d(L, n), **kw).rhs(_conserved(rho, *vel, p))
    tp = lambda f: np.transpose(f, perm)  # noqa: E731
    n2 = tuple(n[k] for k in perm)
    d2 = tuple(_d(L, n)[k] for k in perm)
    vel2 = [tp(vel[k]) for k in perm]
    r2 = Cgrid(n2, d2, **kw).rhs(_conserved(tp(rho), *vel2, tp(p)))
    ref = np.stack([tp(r[..., 0])] + [tp(r[..., 1 + k]) for k in perm] + [tp(r[..., 4])], axis=-1)
    assert np.abs(r2 - ref).max() <= 1e-11 * np.abs(ref).max()
