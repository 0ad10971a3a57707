"""Pins of the SGS oracle (oracle/sgs.py): closed forms of the Smagorinsky and
AMD kernels, an independent tensor formulation of both, and invariances of
the dynamic procedure (an identity test filter gives no resolved stress; the
coefficient does not depend on the velocity scale).  The sigma model against
numpy's singular values and the flows where it must vanish; the Moeng wall
stress against the log law of a uniform first-level velocity."""
import numpy as np

import oracle.sgs as G
from oracle.spectral import Spectral


def _grad(shape, seed):
    return np.random.default_rng(seed).standard_normal(shape + (9,))


def test_sij_is_symmetric_part():
    g = _grad((5, 4, 3), 0)
    S = G.sij(g)
    T = g.reshape(g.shape[:-1] + (3, 3))          # T[..., i, j] = d u_i / d x_j
    Sym = 0.5 * (T + np.swapaxes(T, -1, -2))
    idx = [(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)]
    for c, (i, j) in enumerate(idx):
        assert np.allclose(S[..., c], Sym[..., i, j], rtol=0, atol=1e-15)


def test_smagorinsky_closed_forms():
    g = np.zeros((1, 9))
    g[0, 1] = 3.0                                  # pure shear du/dy = 3: |S| = 3
    assert abs(G.smagorinsky(G.sij(g))[0] - 3.0) < 1e-15
    g = _grad((6, 5, 4), 1)
    T = g.reshape(g.shape[:-1] + (3, 3))
    Sym = 0.5 * (T + np.swapaxes(T, -1, -2))
    assert np.allclose(G.smagorinsky(G.sij(g)), np.sqrt(2 * np.einsum("...ij,...ij", Sym, Sym)), rtol=1e-14)


def test_amd_closed_forms():
    cx, cy, cz = 0.3, 0.2, 0.5
    g = np.zeros((1, 9))
    g[0, 1] = 2.0                                  # pure shear: AMD vanishes
    assert G.amd(g, G.sij(g), cx, cy, cz)[0] == 0.0
    a = 0.7                                        # axisymmetric strain diag(-a, -a, 2a)
    g = np.zeros((1, 9))
    g[0, 0], g[0, 4], g[0, 8] = -a, -a, 2 * a
    num = -a ** 3 * (cx ** 2 + cy ** 2) + 8 * a ** 3 * cz ** 2
    assert abs(G.amd(g, G.sij(g), cx, cy, cz)[0] - max(-num / (6 * a * a), 0.0)) < 1e-15
    g = _grad((4, 4, 4), 2)
    T = g.reshape(g.shape[:-1] + (3, 3))
    Sym = 0.5 * (T + np.swapaxes(T, -1, -2))
    C2 = np.array([cx, cy, cz]) ** 2
    ref = np.maximum(-np.einsum("...ik,...jk,k,...ij->...", T, T, C2, Sym) / (np.einsum("...ij,...ij", T, T) + 1e-32),
                     0.0)
    assert np.allclose(G.amd(g, G.sij(g), cx, cy, cz), ref, rtol=1e-13, atol=1e-15)


def _dyn_fields(scale=1.0, n=16, nz=4, seed=3):
    rng = np.random.default_rng(seed)
    dx = 2 * np.pi / n
    sp = Spectral(n, n, nz, dx, dx, 0.1)
    u, v, w = (scale * rng.standard_normal((n, n, nz)) for _ in range(3))
    g = scale * rng.standard_normal((n, n, nz, 9))
    S = G.sij(g)
    return sp, dx, u, v, w, S, G.smagorinsky(S)


def test_dynamic_identity_filter_gives_zero():
    sp, dx, u, v, w, S, D = _dyn_fields()
    lam = G.dynamic_procedure(sp, dx, dx, 1e-3, u, v, w, sp.fft(u), sp.fft(v), sp.fft(w), S, D)
    lam2 = G.dynamic_procedure(sp, dx, dx, 2.0, u, v, w, sp.fft(u), sp.fft(v), sp.fft(w), S, D)
    # the resolved stress L_ij = F(u_i u_j) - F(u_i) F(u_j) cancels to round-off
    assert np.all(lam <= 1e-12 * lam2.max()) and lam2.max() > 0


def test_dynamic_scale_invariance():
    sp, dx, u, v, w, S, D = _dyn_fields(1.0)
    lam1 = G.dynamic_procedure(sp, dx, dx, 2.0, u, v, w, sp.fft(u), sp.fft(v), sp.fft(w), S, D)
    sp, dx, u, v, w, S, D = _dyn_fields(3.0)
    lam3 = G.dynamic_procedure(sp, dx, dx, 2.0, u, v, w, sp.fft(u), sp.fft(v), sp.fft(w), S, D)
    assert np.any(lam1 > 0)
    assert np.allclose(lam1, lam3, rtol=1e-10, atol=0)


def test_dynamic_coefficient_closed_form():
    """Germano identity with the sharp test filter, in closed form.

    u = A cos(2x) + B_z cos(5x), v = C cos(5x), w = 0 on a 16-point periodic
    grid (k integer), uniform strain S = [[s, t, 0], [t, -s, 0], [0, 0, 0]]
    and uniform kernel D.  With deltaRat = 2 the test filter keeps |k| < 8/3
    (standardDynamicProcedure.F90 via spectral testFilter), so
    F(u) = A cos 2x, F(v) = 0 and the planar means of the resolved stress are
    <L11> = B^2/2, <L22> = C^2/2, <L12> = B C/2 (every other product mode,
    |k| = 3, 4, 6, 7, 10, is removed).  The filtered strain is S itself and
    its Smagorinsky kernel |S| = sqrt(2 S:S) = 2 sqrt(s^2 + t^2), so
    M_ij = F(D S_ij) - deltaRat^2 |S| S_ij = (D - 4 |S|) S_ij = Dm S_ij, and
    per plane
        lambda = max(<L:M>, 0) / (2 <M:M>)
               = Dm (s (B^2 - C^2)/2 + B C t) / (4 Dm^2 (s^2 + t^2)).
    A dropped factor 2 on an off-diagonal term, a wrong deltaRat power or a
    mean taken over the wrong axes changes the value."""
    n, nz = 16, 3
    dx = 2 * np.pi / n
    sp = Spectral(n, n, nz, dx, dx, 0.1)
    x = np.arange(n) * dx
    Az, C, s, t, D = 0.9, 0.7, -0.5, -0.3, 1.7
    Bz = np.array([1.2, 1.8, 2.4])
    u = Az * np.cos(2 * x)[:, None, None] + Bz[None, None, :] * np.cos(5 * x)[:, None, None] * np.ones((1, n, 1))
    v = C * np.cos(5 * x)[:, None, None] * np.ones((1, n, nz))
    w = np.zeros((n, n, nz))
    S = np.zeros((n, n, nz, 6))
    S[..., 0], S[..., 3], S[..., 1] = s, -s, t
    Dsgs = np.full((n, n, nz), D)
    lam = G.dynamic_procedure(sp, dx, dx, 2.0, u, v, w, sp.fft(u), sp.fft(v), sp.fft(w), S, Dsgs)
    Dm = D - 4 * 2 * np.sqrt(s * s + t * t)
    num = Dm * (s * (Bz ** 2 - C ** 2) / 2 + Bz * C * t)
    expect = np.maximum(num, 0) / (4 * Dm * Dm * (s * s + t * t))
    assert np.all(expect > 0)
    assert np.allclose(lam, expect, rtol=1e-12, atol=0)
    # the opposite strain gives <L:M> < 0 in every plane: clipped to zero
    lam = G.dynamic_procedure(sp, dx, dx, 2.0, u, v, w, sp.fft(u), sp.fft(v), sp.fft(w), -S, Dsgs)
    assert np.all(lam == 0.0)


def test_sigma_is_nicouds_model_of_the_singular_values():
    """The closed-form roots equal numpy's singular values of A: for random
    gradients nu = s3 (s1 - s2) (s2 - s3) / s1^2 with s = svd(A)."""
    rng = np.random.default_rng(21)
    A = rng.standard_normal((500, 3, 3)) * rng.uniform(0.1, 5.0, (500, 1, 1))
    s = np.linalg.svd(A, compute_uv=False)
    ref = s[:, 2] * (s[:, 0] - s[:, 1]) * (s[:, 1] - s[:, 2]) / (s[:, 0] ** 2)
    got = G.sigma(A.reshape(500, 9))
    scale = np.abs(A).reshape(500, -1).max(axis=1)
    assert np.abs(got - ref).max() <= 1e-9 * scale.max()
    assert (got >= 0).all()


def test_sigma_vanishes_where_nicoud_requires():
    """Two-dimensional flow (s3 = 0), solid rotation (s1 = s2) and isotropic
    expansion (all equal) give no eddy viscosity; a pure shear is 2D."""
    rng = np.random.default_rng(22)
    twod = np.zeros((50, 3, 3))
    twod[:, :2, :2] = rng.standard_normal((50, 2, 2))
    w = rng.standard_normal((50, 3))
    rot = np.zeros((50, 3, 3))
    rot[:, 0, 1], rot[:, 1, 0] = -w[:, 2], w[:, 2]
    rot[:, 0, 2], rot[:, 2, 0] = w[:, 1], -w[:, 1]
    rot[:, 1, 2], rot[:, 2, 1] = -w[:, 0], w[:, 0]
    iso = np.eye(3)[None] * rng.uniform(0.5, 2.0, (50, 1, 1))
    for A in (twod, rot, iso):
        nu = G.sigma(A.reshape(50, 9))
        assert np.abs(nu).max() <= 1e-6 * np.abs(A).max() ** 1, np.abs(nu).max()


def test_moeng_wall_stress_is_the_log_law():
    """A uniform first-level velocity (U, V) gives
    ustar = kappa |U| / ln(dz / (2 z0)) and a wall stress of magnitude ustar^2
    against the velocity: tau13 = -ustar^2 U/|U|, tau23 = -ustar^2 V/|U|."""
    nx, ny, nz, dz, z0 = 16, 8, 8, 0.05, 1e-4
    sp = Spectral(nx, ny, nz, 2 * np.pi / nx, 2 * np.pi / ny, dz)
    U, V = 3.0, -1.2
    rng = np.random.default_rng(23)
    u = np.empty((nx, ny, nz))
    v = np.empty((nx, ny, nz))
    u[:, :, 0], v[:, :, 0] = U, V
    u[:, :, 1:] = rng.standard_normal((nx, ny, nz - 1))
    v[:, :, 1:] = rng.standard_normal((nx, ny, nz - 1))
    Uspmn, ustar, M, Umn = G.wall_model(sp, u, v, sp.fft(u), dz, z0)
    mag = np.hypot(U, V)
    assert abs(Uspmn - mag) <= 1e-13 * mag and abs(Umn - U) <= 1e-13 * U
    assert abs(ustar - 0.4 * mag / np.log(dz / (2 * z0))) <= 1e-13 * ustar
    assert abs(M * U + ustar ** 2 * U / mag) <= 1e-12 * ustar ** 2
    assert abs(M * V + ustar ** 2 * V / mag) <= 1e-12 * ustar ** 2
