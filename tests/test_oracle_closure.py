"""CPU pins of the oracle's learned-closure algebra (oracle/closure.py), the
checker of the closure kernels.

  * Frame indifference: for a rotation Q, A -> Q A Q^T leaves the features
    unchanged and rotates the stress, tau -> Q tau Q^T.  A wrong sign in a
    basis tensor or an invariant that is not one breaks it.
  * The principal strains are numpy.linalg.eigvalsh of s (descending).
  * Known invariants: -tr r^2 = |w|^2 / 2 and tr(r^2 s) = w.s.w / 4 with the
    vorticity w (vortex stretching).
  * Clark's gradient model Delta^2/12 (A A^T)_dev is the basis combination
    G = (0, -1/12, 1/12, -1/12, 0, 0): fixes T2's sign and T3, T4.
  * The basis tensors are symmetric and traceless.
  * dL/dG against central differences of L.
"""
import numpy as np

import oracle.closure as C


def _rot(rng):
    q, _ = np.linalg.qr(rng.standard_normal((3, 3)))
    return q * np.sign(np.linalg.det(q))


def _grads(n, seed, traceless=True):
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((n, 3, 3))
    if traceless:
        A -= (np.trace(A, axis1=1, axis2=2) / 3)[:, None, None] * np.eye(3)
    return A.reshape(n, 9)


def _full(t6):
    n = len(t6)
    M = np.zeros((n, 3, 3))
    for c, (i, j) in enumerate(C.IDX6):
        M[:, i, j] = t6[:, c]
        M[:, j, i] = t6[:, c]
    return M


def test_frame_indifference():
    rng = np.random.default_rng(1)
    A = _grads(200, 2)
    G = rng.standard_normal((200, 6))
    Q = _rot(rng)
    Ar = (Q @ A.reshape(-1, 3, 3) @ Q.T).reshape(-1, 9)
    X, Xr = C.features(A, 0.5), C.features(Ar, 0.5)
    assert np.abs(X - Xr).max() <= 1e-10 * np.abs(X).max()
    t, tr = C.stress(A, G, 0.3), C.stress(Ar, G, 0.3)
    assert np.abs(Q @ _full(t) @ Q.T - _full(tr)).max() <= 1e-12 * np.abs(t).max()


def test_principal_strains_and_invariants():
    A = _grads(100, 3)
    g = C.scale(A)
    X = C.features(A, 1.0)
    s, r = C.split(A / g)
    lam = np.linalg.eigvalsh(s)[:, ::-1]
    assert np.abs(X[:, 6:9] - lam).max() <= 1e-12
    assert (X[:, 6] >= X[:, 7]).all() and (X[:, 7] >= X[:, 8]).all()
    # vorticity of the normalised gradient: w_i = eps_ijk a_kj
    a = (A / g).reshape(-1, 3, 3)
    w = np.stack([a[:, 2, 1] - a[:, 1, 2], a[:, 0, 2] - a[:, 2, 0], a[:, 1, 0] - a[:, 0, 1]], axis=1)
    assert np.abs(X[:, 1] - np.sqrt(0.5 * (w * w).sum(axis=1))).max() <= 1e-12
    stretch = np.einsum("pi,pij,pj->p", w, s, w) / 4.0
    assert np.abs(X[:, 3] - np.cbrt(stretch)).max() <= 1e-12
    assert np.abs(X[:, 0] - np.sqrt((s * s).sum(axis=(1, 2)))).max() <= 1e-12
    assert (X[:, 9] == 1.0).all()


def test_gradient_model_is_a_basis_combination():
    A = _grads(300, 4)
    G = np.tile([0.0, -1.0 / 12, 1.0 / 12, -1.0 / 12, 0.0, 0.0], (300, 1))
    delta = 0.7
    t = C.stress(A, G, delta)
    ref = C.gradient_model(A, delta)
    assert np.abs(t - ref).max() <= 1e-12 * np.abs(ref).max()
    # a wrong sign of T2 would not reproduce it
    G2 = G.copy()
    G2[:, 1] *= -1
    assert np.abs(C.stress(A, G2, delta) - ref).max() > 1e-3 * np.abs(ref).max()


def test_basis_symmetric_traceless():
    A = _grads(50, 5, traceless=False)
    s, r = C.split(A / C.scale(A))
    T = C.basis(s, r)
    tr = T[:, :, 0] + T[:, :, 3] + T[:, :, 5]
    assert np.abs(tr).max() <= 1e-12
    assert np.abs(np.trace(s, axis1=1, axis2=2)).max() <= 1e-14


def test_loss_gradient_finite_differences():
    rng = np.random.default_rng(6)
    A9 = _grads(40, 7) / 2.0
    G = rng.standard_normal((40, 6))
    T = rng.standard_normal((40, 6))
    L, dG = C.loss(G, A9, T)
    h = 1e-6
    num = np.zeros_like(G)
    for p in range(40):
        for k in range(6):
            Gp, Gm = G.copy(), G.copy()
            Gp[p, k] += h
            Gm[p, k] -= h
            num[p, k] = (C.loss(Gp, A9, T)[0] - C.loss(Gm, A9, T)[0]) / (2 * h)
    assert np.abs(num - dG).max() <= 1e-7 * np.abs(dG).max()
    assert L > 0


def test_dissipation_sign_of_eddy_viscosity():
    """G = (-c |s|, 0 ...) is an eddy viscosity: dissipation 2 c |s| s:s >= 0."""
    A = _grads(100, 8)
    s, _ = C.split(A / C.scale(A))
    smag = np.sqrt(2 * (s * s).sum(axis=(1, 2)))
    G = np.zeros((100, 6))
    G[:, 0] = -0.1 * smag
    assert (C.dissipation_sign(A, G) >= 0).all()
    assert (C.dissipation_sign(A, -G) <= 0).all()
