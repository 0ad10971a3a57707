"""TEST INFRASTRUCTURE ONLY (the checker): numpy restatement of the learned
closure's pointwise algebra (padeops_amd/csrc/pdo_closure.hip).  There is no
reference counterpart: the closure is this library's own part; the pins in
tests/test_oracle_closure.py fix it by mathematics (frame invariance, the
eigenvalues of numpy.linalg.eigvalsh, the gradient model as an exact
combination of the basis, finite differences of the loss).

A (n, 9) velocity gradient, component 3 i + j = du_i/dx_j; scale
g = sqrt(mean |A|^2); a = A / g; s = sym(a) - tr(a)/3 I; r = skew(a).
"""
import numpy as np

W6 = np.array([1.0, 2.0, 2.0, 1.0, 2.0, 1.0])
IDX6 = ((0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2))


def scale(A):
    return np.sqrt((A.astype(np.float64) ** 2).sum(axis=1).mean())


def split(a):
    """(s, r) of the normalised gradients a (n, 9)."""
    M = a.reshape(-1, 3, 3).astype(np.float64)
    t3 = np.trace(M, axis1=1, axis2=2) / 3.0
    s = 0.5 * (M + M.transpose(0, 2, 1)) - t3[:, None, None] * np.eye(3)
    r = 0.5 * (M - M.transpose(0, 2, 1))
    return s, r


def _tr(x):
    return np.trace(x, axis1=1, axis2=2)


def _sroot(v, p):
    return np.sign(v) * np.abs(v) ** (1.0 / p)


def features(A, dratio, g=None):
    """X (n, 10) of the gradients A with scale g (default: A's own)."""
    g = scale(A) if g is None else g
    s, r = split(A / g)
    s2, r2 = s @ s, r @ r
    i1, i2, i3 = _tr(s2), -_tr(r2), _tr(s2 @ s)
    i4, i5, i6 = _tr(r2 @ s), _tr(r2 @ s2), _tr(r2 @ s @ r @ s2)
    lam = np.linalg.eigvalsh(s)[:, ::-1]          # descending principal strains
    X = np.stack([np.sqrt(np.maximum(i1, 0)), np.sqrt(np.maximum(i2, 0)), np.cbrt(i3), np.cbrt(i4),
                  np.maximum(-i5, 0) ** 0.25, _sroot(i6, 6.0), lam[:, 0], lam[:, 1], lam[:, 2],
                  np.full(len(A), float(dratio))], axis=1)
    return X


def basis(s, r):
    """T (n, 6 tensors, 6 components 11 12 13 22 23 33)."""
    I = np.eye(3)
    s2, r2 = s @ s, r @ r
    Ts = [s, s @ r - r @ s, s2 - (_tr(s2) / 3)[:, None, None] * I, r2 - (_tr(r2) / 3)[:, None, None] * I,
          r @ s2 - s2 @ r, r2 @ s + s @ r2 - (2 * _tr(s @ r2) / 3)[:, None, None] * I]
    return np.stack([np.stack([0.5 * (t[:, i, j] + t[:, j, i]) for i, j in IDX6], axis=1) for t in Ts], axis=1)


def stress(A, G, delta, g=None):
    """tau (n, 6) = Delta^2 g^2 sum_n G_n T_n(a)."""
    g = scale(A) if g is None else g
    s, r = split(A / g)
    return delta * delta * g * g * np.einsum("pn,pnc->pc", G.astype(np.float64), basis(s, r))


def dissipation_sign(A, G, g=None):
    """-tau_ij s_ij of the normalised stress (its sign decides the fallback)."""
    g = scale(A) if g is None else g
    s, r = split(A / g)
    t = np.einsum("pn,pnc->pc", G.astype(np.float64), basis(s, r))
    S6 = np.stack([s[:, i, j] for i, j in IDX6], axis=1)
    return -(W6 * t * S6).sum(axis=1)


def loss(G, A9, T):
    """(L, dL/dG) of L = sum w_c (p_c - T_c)^2 / (9 B), p = sum_n G_n T_n(A9)."""
    s, r = split(A9)
    Tb = basis(s, r)
    p = np.einsum("pn,pnc->pc", G.astype(np.float64), Tb)
    e = p - T
    B = len(G)
    L = (W6 * e * e).sum() / (9.0 * B)
    dG = np.einsum("pc,pnc->pn", 2 * W6 * e, Tb) / (9.0 * B)
    return L, dG


def gradient_model(A, delta):
    """Clark's model Delta^2/12 (A A^T) deviatoric, components 11 12 13 22 23 33."""
    M = A.reshape(-1, 3, 3).astype(np.float64)
    M = M - (np.trace(M, axis1=1, axis2=2) / 3)[:, None, None] * np.eye(3)
    t = M @ M.transpose(0, 2, 1) * delta * delta / 12.0
    t = t - (_tr(t) / 3)[:, None, None] * np.eye(3)
    return np.stack([t[:, i, j] for i, j in IDX6], axis=1)
