"""Pins of the explicit-filter and Fourier-derivative oracle (oracle/explicit.py)
by what the mathematics fixes: transfer functions on Fourier modes, moment
conditions of the closure rows, the reflection identities of the bc = +-1
rows, and exact derivatives of resolved Fourier modes."""
import numpy as np
import pytest

import oracle.explicit as X

KINDS = ["gaussian", "lstsq"]


def _transfer(c, k):
    return c[0] + 2 * sum(c[q] * np.cos(q * k) for q in range(1, 5))


@pytest.mark.parametrize("kind", KINDS)
@pytest.mark.parametrize("n", [8, 16, 33])
def test_periodic_transfer_function(kind, n):
    x = np.arange(n)
    c = X.coeffs(kind)
    for m in range(0, n // 2 + 1):
        k = 2 * np.pi * m / n
        for f in (np.cos(k * x), np.sin(k * x)):
            got = X.filter_axis0(f[:, None], kind, True)[:, 0]
            assert np.allclose(got, _transfer(c, k) * f, atol=1e-14)


def test_transfer_at_zero_and_nyquist():
    for kind in KINDS:
        c = X.coeffs(kind)
        assert abs(_transfer(c, 0.0) - 1.0) < 1e-7       # constants kept
        assert abs(_transfer(c, np.pi)) < 1e-5           # the grid-scale mode removed
    # the Gaussian filter of width 4 dx: T(k) ~ exp(-k^2 Delta^2 / 24) at small k
    c = X.coeffs("gaussian")
    k = 0.2
    assert abs(_transfer(c, k) - np.exp(-k * k * 16 / 24)) < 1e-4


@pytest.mark.parametrize("kind", KINDS)
def test_closure_rows_moments(kind):
    n = 12
    ones = np.ones((n, 1))
    lin = np.arange(n, dtype=float)[:, None]
    out1 = X.filter_axis0(ones, kind, False, 0, 0)
    assert np.allclose(out1, 1.0, atol=1e-15)  # every row sums to one
    outl = X.filter_axis0(lin, kind, False, 0, 0)[:, 0]
    # symmetric rows 2..n-1 keep linear functions; the one-sided first and last rows do not
    assert np.allclose(outl[1:n - 1], lin[1:n - 1, 0], atol=1e-13)
    assert abs(outl[0] - (5 / 6 * 0 + 1 / 6 * 1)) < 1e-15
    assert abs(outl[n - 1] - (5 / 6 * (n - 1) + 1 / 6 * (n - 2))) < 1e-13


@pytest.mark.parametrize("bc", [1, -1])
def test_reflection_rows_equal_periodic_extension(bc):
    # rows near an end with bc = +-1 equal the periodic filter of the field
    # extended by f(-m) = bc f(m) (left) / f(n-1+m) = bc f(n-1-m) (right)
    rng = np.random.default_rng(3)
    n = 20
    f = rng.standard_normal(n)
    g = np.concatenate([f, bc * f[n - 2:0:-1]])  # period 2n - 2, both ends reflected with bc
    per = X.filter_axis0(g[:, None], "gaussian", True)[:n, 0]
    for b1, bn in ((bc, 0), (0, bc), (bc, bc)):
        got = X.filter_axis0(f[:, None], "gaussian", False, b1, bn)[:, 0]
        if b1 == bc:
            assert np.allclose(got[:4], per[:4], atol=1e-14)
        if bn == bc:
            assert np.allclose(got[n - 4:], per[n - 4:], atol=1e-14)
        assert np.allclose(got[4:n - 4], per[4:n - 4], atol=1e-14)


def test_lstsq_ignores_bc():
    f = np.random.default_rng(4).standard_normal((16, 2))
    a = X.filter_axis0(f, "lstsq", False, 1, -1)
    b = X.filter_axis0(f, "lstsq", False, 0, 0)
    assert np.array_equal(a, b)


def test_axis_and_n1():
    f = np.random.default_rng(5).standard_normal((9, 10, 11))
    for ax in range(3):
        g = X.filter_axis(f, ax, "gaussian", True)
        h = np.moveaxis(X.filter_axis0(np.moveaxis(f, ax, 0), "gaussian", True), 0, ax)
        assert np.array_equal(g, h)
    one = np.random.default_rng(6).standard_normal((1, 4))
    assert np.array_equal(X.filter_axis0(one, "gaussian", False, 1, 1), one)


def test_wavenums():
    n, dx = 8, 0.5
    k = X.wavenums(n, dx)
    assert np.allclose(k, np.array([0, 1, 2, 3, -4]) * 2 * np.pi / (n * dx))
    k7 = X.wavenums(7, 1.0)  # odd n: (0, 1, 2, 3) * 2 pi / 6 ... as the reference's dummy = n - 1
    assert np.allclose(k7, np.array([0, 1, 2, 3]) * 2 * np.pi / 6)


@pytest.mark.parametrize("n", [16, 32])
def test_fourier_derivative_exact(n):
    L = 2 * np.pi
    dx = L / n
    x = np.arange(n) * dx
    f = np.zeros((n, 3))
    d1 = np.zeros_like(f)
    d2 = np.zeros_like(f)
    for m, (a, b) in zip(range(1, n // 2), np.random.default_rng(n).standard_normal((n // 2 - 1, 2))):
        f[:, 0] += a * np.cos(m * x) + b * np.sin(m * x)
        d1[:, 0] += m * (-a * np.sin(m * x) + b * np.cos(m * x))
        d2[:, 0] += -m * m * (a * np.cos(m * x) + b * np.sin(m * x))
    # the Nyquist mode is removed by both operators
    f[:, 1] = np.cos(np.pi * np.arange(n))
    g1 = X.fourier_deriv(f, 0, dx, 1)
    g2 = X.fourier_deriv(f, 0, dx, 2)
    assert np.allclose(g1[:, 0], d1[:, 0], atol=1e-10 * n * n)
    assert np.allclose(g2[:, 0], d2[:, 0], atol=1e-9 * n * n)
    assert np.allclose(g1[:, 1], 0, atol=1e-12) and np.allclose(g2[:, 1], 0, atol=1e-12)
    assert np.allclose(g1[:, 2], 0) and np.allclose(g2[:, 2], 0)


# the library's host tables (the matrix its kernel applies) against the oracle
@pytest.mark.parametrize("kind", KINDS)
@pytest.mark.parametrize("periodic", [True, False])
@pytest.mark.parametrize("n", [8, 9, 16, 31])
def test_library_tables_match_oracle(kind, periodic, n):
    import padeops_amd as P
    o = P.gaussian() if kind == "gaussian" else P.lstsq()
    o.init(n, periodic)
    eye = np.eye(n)
    for bc1, bcn in ([(0, 0)] if periodic else [(0, 0), (1, 1), (-1, -1), (1, -1), (-1, 0), (0, 1)]):
        W = o.dense_matrix(bc1, bcn)
        ref = X.filter_axis0(eye, kind, periodic, bc1, bcn)  # column k = filter of e_k
        assert np.allclose(W, ref, atol=1e-16, rtol=0), (bc1, bcn)
