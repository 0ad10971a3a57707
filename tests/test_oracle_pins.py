"""Pins of the CPU oracle (oracle/padeops_oracle.c) -- CPU only.

The reference ships no golden vectors and its Fortran cannot be built here
(every hot-path module imports `mpi` / `decomp_2d`, absent for amdflang), so
the oracle is pinned by what the mathematics fixes:
  * periodic operators are circulant: a single Fourier mode is scaled by the
    closed-form symbol of the scheme (a wrong coefficient, sign or offset
    changes the symbol);
  * non-periodic closures are exact on low-degree polynomials;
  * symmetric / antisymmetric boundary variants equal the periodic operator
    applied to the even / odd extension of the data;
  * the bordered-LU, Thomas and penta solves equal a dense solve of the same
    matrices;
  * the FFT equals numpy.fft, and the Poisson solve reproduces the analytic
    field of the reference's tests/test_PoissonPeriodic.F90;
  * the reference's own test programs (test_cd06, test_cd10, test_cf90),
    which print errors against analytic answers, are rerun as
    manufactured-solution checks.
"""
import itertools

import numpy as np
import pytest

import oracle as O

TWO_PI = 2 * np.pi


def symbol(op, w, dx):
    """Closed-form Fourier symbol s(w) of a periodic operator, w = k dx:
    D f = s f for f = exp(i k x)."""
    if op == O.CD06_D1:
        a, b, al = 14 / 9, 1 / 9, 1 / 3
        return 1j * (a * np.sin(w) + b / 2 * np.sin(2 * w)) / (1 + 2 * al * np.cos(w)) / dx
    if op == O.CD06_D2:
        a, b, al = 12 / 11, 3 / 11, 2 / 11
        return -(2 * a * (1 - np.cos(w)) + b / 2 * (1 - np.cos(2 * w))) / (1 + 2 * al * np.cos(w)) / dx ** 2
    if op == O.CD10_D1:
        a, b, c, al, be = 17 / 12, 101 / 150, 1 / 100, 1 / 2, 1 / 20
        return 1j * (a * np.sin(w) + b / 2 * np.sin(2 * w) + c / 3 * np.sin(3 * w)) / (
            1 + 2 * al * np.cos(w) + 2 * be * np.cos(2 * w)) / dx
    if op == O.CD10_D2:
        a, b, c, al, be = 1065 / 1798, 1038 / 899, 79 / 1798, 334 / 899, 43 / 1798
        return -(2 * a * (1 - np.cos(w)) + b / 2 * (1 - np.cos(2 * w)) + 2 * c / 9 * (1 - np.cos(3 * w))) / (
            1 + 2 * al * np.cos(w) + 2 * be * np.cos(2 * w)) / dx ** 2
    if op == O.CF90:
        # transfer function of tests/test_cf90.F90:108-116
        a90, b90, c90, d90, e90 = 9.9965e-1, 6.6652e-1, 1.6674e-1, 4.0e-5, -5.0e-6
        al, be = 6.6624e-1, 1.6688e-1
        return (a90 + 2 * b90 * np.cos(w) + 2 * c90 * np.cos(2 * w) + 2 * d90 * np.cos(3 * w) +
                2 * e90 * np.cos(4 * w)) / (1 + 2 * al * np.cos(w) + 2 * be * np.cos(2 * w))
    raise ValueError(op)


@pytest.mark.parametrize("op", [O.CD06_D1, O.CD06_D2, O.CD10_D1, O.CD10_D2, O.CF90])
@pytest.mark.parametrize("n", [16, 64])
def test_periodic_symbols(op, n):
    dx = 0.7
    x = np.arange(n) * dx
    for k in range(0, n // 2 + 1):
        w = TWO_PI * k / n
        f = np.exp(1j * k * TWO_PI / n * np.arange(n))
        re = O.apply1d(op, f.real, dx)
        im = O.apply1d(op, f.imag, dx)
        s = symbol(op, w, dx)
        expect = s * f
        scale = max(1.0, abs(s))
        assert np.abs(re - expect.real).max() <= 2e-13 * scale, (op, k)
        assert np.abs(im - expect.imag).max() <= 2e-13 * scale, (op, k)
    del x


def stagg_symbol(op, w, dz):
    """Staggered operators map a mode sampled at edges (z = (j-1) dz) to cells
    (z = (j-1/2) dz) or back; symbol relative to the target points."""
    if op in (O.STG_D1_E2C, O.STG_D1_C2E):
        a, b, al = 63 / 62, 17 / 62, 9 / 62
        return 1j * (2 * a * np.sin(w / 2) + (2 / 3) * b * np.sin(3 * w / 2)) / (1 + 2 * al * np.cos(w)) / dz
    if op in (O.STG_IN_E2C, O.STG_IN_C2E):
        return (1.5 * np.cos(w / 2) + 0.1 * np.cos(3 * w / 2)) / (1 + 0.6 * np.cos(w))
    a, b, al = 12 / 11, 3 / 11, 2 / 11
    return -(2 * a * (1 - np.cos(w)) + b / 2 * (1 - np.cos(2 * w))) / (1 + 2 * al * np.cos(w)) / dz ** 2


@pytest.mark.parametrize("op", [O.STG_D1_E2C, O.STG_D1_C2E, O.STG_D2_C2C, O.STG_D2_E2E, O.STG_IN_E2C,
                                O.STG_IN_C2E])
def test_staggered_symbols(op):
    n, dz = 32, 0.3
    e2c = op in (O.STG_D1_E2C, O.STG_IN_E2C)
    c2e = op in (O.STG_D1_C2E, O.STG_IN_C2E)
    for k in range(1, n // 2):
        kk = TWO_PI * k / (n * dz)
        w = kk * dz
        # source points: edges z_j = j dz (j = 0..n) or cells (j + 1/2) dz
        if op in (O.STG_D1_E2C, O.STG_IN_E2C, O.STG_D2_E2E):
            zs = np.arange(n + 1) * dz
        else:
            zs = (np.arange(n) + 0.5) * dz
        if e2c:
            zt = (np.arange(n) + 0.5) * dz
        elif c2e:
            zt = np.arange(n + 1) * dz
        else:
            zt = zs
        for part in (np.cos, np.sin):
            f = part(kk * zs)
            g = O.apply(op, f.reshape(1, 1, -1, order="F"), 3, dz, True, n=n).reshape(-1)
            s = stagg_symbol(op, w, dz)
            exact = (s * np.exp(1j * kk * zt))
            exact = exact.real if part is np.cos else exact.imag
            assert np.abs(g - exact).max() <= 1e-12 * max(1, abs(s)), (op, k)


@pytest.mark.parametrize("op,deg", [(O.CD06_D1, 4), (O.CD10_D1, 4), (O.CD10_D2, 4), (O.CF90, 1)])
def test_nonperiodic_closures_polynomial_exact(op, deg):
    """Closure rows are exact on polynomials up to their order; one degree
    higher they are not (so the check has teeth)."""
    n, dx = 24, 0.1
    x = np.arange(n) * dx - 1.0
    d = {O.CD06_D1: 1, O.CD10_D1: 1, O.CD10_D2: 2, O.CF90: 0}[op]
    for p in range(deg + 2):
        f = x ** p
        if d == 0:
            exact = f
        else:
            c = np.prod([p - q for q in range(d)]) if p >= d else 0.0
            exact = c * x ** max(p - d, 0) if p >= d else np.zeros(n)
        g = O.apply1d(op, f, dx, periodic=False)
        err = np.abs(g - exact).max()
        if p <= deg:
            assert err < 1e-9, (op, p, err)
        else:
            assert err > 1e-8, (op, p, err)


def _extension(f, p1, pn):
    """Periodic data g of period 4(n-1) built from f by reflecting about point
    1 with parity p1 and about point n with parity pn."""
    n = len(f)
    N = 4 * (n - 1)
    g = np.zeros(N)
    for j in range(N):
        # map index j (0-based, point 1 at 0) back into [0, n-1] by reflections
        idx, s = j, 1.0
        for _ in range(8):
            if idx < 0:
                idx, s = -idx, s * p1
            elif idx > n - 1:
                idx, s = 2 * (n - 1) - idx, s * pn
            else:
                break
        g[j] = s * f[idx]
    # the walk above covers j in [0, 2(n-1)]; continue periodically with the
    # left reflection for the rest
    for j in range(2 * (n - 1) + 1, N):
        g[j] = p1 * g[N - j]
    return g


@pytest.mark.parametrize("op", [O.CD10_D1, O.CD10_D2, O.CF90])
@pytest.mark.parametrize("bc1,bcn", [b for b in itertools.product([1, -1], repeat=2)])
def test_symmetric_antisymmetric_ends_equal_extension(op, bc1, bcn):
    n, dx = 21, 0.2
    rng = np.random.default_rng(5)
    f = rng.standard_normal(n)
    if bc1 == -1:
        f[0] = 0.0
    if bcn == -1:
        f[-1] = 0.0
    g = _extension(f, bc1, bcn)
    per = O.apply1d(op, g, dx, periodic=True)[:n]
    nonper = O.apply1d(op, f, dx, periodic=False, bc1=bc1, bcn=bcn)
    assert np.abs(per - nonper).max() <= 1e-11 * max(1, np.abs(per).max())


CASES = [(O.CD06_D1, True, 0, 0), (O.CD06_D1, False, 0, 0), (O.CD06_D2, True, 0, 0)] + \
        [(op, True, 0, 0) for op in (O.CD10_D1, O.CD10_D2, O.CF90)] + \
        [(op, False, b1, bn) for op in (O.CD10_D1, O.CD10_D2, O.CF90) for b1 in (0, 1, -1) for bn in (0, 1, -1)] + \
        [(op, True, 0, 0) for op in (O.STG_D1_E2C, O.STG_D1_C2E, O.STG_D2_C2C, O.STG_D2_E2E, O.STG_IN_E2C,
                                     O.STG_IN_C2E)]


@pytest.mark.parametrize("op,periodic,bc1,bcn", CASES)
def test_factored_solve_equals_dense_solve(op, periodic, bc1, bcn):
    """The bordered-LU / Thomas / penta solves equal numpy.linalg.solve of the
    matrices the oracle factorizes."""
    n, dx = 19, 0.3
    A, B = O.dense(op, n, dx, periodic, bc1, bcn)
    rng = np.random.default_rng(n)
    f = rng.standard_normal(B.shape[1])
    stag = op >= 10
    g = O.apply(op, f.reshape(1, 1, -1, order="F"), 3, dx, periodic, bc1, bcn, n=n).reshape(-1)
    ref = np.linalg.solve(A, B @ f)
    assert np.abs(g[:n] - ref).max() <= 1e-12 * max(1, np.abs(ref).max())
    if stag and len(g) == n + 1:
        assert g[n] == g[0]


def test_fft_matches_numpy():
    rng = np.random.default_rng(1)
    for shape in [(8, 4, 16), (12, 6, 5)]:
        z = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
        z = np.asfortranarray(z)
        for d in (1, 2, 3):
            got = O.fft_c2c(z, d, -1)
            assert np.abs(got - np.fft.fft(z, axis=d - 1)).max() < 1e-11
            got = O.fft_c2c(z, d, +1)
            assert np.abs(got - np.fft.ifft(z, axis=d - 1) * shape[d - 1]).max() < 1e-11


def test_poisson_analytic():
    """tests/test_PoissonPeriodic.F90:12-25,109-118: 64x32x16, sin(lx)cos(my)sin(nz)."""
    nx, ny, nz = 64, 32, 16
    L = TWO_PI
    x, y, z = (np.arange(m) * L / m for m in (nx, ny, nz))
    X, Y, Z = np.meshgrid(x, y, z, indexing="ij")
    l, m, k = 3, 2, 4
    f = np.sin(l * X) * np.cos(m * Y) * np.sin(k * Z)
    rhs = -(l * l + m * m + k * k) * f
    out = O.poisson_periodic(np.asfortranarray(rhs), L / nx, L / ny, L / nz, 0, 0)
    assert np.abs(out - f).max() < 1e-12
    # modified wavenumbers: exact for the discrete cd06 operator applied twice
    dx = L / nx
    D = lambda a, d: O.apply(O.CD06_D1, a, d, L / (nx, ny, nz)[d - 1], True)
    p = np.asfortranarray(np.random.default_rng(2).standard_normal((nx, ny, nz)))
    lap = D(D(p, 1), 1) + D(D(p, 2), 2) + D(D(p, 3), 3)
    p2 = O.poisson_periodic(lap, dx, L / ny, L / nz, 1, 1)
    # p is recovered up to its components on the null modes of D (mean and the
    # Nyquist modes): apply D once to compare
    for d in (1, 2, 3):
        assert np.abs(D(p2, d) - D(p, d)).max() < 1e-10


def test_wavenumber_ordering():
    """GetWaveNums (PoissonPeriodic.F90:226-261): Nyquist is negative."""
    k = O.modk(0, 8, 1.0)
    assert np.allclose(k * 8 / TWO_PI, [0, 1, 2, 3, -4, -3, -2, -1])


def test_reference_test_cd10_periodic():
    """tests/test_cd10.F90: 32^3 periodic sin/cos, d1 and d2 -- errors at the
    10th-order level."""
    n = 32
    dx = TWO_PI / n
    x = np.arange(n) * dx
    f = np.sin(x).reshape(-1, 1, 1, order="F") * np.ones((1, 2, 2))
    d1 = O.apply(O.CD10_D1, f, 1, dx, True)
    d2 = O.apply(O.CD10_D2, f, 1, dx, True)
    assert np.abs(d1[:, 0, 0] - np.cos(x)).max() < 1e-11
    assert np.abs(d2[:, 0, 0] + np.sin(x)).max() < 1e-9


def test_reference_test_cd06():
    """tests/test_cd06.F90:40: sin(x) differentiated with cd06 (sixth order)."""
    n = 256
    dx = TWO_PI / n
    x = np.arange(n) * dx
    d1 = O.apply1d(O.CD06_D1, np.sin(x), dx)
    assert np.abs(d1 - np.cos(x)).max() < 1e-11


def test_reference_test_cf90_nonperiodic():
    """tests/test_cf90.F90: 32^3 non-periodic, omega = 12, against the
    transfer function T(omega dx); the closures only perturb a few rows."""
    n = 32
    dx = TWO_PI / (n - 1)
    x = np.arange(n) * dx
    f = np.cos(12 * x)
    T = symbol(O.CF90, 12 * dx, dx)
    g = O.apply1d(O.CF90, f, dx, periodic=False)
    err = np.abs(g - T * f)
    assert err[6:-6].max() < 0.05 and err.max() < 0.2


def test_error_codes():
    with pytest.raises(O.OracleError) as e:
        O.apply1d(O.CD10_D1, np.zeros(7), 1.0)
    assert e.value.ierr == 2
    with pytest.raises(O.OracleError) as e:
        O.apply1d(O.CF90, np.zeros(9), 1.0)
    assert e.value.ierr == 7
    with pytest.raises(O.OracleError) as e:
        O.apply1d(O.CD06_D1, np.zeros(5), 1.0)
    assert e.value.ierr == 3
    with pytest.raises(O.OracleError) as e:
        O.apply1d(O.CD10_D1, np.zeros(16), 1.0, periodic=False, bc1=2)
    assert e.value.ierr == 324
    assert np.all(O.apply1d(O.CD06_D1, np.ones(1), 1.0) == 0)
    assert np.all(O.apply1d(O.CF90, np.full(1, 3.0), 1.0) == 3.0)


@pytest.mark.parametrize("kind,op", [(1, O.CD06_D1), (2, O.STG_D1_E2C), (3, O.CD10_D1)])
@pytest.mark.parametrize("n", [16, 24])
def test_modified_wavenumbers_are_the_operators_eigenvalues(kind, op, n):
    """The modified wavenumber of each kind (cd06, cd06stagg, cd10; the
    Get_ModK callbacks of PoissonPeriodic.F90:188-204) is the eigenvalue of
    the discrete periodic first-derivative operator on the wavenumbers the
    Poisson solver uses: D exp(i k x_in) = i k' exp(i k x_out).  For the
    staggered operator the input sits on edges and the output on cells."""
    dx = 0.37
    kp = O.modk(kind, n, dx)
    k = O.modk(0, n, dx)
    xin = np.arange(n + (1 if op == O.STG_D1_E2C else 0)) * dx
    xout = np.arange(n) * dx + (0.5 * dx if op == O.STG_D1_E2C else 0.0)
    for m in range(n):
        re = O.apply(op, np.cos(k[m] * xin).reshape(1, 1, -1), 3, dx, True, n=n).ravel()[:n]
        im = O.apply(op, np.sin(k[m] * xin).reshape(1, 1, -1), 3, dx, True, n=n).ravel()[:n]
        got = re + 1j * im
        expect = 1j * kp[m] * np.exp(1j * k[m] * xout)
        assert np.abs(got - expect).max() <= 1e-12 * max(1.0, abs(kp[m])), (kind, m)
