"""Non-periodic cd06stagg (cd06stagg.F90:197-232, STAGG_CD06_files/*): pins of
the oracle and CPU checks of the library's tables (no device work).

Oracle pins:
* the even / odd variants equal the periodic staggered (or collocated cd06)
  operator applied to the even / odd reflection of the field about both
  walls (a 4n-periodic extension), restricted to the original grid;
* the one-sided ("sided") closures are exact for polynomials up to the degree
  they are built for.
Library: the system A y = B f of every operator and boundary combination of
the Pade6stagg object set equals the oracle's, and the partitioned-solve
tables reproduce the oracle's output in a numpy model of the kernels."""
import ctypes
import itertools

import numpy as np
import pytest

import oracle as O
import padeops_amd as P
import spike_model as S

# (operator, input grid, output grid, derivative order) ; periodic equivalent
OPS = {
    O.STG_D1_E2C: ("e", "c", 1, O.STG_D1_E2C), O.STG_D1_C2E: ("c", "e", 1, O.STG_D1_C2E),
    O.STG_D1_C2C: ("c", "c", 1, O.CD06_D1), O.STG_D1_E2E: ("e", "e", 1, O.CD06_D1),
    O.STG_IN_E2C: ("e", "c", 0, O.STG_IN_E2C), O.STG_IN_C2E: ("c", "e", 0, O.STG_IN_C2E),
    O.STG_D2_C2C: ("c", "c", 2, O.STG_D2_C2C), O.STG_D2_E2E: ("e", "e", 2, O.STG_D2_E2E),
}
KIND = {O.STG_D1_E2C: 10, O.STG_D1_C2E: 11, O.STG_D2_C2C: 12, O.STG_D2_E2E: 13, O.STG_IN_E2C: 14,
        O.STG_IN_C2E: 15, O.STG_D1_C2C: 16, O.STG_D1_E2E: 17}
# the nine objects of Pade6stagg (PadeDerOps.F90:90-110): (isTopEven, isBotEven, isTopSided, isBotSided)
PADE6 = {"ES": (1, 1, 0, 1), "SE": (1, 1, 1, 0), "SS": (1, 1, 1, 1), "SO": (0, 0, 1, 0), "OS": (0, 0, 0, 1),
         "OO": (0, 0, 0, 0), "EE": (1, 1, 0, 0), "OE": (0, 1, 0, 0), "EO": (1, 0, 0, 0)}


def _flags(te, be, ts, bs):
    return te * O.TOP_EVEN + be * O.BOT_EVEN + ts * O.TOP_SIDED + bs * O.BOT_SIDED


def _extend(f, grid, n, pb, pt):
    """4n-periodic extension of f (cells: n values, edges: n+1 values) that is
    even (+1) / odd (-1) about the bottom wall (pb) and the top wall (pt)."""
    if grid == "c":
        g = np.empty(4 * n)
        for k in range(4 * n):
            if k < n:
                g[k] = f[k]
            elif k < 2 * n:
                g[k] = pt * f[2 * n - 1 - k]
            elif k < 3 * n:
                g[k] = pt * pb * f[k - 2 * n]
            else:
                g[k] = pb * f[4 * n - 1 - k]
        return g
    g = np.empty(4 * n + 1)
    for k in range(4 * n + 1):
        if k <= n:
            g[k] = f[k]
        elif k <= 2 * n:
            g[k] = pt * f[2 * n - k]
        elif k <= 3 * n:
            g[k] = pt * pb * f[k - 2 * n]
        else:
            g[k] = pb * f[4 * n - k]
    return g


@pytest.mark.parametrize("op", list(OPS))
@pytest.mark.parametrize("pb,pt", list(itertools.product([1, -1], repeat=2)))
def test_oracle_even_odd_is_reflected_periodic(op, pb, pt):
    n, dz = 24, 0.07
    gi, go, d, per = OPS[op]
    rng = np.random.default_rng(op * 7 + pb + 3 * pt)
    f = rng.standard_normal(n + 1 if gi == "e" else n)
    if gi == "e":
        # an odd edge field vanishes on its wall
        if pb < 0:
            f[0] = 0.0
        if pt < 0:
            f[n] = 0.0
    y = O.stagg_np(op, f.reshape(1, 1, -1), dz, _flags(pt > 0, pb > 0, 0, 0), n).ravel()
    g = _extend(f, gi, n, pb, pt)
    if per == O.CD06_D1:
        gp = g[: 4 * n]
        yp = O.apply(O.CD06_D1, gp.reshape(1, 1, -1), 3, dz, True).ravel()
        if go == "e":
            yp = np.append(yp, yp[0])
    else:
        yp = O.apply(per, g.reshape(1, 1, -1), 3, dz, True, n=4 * n).ravel()
    ref = yp[: n + 1 if go == "e" else n]
    assert np.abs(y - ref).max() <= 1e-12 * max(1.0, np.abs(ref).max())


# degree up to which each one-sided closure (with the interior scheme) is exact
SIDED_DEGREE = {O.STG_D1_E2C: 3, O.STG_D1_C2E: 3, O.STG_D1_C2C: 4, O.STG_IN_E2C: 3, O.STG_IN_C2E: 2}


@pytest.mark.parametrize("op", list(SIDED_DEGREE))
def test_oracle_sided_closures_polynomial_exact(op):
    n, L = 32, 1.0
    dz = L / n
    gi, go, d, _ = OPS[op]
    zc, ze = (np.arange(n) + 0.5) * dz, np.arange(n + 1) * dz
    zi, zo = (ze if gi == "e" else zc), (ze if go == "e" else zc)
    flags = _flags(1, 1, 1, 1)
    for deg in range(SIDED_DEGREE[op] + 1):
        f = (zi - 0.37) ** deg
        if d == 1:
            ex = deg * (zo - 0.37) ** (deg - 1) if deg >= 1 else 0 * zo
        else:
            ex = (zo - 0.37) ** deg
        y = O.stagg_np(op, f.reshape(1, 1, -1), dz, flags, n).ravel()
        assert np.abs(y - ex).max() < 1e-11, (deg, np.abs(y - ex).max())
    # and not beyond: the closure is what limits the order
    deg = SIDED_DEGREE[op] + 1
    f = (zi - 0.37) ** deg
    ex = deg * (zo - 0.37) ** (deg - 1) if d == 1 else (zo - 0.37) ** deg
    y = O.stagg_np(op, f.reshape(1, 1, -1), dz, flags, n).ravel()
    assert np.abs(y - ex).max() > 1e-9


def _np_obj(n, dz, te, be, ts, bs):
    h = ctypes.c_void_p()
    assert P.lib().pdo_cd06stagg_init(ctypes.byref(h), n, dz, te, be, ts, bs) == 0
    o = P.cd06stagg()
    o._h = h
    o.n = n
    return o


@pytest.mark.parametrize("combo", list(PADE6))
@pytest.mark.parametrize("n", [32, 40])
def test_library_system_matches_oracle(combo, n):
    dz = 0.13
    te, be, ts, bs = PADE6[combo]
    o = _np_obj(n, dz, te, be, ts, bs)
    for op, kind in KIND.items():
        A, B = o.dense(kind)
        Ao, Bo = O.stagg_np_dense(op, n, dz, _flags(te, be, ts, bs))
        M, Mo = np.linalg.solve(A, B), np.linalg.solve(Ao, Bo)
        assert np.abs(M - Mo).max() <= 1e-12 * max(1, np.abs(Mo).max()), (combo, kind)


# n = 64 and 256 cells: the edge systems (ddz_C2E, ddz_E2E, d2dz2_E2E,
# InterpZ_C2E) have n + 1 rows, which the register path holds as n/32
# segments with one row more in the last
@pytest.mark.parametrize("n", [64, 256])
@pytest.mark.parametrize("combo", ["SS", "OE", "ES", "EE", "OO"])
def test_library_partitioned_tables(combo, n):
    dz = 0.11
    te, be, ts, bs = PADE6[combo]
    o = _np_obj(n, dz, te, be, ts, bs)
    for op, kind in KIND.items():
        A, B = o.dense(kind)
        f = np.random.default_rng(kind).standard_normal(B.shape[1])
        ref = O.stagg_np(op, f.reshape(1, 1, -1), dz, _flags(te, be, ts, bs), n).ravel()
        for general in (False, True):
            t = S.tables(o, kind, 0, 0, general)
            if not general:
                assert (t["m"], t["P"], t["xr"]) == (32, n // 32, A.shape[0] - n), (combo, kind)
            d = B @ f
            for r, v in S.closure_rows(t, f, B.shape[1]).items():
                assert abs(v - d[r]) <= 1e-10 * (1 + abs(d[r]))
            y = S.solve(t, d, False)
            assert np.linalg.norm(y - ref) <= 1e-12 * np.linalg.norm(ref), (combo, kind, general)


def test_init_codes():
    h = ctypes.c_void_p()
    assert P.lib().pdo_cd06stagg_init(ctypes.byref(h), 4, 0.1, 1, 1, 0, 0) == 21
    # the collocated derivatives do not exist on a periodic object
    o = P.cd06stagg()
    assert o.init(16, 0.1) == 0
    with pytest.raises(P.PdoError):
        o.dense(16)


def test_pade6stagg_modified_wavenumbers():
    """getmodCD06stagg equals the oracle's modified wavenumber and the closed
    form of the staggered symbol; fd02 is 2 sin(w/2)/dz; Fourier is k."""
    from oracle.spectral import getwavenums
    n, dz = 32, 2 * np.pi / 32
    k = getwavenums(n, dz)
    kp = P.Pade6stagg.getModifiedWavenumbers(k, dz, "cd06")
    assert np.abs(kp - O.modk(2, n, dz)).max() < 1e-13
    w = k * dz
    sym = (2 * (63 / 62) * np.sin(w / 2) + (2 / 3) * (17 / 62) * np.sin(1.5 * w)) / (1 + 2 * (9 / 62) * np.cos(w)) / dz
    assert np.abs(kp - sym).max() < 1e-13
    assert np.abs(P.Pade6stagg.getModifiedWavenumbers(k, dz, "fd02") - 2 * np.sin(w / 2) / dz).max() < 1e-13
    assert np.array_equal(P.Pade6stagg.getModifiedWavenumbers(k, dz, "fourierColl"), k)
