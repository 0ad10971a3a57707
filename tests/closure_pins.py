"""Defining conditions of the non-periodic closure rows (CPU, test helper).

Every closure row of the reference's compact schemes is a member of a family
whose free parameters were fixed by one extra condition.  These functions
check those conditions on the dense (A, B) pair of an operator, A y = B f:

* order of accuracy, row by row -- the number of coefficients a row has fixes
  the degree of the polynomials it differentiates (interpolates, filters)
  exactly; rows whose family parameter was chosen to gain a degree show that
  degree, and one degree more fails;
* global conservation (Lele 1992, section 4.2; the weights of
  cd06.F90:48-53, cd10.F90:72-78): the weighted closure rows of B telescope,
  so every column sum of B vanishes except the first and the last, which are
  -/+ the interior row sum of A;
* compact stencils: the cd10 second-derivative row 1 (cd10.F90:88-94) and the
  staggered one-sided rows with alpha0 (STAGG_CD06_files/D1RHS_E2C_common.F90,
  InterpRHS_E2C_common.F90) pick the family member whose farthest RHS
  coefficient vanishes (e = (11 - alpha)/12 = 0 at alpha = 11,
  d0 = -(alpha0 + 1)/24 = 0 at alpha0 = -1, d0 = (1 - alpha0)/16 = 0 at
  alpha0 = 1);
* cf90 closure rows (cf90.F90:27-48): each preserves constants (B row sum =
  A row sum) and removes the grid-scale mode (-1)^j (B row times it = 0).

`failures(library)` runs all of them on one build of the oracle and returns
the names of the conditions that do not hold, so a deliberately perturbed
build can be shown to fail at least one.
"""
import numpy as np

import oracle as O

# ---------------------------------------------------------------------------
# row-by-row degree of exactness
# ---------------------------------------------------------------------------


def _deriv(p, x, d):
    if d == 0:
        return x ** p
    if d == 1:
        return p * x ** (p - 1) if p >= 1 else np.zeros_like(x)
    return p * (p - 1) * x ** (p - 2) if p >= 2 else np.zeros_like(x)


def row_degree(A, B, i, d, xin, xout, pmax=13):
    """Largest p such that row i of A y = B f holds for f = x^q, y = d-th
    derivative of f, for every q <= p (positions centred near the row)."""
    c = xout[i] + 0.25
    xi, xo = xin - c, xout - c
    deg = -1
    for p in range(pmax + 1):
        f, g = xi ** p, _deriv(p, xo, d)
        lhs, rhs = A[i] @ g, B[i] @ f
        scale = np.abs(A[i]) @ np.abs(g) + np.abs(B[i]) @ np.abs(f)
        if abs(lhs - rhs) > 1e-11 * max(scale, 1e-300):
            break
        deg = p
    return deg


# Expected degree of each closure row (0-based row index from the bottom end).
# The reasoning, row by row:
#  cd06 d1: row 0 has 5 coefficients (alpha=3, p, q, r, s; cd06.F90:20-24),
#    which fix degree 4 (Lele's fourth-order one-sided scheme); row 1 is the
#    standard fourth-order Pade scheme (alpha=1/4, q=3/4); row 2 is the
#    one-parameter fourth-order tridiagonal family (alpha'', q'', r''),
#    degree 4 for any alpha'' -- alpha'' is fixed by conservation instead.
#  cd10 d1: rows 0, 1 as cd06; row 2 the pentadiagonal sixth-order family
#    in alpha'' (degree 6); row 3 the eighth-order family in alpha''' (degree
#    8); both parameters are fixed by conservation.
#  cd10 d2: row 0 is the third-order family in alpha (degree 4 for any
#    alpha, cd10.F90:88-94); row 1 is the fourth-order Pade second derivative
#    (alpha = 1/10 is the member that reaches degree 5); row 2 has four
#    coefficients (alpha, beta, a, b) and symmetric stencils, so the
#    conditions at degree 2, 4, 6, 8 fix all four: degree 9.
#  cf90: row 1 has (alpha, a, b) and row 2/3 (alpha, beta, a, b, c[, d]);
#    symmetric rows are exact on odd degrees about their centre, the
#    constant-preservation condition gives degree 1, and row 2/3 also match
#    the x^2 term (degree 3); the remaining freedom goes to the transfer
#    function (Nyquist mode removed).
COLLOCATED_DEGREES = {
    O.CD06_D1: (1, {0: 4, 1: 4, 2: 4}),
    O.CD10_D1: (1, {0: 4, 1: 4, 2: 6, 3: 8}),
    O.CD10_D2: (2, {0: 4, 1: 5, 2: 9}),
    O.CF90: (0, {1: 1, 2: 3, 3: 3}),
}

# Staggered one-sided closures (cd06stagg sided ends, STAGG_CD06_files):
#  D1 E2C row 0: family in alpha0 with 4 RHS coefficients, degree 3; row 1:
#    the fourth-order staggered family in alpha1, degree 4.
#  D1 C2E row 0: explicit 4-point one-sided formula (4 coefficients: degree
#    3); row 1: alpha (f'_{i-1} + f'_{i+1}) + f'_i = a (f_{i+1/2} - f_{i-1/2})
#    with a = 1 + 2 alpha is fourth order only at alpha = 1/22: degree 4.
#  D1 C2C: the collocated cd06 closure on the cell grid (as cd06 d1).
#  Interp E2C row 0: family in alpha0 with 4 coefficients, degree 3.
#  Interp C2E row 0: 3-point extrapolation (degree 2); row 1: alpha (y_{i-1}
#    + y_{i+1}) + y_i = a1/2 (f_{i-1/2} + f_{i+1/2}) matches x^2 only at
#    alpha = 1/6: degree 3.
STAGGERED_DEGREES = {
    O.STG_D1_E2C: ("e", "c", 1, {0: 3, 1: 4}),
    O.STG_D1_C2E: ("c", "e", 1, {0: 3, 1: 4}),
    O.STG_D1_C2C: ("c", "c", 1, {0: 4, 1: 4, 2: 4}),
    O.STG_IN_E2C: ("e", "c", 0, {0: 3}),
    O.STG_IN_C2E: ("c", "e", 0, {0: 2, 1: 3}),
}

N = 24


def _both_ends(A, B, d, xin, xout, rows, tag, out):
    # the top end is the mirror image: reverse rows and columns (x -> -x
    # flips the sign of odd derivatives)
    Ar, Br = A[::-1, ::-1], B[::-1, ::-1] * (-1.0) ** d
    xin_r, xout_r = -xin[::-1], -xout[::-1]
    for r, deg in rows.items():
        for end, (AA, BB, xi, xo) in (("bottom", (A, B, xin, xout)), ("top", (Ar, Br, xin_r, xout_r))):
            got = row_degree(AA, BB, r, d, xi, xo)
            if got != deg:
                out.append("%s row %d (%s): degree %d, expected %d" % (tag, r, end, got, deg))


def degree_failures(library=None):
    out = []
    x = np.arange(N, dtype=float)
    for op, (d, rows) in COLLOCATED_DEGREES.items():
        A, B = O.dense(op, N, 1.0, False, 0, 0, library=library)
        _both_ends(A, B, d, x, x, rows, "op%d" % op, out)
    ze, zc = np.arange(N + 1, dtype=float), np.arange(N) + 0.5
    flags = O.TOP_SIDED | O.BOT_SIDED | O.TOP_EVEN | O.BOT_EVEN
    for op, (gi, go, d, rows) in STAGGERED_DEGREES.items():
        A, B = O.stagg_np_dense(op, N, 1.0, flags, library=library)
        xin, xout = (ze if gi == "e" else zc), (ze if go == "e" else zc)
        _both_ends(A, B, d, xin, xout, rows, "stagg%d" % op, out)
    return out


# ---------------------------------------------------------------------------
# global conservation of the first-derivative closures
# ---------------------------------------------------------------------------


def telescoping_failures(library=None):
    """Column sums of B (dx = 1) vanish except at the two end columns, which
    equal -/+ the interior row sum of A (1 + 2 alpha + 2 beta)."""
    out = []
    cases = [("cd06", *O.dense(O.CD06_D1, N, 1.0, False, 0, 0, library=library)),
             ("cd10", *O.dense(O.CD10_D1, N, 1.0, False, 0, 0, library=library)),
             ("stagg C2C", *O.stagg_np_dense(O.STG_D1_C2C, N, 1.0, O.TOP_SIDED | O.BOT_SIDED, library=library))]
    for tag, A, B in cases:
        cs = B.sum(axis=0)
        interior = A[N // 2].sum()
        scale = np.abs(B).sum(axis=0).max()
        if np.abs(cs[1:-1]).max() > 1e-13 * scale:
            out.append("%s: interior column sums of B do not vanish (%.3g)" % (tag, np.abs(cs[1:-1]).max()))
        if abs(cs[0] + interior) > 1e-13 * scale or abs(cs[-1] - interior) > 1e-13 * scale:
            out.append("%s: end column sums %.15g %.15g, expected -/+%.15g" % (tag, cs[0], cs[-1], interior))
    return out


# ---------------------------------------------------------------------------
# compact stencils chosen by a family parameter
# ---------------------------------------------------------------------------


def stencil_failures(library=None):
    out = []
    A, B = O.dense(O.CD10_D2, N, 1.0, False, 0, 0, library=library)
    for end, row in (("bottom", B[0]), ("top", B[-1][::-1])):
        nz = np.flatnonzero(np.abs(row) > 1e-14 * np.abs(row).max())
        if list(nz) != [0, 1, 2, 3]:
            out.append("cd10 d2 row 0 (%s): RHS columns %s, expected 4 points" % (end, list(nz)))
    flags = O.TOP_SIDED | O.BOT_SIDED | O.TOP_EVEN | O.BOT_EVEN
    for op, tag in ((O.STG_D1_E2C, "stagg D1 E2C"), (O.STG_IN_E2C, "stagg interp E2C")):
        A, B = O.stagg_np_dense(op, N, 1.0, flags, library=library)
        for end, row in (("bottom", B[0]), ("top", B[-1][::-1])):
            nz = np.flatnonzero(np.abs(row) > 1e-14 * np.abs(row).max())
            if list(nz) != [0, 1, 2]:
                out.append("%s row 0 (%s): RHS columns %s, expected 3 points" % (tag, end, list(nz)))
    return out


# ---------------------------------------------------------------------------
# cf90 closure rows: constants kept, grid-scale mode removed
# ---------------------------------------------------------------------------


def cf90_failures(library=None):
    out = []
    A, B = O.dense(O.CF90, N, 1.0, False, 0, 0, library=library)
    one = np.ones(N)
    nyq = (-1.0) ** np.arange(N)
    for r in (1, 2, 3, N - 4, N - 3, N - 2):
        if abs(B[r] @ one - A[r] @ one) > 1e-14 * (A[r] @ one):
            out.append("cf90 row %d: constants not preserved (%.3g)" % (r, B[r] @ one - A[r] @ one))
        if abs(B[r] @ nyq) > 1e-14 * np.abs(B[r]).sum():
            out.append("cf90 row %d: Nyquist mode not removed (%.3g)" % (r, B[r] @ nyq))
    # row 0 (and n-1) is the identity (cf90.F90:27)
    for r in (0, N - 1):
        e = np.zeros(N)
        e[r] = 1.0
        if np.abs(A[r] - e).max() or np.abs(B[r] - e).max():
            out.append("cf90 row %d: not the identity" % r)
    return out


def failures(library=None):
    return (degree_failures(library) + telescoping_failures(library) + stencil_failures(library) +
            cf90_failures(library))
