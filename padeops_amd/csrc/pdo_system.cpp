// Operator systems (A y = B f) of the PadeOps compact schemes and the
// partitioned-solve tables built from them.  Coefficients restate the
// reference modules cited at each builder; the algorithm that solves the
// systems on the GPU is this library's own (see pdo_compact.hip).
#include "pdo_system.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <stdexcept>

namespace pdo {

FamShape family_shape(int fam)
{
    switch (fam) {
    case FAM_CD06_D1: return {1, 2, -1, 0, 0};
    case FAM_SYM2C: return {1, 2, +1, 0, 1};
    case FAM_CD10_D1: return {2, 3, -1, 0, 0};
    case FAM_CD10_D2: return {2, 3, +1, 0, 1};
    case FAM_CF90: return {2, 4, +1, 0, 1};
    case FAM_STG_E2C_A: return {1, 2, -1, +1, 0};
    case FAM_STG_C2E_A: return {1, 2, -1, -1, 0};
    case FAM_STG_E2C_S: return {1, 2, +1, +1, 0};
    case FAM_STG_C2E_S: return {1, 2, +1, -1, 0};
    }
    throw std::runtime_error("bad family");
}

// input offsets (relative to the output row) of pair q of a family
static void pair_offsets(const FamShape &f, int q, int &hi, int &lo)
{
    hi = q + (f.shift < 0 ? -1 : 0);
    lo = -q + (f.shift > 0 ? 1 : 0);
}

static void set_band(SystemDesc &s, double be, double al)
{
    s.lhs.assign(s.n_sys, {be, al, 1.0, al, be});
}

// ---------------------------------------------------------------------------
// cd06 (src/derivatives/cd06.F90)
// ---------------------------------------------------------------------------
int build_cd06(SystemDesc &s, int kind, int n, double dx, bool periodic, int bc1, int bcn)
{
    // coefficients cd06.F90:14-21
    const double alpha06d1 = 1.0 / 3.0, a06d1 = (14.0 / 9.0) / 2.0, b06d1 = (1.0 / 9.0) / 4.0;
    const double alpha06d2 = 2.0 / 11.0, a06d2 = 12.0 / 11.0, b06d2 = (3.0 / 11.0) / 4.0;
    if (periodic && n < 6) return 3;     // cd06.F90:146-158
    if (!periodic && n < 8) return 2;    // cd06.F90:172-181
    s = SystemDesc();
    s.n_sys = s.n_in = s.n_out = n;
    s.periodic = periodic;
    s.K = 1;
    const double odx = 1.0 / dx;
    if (kind == KIND_D2) {
        if (!periodic) return -3;  // cd06.F90:330-343 / :492-513: not implemented in the reference
        s.family = FAM_SYM2C;
        s.RB = 2;
        const double A = a06d2 * odx * odx, B = b06d2 * odx * odx;
        s.c[0] = -2.0 * (A + B); s.c[1] = A; s.c[2] = B;
        set_band(s, 0.0, alpha06d2);
        return 0;
    }
    s.family = FAM_CD06_D1;
    s.c[1] = a06d1 * odx; s.c[2] = b06d1 * odx;
    set_band(s, 0.0, alpha06d1);
    if (periodic) { s.RB = 2; return 0; }
    // non-periodic closures, cd06.F90:27-59 (edge scheme of Lele 1992 with weights w1..w3)
    const double ah = alpha06d1, al = 3.0, p = 17.0 / 6.0, q = 1.5, r = 1.5, ss = -1.0 / 6.0;
    const double qp = 0.75, ap = 0.25;
    const double app = ((40 * ah - 1) * q + 7 * (4 * ah - 1) * ss) / (16 * (ah + 2) * q + 8 * (1 - 4 * ah) * ss);
    const double qpp = (app + 2) / 3.0, rpp = (4 * app - 1) / 12.0;
    const double w1 = (2 * ah + 1) / (2 * (q + ss));
    const double w2 = ((8 * ah + 7) * q - 6 * (2 * ah + 1) * r + (8 * ah + 7) * ss) / (9 * (q + ss));
    const double w3 = (4 * (ah + 2) * q + 2 * (1 - 4 * ah) * ss) / (9 * (q + ss));
    s.RB = 3;
    // LHS rows, cd06.F90:271-307.  Only bc = 0 has closure rows; any other
    // value leaves the interior rows in place (the empty "case(1)" branches
    // of cd06.F90:288-292,318-324), truncated at the end of the domain.
    if (bc1 == 0) {
        s.lhs[0] = {0, 0, w1, w1 * al, 0};
        s.lhs[1] = {0, w2 * ap, w2, w2 * ap, 0};
        s.lhs[2] = {0, w3 * app, w3, w3 * app, 0};
    } else {
        s.lhs[0][1] = 0.0;
    }
    if (bcn == 0) {
        s.lhs[n - 1] = {0, w1 * al, w1, 0, 0};
        s.lhs[n - 2] = {0, w2 * ap, w2, w2 * ap, 0};
        s.lhs[n - 3] = {0, w3 * app, w3, w3 * app, 0};
    } else {
        s.lhs[n - 1][3] = 0.0;
    }
    // RHS rows, cd06.F90:578-598
    const double a1 = -w1 * p * odx, b1 = w1 * q * odx, c1 = w1 * r * odx, d1 = w1 * ss * odx;
    const double a2 = w2 * qp * odx, a3 = w3 * qpp * odx, b3 = w3 * rpp * odx;
    s.special[0] = {{0, a1}, {1, b1}, {2, c1}, {3, d1}};
    s.special[1] = {{0, -a2}, {2, a2}};
    s.special[2] = {{0, -b3}, {1, -a3}, {3, a3}, {4, b3}};
    s.special[n - 3] = {{n - 5, -b3}, {n - 4, -a3}, {n - 2, a3}, {n - 1, b3}};
    s.special[n - 2] = {{n - 3, -a2}, {n - 1, a2}};
    s.special[n - 1] = {{n - 1, -a1}, {n - 2, -b1}, {n - 3, -c1}, {n - 4, -d1}};
    return 0;
}

// Symmetric (p=+1) / antisymmetric (p=-1) ends: the interior row folded about
// the end point, with ghost values f(1-j) = p f(1+j) (resp. about n).  This is
// the construction behind the bc = +-1 rows of cd10.F90 and cf90.F90.
static void fold_rhs_rows(SystemDesc &s, int side, int p_in, int nrows)
{
    const FamShape f = family_shape(s.family);
    const int n = s.n_sys;
    for (int k = 0; k < nrows; k++) {
        const int i = side == 0 ? k : n - 1 - k;
        std::map<int, double> row;
        auto add = [&](int col, double w) {
            double sg = 1.0;
            if (col < 0) { col = -col; sg = p_in; }
            if (col > n - 1) { col = 2 * (n - 1) - col; sg = p_in; }
            row[col] += sg * w;
        };
        if (f.center) add(i, s.c[0]);
        for (int q = 1; q <= f.R; q++) {
            int hi, lo;
            pair_offsets(f, q, hi, lo);
            add(i + hi, s.c[q]);
            add(i + lo, f.sigma * s.c[q]);
        }
        std::vector<std::pair<int, double>> v(row.begin(), row.end());
        s.special[i] = v;
    }
}

static void fold_lhs_rows(SystemDesc &s, int side, int p_out, double be, double al)
{
    const int n = s.n_sys;
    if (side == 0) {
        s.lhs[0] = {0, 0, 1.0, al + p_out * al, be + p_out * be};
        s.lhs[1] = {0, al, 1.0 + p_out * be, al, be};
    } else {
        s.lhs[n - 1] = {be + p_out * be, al + p_out * al, 1.0, 0, 0};
        s.lhs[n - 2] = {be, al, 1.0 + p_out * be, al, 0};
    }
}

// ---------------------------------------------------------------------------
// cd10 (src/derivatives/cd10.F90)
// ---------------------------------------------------------------------------
int build_cd10(SystemDesc &s, int kind, int n, double dx, bool periodic, int bc1, int bcn)
{
    // coefficients cd10.F90:15-27
    const double alpha10d1 = 0.5, beta10d1 = 1.0 / 20.0;
    const double a10d1 = (17.0 / 12.0) / 2.0, b10d1 = (101.0 / 150.0) / 4.0, c10d1 = (1.0 / 100.0) / 6.0;
    const double alpha10d2 = 334.0 / 899.0, beta10d2 = 43.0 / 1798.0;
    const double a10d2 = 1065.0 / 1798.0, b10d2 = (1038.0 / 899.0) / 4.0, c10d2 = (79.0 / 1798.0) / 9.0;
    if (bc1 < -1 || bc1 > 1 || bcn < -1 || bcn > 1) return 324;  // cd10.F90:2042-2058
    if (n < 8) return 2;                                          // cd10.F90:216-224
    s = SystemDesc();
    s.n_sys = s.n_in = s.n_out = n;
    s.periodic = periodic;
    s.K = 2;
    const double odx = 1.0 / dx, odx2 = odx * odx;
    double al, be;
    if (kind == KIND_D1) {
        s.family = FAM_CD10_D1;
        s.c[1] = a10d1 * odx; s.c[2] = b10d1 * odx; s.c[3] = c10d1 * odx;
        al = alpha10d1; be = beta10d1;
    } else {
        s.family = FAM_CD10_D2;
        s.c[1] = a10d2 * odx2; s.c[2] = b10d2 * odx2; s.c[3] = c10d2 * odx2;
        s.c[0] = -2.0 * (s.c[1] + s.c[2] + s.c[3]);
        al = alpha10d2; be = beta10d2;
    }
    set_band(s, be, al);
    s.RB = 3;
    if (periodic) return 0;
    if (kind == KIND_D1) {
        // first-derivative closures, cd10.F90:34-78 (rows 1-4 / n-3..n)
        const double qh = a10d1, rh = b10d1, sh = c10d1, ah = alpha10d1;
        (void)ah;
        const double a = 3.0, p = -17.0 / 6.0, q = 1.5, r = 1.5, ss = -1.0 / 6.0, qp = 0.75, ap = 0.25;
        const double appp = (8 * rh - 175 * sh) / (18 * rh - 550 * sh);
        const double bppp = (-3 + 8 * appp) / 20.0, qppp = (12 - 7 * appp) / 12.0;
        const double rppp = (568 * appp - 183) / 600.0, sppp = (9 * appp - 4) / 300.0;
        const double t = qh + rh - sh * (qppp / sppp - 1), u = ss * (rh + 2 * sh) - q * (qh + rh + sh);
        const double app = ((17 * u) / (72 * (q + ss) * t) - 8.0 / 9.0) / ((19 * u) / (24 * (q + ss) * t) - 1.0 / 3.0);
        const double bpp = (-1 + 3 * app) / 12.0, qpp = 2.0 * (8 - 3 * app) / 18.0, rpp = (-17 + 57 * app) / 72.0;
        const double w1 = (qh + 2 * rh + 3 * sh) / (q + ss);
        const double w2 = (1 / qp) * (rh + sh * (1 + qppp / sppp) - r * (qh + 2 * rh + 3 * sh) / (q + ss));
        const double w3 = (qh + rh + sh * (1 - qppp / sppp)) / rpp;
        const double w4 = sh / sppp;
        const double A1 = w1 * p * odx, B1 = w1 * q * odx, C1 = w1 * r * odx, D1 = w1 * ss * odx;
        const double A2 = w2 * qp * odx, A3 = w3 * qpp * odx, B3 = w3 * rpp * odx;
        const double A4 = w4 * qppp * odx, B4 = w4 * rppp * odx, C4 = w4 * sppp * odx;
        if (bc1 == 0) {
            // cd10.F90:445-465 (LHS), :1164-1177 (RHS)
            s.lhs[0] = {0, 0, w1, w1 * a, 0};
            s.lhs[1] = {0, w2 * ap, w2, w2 * ap, 0};
            s.lhs[2] = {w3 * bpp, w3 * app, w3, w3 * app, w3 * bpp};
            s.lhs[3] = {w4 * bppp, w4 * appp, w4, w4 * appp, w4 * bppp};
            s.special[0] = {{0, A1}, {1, B1}, {2, C1}, {3, D1}};
            s.special[1] = {{0, -A2}, {2, A2}};
            s.special[2] = {{0, -B3}, {1, -A3}, {3, A3}, {4, B3}};
            s.special[3] = {{0, -C4}, {1, -B4}, {2, -A4}, {4, A4}, {5, B4}, {6, C4}};
        } else {
            fold_lhs_rows(s, 0, -bc1, be, al);
            fold_rhs_rows(s, 0, bc1, 3);
        }
        if (bcn == 0) {
            // cd10.F90:504-523 (LHS), :1214-1226 (RHS)
            s.lhs[n - 1] = {0, w1 * a, w1, 0, 0};
            s.lhs[n - 2] = {0, w2 * ap, w2, w2 * ap, 0};
            s.lhs[n - 3] = {w3 * bpp, w3 * app, w3, w3 * app, w3 * bpp};
            s.lhs[n - 4] = {w4 * bppp, w4 * appp, w4, w4 * appp, w4 * bppp};
            s.special[n - 4] = {{n - 7, -C4}, {n - 6, -B4}, {n - 5, -A4}, {n - 3, A4}, {n - 2, B4}, {n - 1, C4}};
            s.special[n - 3] = {{n - 5, -B3}, {n - 4, -A3}, {n - 2, A3}, {n - 1, B3}};
            s.special[n - 2] = {{n - 3, -A2}, {n - 1, A2}};
            s.special[n - 1] = {{n - 1, -A1}, {n - 2, -B1}, {n - 3, -C1}, {n - 4, -D1}};
        } else {
            fold_lhs_rows(s, 1, -bcn, be, al);
            fold_rhs_rows(s, 1, bcn, 3);
        }
    } else {
        // second-derivative closures, cd10.F90:81-99
        const double b1al = 11.0;
        const double b1a = (11 * b1al + 35) / 12, b1b = -(5 * b1al + 26) / 3, b1c = (b1al + 19) / 2;
        const double b1d = (b1al - 14) / 3, b1e = (11 - b1al) / 12;
        const double b2al = 0.1, b2a = 4 * (1 - b2al) / 3;
        const double b3al = 344.0 / 1179.0, b3be = (38.0 * b3al - 9.0) / 214.0;
        const double b3a = (696 - 1191 * b3al) / 428, b3b = ((2454 * b3al - 294) / 535) / 4.0;
        const double A1 = b1a * odx2, B1 = b1b * odx2, C1 = b1c * odx2, D1 = b1d * odx2, E1 = b1e * odx2;
        const double A2 = b2a * odx2, A3 = b3a * odx2, B3 = b3b * odx2;
        s.RB = 4;
        if (bc1 == 0) {
            // cd10.F90:621-641 (LHS), :1653-1661 (RHS)
            s.lhs[0] = {0, 0, 1, b1al, 0};
            s.lhs[1] = {0, b2al, 1, b2al, 0};
            s.lhs[2] = {b3be, b3al, 1, b3al, b3be};
            s.special[0] = {{0, A1}, {1, B1}, {2, C1}, {3, D1}, {4, E1}};
            s.special[1] = {{0, A2}, {1, -2 * A2}, {2, A2}};
            s.special[2] = {{0, B3}, {1, A3}, {2, -2 * (A3 + B3)}, {3, A3}, {4, B3}};
        } else {
            fold_lhs_rows(s, 0, bc1, be, al);
            fold_rhs_rows(s, 0, bc1, 3);
        }
        if (bcn == 0) {
            // cd10.F90:667-685 (LHS), :1693-1701 (RHS)
            s.lhs[n - 3] = {b3be, b3al, 1, b3al, b3be};
            s.lhs[n - 2] = {0, b2al, 1, b2al, 0};
            s.lhs[n - 1] = {0, b1al, 1, 0, 0};
            s.special[n - 3] = {{n - 5, B3}, {n - 4, A3}, {n - 3, -2 * (A3 + B3)}, {n - 2, A3}, {n - 1, B3}};
            s.special[n - 2] = {{n - 3, A2}, {n - 2, -2 * A2}, {n - 1, A2}};
            s.special[n - 1] = {{n - 1, A1}, {n - 2, B1}, {n - 3, C1}, {n - 4, D1}, {n - 5, E1}};
        } else {
            fold_lhs_rows(s, 1, bcn, be, al);
            fold_rhs_rows(s, 1, bcn, 3);
        }
    }
    return 0;
}

// ---------------------------------------------------------------------------
// cf90 (src/filters/cf90.F90)
// ---------------------------------------------------------------------------
int build_cf90(SystemDesc &s, int n, bool periodic, int bc1, int bcn)
{
    // cf90.F90:15-48
    const double alpha90 = 6.6624e-1, beta90 = 1.6688e-1;
    const double a90 = 9.9965e-1, b90 = 6.6652e-1, c90 = 1.6674e-1, d90 = 4.0e-5, e90 = -5.0e-6;
    const double b2al = 4.997e-1, b2a = 9.997e-1, b2b = 4.9985e-1;
    const double b3al = 6.6624e-1, b3be = 1.6688e-1, b3a = 9.9952e-1, b3b = 6.6656e-1, b3c = 1.668e-1;
    const double b4al = 6.6624e-1, b4be = 1.6688e-1, b4a = 9.9968e-1, b4b = 6.6652e-1, b4c = 1.6672e-1,
                 b4d = 4.0e-5;
    if (bc1 < -1 || bc1 > 1 || bcn < -1 || bcn > 1) return 324;  // cf90.F90:1036-1052
    if (n < 10) return 7;                                         // cf90.F90:123-131
    s = SystemDesc();
    s.n_sys = s.n_in = s.n_out = n;
    s.periodic = periodic;
    s.K = 2;
    s.RB = 4;
    s.family = FAM_CF90;
    s.c[0] = a90; s.c[1] = b90; s.c[2] = c90; s.c[3] = d90; s.c[4] = e90;
    set_band(s, beta90, alpha90);
    if (periodic) return 0;
    if (bc1 == 0) {
        // cf90.F90:292-316 (LHS), :653-667 (RHS)
        s.lhs[0] = {0, 0, 1, 0, 0};
        s.lhs[1] = {0, b2al, 1, b2al, 0};
        s.lhs[2] = {b3be, b3al, 1, b3al, b3be};
        s.lhs[3] = {b4be, b4al, 1, b4al, b4be};
        s.special[0] = {{0, 1.0}};
        s.special[1] = {{0, b2b}, {1, b2a}, {2, b2b}};
        s.special[2] = {{0, b3c}, {1, b3b}, {2, b3a}, {3, b3b}, {4, b3c}};
        s.special[3] = {{0, b4d}, {1, b4c}, {2, b4b}, {3, b4a}, {4, b4b}, {5, b4c}, {6, b4d}};
    } else {
        fold_lhs_rows(s, 0, bc1, beta90, alpha90);
        fold_rhs_rows(s, 0, bc1, 4);
    }
    if (bcn == 0) {
        // cf90.F90:341-365 (LHS), :721-735 (RHS)
        s.lhs[n - 1] = {0, 0, 1, 0, 0};
        s.lhs[n - 2] = {0, b2al, 1, b2al, 0};
        s.lhs[n - 3] = {b3be, b3al, 1, b3al, b3be};
        s.lhs[n - 4] = {b4be, b4al, 1, b4al, b4be};
        s.special[n - 1] = {{n - 1, 1.0}};
        s.special[n - 2] = {{n - 3, b2b}, {n - 2, b2a}, {n - 1, b2b}};
        s.special[n - 3] = {{n - 5, b3c}, {n - 4, b3b}, {n - 3, b3a}, {n - 2, b3b}, {n - 1, b3c}};
        s.special[n - 4] = {{n - 7, b4d}, {n - 6, b4c}, {n - 5, b4b}, {n - 4, b4a}, {n - 3, b4b}, {n - 2, b4c},
                            {n - 1, b4d}};
    } else {
        fold_lhs_rows(s, 1, bcn, beta90, alpha90);
        fold_rhs_rows(s, 1, bcn, 4);
    }
    return 0;
}

// ---------------------------------------------------------------------------
// cd06stagg periodic (src/derivatives/cd06stagg.F90:170-195, 301-629)
// ---------------------------------------------------------------------------
int build_stagg_periodic(SystemDesc &s, int kind, int n, double dz)
{
    if (n <= 4) return 21;  // cd06stagg.F90:181-183
    s = SystemDesc();
    s.n_sys = n;
    s.periodic = true;
    s.K = 1;
    s.RB = 2;
    const double odz = 1.0 / dz;
    const double aD1 = (63.0 / 62.0) * odz, bD1 = ((17.0 / 62.0) / 3.0) * odz;  // :308
    const double aD2 = (12.0 / 11.0) * odz * odz, bD2 = ((3.0 / 11.0) / 4.0) * odz * odz;  // :412
    const double aI = 1.5 * 0.5, bI = 0.1 * 0.5;  // :533-536
    switch (kind) {
    case KIND_STG_D1_E2C:
        s.family = FAM_STG_E2C_A; s.n_in = n + 1; s.n_out = n; s.c[1] = aD1; s.c[2] = bD1;
        set_band(s, 0, 9.0 / 62.0); break;
    case KIND_STG_D1_C2E:
        s.family = FAM_STG_C2E_A; s.n_in = n; s.n_out = n + 1; s.c[1] = aD1; s.c[2] = bD1;
        set_band(s, 0, 9.0 / 62.0); break;
    case KIND_STG_IN_E2C:
        s.family = FAM_STG_E2C_S; s.n_in = n + 1; s.n_out = n; s.c[1] = aI; s.c[2] = bI;
        set_band(s, 0, 3.0 / 10.0); break;
    case KIND_STG_IN_C2E:
        s.family = FAM_STG_C2E_S; s.n_in = n; s.n_out = n + 1; s.c[1] = aI; s.c[2] = bI;
        set_band(s, 0, 3.0 / 10.0); break;
    case KIND_STG_D2_C2C:
        s.family = FAM_SYM2C; s.n_in = n; s.n_out = n; s.c[0] = -2 * (aD2 + bD2); s.c[1] = aD2; s.c[2] = bD2;
        set_band(s, 0, 2.0 / 11.0); break;
    case KIND_STG_D2_E2E:
        // reads edges 1..n with period n (the edge n+1 is not read), writes n+1
        s.family = FAM_SYM2C; s.n_in = n + 1; s.n_out = n + 1; s.c[0] = -2 * (aD2 + bD2); s.c[1] = aD2;
        s.c[2] = bD2; set_band(s, 0, 2.0 / 11.0); break;
    default: return -1;
    }
    return 0;
}

// ---------------------------------------------------------------------------
// cd06stagg, non-periodic (cd06stagg.F90:197-232, STAGG_CD06_files/*.F90)
//
// The even / odd boundary variants are the periodic stencils applied to the
// even / odd reflection of the field about the wall: a cell grid (n points)
// mirrors about the half-cell walls, an edge grid (n+1 points) about its end
// points.  First derivatives flip the parity between input and output,
// second derivatives and interpolation keep it.  The ghost values are folded
// into explicit closure rows (RHS) and into the tridiagonal rows (LHS).  The
// "sided" variants replace the end rows by the one-sided closures of the
// reference (ComputeTri_allRoutines.F90, D1RHS_*_common.F90,
// InterpRHS_*_common.F90); D2 and the E2E first derivative have none.
// ---------------------------------------------------------------------------
namespace {
struct Reflect {
    bool edge;   // edge grid (n+1 points) or cell grid (n points)
    int n;       // cells
    int pb, pt;  // parity (+1 even, -1 odd) at the bottom and top wall
    // map index r (any integer near the range) to (index in range, sign)
    void map(int r, int &j, double &sg) const
    {
        sg = 1.0;
        const int last = edge ? n : n - 1;
        for (int it = 0; it < 4 && (r < 0 || r > last); it++) {
            if (r < 0) { r = edge ? -r : -1 - r; sg *= pb; }
            else { r = edge ? 2 * n - r : 2 * n - 1 - r; sg *= pt; }
        }
        j = r;
    }
};
}  // namespace

int build_stagg_nonperiodic(SystemDesc &s, int kind, int n, double dz, int flags)
{
    if (n <= 4) return 21;  // cd06stagg.F90:216-218
    const bool te = flags & 1, be = flags & 2, ts = flags & 4, bs = flags & 8;
    s = SystemDesc();
    s.periodic = false;
    s.K = 1;
    s.RB = 2;
    const double odz = 1.0 / dz;
    const double aD1 = (63.0 / 62.0) * odz, bD1 = ((17.0 / 62.0) / 3.0) * odz;
    const double aD2 = (12.0 / 11.0) * odz * odz, bD2 = ((3.0 / 11.0) / 4.0) * odz * odz;
    const double aI = 1.5 * 0.5, bI = 0.1 * 0.5;
    const double aC = ((14.0 / 9.0) / 2.0) * odz, bC = ((1.0 / 9.0) / 4.0) * odz;
    bool in_edge, out_edge, flip, has_sided;
    double al;
    switch (kind) {
    case KIND_STG_D1_E2C:
        s.family = FAM_STG_E2C_A; s.c[1] = aD1; s.c[2] = bD1; al = 9.0 / 62.0;
        in_edge = true; out_edge = false; flip = true; has_sided = true; break;
    case KIND_STG_D1_C2E:
        s.family = FAM_STG_C2E_A; s.c[1] = aD1; s.c[2] = bD1; al = 9.0 / 62.0;
        in_edge = false; out_edge = true; flip = true; has_sided = true; break;
    case KIND_STG_IN_E2C:
        s.family = FAM_STG_E2C_S; s.c[1] = aI; s.c[2] = bI; al = 3.0 / 10.0;
        in_edge = true; out_edge = false; flip = false; has_sided = true; break;
    case KIND_STG_IN_C2E:
        s.family = FAM_STG_C2E_S; s.c[1] = aI; s.c[2] = bI; al = 3.0 / 10.0;
        in_edge = false; out_edge = true; flip = false; has_sided = true; break;
    case KIND_STG_D2_C2C:
        s.family = FAM_SYM2C; s.c[0] = -2 * (aD2 + bD2); s.c[1] = aD2; s.c[2] = bD2; al = 2.0 / 11.0;
        in_edge = false; out_edge = false; flip = false; has_sided = false; break;
    case KIND_STG_D2_E2E:
        s.family = FAM_SYM2C; s.c[0] = -2 * (aD2 + bD2); s.c[1] = aD2; s.c[2] = bD2; al = 2.0 / 11.0;
        in_edge = true; out_edge = true; flip = false; has_sided = false; break;
    case KIND_STG_D1_C2C:
        s.family = FAM_CD06_D1; s.c[1] = aC; s.c[2] = bC; al = 1.0 / 3.0;
        in_edge = false; out_edge = false; flip = true; has_sided = true; break;
    case KIND_STG_D1_E2E:
        s.family = FAM_CD06_D1; s.c[1] = aC; s.c[2] = bC; al = 1.0 / 3.0;
        in_edge = true; out_edge = true; flip = true; has_sided = false; break;
    default: return -1;
    }
    s.n_in = in_edge ? n + 1 : n;
    s.n_sys = s.n_out = out_edge ? n + 1 : n;
    const int N = s.n_sys, NI = s.n_in;
    set_band(s, 0, al);
    const Reflect rin{in_edge, n, be ? 1 : -1, te ? 1 : -1};
    const int ob = flip ? -1 : 1;
    const Reflect rout{out_edge, n, ob * rin.pb, ob * rin.pt};
    const FamShape fs = family_shape(s.family);

    // the interior stencil of row i, folded into range
    auto folded_row = [&](int i) {
        std::map<int, double> row;
        auto add = [&](int r, double w) {
            int j; double sg;
            rin.map(r, j, sg);
            row[j] += sg * w;
        };
        if (fs.center) add(i, s.c[0]);
        for (int q = 1; q <= fs.R; q++) {
            int hi, lo;
            pair_offsets(fs, q, hi, lo);
            add(i + hi, s.c[q]);
            add(i + lo, fs.sigma * s.c[q]);
        }
        return row;
    };
    auto reaches_out = [&](int i) {
        for (int q = 1; q <= fs.R; q++) {
            int hi, lo;
            pair_offsets(fs, q, hi, lo);
            if (i + lo < 0 || i + hi >= NI) return true;
        }
        return false;
    };
    auto set_special = [&](int i, const std::map<int, double> &row) {
        std::vector<std::pair<int, double>> v;
        for (auto &kv : row) if (kv.second != 0.0) v.push_back({kv.first, kv.second});
        s.special[i] = v;
    };
    // even / odd closures (RHS and LHS folding)
    for (int i = 0; i < N; i++) {
        if (reaches_out(i)) set_special(i, folded_row(i));
        for (int o = -1; o <= 1; o += 2) {
            const int jj = i + o;
            if (jj >= 0 && jj < N) continue;
            int j; double sg;
            rout.map(jj, j, sg);
            const double w = s.lhs[i][o + 2];
            s.lhs[i][o + 2] = 0.0;
            s.lhs[i][j - i + 2] += sg * w;
        }
    }
    // one-sided closures
    if (has_sided && (bs || ts)) {
        const double w0s = 223.0 / 186.0, w1s = 61.0 / 62.0;
        typedef std::vector<std::pair<int, double>> Row;
        auto put = [&](bool bottom, int k, const Row &rhs, double lo, double dg, double up) {
            // bottom: row k from the start, columns as given; top: mirrored
            // row and columns, LHS neighbours swapped (the reference's
            // top rows are the mirror images of the bottom ones)
            Row r;
            for (auto &cw : rhs) r.push_back({bottom ? cw.first : NI - 1 - cw.first, cw.second});
            const int i = bottom ? k : N - 1 - k;
            s.special[i] = r;
            s.lhs[i] = {0.0, bottom ? lo : up, dg, bottom ? up : lo, 0.0};
        };
        for (int side = 0; side < 2; side++) {
            const bool bottom = side == 0;
            if (bottom ? !bs : !ts) continue;
            const double sgn = bottom ? 1.0 : -1.0;  // first derivatives change sign under mirroring
            switch (kind) {
            case KIND_STG_D1_E2C: {  // D1RHS_E2C_common.F90, ComputeTriD1_E2C
                const double alpha1 = 37.0 / 183.0, alpha0 = -1.0;
                const double a0 = (1.0 / 24.0) * (alpha0 - 23.0), b0 = (1.0 / 8.0) * (-9.0 * alpha0 + 7.0);
                const double c0 = (1.0 / 8.0) * (9.0 * alpha0 + 1.0), d0 = -(1.0 / 24.0) * (alpha0 + 1.0);
                const double a1 = (3.0 / 8.0) * (3.0 - 2.0 * alpha1), b1 = (1.0 / 8.0) * (-1.0 + 22.0 * alpha1);
                const double f0 = sgn * w0s * odz, f1 = sgn * w1s * odz;
                put(bottom, 0, {{0, f0 * a0}, {1, f0 * b0}, {2, f0 * c0}, {3, f0 * d0}}, 0.0, w0s, w0s * alpha0);
                put(bottom, 1, {{0, f1 * (-b1 / 3.0)}, {1, f1 * (-a1)}, {2, f1 * a1}, {3, f1 * (b1 / 3.0)}},
                    w1s * alpha1, w1s, w1s * alpha1);
                break;
            }
            case KIND_STG_D1_C2E: {  // D1RHS_C2E_common.F90, ComputeTriD1_C2E
                const double a0 = -71.0 / 24.0, b0 = 47.0 / 8.0, c0 = -31.0 / 8.0, d0 = 23.0 / 24.0;
                const double a1 = 12.0 / 11.0;
                const double f0 = sgn * odz, f1 = sgn * a1 * odz;
                put(bottom, 0, {{0, f0 * a0}, {1, f0 * b0}, {2, f0 * c0}, {3, f0 * d0}}, 0.0, 1.0, 0.0);
                put(bottom, 1, {{0, -f1}, {1, f1}}, 1.0 / 22.0, 1.0, 1.0 / 22.0);
                break;
            }
            case KIND_STG_D1_C2C: {  // D1RHS_C2C_common.F90, ComputeTriD1_C2C (Lele closures)
                const double ah = 1.0 / 3.0, p = 17.0 / 6.0, q = 1.5, r = 1.5, sl = -1.0 / 6.0, aLOW = 3.0;
                const double q_p = 0.75, alpha_p = 0.25;
                const double alpha_pp = ((40 * ah - 1) * q + 7 * (4 * ah - 1) * sl) /
                                        (16 * (ah + 2) * q + 8 * (1 - 4 * ah) * sl);
                const double q_pp = (1.0 / 3.0) * (alpha_pp + 2), r_pp = (1.0 / 12.0) * (4 * alpha_pp - 1);
                const double w1 = (2 * ah + 1) / (2 * (q + sl));
                const double w2 = ((8 * ah + 7) * q - 6 * (2 * ah + 1) * r + (8 * ah + 7) * sl) / (9 * (q + sl));
                const double w3 = (4 * (ah + 2) * q + 2 * (1 - 4 * ah) * sl) / (9 * (q + sl));
                const double g = sgn * odz;
                put(bottom, 0, {{0, g * w1 * (-p)}, {1, g * w1 * q}, {2, g * w1 * r}, {3, g * w1 * sl}}, 0.0, w1,
                    w1 * aLOW);
                put(bottom, 1, {{0, -g * w2 * q_p}, {2, g * w2 * q_p}}, w2 * alpha_p, w2, w2 * alpha_p);
                put(bottom, 2, {{0, -g * w3 * r_pp}, {1, -g * w3 * q_pp}, {3, g * w3 * q_pp}, {4, g * w3 * r_pp}},
                    w3 * alpha_pp, w3, w3 * alpha_pp);
                break;
            }
            case KIND_STG_IN_E2C: {  // InterpRHS_E2C_common.F90, ComputeTriInterp_E2C
                const double alpha0 = 1.0;
                const double a0 = (1.0 / 16.0) * (5.0 - alpha0), b0 = (1.0 / 16.0) * (9.0 * alpha0 + 15.0);
                const double c0 = (1.0 / 16.0) * (9.0 * alpha0 - 5.0), d0 = (1.0 / 16.0) * (1.0 - alpha0);
                put(bottom, 0, {{0, a0}, {1, b0}, {2, c0}, {3, d0}}, 0.0, 1.0, alpha0);
                break;
            }
            case KIND_STG_IN_C2E: {  // InterpRHS_C2E_common.F90, ComputeTriInterp_C2E
                const double a0 = 15.0 / 8.0, b0 = -5.0 / 4.0, c0 = 3.0 / 8.0;
                const double alpha1 = 1.0 / 6.0, a1 = (1.0 / 8.0) * (9.0 + 10.0 * alpha1);
                put(bottom, 0, {{0, a0}, {1, b0}, {2, c0}}, 0.0, 1.0, 0.0);
                put(bottom, 1, {{0, a1 / 2.0}, {1, a1 / 2.0}}, alpha1, 1.0, alpha1);
                break;
            }
            }
        }
    }
    // every row up to the last closure row of a side is explicit (the kernels
    // read those rows from the closure table)
    int lastL = -1, firstR = N;
    for (auto &kv : s.special) {
        if (kv.first < N / 2) lastL = std::max(lastL, kv.first);
        else firstR = std::min(firstR, kv.first);
    }
    for (int i = 0; i <= lastL; i++)
        if (!s.special.count(i)) set_special(i, folded_row(i));
    for (int i = firstR; i < N; i++)
        if (!s.special.count(i)) set_special(i, folded_row(i));
    return 0;
}

// number of input points read directly before wrapping (see kernels)
static int n_read(const SystemDesc &s)
{
    if (s.family == FAM_STG_E2C_A || s.family == FAM_STG_E2C_S) return s.n_sys + 1;
    return s.n_sys;
}

static int wrap_in(const SystemDesc &s, int r)
{
    if (r < 0) return r + s.n_sys;
    if (r >= n_read(s)) return r - s.n_sys;
    return r;
}

double rhs_row(const SystemDesc &s, int i, const double *f)
{
    auto it = s.special.find(i);
    if (it != s.special.end()) {
        double acc = 0;
        for (auto &cw : it->second) acc += cw.second * f[cw.first];
        return acc;
    }
    const FamShape fs = family_shape(s.family);
    auto at = [&](int r) { return s.periodic ? f[wrap_in(s, r)] : f[r]; };
    double acc = fs.center ? s.c[0] * at(i) : 0.0;
    for (int q = 1; q <= fs.R; q++) {
        int hi, lo;
        pair_offsets(fs, q, hi, lo);
        acc += s.c[q] * (at(i + hi) + fs.sigma * at(i + lo));
    }
    return acc;
}

void dense(const SystemDesc &s, double *A, double *B)
{
    const int n = s.n_sys, ni = s.n_in;
    std::memset(A, 0, sizeof(double) * n * n);
    std::memset(B, 0, sizeof(double) * n * ni);
    for (int i = 0; i < n; i++)
        for (int o = -2; o <= 2; o++) {
            int j = i + o;
            if (s.periodic) j = ((j % n) + n) % n;
            else if (j < 0 || j >= n) continue;
            A[i * n + j] += s.lhs[i][o + 2];
        }
    std::vector<double> e(ni, 0.0);
    for (int c = 0; c < ni; c++) {
        std::fill(e.begin(), e.end(), 0.0);
        e[c] = 1.0;
        for (int i = 0; i < n; i++) B[i * ni + c] = rhs_row(s, i, e.data());
    }
}

// ---------------------------------------------------------------------------
// Partitioned solve tables
// ---------------------------------------------------------------------------
bool choose_partition(int n_sys, int K, int nb, bool periodic, int &m, int &P, int &xr)
{
    // the register-resident kernels hold MAXM rows per lane; a non-periodic
    // system of 32 P + 1 rows (the edge systems of a staggered grid of 32 P
    // cells) gives its last segment one row more
    (void)K; (void)nb;
    m = MAXM;
    P = n_sys / MAXM;
    xr = n_sys % MAXM;
    if (xr > 1 || (xr == 1 && periodic)) return false;
    return P >= 1 && P <= PMAX;
}

// dense Gauss-Jordan inverse with partial pivoting (small matrices)
static void invert(std::vector<double> &M, int N)
{
    std::vector<double> I(N * N, 0.0);
    for (int i = 0; i < N; i++) I[i * N + i] = 1.0;
    for (int c = 0; c < N; c++) {
        int piv = c;
        for (int r = c + 1; r < N; r++)
            if (std::fabs(M[r * N + c]) > std::fabs(M[piv * N + c])) piv = r;
        if (M[piv * N + c] == 0.0) throw std::runtime_error("singular reduced system");
        if (piv != c)
            for (int k = 0; k < N; k++) { std::swap(M[c * N + k], M[piv * N + k]); std::swap(I[c * N + k], I[piv * N + k]); }
        const double d = 1.0 / M[c * N + c];
        for (int k = 0; k < N; k++) { M[c * N + k] *= d; I[c * N + k] *= d; }
        for (int r = 0; r < N; r++) {
            if (r == c) continue;
            const double f = M[r * N + c];
            if (f == 0.0) continue;
            for (int k = 0; k < N; k++) { M[r * N + k] -= f * M[c * N + k]; I[r * N + k] -= f * I[c * N + k]; }
        }
    }
    M = I;
}

int build_spike(const SystemDesc &s, SpikeTables &t, bool force_general)
{
    const int n = s.n_sys, K = s.K;
    int nbl = 0, nbr = 0, wb = 0;
    for (auto &kv : s.special) {
        int i = kv.first;
        if (i < n / 2) nbl = std::max(nbl, i + 1);
        else nbr = std::max(nbr, n - i);
        for (auto &cw : kv.second) {
            int col = cw.first;
            if (i < n / 2) wb = std::max(wb, col + 1);
            else wb = std::max(wb, s.n_in - col);
        }
    }
    if (nbl > NBMAX || nbr > NBMAX || wb > WBMAX) return -1;
    if (!s.special.empty() && s.n_in < WBMAX) return -1;
    int m, P, xr;
    bool fast = !force_general && choose_partition(n, K, std::max(nbl, nbr), s.periodic, m, P, xr);
    if (!fast) { m = n; P = 1; xr = 0; }
    // rows of a segment of type ty (the last, or the only, segment holds xr
    // more), and the row stride of the per-type tables
    const int mt = m + xr;
    auto rows_of = [&](int ty) { return (ty == 2 || ty == 3) ? m + xr : m; };
    t = SpikeTables();
    t.m = m; t.P = P; t.xr = xr; t.general = !fast; t.K = K; t.nbl = nbl; t.nbr = nbr; t.wb = wb;
    t.fac.assign(4 * mt * 5, 0.0);
    t.spk.assign(4 * mt * 4, 0.0);
    t.lhs.assign(4 * mt * 5, 0.0);
    t.G.assign((size_t)P * 2 * K * 2 * K * P, 0.0);
    t.sp.assign(2 * NBMAX * WBMAX, 0.0);

    auto seg_type = [&](int sg) {
        if (s.periodic) return 1;
        if (P == 1) return 3;
        if (sg == 0) return 0;
        if (sg == P - 1) return 2;
        return 1;
    };
    // representative segment of each type
    int rep[4] = {-1, -1, -1, -1};
    for (int sg = 0; sg < P; sg++)
        if (rep[seg_type(sg)] < 0) rep[seg_type(sg)] = sg;

    // per type: local LU, spikes, band
    for (int ty = 0; ty < 4; ty++) {
        const int sg = rep[ty];
        if (sg < 0) continue;
        const int r0 = sg * m;  // segments start every m rows
        const int m = rows_of(ty);
        std::vector<double> l1(m, 0), l2(m, 0), u0(m, 0), u1(m, 0), u2(m, 0);
        auto Aloc = [&](int j, int o) -> double {  // band entry of row j (local), offset o, inside the block
            int jj = j + o;
            if (jj < 0 || jj >= m) return 0.0;
            return s.lhs[r0 + j][o + 2];
        };
        for (int j = 0; j < m; j++) {
            double L2 = 0, L1 = 0;
            if (j >= 2) L2 = Aloc(j, -2) / u0[j - 2];
            if (j >= 1) L1 = (Aloc(j, -1) - (j >= 2 ? L2 * u1[j - 2] : 0.0)) / u0[j - 1];
            double U0 = Aloc(j, 0) - (j >= 1 ? L1 * u1[j - 1] : 0.0) - (j >= 2 ? L2 * u2[j - 2] : 0.0);
            double U1 = Aloc(j, 1) - (j >= 1 ? L1 * u2[j - 1] : 0.0);
            double U2 = Aloc(j, 2);
            if (U0 == 0.0) return -2;
            l1[j] = L1; l2[j] = L2; u0[j] = U0; u1[j] = U1; u2[j] = U2;
        }
        for (int j = 0; j < m; j++) {
            double *F = &t.fac[(ty * mt + j) * 5];
            F[0] = l1[j]; F[1] = l2[j]; F[2] = 1.0 / u0[j]; F[3] = u1[j] / u0[j]; F[4] = u2[j] / u0[j];
            double *Lb = &t.lhs[(ty * mt + j) * 5];
            for (int o = 0; o < 5; o++) Lb[o] = s.lhs[r0 + j][o];
            // entries leaving the domain of a non-periodic system are zero by construction
        }
        auto solve_local = [&](std::vector<double> &v) {
            for (int j = 0; j < m; j++) {
                if (j >= 1) v[j] -= l1[j] * v[j - 1];
                if (j >= 2) v[j] -= l2[j] * v[j - 2];
            }
            for (int j = m - 1; j >= 0; j--) {
                double x = v[j];
                if (j + 1 < m) x -= u1[j] * v[j + 1];
                if (j + 2 < m) x -= u2[j] * v[j + 2];
                v[j] = x / u0[j];
            }
        };
        const bool has_prev = s.periodic || ty == 1 || ty == 2;
        const bool has_next = s.periodic || ty == 0 || ty == 1;
        for (int r = 0; r < K; r++) {
            // coupling columns: previous segment's row m-K+r, next segment's row r
            std::vector<double> bv(m, 0.0), cv(m, 0.0);
            for (int j = 0; j < m; j++)
                for (int o = -2; o <= 2; o++) {
                    const int col = j + o;  // local column
                    if (col < 0 && has_prev && col == -K + r) bv[j] += s.lhs[r0 + j][o + 2];
                    if (col >= m && has_next && col == m + r) cv[j] += s.lhs[r0 + j][o + 2];
                }
            solve_local(bv);
            solve_local(cv);
            for (int j = 0; j < m; j++) {
                t.spk[(ty * mt + j) * 4 + r] = bv[j];
                t.spk[(ty * mt + j) * 4 + 2 + r] = cv[j];
            }
        }
    }
    // special RHS rows
    for (auto &kv : s.special) {
        const int i = kv.first;
        const bool left = i < n / 2;
        const int jr = left ? i : i - (n - nbr);
        for (auto &cw : kv.second) {
            const int c = left ? cw.first : cw.first - (s.n_in - WBMAX);
            if (c < 0 || c >= WBMAX) return -1;
            t.sp[((left ? 0 : 1) * NBMAX + jr) * WBMAX + c] += cw.second;
        }
    }
    // reduced system over the interface unknowns u_s = (y_s[0..K), y_s[m-K..m))
    const int NR = 2 * K * P;
    std::vector<double> R(NR * NR, 0.0);
    for (int i = 0; i < NR; i++) R[i * NR + i] = 1.0;
    auto prev_of = [&](int sg) { return sg > 0 ? sg - 1 : (s.periodic ? P - 1 : -1); };
    auto next_of = [&](int sg) { return sg < P - 1 ? sg + 1 : (s.periodic ? 0 : -1); };
    for (int sg = 0; sg < P; sg++) {
        const int ty = seg_type(sg);
        const int pv = prev_of(sg), nx = next_of(sg);
        for (int q = 0; q < 2 * K; q++) {
            const int j = q < K ? q : rows_of(ty) - 2 * K + q;  // local row of this unknown
            const int row = sg * 2 * K + q;
            for (int r = 0; r < K; r++) {
                if (pv >= 0) R[row * NR + pv * 2 * K + K + r] += t.spk[(ty * mt + j) * 4 + r];
                if (nx >= 0) R[row * NR + nx * 2 * K + r] += t.spk[(ty * mt + j) * 4 + 2 + r];
            }
        }
    }
    invert(R, NR);
    // decay of the (block circulant, periodic) coupling with segment offset
    t.D32 = t.D64 = t.D32R = t.D32C = P;
    if (s.periodic && P > 1) {
        std::vector<double> blk(P, 0.0);
        for (int q = 0; q < 2 * K; q++)
            for (int dd = 0; dd < P; dd++)
                for (int r = 0; r < 2 * K; r++)
                    blk[dd] = std::max(blk[dd], std::fabs(R[((P - 1) * 2 * K + q) * NR + ((dd + P - 1) % P) * 2 * K + r]));
        const double top = *std::max_element(blk.begin(), blk.end());
        auto reach = [&](double tol) {
            int D = 0;
            for (int dd = 1; dd <= P / 2; dd++)
                if (blk[dd] > tol * top || blk[P - dd] > tol * top) D = dd;
            return D + 1;  // G rows of segment s are R rows of its neighbours (offset +-1)
        };
        t.D32 = reach(1e-9);
        t.D64 = reach(1e-18);
        t.D32R = reach(1e-6);
        t.D32C = reach(1e-3);
    }
    for (int sg = 0; sg < P; sg++) {
        const int pv = prev_of(sg), nx = next_of(sg);
        for (int r = 0; r < K; r++) {
            for (int jj = 0; jj < NR; jj++) {
                if (pv >= 0) t.G[((size_t)sg * 2 * K + r) * NR + jj] = R[(pv * 2 * K + K + r) * NR + jj];
                if (nx >= 0) t.G[((size_t)sg * 2 * K + K + r) * NR + jj] = R[(nx * 2 * K + r) * NR + jj];
            }
        }
    }
    return 0;
}

}  // namespace pdo
