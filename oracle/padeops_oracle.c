/*
 * padeops_oracle.c -- CPU FP64 restatement of the PadeOps compact-operator /
 * cf90-filter / periodic FFT-Poisson hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker that the HIP library is
 * compared against (tests/, __graft_entry__.smoke(), bench.py cpu_baseline).
 * Nothing in the product library links, loads or calls it.
 *
 * Every routine restates one reference routine and cites it (paths relative to
 * the PadeOps source tree).  Arrays follow the Fortran convention of the
 * reference: a field f(n1,n2,n3) is stored with the first index fastest.
 * Inside the 1-D pencil routines indices are 1-based (element 0 unused) so the
 * recurrences read exactly like the Fortran they restate.
 *
 * Pinning (see tests/test_oracle_*.py): periodic operators against their
 * closed-form Fourier symbols; non-periodic closures by polynomial exactness;
 * the symmetric/antisymmetric boundary variants by the even/odd-extension
 * identity with the periodic operator; the bordered-LU / Thomas / penta solves
 * against a dense solve of the same matrices; FFT against numpy.fft; Poisson
 * against the analytic field of tests/test_PoissonPeriodic.F90.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <complex.h>

#ifdef _OPENMP
#include <omp.h>
#endif

/* operator codes (oracle-private numbering) */
enum {
    ORC_CD06_D1 = 1, ORC_CD06_D2 = 2, ORC_CD10_D1 = 3, ORC_CD10_D2 = 4, ORC_CF90 = 5,
    ORC_STG_D1_E2C = 10, ORC_STG_D1_C2E = 11, ORC_STG_D2_C2C = 12, ORC_STG_D2_E2E = 13,
    ORC_STG_IN_E2C = 14, ORC_STG_IN_C2E = 15
};

/* ------------------------------------------------------------------------ */
/* Scheme constants                                                          */
/* ------------------------------------------------------------------------ */

/* cd06.F90:14-21 */
static const double alpha06d1 = 1.0 / 3.0;
static const double a06d1 = (14.0 / 9.0) / 2.0;
static const double b06d1 = (1.0 / 9.0) / 4.0;
static const double alpha06d2 = 2.0 / 11.0;
static const double a06d2 = (12.0 / 11.0) / 1.0;
static const double b06d2 = (3.0 / 11.0) / 4.0;

/* cd10.F90:15-27 */
static const double alpha10d1 = 1.0 / 2.0;
static const double beta10d1 = 1.0 / 20.0;
static const double a10d1 = (17.0 / 12.0) / 2.0;
static const double b10d1 = (101.0 / 150.0) / 4.0;
static const double c10d1 = (1.0 / 100.0) / 6.0;
static const double alpha10d2 = 334.0 / 899.0;
static const double beta10d2 = 43.0 / 1798.0;
static const double a10d2 = (1065.0 / 1798.0) / 1.0;
static const double b10d2 = (1038.0 / 899.0) / 4.0;
static const double c10d2 = (79.0 / 1798.0) / 9.0;

/* cf90.F90:15-48 */
static const double alpha90 = 6.6624e-1, beta90 = 1.6688e-1;
static const double a90 = 9.9965e-1, b90 = 6.6652e-1, c90 = 1.6674e-1, d90 = 4.0e-5, e90 = -5.0e-6;
static const double b2_alpha90 = 4.997e-1, b2_a90 = 9.997e-1, b2_b90 = 4.9985e-1;
static const double b3_alpha90 = 6.6624e-1, b3_beta90 = 1.6688e-1, b3_a90 = 9.9952e-1,
                    b3_b90 = 6.6656e-1, b3_c90 = 1.668e-1;
static const double b4_alpha90 = 6.6624e-1, b4_beta90 = 1.6688e-1, b4_a90 = 9.9968e-1,
                    b4_b90 = 6.6652e-1, b4_c90 = 1.6672e-1, b4_d90 = 4.0e-5;

/* cd06 non-periodic closure family, cd06.F90:27-59 */
typedef struct {
    double alpha, p, q, r, s, alpha_p, q_p, alpha_pp, q_pp, r_pp, w1, w2, w3;
} cd06_np_t;

static cd06_np_t cd06_np(void)
{
    cd06_np_t c;
    const double ah = alpha06d1;
    c.alpha = 3.0; c.p = 17.0 / 6.0; c.q = 3.0 / 2.0; c.r = 3.0 / 2.0; c.s = -1.0 / 6.0;
    c.q_p = 3.0 / 4.0; c.alpha_p = 1.0 / 4.0;
    c.alpha_pp = ((40 * ah - 1) * c.q + 7 * (4 * ah - 1) * c.s) /
                 (16 * (ah + 2) * c.q + 8 * (1 - 4 * ah) * c.s);
    c.q_pp = (1.0 / 3.0) * (c.alpha_pp + 2);
    c.r_pp = (1.0 / 12.0) * (4 * c.alpha_pp - 1);
    c.w1 = (2 * ah + 1) / (2 * (c.q + c.s));
    c.w2 = ((8 * ah + 7) * c.q - 6 * (2 * ah + 1) * c.r + (8 * ah + 7) * c.s) / (9 * (c.q + c.s));
    c.w3 = (4 * (ah + 2) * c.q + 2 * (1 - 4 * ah) * c.s) / (9 * (c.q + c.s));
    return c;
}

/* cd10 non-periodic closure family, cd10.F90:34-99 */
typedef struct {
    double alpha, p, q, r, s, q_p, alpha_p;
    double alpha_ppp, beta_ppp, q_ppp, r_ppp, s_ppp;
    double alpha_pp, beta_pp, q_pp, r_pp;
    double w1, w2, w3, w4;
    double b1_alpha, b1_a, b1_b, b1_c, b1_d, b1_e, b2_alpha, b2_a, b3_alpha, b3_beta, b3_a, b3_b;
} cd10_np_t;

static cd10_np_t cd10_np(void)
{
    cd10_np_t c;
    const double qh = a10d1, rh = b10d1, sh = c10d1;
    c.alpha = 3.0; c.p = -17.0 / 6.0; c.q = 3.0 / 2.0; c.r = 3.0 / 2.0; c.s = -1.0 / 6.0;
    c.q_p = 3.0 / 4.0; c.alpha_p = 1.0 / 4.0;
    c.alpha_ppp = (8 * rh - 175 * sh) / (18 * rh - 550 * sh);
    c.beta_ppp = (1.0 / 20.0) * (-3 + 8 * c.alpha_ppp);
    c.q_ppp = (1.0 / 12.0) * (12 - 7 * c.alpha_ppp);
    c.r_ppp = (1.0 / 600.0) * (568 * c.alpha_ppp - 183);
    c.s_ppp = (1.0 / 300.0) * (9 * c.alpha_ppp - 4);
    {
        const double q = c.q, s = c.s;
        const double t = qh + rh - sh * (c.q_ppp / c.s_ppp - 1);
        const double u = s * (rh + 2 * sh) - q * (qh + rh + sh);
        c.alpha_pp = ((17 * u) / (72 * (q + s) * t) - 8.0 / 9.0) /
                     ((19 * u) / (24 * (q + s) * t) - 1.0 / 3.0);
    }
    c.beta_pp = (1.0 / 12.0) * (-1 + 3 * c.alpha_pp);
    c.q_pp = (2.0 / 18.0) * (8 - 3 * c.alpha_pp);
    c.r_pp = (1.0 / 72.0) * (-17 + 57 * c.alpha_pp);
    c.w1 = (qh + 2 * rh + 3 * sh) / (c.q + c.s);
    c.w2 = (1 / c.q_p) * (rh + sh * (1 + c.q_ppp / c.s_ppp) - c.r * (qh + 2 * rh + 3 * sh) / (c.q + c.s));
    c.w3 = (qh + rh + sh * (1 - c.q_ppp / c.s_ppp)) / (c.r_pp);
    c.w4 = sh / c.s_ppp;
    /* second-derivative closures, cd10.F90:81-99 */
    c.b1_alpha = 11.0;
    c.b1_a = (11 * c.b1_alpha + 35) / 12;
    c.b1_b = -(5 * c.b1_alpha + 26) / 3;
    c.b1_c = (c.b1_alpha + 19) / 2;
    c.b1_d = (c.b1_alpha - 14) / 3;
    c.b1_e = (11 - c.b1_alpha) / 12;
    c.b2_alpha = 1.0 / 10.0;
    c.b2_a = (4 * (1 - c.b2_alpha) / 3) / 1.0;
    c.b3_alpha = 344.0 / 1179.0;
    c.b3_beta = (38.0 * c.b3_alpha - 9.0) / 214.0;
    c.b3_a = ((696 - 1191 * c.b3_alpha) / 428) / 1.0;
    c.b3_b = ((2454 * c.b3_alpha - 294) / 535) / 4.0;
    return c;
}

/* ------------------------------------------------------------------------ */
/* Factorizations and solves (1-based arrays, column j of LU at LU[j][i])    */
/* ------------------------------------------------------------------------ */

/* Bordered LU of the cyclic tridiagonal [b d a]; cd06.F90:221-262
 * (also cd06stagg.F90:631-672).  Columns: 1 bc, 2 h, 3 1/c, 4 a/c, 5 v/c. */
static void lu_cyclic_tri(double *LU[6], int n, double b, double d, double a)
{
    double *bc = LU[1], *h = LU[2], *c = LU[3], *aa = LU[4], *v = LU[5];
    for (int j = 1; j <= 5; j++) memset(LU[j], 0, sizeof(double) * (n + 1));
    c[1] = d; v[1] = b; h[1] = a / c[1];
    for (int i = 2; i <= n - 1; i++) c[i] = d - (b / c[i - 1]) * a;
    for (int i = 2; i <= n - 2; i++) {
        v[i] = -(b / c[i - 1]) * v[i - 1];
        h[i] = -(a / c[i]) * h[i - 1];
    }
    v[n - 1] = a - (b / c[n - 2]) * v[n - 2];
    h[n - 1] = (b - h[n - 2] * a) / c[n - 1];
    {
        double sum = 0.0;
        for (int i = 1; i <= n - 1; i++) sum += h[i] * v[i];
        c[n] = d - sum;
    }
    for (int i = 2; i <= n - 1; i++) bc[i] = b / c[i - 1];
    for (int i = 1; i <= n - 2; i++) aa[i] = a;
    for (int i = 1; i <= n; i++) c[i] = 1.0 / c[i];
    for (int i = 1; i <= n; i++) aa[i] = aa[i] * c[i];
    for (int i = 1; i <= n; i++) v[i] = v[i] * c[i];
}

/* cd06.F90:345-371 (SolveXLU1), identical per pencil to SolveYLU1/ZLU1 */
static void solve_cyclic_tri(double *const LU[6], int n, double *y)
{
    double sum1 = LU[2][1] * y[1];
    for (int i = 2; i <= n - 1; i++) {
        y[i] = y[i] - LU[1][i] * y[i - 1];
        sum1 = sum1 + LU[2][i] * y[i];
    }
    y[n] = y[n] - sum1;
    y[n] = y[n] * LU[3][n];
    y[n - 1] = y[n - 1] * LU[3][n - 1] - y[n] * LU[5][n - 1];
    for (int i = n - 2; i >= 1; i--)
        y[i] = y[i] * LU[3][i] - y[i + 1] * LU[4][i] - y[n] * LU[5][i];
}

/* Thomas factors from sub/diag/super (a,b,c); cd06.F90:308-325 and
 * STAGG_CD06_files/ComputeTri_allRoutines.F90:36-45.  Tri: 1 a*den, 2 den, 3 cp */
static void tri_factor(double *Tri[4], int n, const double *a, const double *b, const double *c)
{
    double *cp = Tri[3], *den = Tri[2];
    cp[n] = 0.0;
    cp[1] = c[1] / b[1];
    for (int i = 2; i <= n - 1; i++) cp[i] = c[i] / (b[i] - a[i] * cp[i - 1]);
    den[1] = 1.0 / b[1];
    for (int i = 2; i <= n; i++) den[i] = 1.0 / (b[i] - a[i] * cp[i - 1]);
    for (int i = 1; i <= n; i++) Tri[1][i] = a[i] * den[i];
}

/* cd06.F90:432-452 (SolveXTri1); TridiagSolver_allRoutines.F90:1-16 */
static void solve_tri(double *const Tri[4], int n, double *y)
{
    y[1] = y[1] * Tri[2][1];
    for (int i = 2; i <= n; i++) y[i] = y[i] * Tri[2][i] - y[i - 1] * Tri[1][i];
    for (int i = n - 1; i >= 1; i--) y[i] = y[i] - Tri[3][i] * y[i + 1];
}

/* Bordered LU of the cyclic pentadiagonal [e a d c f]; cd10.F90:351-427 and
 * cf90.F90:198-274.  Columns 1 b, 2 eg, 3 k, 4 l, 5 1/g, 6 h, 7 ff, 8 v, 9 w. */
static void lu_cyclic_penta(double *LU[10], int n, double e, double a, double d, double c, double f)
{
    double *b = LU[1], *eg = LU[2], *k = LU[3], *l = LU[4], *g = LU[5], *h = LU[6], *ff = LU[7],
           *v = LU[8], *w = LU[9];
    for (int j = 1; j <= 9; j++) memset(LU[j], 0, sizeof(double) * (n + 1));
    g[1] = d; b[2] = a / g[1]; h[1] = c; k[1] = f / g[1]; w[1] = a; v[1] = e; l[1] = c / g[1];
    g[2] = d - b[2] * h[1]; k[2] = -k[1] * h[1] / g[2]; w[2] = e - b[2] * w[1]; v[2] = -b[2] * v[1];
    l[2] = (f - l[1] * h[1]) / g[2]; h[2] = c - b[2] * f;
    for (int i = 3; i <= n - 3; i++) {
        b[i] = (a - (e / g[i - 2]) * h[i - 2]) / g[i - 1];
        h[i] = c - b[i] * f;
        g[i] = d - (e / g[i - 2]) * f - b[i] * h[i - 1];
    }
    b[n - 2] = (a - (e / g[n - 4]) * h[n - 4]) / g[n - 3];
    g[n - 2] = d - (e / g[n - 4]) * f - b[n - 2] * h[n - 3];
    for (int i = 3; i <= n - 4; i++) {
        k[i] = -(k[i - 2] * f + k[i - 1] * h[i - 1]) / g[i];
        v[i] = -(e / g[i - 2]) * v[i - 2] - b[i] * v[i - 1];
    }
    k[n - 3] = (e - k[n - 5] * f - k[n - 4] * h[n - 4]) / g[n - 3];
    k[n - 2] = (a - k[n - 4] * f - k[n - 3] * h[n - 3]) / g[n - 2];
    v[n - 3] = f - (e / g[n - 5]) * v[n - 5] - b[n - 3] * v[n - 4];
    v[n - 2] = c - (e / g[n - 4]) * v[n - 4] - b[n - 2] * v[n - 3];
    {
        double s = 0.0;
        for (int i = 1; i <= n - 2; i++) s += k[i] * v[i];
        g[n - 1] = d - s;
    }
    for (int i = 3; i <= n - 3; i++) {
        w[i] = -(e / g[i - 2]) * w[i - 2] - b[i] * w[i - 1];
        l[i] = -(l[i - 2] * f + l[i - 1] * h[i - 1]) / g[i];
    }
    w[n - 2] = f - (e / g[n - 4]) * w[n - 4] - b[n - 2] * w[n - 3];
    {
        double s = 0.0;
        for (int i = 1; i <= n - 2; i++) s += k[i] * w[i];
        w[n - 1] = c - s;
    }
    l[n - 2] = (e - l[n - 4] * f - l[n - 3] * h[n - 3]) / g[n - 2];
    {
        double s = 0.0;
        for (int i = 1; i <= n - 2; i++) s += l[i] * v[i];
        l[n - 1] = (a - s) / g[n - 1];
    }
    {
        double s = 0.0;
        for (int i = 1; i <= n - 1; i++) s += l[i] * w[i];
        g[n] = d - s;
    }
    for (int i = 3; i <= n - 2; i++) eg[i] = e / g[i - 2];
    for (int i = 1; i <= n - 4; i++) ff[i] = f;
    for (int i = 1; i <= n; i++) g[i] = 1.0 / g[i];
}

/* cd10.F90:712-749 (SolveXLU1); cf90.F90:420-453 */
static void solve_cyclic_penta(double *const LU[10], int n, double *y)
{
    double sum1, sum2;
    y[2] = y[2] - LU[1][2] * y[1];
    sum1 = LU[3][1] * y[1] + LU[3][2] * y[2];
    sum2 = LU[4][1] * y[1] + LU[4][2] * y[2];
    for (int i = 3; i <= n - 2; i++) {
        y[i] = y[i] - LU[1][i] * y[i - 1] - LU[2][i] * y[i - 2];
        sum1 = sum1 + LU[3][i] * y[i];
        sum2 = sum2 + LU[4][i] * y[i];
    }
    y[n - 1] = y[n - 1] - sum1;
    y[n] = (y[n] - sum2 - LU[4][n - 1] * y[n - 1]) * LU[5][n];
    y[n - 1] = (y[n - 1] - LU[9][n - 1] * y[n]) * LU[5][n - 1];
    y[n - 2] = (y[n - 2] - LU[8][n - 2] * y[n - 1] - LU[9][n - 2] * y[n]) * LU[5][n - 2];
    y[n - 3] = (y[n - 3] - LU[6][n - 3] * y[n - 2] - LU[8][n - 3] * y[n - 1] - LU[9][n - 3] * y[n]) *
               LU[5][n - 3];
    for (int i = n - 4; i >= 1; i--)
        y[i] = (y[i] - LU[6][i] * y[i + 1] - LU[7][i] * y[i + 2] - LU[8][i] * y[n - 1] -
                LU[9][i] * y[n]) * LU[5][i];
}

/* Non-periodic penta factorization from rows (bt,b,d,a,at);
 * cd10.F90:577-595 (ComputePenta1), cf90.F90:396-415.
 * penta columns: 1 bt 2 b 3 d 4 a 5 at 6 e 7 obc 8 f 9 g 10 eobc */
static void penta_factor(double *P[11], int n)
{
    double *bt = P[1], *b = P[2], *d = P[3], *a = P[4], *at = P[5], *e = P[6], *obc = P[7],
           *f = P[8], *g = P[9], *eobc = P[10];
    obc[1] = 1.0 / d[1];
    obc[2] = 1.0 / (d[2] - b[2] * a[1] * obc[1]);
    e[1] = a[1];
    f[2] = b[2] * obc[1];
    for (int i = 3; i <= n; i++) {
        g[i] = bt[i] * obc[i - 2];
        e[i - 1] = a[i - 1] - f[i - 1] * at[i - 2];
        f[i] = (b[i] - g[i] * e[i - 2]) * obc[i - 1];
        obc[i] = 1.0 / (d[i] - f[i] * e[i - 1] - g[i] * at[i - 2]);
    }
    for (int i = 1; i <= n; i++) eobc[i] = e[i] * obc[i];
}

/* cd10.F90:823-850 (SolveXPenta1); cf90.F90:485-508 */
static void solve_penta(double *const P[11], int n, double *y)
{
    y[2] = y[2] - P[8][2] * y[1];
    for (int i = 3; i <= n; i++) y[i] = y[i] - P[9][i] * y[i - 2] - P[8][i] * y[i - 1];
    y[n] = y[n] * P[7][n];
    y[n - 1] = y[n - 1] * P[7][n - 1] - P[10][n - 1] * y[n];
    for (int i = n - 2; i >= 1; i--)
        y[i] = y[i] * P[7][i] - y[i + 2] * P[5][i] * P[7][i] - y[i + 1] * P[10][i];
}

/* ------------------------------------------------------------------------ */
/* Operator context: per-call tables                                          */
/* ------------------------------------------------------------------------ */

typedef struct {
    int op, n, n_in, n_out, periodic, bc1, bcn;
    double dx;
    int kind;                 /* 1 cyclic tri, 2 cyclic penta, 3 tri (Thomas), 4 penta */
    double *col[12];          /* factor columns, 1-based */
    double *buf;
    /* matrix rows before factorization (for the dense export) */
    double *rows[6];          /* 1 bt 2 b 3 d 4 a 5 at (penta) / 2 sub 3 diag 4 sup (tri) */
    double cyc[5];            /* cyclic band e a d c f */
} orc_ctx;

static void ctx_free(orc_ctx *c) { free(c->buf); }

/* fold helper for the symmetric/antisymmetric boundary variants: returns the
 * value of the input at 1-based index j with ghost values reflected about 1
 * (parity p1) and about n (parity pn). */
static inline double ghost(const double *f, int n, int j, int p1, int pn)
{
    if (j < 1) return p1 * f[2 - j];
    if (j > n) return pn * f[2 * n - j];
    return f[j];
}

/* Build LHS rows and factor them.  Returns the reference ierr. */
static int ctx_init(orc_ctx *c, int op, int n, double dx, int periodic, int bc1, int bcn)
{
    memset(c, 0, sizeof(*c));
    c->op = op; c->n = n; c->dx = dx; c->periodic = periodic; c->bc1 = bc1; c->bcn = bcn;
    c->n_in = c->n_out = n;
    if (op == ORC_STG_D1_E2C || op == ORC_STG_IN_E2C) c->n_in = n + 1;
    if (op == ORC_STG_D1_C2E || op == ORC_STG_IN_C2E) c->n_out = n + 1;
    if (op == ORC_STG_D2_E2E) { c->n_in = n + 1; c->n_out = n + 1; }
    if (op != ORC_CD06_D1 && op != ORC_CD06_D2 && ((bc1 < -1 || bc1 > 1) || (bcn < -1 || bcn > 1))) return 324;
    /* n == 1: derivatives return zero and the filter returns its input
     * (cd06.F90:780-783, cd10.F90:2038-2041, cf90.F90:1027-1030) */
    if (n == 1 && op <= ORC_CF90) { c->kind = 0; return 0; }

    /* minimum sizes: cd06.F90:146-158,172-181 ; cd10.F90:216-224,236-262 ;
     * cf90.F90:123-131,147-170 ; cd06stagg.F90:181-183 */
    if (op == ORC_CD06_D1 || op == ORC_CD06_D2) {
        if (periodic && n < 6) return 3;
        if (!periodic && n < 8) return 2;
        if (!periodic && op == ORC_CD06_D2) return 324;      /* cd06.F90:330-343: incomplete */
    } else if (op == ORC_CD10_D1 || op == ORC_CD10_D2) {
        if (n < 8) return 2;
    } else if (op == ORC_CF90) {
        if (n < 10) return 7;
    } else {
        if (n <= 4) return 21;
        if (!periodic) return 21; /* staggered non-periodic closures: not restated here */
    }

    const int N = n + 2;
    c->buf = calloc((size_t)N * 18, sizeof(double));
    for (int j = 0; j < 12; j++) c->col[j] = c->buf + (size_t)j * N;
    for (int j = 0; j < 6; j++) c->rows[j] = c->buf + (size_t)(12 + j) * N;

    if (periodic) {
        double al = 0, be = 0;
        switch (op) {
        case ORC_CD06_D1: al = alpha06d1; break;
        case ORC_CD06_D2: al = alpha06d2; break;
        case ORC_CD10_D1: al = alpha10d1; be = beta10d1; break;
        case ORC_CD10_D2: al = alpha10d2; be = beta10d2; break;
        case ORC_CF90: al = alpha90; be = beta90; break;
        case ORC_STG_D1_E2C: case ORC_STG_D1_C2E: al = 9.0 / 62.0; break;   /* cd06stagg.F90:174 */
        case ORC_STG_D2_C2C: case ORC_STG_D2_E2E: al = 2.0 / 11.0; break;   /* cd06stagg.F90:175 */
        case ORC_STG_IN_E2C: case ORC_STG_IN_C2E: al = 3.0 / 10.0; break;   /* cd06stagg.F90:176 */
        }
        c->cyc[0] = be; c->cyc[1] = al; c->cyc[2] = 1.0; c->cyc[3] = al; c->cyc[4] = be;
        if (be == 0.0) {
            c->kind = 1;
            lu_cyclic_tri(c->col, n, al, 1.0, al);
        } else {
            c->kind = 2;
            lu_cyclic_penta(c->col, n, be, al, 1.0, al, be);
        }
        return 0;
    }

    if (op == ORC_CD06_D1) {
        /* cd06.F90:264-327 (bc 0 at both ends) */
        cd06_np_t k = cd06_np();
        double *a = c->rows[2], *b = c->rows[3], *cc = c->rows[4];
        for (int i = 1; i <= n; i++) { a[i] = alpha06d1; b[i] = 1.0; cc[i] = alpha06d1; }
        /* only bc 0 sets closure rows; other values fall through the empty
         * "case(1)" branches and keep the interior rows (cd06.F90:288-292,318-324) */
        if (bc1 == 0) {
            a[1] = k.w1 * 0.0; a[2] = k.w2 * k.alpha_p; a[3] = k.w3 * k.alpha_pp;
            b[1] = k.w1; b[2] = k.w2; b[3] = k.w3;
            cc[1] = k.w1 * k.alpha; cc[2] = k.w2 * k.alpha_p; cc[3] = k.w3 * k.alpha_pp;
        }
        if (bcn == 0) {
            cc[n] = k.w1 * 0.0; cc[n - 1] = k.w2 * k.alpha_p; cc[n - 2] = k.w3 * k.alpha_pp;
            b[n] = k.w1; b[n - 1] = k.w2; b[n - 2] = k.w3;
            a[n] = k.w1 * k.alpha; a[n - 1] = k.w2 * k.alpha_p; a[n - 2] = k.w3 * k.alpha_pp;
        }
        c->kind = 3;
        tri_factor(c->col, n, a, b, cc);
        return 0;
    }

    /* penta families */
    double *bt = c->rows[1], *b = c->rows[2], *d = c->rows[3], *a = c->rows[4], *at = c->rows[5];
    double al, be;
    int pout1, poutn; /* parity of the output under reflection for bc = +-1 */
    if (op == ORC_CD10_D1) { al = alpha10d1; be = beta10d1; pout1 = -bc1; poutn = -bcn; }
    else if (op == ORC_CD10_D2) { al = alpha10d2; be = beta10d2; pout1 = bc1; poutn = bcn; }
    else { al = alpha90; be = beta90; pout1 = bc1; poutn = bcn; }
    for (int i = 1; i <= n; i++) { bt[i] = be; b[i] = al; d[i] = 1.0; a[i] = al; at[i] = be; }

    if (op == ORC_CD10_D1) {
        cd10_np_t k = cd10_np();
        /* cd10.F90:445-465 and :504-523 */
        if (bc1 == 0) {
            bt[1] = 0; b[1] = 0; d[1] = k.w1; a[1] = k.w1 * k.alpha; at[1] = 0;
            bt[2] = 0; b[2] = k.w2 * k.alpha_p; d[2] = k.w2; a[2] = k.w2 * k.alpha_p; at[2] = 0;
            bt[3] = k.w3 * k.beta_pp; b[3] = k.w3 * k.alpha_pp; d[3] = k.w3; a[3] = k.w3 * k.alpha_pp;
            at[3] = k.w3 * k.beta_pp;
            bt[4] = k.w4 * k.beta_ppp; b[4] = k.w4 * k.alpha_ppp; d[4] = k.w4;
            a[4] = k.w4 * k.alpha_ppp; at[4] = k.w4 * k.beta_ppp;
        }
        if (bcn == 0) {
            bt[n] = 0; b[n] = k.w1 * k.alpha; d[n] = k.w1; a[n] = 0; at[n] = 0;
            bt[n - 1] = 0; b[n - 1] = k.w2 * k.alpha_p; d[n - 1] = k.w2; a[n - 1] = k.w2 * k.alpha_p;
            at[n - 1] = 0;
            bt[n - 2] = k.w3 * k.beta_pp; b[n - 2] = k.w3 * k.alpha_pp; d[n - 2] = k.w3;
            a[n - 2] = k.w3 * k.alpha_pp; at[n - 2] = k.w3 * k.beta_pp;
            bt[n - 3] = k.w4 * k.beta_ppp; b[n - 3] = k.w4 * k.alpha_ppp; d[n - 3] = k.w4;
            a[n - 3] = k.w4 * k.alpha_ppp; at[n - 3] = k.w4 * k.beta_ppp;
        }
    } else if (op == ORC_CD10_D2) {
        cd10_np_t k = cd10_np();
        /* cd10.F90:621-641 and :667-685 */
        if (bc1 == 0) {
            bt[1] = 0; b[1] = 0; d[1] = 1; a[1] = k.b1_alpha; at[1] = 0;
            bt[2] = 0; b[2] = k.b2_alpha; d[2] = 1; a[2] = k.b2_alpha; at[2] = 0;
            bt[3] = k.b3_beta; b[3] = k.b3_alpha; d[3] = 1; a[3] = k.b3_alpha; at[3] = k.b3_beta;
        }
        if (bcn == 0) {
            bt[n - 2] = k.b3_beta; b[n - 2] = k.b3_alpha; d[n - 2] = 1; a[n - 2] = k.b3_alpha;
            at[n - 2] = k.b3_beta;
            bt[n - 1] = 0; b[n - 1] = k.b2_alpha; d[n - 1] = 1; a[n - 1] = k.b2_alpha; at[n - 1] = 0;
            bt[n] = 0; b[n] = k.b1_alpha; d[n] = 1; a[n] = 0; at[n] = 0;
        }
    } else {
        /* cf90.F90:292-316 and :341-365 */
        if (bc1 == 0) {
            bt[1] = 0; b[1] = 0; d[1] = 1; a[1] = 0; at[1] = 0;
            bt[2] = 0; b[2] = b2_alpha90; d[2] = 1; a[2] = b2_alpha90; at[2] = 0;
            bt[3] = b3_beta90; b[3] = b3_alpha90; d[3] = 1; a[3] = b3_alpha90; at[3] = b3_beta90;
            bt[4] = b4_beta90; b[4] = b4_alpha90; d[4] = 1; a[4] = b4_alpha90; at[4] = b4_beta90;
        }
        if (bcn == 0) {
            bt[n] = 0; b[n] = 0; d[n] = 1; a[n] = 0; at[n] = 0;
            bt[n - 1] = 0; b[n - 1] = b2_alpha90; d[n - 1] = 1; a[n - 1] = b2_alpha90; at[n - 1] = 0;
            bt[n - 2] = b3_beta90; b[n - 2] = b3_alpha90; d[n - 2] = 1; a[n - 2] = b3_alpha90;
            at[n - 2] = b3_beta90;
            bt[n - 3] = b4_beta90; b[n - 3] = b4_alpha90; d[n - 3] = 1; a[n - 3] = b4_alpha90;
            at[n - 3] = b4_beta90;
        }
    }
    /* symmetric / antisymmetric ends: the interior rows folded about the end
     * point with the output parity (cd10.F90:467-501,525-557,643-666,686-710;
     * cf90.F90:317-340,366-389). */
    if (bc1 != 0) {
        bt[1] = 0; b[1] = 0; d[1] = 1; a[1] = al + pout1 * al; at[1] = be + pout1 * be;
        bt[2] = 0; b[2] = al; d[2] = 1 + pout1 * be; a[2] = al; at[2] = be;
    }
    if (bcn != 0) {
        bt[n] = be + poutn * be; b[n] = al + poutn * al; d[n] = 1; a[n] = 0; at[n] = 0;
        bt[n - 1] = be; b[n - 1] = al; d[n - 1] = 1 + poutn * be; a[n - 1] = al; at[n - 1] = 0;
    }
    /* factor on a copy of the rows so rows[] keep the matrix */
    for (int j = 1; j <= 5; j++) memcpy(c->col[j], c->rows[j], sizeof(double) * (n + 1));
    c->kind = 4;
    penta_factor(c->col, n);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Right-hand sides (1-based pencils)                                         */
/* ------------------------------------------------------------------------ */

static inline int wrap(int j, int n) { return ((j - 1) % n + n) % n + 1; }

static void rhs_1d(const orc_ctx *c, const double *f, double *R)
{
    const int n = c->n;
    const double odx = 1.0 / c->dx, odx2 = odx / c->dx;
    const int p1 = c->bc1, pn = c->bcn;
    switch (c->op) {
    case ORC_CD06_D1: {
        /* cd06.F90:515-600 (ComputeXD1RHS) */
        const double A = a06d1 * odx, B = b06d1 * odx;
        if (c->periodic) {
            for (int i = 1; i <= n; i++)
                R[i] = A * (f[wrap(i + 1, n)] - f[wrap(i - 1, n)]) + B * (f[wrap(i + 2, n)] - f[wrap(i - 2, n)]);
        } else {
            cd06_np_t k = cd06_np();
            const double a3 = k.w3 * k.q_pp * odx, b3 = k.w3 * k.r_pp * odx, a2 = k.w2 * k.q_p * odx;
            const double a1 = k.w1 * (-k.p * odx), b1 = k.w1 * (k.q * odx), c1 = k.w1 * (k.r * odx),
                         d1 = k.w1 * (k.s * odx);
            R[1] = a1 * f[1] + b1 * f[2] + c1 * f[3] + d1 * f[4];
            R[2] = a2 * (f[3] - f[1]);
            R[3] = a3 * (f[4] - f[2]) + b3 * (f[5] - f[1]);
            for (int i = 4; i <= n - 3; i++) R[i] = B * (f[i + 2] - f[i - 2]) + A * (f[i + 1] - f[i - 1]);
            R[n - 2] = a3 * (f[n - 1] - f[n - 3]) + b3 * (f[n] - f[n - 4]);
            R[n - 1] = a2 * (f[n] - f[n - 2]);
            R[n] = -a1 * f[n] - b1 * f[n - 1] - c1 * f[n - 2] - d1 * f[n - 3];
        }
        break;
    }
    case ORC_CD06_D2: {
        /* cd06.F90:745-773 (periodic RHS, scaled by 1/dx^2 as the staggered
         * C2C second derivative cd06stagg.F90:405-432 with the same a,b) */
        const double A = a06d2 * odx2, B = b06d2 * odx2;
        for (int i = 1; i <= n; i++)
            R[i] = A * (f[wrap(i + 1, n)] - 2 * f[i] + f[wrap(i - 1, n)]) +
                   B * (f[wrap(i + 2, n)] - 2 * f[i] + f[wrap(i - 2, n)]);
        break;
    }
    case ORC_CD10_D1: {
        /* cd10.F90:1098-1262 (ComputeXD1RHS) */
        const double A = a10d1 * odx, B = b10d1 * odx, C = c10d1 * odx;
        if (c->periodic) {
            for (int i = 1; i <= n; i++)
                R[i] = A * (f[wrap(i + 1, n)] - f[wrap(i - 1, n)]) + B * (f[wrap(i + 2, n)] - f[wrap(i - 2, n)]) +
                       C * (f[wrap(i + 3, n)] - f[wrap(i - 3, n)]);
            break;
        }
        cd10_np_t k = cd10_np();
        const double a4 = k.w4 * k.q_ppp * odx, b4 = k.w4 * k.r_ppp * odx, c4 = k.w4 * k.s_ppp * odx;
        const double a3 = k.w3 * k.q_pp * odx, b3 = k.w3 * k.r_pp * odx, a2 = k.w2 * k.q_p * odx;
        const double a1 = k.w1 * (k.p * odx), b1 = k.w1 * (k.q * odx), c1 = k.w1 * (k.r * odx),
                     d1 = k.w1 * (k.s * odx);
        for (int i = 1; i <= n; i++) {
            if (p1 == 0 && i <= 4) continue;
            if (pn == 0 && i >= n - 3) continue;
            R[i] = A * (ghost(f, n, i + 1, p1, pn) - ghost(f, n, i - 1, p1, pn)) +
                   B * (ghost(f, n, i + 2, p1, pn) - ghost(f, n, i - 2, p1, pn)) +
                   C * (ghost(f, n, i + 3, p1, pn) - ghost(f, n, i - 3, p1, pn));
        }
        if (p1 == 0) {
            R[1] = a1 * f[1] + b1 * f[2] + c1 * f[3] + d1 * f[4];
            R[2] = a2 * (f[3] - f[1]);
            R[3] = a3 * (f[4] - f[2]) + b3 * (f[5] - f[1]);
            R[4] = a4 * (f[5] - f[3]) + b4 * (f[6] - f[2]) + c4 * (f[7] - f[1]);
        }
        if (pn == 0) {
            R[n - 3] = a4 * (f[n - 2] - f[n - 4]) + b4 * (f[n - 1] - f[n - 5]) + c4 * (f[n] - f[n - 6]);
            R[n - 2] = a3 * (f[n - 1] - f[n - 3]) + b3 * (f[n] - f[n - 4]);
            R[n - 1] = a2 * (f[n] - f[n - 2]);
            R[n] = -a1 * f[n] - b1 * f[n - 1] - c1 * f[n - 2] - d1 * f[n - 3];
        }
        break;
    }
    case ORC_CD10_D2: {
        /* cd10.F90:1587-1760 (ComputeXD2RHS) */
        const double A = a10d2 * odx2, B = b10d2 * odx2, C = c10d2 * odx2;
        if (c->periodic) {
            for (int i = 1; i <= n; i++)
                R[i] = A * (f[wrap(i + 1, n)] - 2 * f[i] + f[wrap(i - 1, n)]) +
                       B * (f[wrap(i + 2, n)] - 2 * f[i] + f[wrap(i - 2, n)]) +
                       C * (f[wrap(i + 3, n)] - 2 * f[i] + f[wrap(i - 3, n)]);
            break;
        }
        cd10_np_t k = cd10_np();
        const double a3 = k.b3_a * odx2, b3 = k.b3_b * odx2, a2 = k.b2_a * odx2;
        const double a1 = k.b1_a * odx2, b1 = k.b1_b * odx2, c1 = k.b1_c * odx2, d1 = k.b1_d * odx2,
                     e1 = k.b1_e * odx2;
        for (int i = 1; i <= n; i++) {
            if (p1 == 0 && i <= 3) continue;
            if (pn == 0 && i >= n - 2) continue;
            R[i] = A * (ghost(f, n, i + 1, p1, pn) - 2 * f[i] + ghost(f, n, i - 1, p1, pn)) +
                   B * (ghost(f, n, i + 2, p1, pn) - 2 * f[i] + ghost(f, n, i - 2, p1, pn)) +
                   C * (ghost(f, n, i + 3, p1, pn) - 2 * f[i] + ghost(f, n, i - 3, p1, pn));
        }
        if (p1 == 0) {
            R[1] = a1 * f[1] + b1 * f[2] + c1 * f[3] + d1 * f[4] + e1 * f[5];
            R[2] = a2 * (f[3] - 2 * f[2] + f[1]);
            R[3] = a3 * (f[4] - 2 * f[3] + f[2]) + b3 * (f[5] - 2 * f[3] + f[1]);
        }
        if (pn == 0) {
            R[n - 2] = a3 * (f[n - 1] - 2 * f[n - 2] + f[n - 3]) + b3 * (f[n] - 2 * f[n - 2] + f[n - 4]);
            R[n - 1] = a2 * (f[n] - 2 * f[n - 1] + f[n - 2]);
            R[n] = a1 * f[n] + b1 * f[n - 1] + c1 * f[n - 2] + d1 * f[n - 3] + e1 * f[n - 4];
        }
        break;
    }
    case ORC_CF90: {
        /* cf90.F90:555-760 (ComputeXRHS) */
        if (c->periodic) {
            for (int i = 1; i <= n; i++)
                R[i] = a90 * f[i] + b90 * (f[wrap(i + 1, n)] + f[wrap(i - 1, n)]) +
                       c90 * (f[wrap(i + 2, n)] + f[wrap(i - 2, n)]) +
                       d90 * (f[wrap(i + 3, n)] + f[wrap(i - 3, n)]) +
                       e90 * (f[wrap(i + 4, n)] + f[wrap(i - 4, n)]);
            break;
        }
        for (int i = 1; i <= n; i++) {
            if (p1 == 0 && i <= 4) continue;
            if (pn == 0 && i >= n - 3) continue;
            R[i] = a90 * f[i] + b90 * (ghost(f, n, i + 1, p1, pn) + ghost(f, n, i - 1, p1, pn)) +
                   c90 * (ghost(f, n, i + 2, p1, pn) + ghost(f, n, i - 2, p1, pn)) +
                   d90 * (ghost(f, n, i + 3, p1, pn) + ghost(f, n, i - 3, p1, pn)) +
                   e90 * (ghost(f, n, i + 4, p1, pn) + ghost(f, n, i - 4, p1, pn));
        }
        if (p1 == 0) {
            R[1] = 1.0 * f[1];
            R[2] = b2_a90 * f[2] + b2_b90 * (f[3] + f[1]);
            R[3] = b3_a90 * f[3] + b3_b90 * (f[4] + f[2]) + b3_c90 * (f[5] + f[1]);
            R[4] = b4_a90 * f[4] + b4_b90 * (f[5] + f[3]) + b4_c90 * (f[6] + f[2]) + b4_d90 * (f[7] + f[1]);
        }
        if (pn == 0) {
            R[n - 3] = b4_a90 * f[n - 3] + b4_b90 * (f[n - 2] + f[n - 4]) + b4_c90 * (f[n - 1] + f[n - 5]) +
                       b4_d90 * (f[n] + f[n - 6]);
            R[n - 2] = b3_a90 * f[n - 2] + b3_b90 * (f[n - 1] + f[n - 3]) + b3_c90 * (f[n] + f[n - 4]);
            R[n - 1] = b2_a90 * f[n - 1] + b2_b90 * (f[n] + f[n - 2]);
            R[n] = 1.0 * f[n];
        }
        break;
    }
    /* staggered periodic RHS, cd06stagg.F90:301-629.  E arrays hold n+1
     * points; row formulas read E(n+1) where the reference does and wrap with
     * period n elsewhere. */
    case ORC_STG_D1_E2C: case ORC_STG_IN_E2C: {
        const int d1 = (c->op == ORC_STG_D1_E2C);
        const double A = d1 ? (63.0 / 62.0) * odx : (3.0 / 2.0) * 0.5;
        const double B = d1 ? ((17.0 / 62.0) / 3.0) * odx : (1.0 / 10.0) * 0.5;
        const double s = d1 ? -1.0 : 1.0;
        for (int i = 1; i <= n; i++) {
            int ip1 = i + 1;                      /* <= n+1, read directly */
            int ip2 = (i + 2 <= n + 1) ? i + 2 : i + 2 - n;
            int im1 = (i - 1 >= 1) ? i - 1 : n;
            R[i] = A * (f[ip1] + s * f[i]) + B * (f[ip2] + s * f[im1]);
        }
        break;
    }
    case ORC_STG_D1_C2E: case ORC_STG_IN_C2E: {
        const int d1 = (c->op == ORC_STG_D1_C2E);
        const double A = d1 ? (63.0 / 62.0) * odx : (3.0 / 2.0) * 0.5;
        const double B = d1 ? ((17.0 / 62.0) / 3.0) * odx : (1.0 / 10.0) * 0.5;
        const double s = d1 ? -1.0 : 1.0;
        for (int i = 1; i <= n; i++)
            R[i] = A * (f[i] + s * f[wrap(i - 1, n)]) + B * (f[wrap(i + 1, n)] + s * f[wrap(i - 2, n)]);
        break;
    }
    case ORC_STG_D2_C2C: case ORC_STG_D2_E2E: {
        const double A = (12.0 / 11.0) * odx2, B = ((3.0 / 11.0) / 4.0) * odx2;
        for (int i = 1; i <= n; i++)
            R[i] = A * (f[wrap(i + 1, n)] - 2 * f[i] + f[wrap(i - 1, n)]) +
                   B * (f[wrap(i + 2, n)] - 2 * f[i] + f[wrap(i - 2, n)]);
        break;
    }
    }
}

static void solve_1d(const orc_ctx *c, double *y)
{
    switch (c->kind) {
    case 1: solve_cyclic_tri(c->col, c->n, y); break;
    case 2: solve_cyclic_penta(c->col, c->n, y); break;
    case 3: solve_tri(c->col, c->n, y); break;
    case 4: solve_penta(c->col, c->n, y); break;
    }
}

/* one pencil: fin[1..n_in] -> out[1..n_out] */
static void apply_1d(const orc_ctx *c, const double *fin, double *out)
{
    if (c->kind == 0) { out[1] = (c->op == ORC_CF90) ? fin[1] : 0.0; return; }
    rhs_1d(c, fin, out);
    solve_1d(c, out);
    /* periodic C2E / E2E: the edge point n+1 repeats point 1,
     * cd06stagg.F90:858,912,1036 */
    if (c->n_out == c->n + 1) out[c->n + 1] = out[1];
}

/* ------------------------------------------------------------------------ */
/* Public entry points                                                        */
/* ------------------------------------------------------------------------ */

/* Apply operator `op` along direction dir (1,2,3) of a column-major field.
 * in dims (n1,n2,n3) with the dir extent equal to the operator's input length;
 * the output has the same dims with the dir extent equal to the output length.
 * Returns the reference ierr (0 ok, 2/3/7/21 too small, 324 bad bc). */
int orc_apply(int op, int n, double dx, int periodic, int bc1, int bcn, int dir, int n1, int n2,
              int n3, const double *f, double *out)
{
    orc_ctx c;
    int ierr = ctx_init(&c, op, n, dx, periodic, bc1, bcn);
    if (ierr) { ctx_free(&c); return ierr; }
    const int din[4] = {0, n1, n2, n3};
    if (din[dir] != c.n_in) { ctx_free(&c); return -1; }
    int dout[4] = {0, n1, n2, n3};
    dout[dir] = c.n_out;
    /* pencil index (a,b) over the two other directions */
    int oa = (dir == 1) ? 2 : 1, ob = (dir == 3) ? 2 : 3;
    long sin_[4] = {0, 1, (long)din[1], (long)din[1] * din[2]};
    long sout[4] = {0, 1, (long)dout[1], (long)dout[1] * dout[2]};
    long npen = (long)din[oa] * din[ob];
#pragma omp parallel
    {
        double *pin = malloc(sizeof(double) * (c.n_in + 2));
        double *pout = malloc(sizeof(double) * (c.n_out + 2));
#pragma omp for schedule(static)
        for (long p = 0; p < npen; p++) {
            long ia = p % din[oa], ib = p / din[oa];
            long bi = ia * sin_[oa] + ib * sin_[ob];
            long bo = ia * sout[oa] + ib * sout[ob];
            for (int i = 1; i <= c.n_in; i++) pin[i] = f[bi + (long)(i - 1) * sin_[dir]];
            apply_1d(&c, pin, pout);
            for (int i = 1; i <= c.n_out; i++) out[bo + (long)(i - 1) * sout[dir]] = pout[i];
        }
        free(pin);
        free(pout);
    }
    ctx_free(&c);
    return 0;
}

/* Dense matrices used by the oracle: A (n x n, the LHS actually factorized)
 * and B (n x n_in, the RHS stencil rows of the solved system), row-major. */
int orc_dense(int op, int n, double dx, int periodic, int bc1, int bcn, double *A, double *B)
{
    orc_ctx c;
    int ierr = ctx_init(&c, op, n, dx, periodic, bc1, bcn);
    if (ierr) { ctx_free(&c); return ierr; }
    /* rows of the solved system: n (the staggered edge row n+1 is a copy) */
    const int no = c.n, ni = c.n_in;
    if (c.kind == 0) { A[0] = 1.0; B[0] = (op == ORC_CF90) ? 1.0 : 0.0; ctx_free(&c); return 0; }
    memset(A, 0, sizeof(double) * no * no);
    if (periodic) {
        for (int i = 0; i < n; i++)
            for (int o = -2; o <= 2; o++) A[i * no + ((i + o) % n + n) % n] += c.cyc[o + 2];
    } else if (c.kind == 3) {
        for (int i = 0; i < n; i++) {
            if (i > 0) A[i * no + i - 1] = c.rows[2][i + 1];
            A[i * no + i] = c.rows[3][i + 1];
            if (i < n - 1) A[i * no + i + 1] = c.rows[4][i + 1];
        }
    } else {
        for (int i = 0; i < n; i++) {
            const double r[5] = {c.rows[1][i + 1], c.rows[2][i + 1], c.rows[3][i + 1], c.rows[4][i + 1],
                                 c.rows[5][i + 1]};
            for (int o = -2; o <= 2; o++)
                if (i + o >= 0 && i + o < n) A[i * no + i + o] = r[o + 2];
        }
    }
    double *e = calloc(ni + 2, sizeof(double)), *R = calloc(c.n_out + 2, sizeof(double));
    for (int j = 0; j < ni; j++) {
        memset(e, 0, sizeof(double) * (ni + 2));
        e[j + 1] = 1.0;
        rhs_1d(&c, e, R);
        for (int i = 0; i < no; i++) B[i * ni + j] = R[i + 1];
    }
    free(e); free(R);
    ctx_free(&c);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* FFT (unnormalized, sign -1 forward) and the periodic Poisson solve          */
/* ------------------------------------------------------------------------ */

/* tw[k] = exp(sign 2 pi i k / n), k < n: the twiddle table of fft_line */
static double complex *twiddles(int n, int sign)
{
    double complex *tw = malloc(sizeof(double complex) * n);
    for (int k = 0; k < n; k++) {
        const double ang = sign * 2.0 * M_PI * (double)k / n;
        tw[k] = cos(ang) + I * sin(ang);
    }
    return tw;
}

static void fft_line(double complex *x, int n, const double complex *tw, double complex *tmp)
{
    if ((n & (n - 1)) == 0) {
        /* iterative radix-2 Cooley-Tukey; stage `len` uses every (n/len)-th twiddle */
        for (int i = 1, j = 0; i < n; i++) {
            int bit = n >> 1;
            for (; j & bit; bit >>= 1) j ^= bit;
            j ^= bit;
            if (i < j) { double complex t = x[i]; x[i] = x[j]; x[j] = t; }
        }
        for (int len = 2; len <= n; len <<= 1) {
            const int step = n / len;
            for (int i = 0; i < n; i += len)
                for (int k = 0; k < len / 2; k++) {
                    const double complex w = tw[k * step];
                    double complex u = x[i + k], v = x[i + k + len / 2] * w;
                    x[i + k] = u + v;
                    x[i + k + len / 2] = u - v;
                }
        }
    } else {
        for (int k = 0; k < n; k++) {
            double complex s = 0;
            for (int j = 0; j < n; j++) s += x[j] * tw[(long)j * k % n];
            tmp[k] = s;
        }
        memcpy(x, tmp, sizeof(double complex) * n);
    }
}

/* in-place c2c transform of all lines along dir of a complex (n1,n2,n3) array */
int orc_fft_c2c_dir(int dir, int n1, int n2, int n3, double *z, int sign)
{
    double complex *a = (double complex *)z;
    const int d[4] = {0, n1, n2, n3};
    long s[4] = {0, 1, n1, (long)n1 * n2};
    int oa = (dir == 1) ? 2 : 1, ob = (dir == 3) ? 2 : 3;
    long npen = (long)d[oa] * d[ob];
    const int n = d[dir];
    double complex *tw = twiddles(n, sign);
#pragma omp parallel
    {
        double complex *line = malloc(sizeof(double complex) * n), *tmp = malloc(sizeof(double complex) * n);
#pragma omp for schedule(static)
        for (long p = 0; p < npen; p++) {
            long base = (p % d[oa]) * s[oa] + (p / d[oa]) * s[ob];
            for (int i = 0; i < n; i++) line[i] = a[base + i * s[dir]];
            fft_line(line, n, tw, tmp);
            for (int i = 0; i < n; i++) a[base + i * s[dir]] = line[i];
        }
        free(line); free(tmp);
    }
    free(tw);
    return 0;
}

/* wavenumbers, PoissonPeriodic.F90:226-261: k = (0..n/2-1, -n/2, .., -1)*2pi/(n dx) */
static void wavenums(int n, double dx, double *k)
{
    int dummy = n - (n % 2);
    double *t = malloc(sizeof(double) * n);
    for (int i = 1; i <= n; i++) t[i - 1] = (-M_PI + (i - 1) * 2.0 * M_PI / (double)dummy) / dx;
    if (n % 2 == 0) {
        for (int i = 0; i < n / 2; i++) { k[i] = t[n / 2 + i]; k[n / 2 + i] = t[i]; }
    } else {
        for (int i = 0; i < (n + 1) / 2; i++) k[i] = t[(n + 1) / 2 - 1 + i];
        for (int i = 0; i < (n - 1) / 2; i++) k[(n + 1) / 2 + i] = t[i];
    }
    free(t);
}

/* modified wavenumber k'(k) for the scheme used by the derivative operators.
 * kind 0 Fourier (exact); 1 cd06 collocated d1; 2 cd06 staggered
 * (PadeDerOps.F90:1034-1053); 3 cd10 collocated d1. */
static double modk(int kind, double k, double dx)
{
    const double w = k * dx;
    switch (kind) {
    case 1: return (2 * a06d1 * sin(w) + 2 * b06d1 * sin(2 * w)) / (1 + 2 * alpha06d1 * cos(w)) / dx;
    case 2: {
        const double al = 9.0 / 62.0, a = 63.0 / 62.0, b = 17.0 / 62.0;
        return (2 * a * sin(w / 2) + (2.0 / 3.0) * b * sin(3 * w / 2)) / (1 + 2 * al * cos(w)) / dx;
    }
    case 3:
        return (2 * a10d1 * sin(w) + 2 * b10d1 * sin(2 * w) + 2 * c10d1 * sin(3 * w)) /
               (1 + 2 * alpha10d1 * cos(w) + 2 * beta10d1 * cos(2 * w)) / dx;
    default: return k;
    }
}

void orc_modk(int kind, int n, double dx, double *kout)
{
    wavenums(n, dx, kout);
    for (int i = 0; i < n; i++) kout[i] = modk(kind, kout[i], dx);
}

/* Periodic Poisson solve, PoissonPeriodic.F90:62-111 (dir_id = 1):
 * 3-D forward FFT, multiply by -1/(kx^2+ky^2+kz^2 [+1e-20]), inverse /N.
 * zero_policy 0: the reference (+1e-20 guard, only mode (1,1,1) zeroed);
 * zero_policy 1: every mode with |k'|^2 <= 1e-14 zeroed (PadePoisson.F90:104-108).
 * The full complex transform is used; the result's real part equals the
 * reference's r2c/c2r result for real input. */
int orc_poisson_periodic(int nx, int ny, int nz, double dx, double dy, double dz, int modk_kind,
                         int zero_policy, const double *rhs, double *out)
{
    const long N = (long)nx * ny * nz;
    double complex *z = malloc(sizeof(double complex) * N);
    for (long i = 0; i < N; i++) z[i] = rhs[i];
    for (int d = 1; d <= 3; d++) orc_fft_c2c_dir(d, nx, ny, nz, (double *)z, -1);
    double *kx = malloc(sizeof(double) * nx), *ky = malloc(sizeof(double) * ny), *kz = malloc(sizeof(double) * nz);
    orc_modk(modk_kind, nx, dx, kx);
    orc_modk(modk_kind, ny, dy, ky);
    orc_modk(modk_kind, nz, dz, kz);
    for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++) {
                long idx = i + (long)nx * (j + (long)ny * k);
                double ks = kx[i] * kx[i] + ky[j] * ky[j] + kz[k] * kz[k];
                if (zero_policy == 0) {
                    z[idx] *= -1.0 / (ks + 1.0e-20);
                } else {
                    z[idx] = (ks <= 1.0e-14) ? 0.0 : z[idx] * (-1.0 / ks);
                }
            }
    if (zero_policy == 0) z[0] = 0.0;
    for (int d = 1; d <= 3; d++) orc_fft_c2c_dir(d, nx, ny, nz, (double *)z, +1);
    for (long i = 0; i < N; i++) out[i] = creal(z[i]) / (double)N;
    free(z); free(kx); free(ky); free(kz);
    return 0;
}

/* Full incompressible RHS evaluation on a triply periodic n^3 grid of side L:
 *   grad[3*i+j] = d u_i / d x_j          (cd06 collocated, 9 fields)
 *   ut_i = F_z F_y F_x u_i               (cf90, periodic)
 *   div  = sum_j d ut_j / d x_j          (cd06)
 *   p    = Poisson(div) with cd06 modified wavenumbers, all |k'|^2<=1e-14 zeroed
 *   uo_i = ut_i - d p / d x_i            (cd06)
 * The composition follows the reference call sites: derivatives%ddx/ddy/ddz
 * (derivatives.F90:447-510), filters%filterx/y/z (filters.F90:220-269),
 * PoissonPeriodic%poisson_solve with Get_ModK callbacks (PoissonPeriodic.F90:62-87,
 * 188-204).  Returns 0, or the first operator error. */
int orc_full_rhs(int n, double L, const double *u, const double *v, const double *w, double *grad,
                 double *uo, double *vo, double *wo, double *p_out)
{
    const long N = (long)n * n * n;
    const double dx = L / n;
    const double *ui[3] = {u, v, w};
    double *uoi[3] = {uo, vo, wo};
    double *t1 = malloc(sizeof(double) * N), *t2 = malloc(sizeof(double) * N);
    double *ut[3] = {malloc(sizeof(double) * N), malloc(sizeof(double) * N), malloc(sizeof(double) * N)};
    double *div = malloc(sizeof(double) * N), *p = malloc(sizeof(double) * N);
    int e = 0;
    for (int i = 0; i < 3 && !e; i++)
        for (int j = 0; j < 3 && !e; j++)
            e = orc_apply(ORC_CD06_D1, n, dx, 1, 0, 0, j + 1, n, n, n, ui[i], grad + (long)(3 * i + j) * N);
    for (int i = 0; i < 3 && !e; i++) {
        e = orc_apply(ORC_CF90, n, dx, 1, 0, 0, 1, n, n, n, ui[i], t1);
        if (!e) e = orc_apply(ORC_CF90, n, dx, 1, 0, 0, 2, n, n, n, t1, t2);
        if (!e) e = orc_apply(ORC_CF90, n, dx, 1, 0, 0, 3, n, n, n, t2, ut[i]);
    }
    memset(div, 0, sizeof(double) * N);
    for (int j = 0; j < 3 && !e; j++) {
        e = orc_apply(ORC_CD06_D1, n, dx, 1, 0, 0, j + 1, n, n, n, ut[j], t1);
        for (long q = 0; q < N; q++) div[q] += t1[q];
    }
    if (!e) e = orc_poisson_periodic(n, n, n, dx, dx, dx, 1, 1, div, p);
    for (int i = 0; i < 3 && !e; i++) {
        e = orc_apply(ORC_CD06_D1, n, dx, 1, 0, 0, i + 1, n, n, n, p, t1);
        for (long q = 0; q < N; q++) uoi[i][q] = ut[i][q] - t1[q];
    }
    if (p_out) memcpy(p_out, p, sizeof(double) * N);
    free(t1); free(t2); free(div); free(p);
    for (int i = 0; i < 3; i++) free(ut[i]);
    return e;
}

int orc_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* ------------------------------------------------------------------------ */
/* cd06stagg, non-periodic (cd06stagg.F90:197-232 + STAGG_CD06_files)        */
/* ------------------------------------------------------------------------ */
/* Extra operator codes of the staggered object (collocated first
 * derivatives on the cell / edge grids, cd06stagg.F90:883-925). */
enum { ORC_STG_D1_C2C = 16, ORC_STG_D1_E2E = 17 };

/* Lele closure constants of cd06stagg.F90:18-57 */
#define SG_ALPHA 3.0
#define SG_P (17.0 / 6.0)
#define SG_Q (3.0 / 2.0)
#define SG_R (3.0 / 2.0)
#define SG_S (-1.0 / 6.0)

static void stg_lele(double *w1, double *w2, double *w3, double *alpha_p, double *q_p, double *alpha_pp,
                     double *q_pp, double *r_pp)
{
    const double ah = 1.0 / 3.0;
    const double q = SG_Q, r = SG_R, s = SG_S;
    *q_p = 3.0 / 4.0;
    *alpha_p = 1.0 / 4.0;
    *alpha_pp = ((40 * ah - 1) * q + 7 * (4 * ah - 1) * s) / (16 * (ah + 2) * q + 8 * (1 - 4 * ah) * s);
    *q_pp = (1.0 / 3.0) * (*alpha_pp + 2);
    *r_pp = (1.0 / 12.0) * (4 * (*alpha_pp) - 1);
    *w1 = (2 * ah + 1) / (2 * (q + s));
    *w2 = ((8 * ah + 7) * q - 6 * (2 * ah + 1) * r + (8 * ah + 7) * s) / (9 * (q + s));
    *w3 = (4 * (ah + 2) * q + 2 * (1 - 4 * ah) * s) / (9 * (q + s));
}

/* Tridiagonal rows (ddn, dg, dup), 1-based, of ComputeTri* in
 * STAGG_CD06_files/ComputeTri_allRoutines.F90. */
static int stg_tri(int op, int n, int te, int be, int ts, int bs, double *ddn, double *dg, double *dup)
{
    const double w0s = 223.0 / 186.0, w1s = 61.0 / 62.0;
    double w1, w2, w3, alpha_p, q_p, alpha_pp, q_pp, r_pp;
    stg_lele(&w1, &w2, &w3, &alpha_p, &q_p, &alpha_pp, &q_pp, &r_pp);
    int N;
    double al;
    switch (op) {
    case ORC_STG_D1_E2C: N = n; al = 9.0 / 62.0; break;
    case ORC_STG_D1_C2E: N = n + 1; al = 9.0 / 62.0; break;
    case ORC_STG_D1_C2C: N = n; al = 1.0 / 3.0; break;
    case ORC_STG_D1_E2E: N = n + 1; al = 1.0 / 3.0; break;
    case ORC_STG_IN_C2E: N = n + 1; al = 3.0 / 10.0; break;
    case ORC_STG_IN_E2C: N = n; al = 3.0 / 10.0; break;
    case ORC_STG_D2_E2E: N = n + 1; al = 2.0 / 11.0; break;
    case ORC_STG_D2_C2C: N = n; al = 2.0 / 11.0; break;
    default: return -1;
    }
    for (int i = 1; i <= N; i++) { ddn[i] = al; dg[i] = 1.0; dup[i] = al; }
    switch (op) {
    case ORC_STG_D1_E2C: {  /* ComputeTriD1_E2C :1-50 */
        const double alpha1 = 37.0 / 183.0, alpha0 = -1.0;
        if (ts) {
            ddn[N] = w0s * alpha0; dg[N] = w0s;
            ddn[N - 1] = w1s * alpha1; dup[N - 1] = w1s * alpha1; dg[N - 1] = w1s;
        } else if (te) { dup[N] = 0; dg[N] = 1 - al; }
        else { dg[N] = 1 + al; dup[N] = 0; }
        if (bs) {
            dup[1] = w0s * alpha0; dg[1] = w0s;
            ddn[2] = w1s * alpha1; dup[2] = w1s * alpha1; dg[2] = w1s;
        } else if (be) { ddn[1] = 0; dg[1] = 1 - al; }
        else { ddn[1] = 0; dg[1] = 1 + al; }
        break;
    }
    case ORC_STG_D1_C2E:  /* ComputeTriD1_C2E :52-100 */
        if (ts) { ddn[N] = 0; dup[N] = 0; ddn[N - 1] = 1.0 / 22.0; dup[N - 1] = 1.0 / 22.0; }
        else if (te) { dup[N] = 0; ddn[N] = 0; }
        else { ddn[N] = 2 * al; dup[N] = 0; }
        if (bs) { dup[1] = 0; ddn[1] = 0; dup[2] = 1.0 / 22.0; ddn[2] = 1.0 / 22.0; }
        else if (be) { ddn[1] = 0; dup[1] = 0; }
        else { ddn[1] = 0; dup[1] = 2 * al; }
        break;
    case ORC_STG_D1_C2C:  /* ComputeTriD1_C2C :102-160 */
        if (ts) {
            dup[N] = w1 * 0; dup[N - 1] = w2 * alpha_p; dup[N - 2] = w3 * alpha_pp;
            dg[N] = w1; dg[N - 1] = w2; dg[N - 2] = w3;
            ddn[N] = w1 * SG_ALPHA; ddn[N - 1] = w2 * alpha_p; ddn[N - 2] = w3 * alpha_pp;
        } else dg[N] = te ? 1 - al : 1 + al;
        if (bs) {
            ddn[1] = w1 * 0; ddn[2] = w2 * alpha_p; ddn[3] = w3 * alpha_pp;
            dg[1] = w1; dg[2] = w2; dg[3] = w3;
            dup[1] = w1 * SG_ALPHA; dup[2] = w2 * alpha_p; dup[3] = w3 * alpha_pp;
        } else dg[1] = be ? 1 - al : 1 + al;
        break;
    case ORC_STG_D1_E2E:  /* ComputeTriD1_E2E :162-195 */
        if (te) { dg[N] = 1; ddn[N] = 0; } else { dg[N] = 1; ddn[N] = 2 * al; }
        if (be) { dg[1] = 1; dup[1] = 0; } else { dg[1] = 1; dup[1] = 2 * al; }
        break;
    case ORC_STG_IN_C2E: {  /* ComputeTriInterp_C2E :197-240 */
        const double alpha1 = 1.0 / 6.0;
        if (ts) { dup[N] = 0; ddn[N] = 0; dup[N - 1] = alpha1; ddn[N - 1] = alpha1; }
        else if (te) { ddn[N] = 2 * al; dg[N] = 1; }
        else { ddn[N] = 0; dg[N] = 1; }
        if (bs) { dup[1] = 0; ddn[1] = 0; dg[1] = 1; dup[2] = alpha1; ddn[2] = alpha1; dg[2] = 1; }
        else if (be) { dup[1] = 2 * al; dg[1] = 1; }
        else { dup[1] = 0; dg[1] = 1; }
        break;
    }
    case ORC_STG_IN_E2C: {  /* ComputeTriInterp_E2C :243-285 */
        const double alpha0 = 1.0;
        if (ts) { ddn[N] = alpha0; dup[N] = 0; }
        else dg[N] = te ? 1 + al : 1 - al;
        if (bs) { dup[1] = alpha0; ddn[1] = 0; }
        else dg[1] = be ? 1 + al : 1 - al;
        break;
    }
    case ORC_STG_D2_E2E:  /* ComputeTriD2_E2E :288-320 */
        if (te) { dg[N] = 1; ddn[N] = 2 * al; } else { dg[N] = 1; ddn[N] = 0; }
        if (be) { dg[1] = 1; dup[1] = 2 * al; } else { dg[1] = 1; dup[1] = 0; }
        break;
    case ORC_STG_D2_C2C:  /* ComputeTriD2_C2C :322-354 */
        dg[N] = te ? 1 + al : 1 - al;
        dg[1] = be ? 1 + al : 1 - al;
        break;
    }
    return N;
}

/* RHS of the non-periodic staggered operators, 1-based f and r
 * (STAGG_CD06_files: D1RHS_*, D2RHS_*, InterpRHS_* common files). */
static void stg_rhs(int op, int n, double dz, int te, int be, int ts, int bs, const double *f, double *r)
{
    const double ob = 1.0 / dz, ob2 = ob / dz;
    const int nE = n + 1;
    const double w0s = 223.0 / 186.0, w1s = 61.0 / 62.0;
    switch (op) {
    case ORC_STG_D1_E2C: {  /* D1RHS_E2C_common.F90 */
        const double a = (63.0 / 62.0), b = (17.0 / 62.0) / 3.0;
        const double alpha1 = 37.0 / 183.0, alpha0 = -1.0;
        const double a0 = (1.0 / 24.0) * (alpha0 - 23.0), b0 = (1.0 / 8.0) * (-9.0 * alpha0 + 7.0);
        const double c0 = (1.0 / 8.0) * (9.0 * alpha0 + 1.0), d0 = -(1.0 / 24.0) * (alpha0 + 1.0);
        const double a1 = (3.0 / 8.0) * (3.0 - 2.0 * alpha1), b1 = (1.0 / 8.0) * (-1.0 + 22.0 * alpha1);
        const double A = a * ob, B = b * ob;
        for (int k = 2; k <= n - 1; k++) r[k] = B * (f[k + 2] - f[k - 1]) + A * (f[k + 1] - f[k]);
        if (bs) {
            r[1] = w0s * ob * (a0 * f[1] + b0 * f[2] + c0 * f[3] + d0 * f[4]);
            r[2] = w1s * ob * ((-b1 / 3.0) * f[1] + (-a1) * f[2] + a1 * f[3] + (b1 / 3.0) * f[4]);
        } else if (be) r[1] = B * (f[3] - f[2]) + A * (f[2] - f[1]);
        else r[1] = B * (f[3] + f[2]) + A * (f[2] - f[1]);
        if (ts) {
            r[n] = -w0s * ob * (a0 * f[nE] + b0 * f[nE - 1] + c0 * f[nE - 2] + d0 * f[nE - 3]);
            r[n - 1] = w1s * ob * ((-b1 / 3.0) * f[nE - 3] + (-a1) * f[nE - 2] + a1 * f[nE - 1] + (b1 / 3.0) * f[nE]);
        } else if (te) r[n] = B * (f[nE - 1] - f[nE - 2]) + A * (f[nE] - f[nE - 1]);
        else r[n] = -B * (f[nE - 1] + f[nE - 2]) + A * (f[nE] - f[nE - 1]);
        break;
    }
    case ORC_STG_D1_C2E: {  /* D1RHS_C2E_common.F90 */
        const double a = (63.0 / 62.0), b = (17.0 / 62.0) / 3.0;
        const double a0 = -71.0 / 24.0, b0 = 47.0 / 8.0, c0 = -31.0 / 8.0, d0 = 23.0 / 24.0, a1 = 12.0 / 11.0;
        const double A = a * ob, B = b * ob;
        for (int k = 2; k <= nE - 1; k++) r[k] = A * (f[k] - f[k - 1]);
        for (int k = 3; k <= nE - 2; k++) r[k] += B * (f[k + 1] - f[k - 2]);
        if (bs) {
            r[1] = (a0 * f[1] + b0 * f[2] + c0 * f[3] + d0 * f[4]) * ob;
            r[2] = (f[2] - f[1]) * (a1 * ob);
        } else if (be) { r[1] = 0; r[2] += B * (f[3] - f[1]); }
        else { r[1] = 2 * B * f[2] + 2 * A * f[1]; r[2] += B * (f[3] + f[1]); }
        if (ts) {
            r[nE - 1] = (f[n] - f[n - 1]) * (a1 * ob);
            r[nE] = (a0 * f[n] + b0 * f[n - 1] + c0 * f[n - 2] + d0 * f[n - 3]) * (-ob);
        } else if (te) { r[nE - 1] += B * (f[n] - f[n - 2]); r[nE] = 0; }
        else { r[nE - 1] -= B * (f[n] + f[n - 2]); r[nE] = -2 * B * f[n - 1] - 2 * A * f[n]; }
        break;
    }
    case ORC_STG_D1_C2C: {  /* D1RHS_C2C_common.F90 */
        double w1, w2, w3, alpha_p, q_p, alpha_pp, q_pp, r_pp;
        stg_lele(&w1, &w2, &w3, &alpha_p, &q_p, &alpha_pp, &q_pp, &r_pp);
        const double A = a06d1 * ob, B = b06d1 * ob;
        const double a3 = w3 * q_pp * ob, b3 = w3 * r_pp * ob, a2 = w2 * q_p * ob;
        const double a_1 = w1 * (-SG_P * ob), b_1 = w1 * (SG_Q * ob), c_1 = w1 * (SG_R * ob), d_1 = w1 * (SG_S * ob);
        if (bs) {
            r[1] = a_1 * f[1] + b_1 * f[2] + c_1 * f[3] + d_1 * f[4];
            r[2] = a2 * (f[3] - f[1]);
            r[3] = a3 * (f[4] - f[2]) + b3 * (f[5] - f[1]);
        } else if (be) {
            r[1] = B * (f[3] - f[2]) + A * (f[2] - f[1]);
            r[2] = B * (f[4] - f[1]) + A * (f[3] - f[1]);
            r[3] = B * (f[5] - f[1]) + A * (f[4] - f[2]);
        } else {
            r[1] = B * (f[3] + f[2]) + A * (f[2] + f[1]);
            r[2] = B * (f[4] + f[1]) + A * (f[3] - f[1]);
            r[3] = B * (f[5] - f[1]) + A * (f[4] - f[2]);
        }
        for (int k = 4; k <= n - 3; k++) r[k] = B * (f[k + 2] - f[k - 2]) + A * (f[k + 1] - f[k - 1]);
        if (ts) {
            r[n - 2] = a3 * (f[n - 1] - f[n - 3]) + b3 * (f[n] - f[n - 4]);
            r[n - 1] = a2 * (f[n] - f[n - 2]);
            r[n] = -a_1 * f[n] - b_1 * f[n - 1] - c_1 * f[n - 2] - d_1 * f[n - 3];
        } else if (te) {
            r[n - 2] = B * (f[n] - f[n - 4]) + A * (f[n - 1] - f[n - 3]);
            r[n - 1] = B * (f[n] - f[n - 3]) + A * (f[n] - f[n - 2]);
            r[n] = B * (f[n - 1] - f[n - 2]) + A * (f[n] - f[n - 1]);
        } else {
            r[n - 2] = B * (f[n] - f[n - 4]) + A * (f[n - 1] - f[n - 3]);
            r[n - 1] = -B * (f[n] + f[n - 3]) + A * (f[n] - f[n - 2]);
            r[n] = -B * (f[n - 1] + f[n - 2]) - A * (f[n] + f[n - 1]);
        }
        break;
    }
    case ORC_STG_D1_E2E: {  /* D1RHS_E2E_common.F90 */
        const double A = a06d1 * ob, B = b06d1 * ob;
        for (int k = 2; k <= nE - 1; k++) r[k] = A * (f[k + 1] - f[k - 1]);
        for (int k = 3; k <= nE - 2; k++) r[k] += B * (f[k + 2] - f[k - 2]);
        if (be) { r[1] = 0; r[2] += B * (f[4] - f[2]); }
        else { r[1] = B * (f[3] + f[3]) + A * (f[2] + f[2]); r[2] += B * (f[4] + f[2]); }
        if (te) { r[nE] = 0; r[nE - 1] += B * (f[nE - 1] - f[nE - 3]); }
        else {
            r[nE] = -B * (f[nE - 2] + f[nE - 2]) - A * (f[nE - 1] + f[nE - 1]);
            r[nE - 1] -= B * (f[nE - 1] + f[nE - 3]);
        }
        break;
    }
    case ORC_STG_IN_E2C: {  /* InterpRHS_E2C_common.F90 */
        const double b = (1.0 / 10.0) / 2.0, a = (3.0 / 2.0) / 2.0, alpha0 = 1.0;
        const double a0 = (1.0 / 16.0) * (5.0 - alpha0), b0 = (1.0 / 16.0) * (9.0 * alpha0 + 15.0);
        const double c0 = (1.0 / 16.0) * (9.0 * alpha0 - 5.0), d0 = (1.0 / 16.0) * (1.0 - alpha0);
        if (bs) r[1] = a0 * f[1] + b0 * f[2] + c0 * f[3] + d0 * f[4];
        else if (be) r[1] = b * (f[3] + f[2]) + a * (f[2] + f[1]);
        else r[1] = b * (f[3] - f[2]) + a * (f[2] + f[1]);
        for (int k = 2; k <= n - 1; k++) r[k] = b * (f[k + 2] + f[k - 1]) + a * (f[k + 1] + f[k]);
        if (ts) r[n] = a0 * f[nE] + b0 * f[nE - 1] + c0 * f[nE - 2] + d0 * f[nE - 3];
        else if (te) r[n] = b * (f[nE - 1] + f[nE - 2]) + a * (f[nE] + f[nE - 1]);
        else r[n] = b * (-f[nE - 1] + f[nE - 2]) + a * (f[nE] + f[nE - 1]);
        break;
    }
    case ORC_STG_IN_C2E: {  /* InterpRHS_C2E_common.F90 */
        const double b = (1.0 / 10.0) / 2.0, a = (3.0 / 2.0) / 2.0;
        const double a0 = 15.0 / 8.0, b0 = -5.0 / 4.0, c0 = 3.0 / 8.0;
        const double alpha1 = 1.0 / 6.0, a1 = (1.0 / 8.0) * (9.0 + 10.0 * alpha1);
        for (int k = 2; k <= nE - 1; k++) r[k] = a * (f[k] + f[k - 1]);
        for (int k = 3; k <= nE - 2; k++) r[k] += b * (f[k + 1] + f[k - 2]);
        if (bs) { r[1] = a0 * f[1] + b0 * f[2] + c0 * f[3]; r[2] = (a1 / 2.0) * (f[2] + f[1]); }
        else if (be) { r[1] = 2 * b * f[2] + 2 * a * f[1]; r[2] += b * (f[3] + f[1]); }
        else { r[1] = 0; r[2] += b * (f[3] - f[1]); }
        if (ts) { r[nE] = a0 * f[n] + b0 * f[n - 1] + c0 * f[n - 2]; r[nE - 1] = (a1 / 2.0) * (f[n] + f[n - 1]); }
        else if (te) { r[nE] = 2 * b * f[n - 1] + 2 * a * f[n]; r[nE - 1] += b * (f[n] + f[n - 2]); }
        else { r[nE] = 0; r[nE - 1] += b * (-f[n] + f[n - 2]); }
        break;
    }
    case ORC_STG_D2_C2C: {  /* D2RHS_C2C_common.F90 */
        const double A = (12.0 / 11.0) * ob2, B = ((3.0 / 11.0) / 4.0) * ob2;
        for (int k = 2; k <= n - 1; k++) r[k] = A * (f[k + 1] + f[k - 1]);
        for (int k = 3; k <= n - 2; k++) r[k] += B * (f[k + 2] + f[k - 2]);
        if (be) { r[1] = B * (f[3] + f[2]) + A * (f[2] + f[1]); r[2] += B * (f[4] + f[1]); }
        else { r[1] = B * (f[3] - f[2]) + A * (f[2] - f[1]); r[2] += B * (f[4] - f[1]); }
        if (te) { r[n] = B * (f[n - 1] + f[n - 2]) + A * (f[n] + f[n - 1]); r[n - 1] += B * (f[n] + f[n - 3]); }
        else { r[n] = B * (-f[n - 1] + f[n - 2]) + A * (-f[n] + f[n - 1]); r[n - 1] += B * (-f[n] + f[n - 3]); }
        for (int k = 1; k <= n; k++) r[k] -= 2 * (B + A) * f[k];
        break;
    }
    case ORC_STG_D2_E2E: {  /* D2RHS_E2E_common.F90 */
        const double A = (12.0 / 11.0) * ob2, B = ((3.0 / 11.0) / 4.0) * ob2;
        for (int k = 2; k <= nE - 1; k++) r[k] = A * (f[k + 1] + f[k - 1]);
        for (int k = 3; k <= nE - 2; k++) r[k] += B * (f[k + 2] + f[k - 2]);
        if (be) { r[1] = B * (f[3] + f[3]) + A * (f[2] + f[2]); r[2] += B * (f[4] + f[2]); }
        else { r[1] = 0; r[2] += B * (f[4] - f[2]); }
        if (te) { r[nE] = 2 * B * f[nE - 2] + 2 * A * f[nE - 1]; r[nE - 1] += B * (f[nE - 1] + f[nE - 3]); }
        else { r[nE] = 0; r[nE - 1] += B * (-f[nE - 1] + f[nE - 3]); }
        for (int k = 1; k <= nE; k++) r[k] -= 2 * (B + A) * f[k];
        break;
    }
    }
}

/* SolveZTriREAL (TridiagSolver_allRoutines.F90:1-16) on one pencil with the
 * factors (ddn*den, den, cp) of ComputeTri_common.F90. */
static void stg_solve(int N, const double *ddn, const double *dg, const double *dup, double *y)
{
    double *cp = malloc(sizeof(double) * (N + 2)), *den = malloc(sizeof(double) * (N + 2));
    cp[1] = dup[1] / dg[1];
    for (int i = 2; i <= N - 1; i++) cp[i] = dup[i] / (dg[i] - ddn[i] * cp[i - 1]);
    den[1] = 1.0 / dg[1];
    for (int i = 2; i <= N; i++) den[i] = 1.0 / (dg[i] - ddn[i] * cp[i - 1]);
    y[1] = y[1] * den[1];
    for (int k = 2; k <= N; k++) y[k] = y[k] * den[k] - y[k - 1] * (ddn[k] * den[k]);
    for (int k = N - 1; k >= 1; k--) y[k] = y[k] - cp[k] * y[k + 1];
    free(cp);
    free(den);
}

static int stg_nin(int op, int n)
{
    return (op == ORC_STG_D1_E2C || op == ORC_STG_IN_E2C || op == ORC_STG_D1_E2E || op == ORC_STG_D2_E2E) ? n + 1 : n;
}

/* Apply a non-periodic cd06stagg operator along z of f(n1, n2, n_in); flags
 * bit 0 isTopEven, 1 isBotEven, 2 isTopSided, 3 isBotSided.  Returns 21 for
 * n <= 4 (cd06stagg.F90:216-218). */
int orc_stagg_np(int op, int n, double dz, int flags, int n1, int n2, const double *f, double *out)
{
    if (n <= 4) return 21;
    const int te = flags & 1, be = (flags >> 1) & 1, ts = (flags >> 2) & 1, bs = (flags >> 3) & 1;
    double *ddn = malloc(sizeof(double) * (n + 3)), *dg = malloc(sizeof(double) * (n + 3)),
           *dup = malloc(sizeof(double) * (n + 3));
    const int N = stg_tri(op, n, te, be, ts, bs, ddn, dg, dup);
    if (N < 0) { free(ddn); free(dg); free(dup); return -1; }
    const int nin = stg_nin(op, n);
    const long pl = (long)n1 * n2;
#pragma omp parallel
    {
        double *pin = malloc(sizeof(double) * (nin + 2)), *y = malloc(sizeof(double) * (N + 2));
#pragma omp for schedule(static)
        for (long p = 0; p < pl; p++) {
            for (int k = 1; k <= nin; k++) pin[k] = f[p + (long)(k - 1) * pl];
            stg_rhs(op, n, dz, te, be, ts, bs, pin, y);
            stg_solve(N, ddn, dg, dup, y);
            for (int k = 1; k <= N; k++) out[p + (long)(k - 1) * pl] = y[k];
        }
        free(pin);
        free(y);
    }
    free(ddn); free(dg); free(dup);
    return 0;
}

/* Dense (A, B) of a non-periodic staggered operator: A N x N, B N x n_in (row-major). */
int orc_stagg_np_dense(int op, int n, double dz, int flags, double *A, double *B)
{
    if (n <= 4) return 21;
    const int te = flags & 1, be = (flags >> 1) & 1, ts = (flags >> 2) & 1, bs = (flags >> 3) & 1;
    double *ddn = malloc(sizeof(double) * (n + 3)), *dg = malloc(sizeof(double) * (n + 3)),
           *dup = malloc(sizeof(double) * (n + 3));
    const int N = stg_tri(op, n, te, be, ts, bs, ddn, dg, dup);
    if (N < 0) { free(ddn); free(dg); free(dup); return -1; }
    const int nin = stg_nin(op, n);
    memset(A, 0, sizeof(double) * N * N);
    for (int i = 1; i <= N; i++) {
        if (i > 1) A[(i - 1) * N + i - 2] = ddn[i];
        A[(i - 1) * N + i - 1] = dg[i];
        if (i < N) A[(i - 1) * N + i] = dup[i];
    }
    double *e = calloc(nin + 2, sizeof(double)), *r = calloc(N + 2, sizeof(double));
    for (int j = 1; j <= nin; j++) {
        memset(e, 0, sizeof(double) * (nin + 2));
        e[j] = 1.0;
        stg_rhs(op, n, dz, te, be, ts, bs, e, r);
        for (int i = 1; i <= N; i++) B[(i - 1) * nin + j - 1] = r[i];
    }
    free(e); free(r); free(ddn); free(dg); free(dup);
    return 0;
}
