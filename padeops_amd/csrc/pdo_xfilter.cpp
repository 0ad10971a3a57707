// Explicit filter objects: gaussianstuff (src/filters/gaussian.F90) and
// lstsqstuff (src/filters/lstsq.F90).  Both are symmetric 9-point stencils
// with 4 closure rows per end on non-periodic pencils; the Gaussian filter
// also takes bc1/bcn = +1 / -1 (even / odd reflection about the end point).
// The closure rows are folded into 8-wide weight tables on the host, once
// per (bc1, bcn) at the call, and travel in the kernel arguments.
#include <cstring>

#include "../../include/padeops_hip.h"
#include "pdo_explicit.hpp"
#include "pdo_ops.hpp"

namespace pdo {

// gaussian.F90:15-44: the width-4dx Gaussian filter and its one-sided rows
static const double GAUSS_C[5] = {3565.0 / 10368.0, 3091.0 / 12960.0, 1997.0 / 25920.0, 149.0 / 12960.0,
                                  107.0 / 103680.0};
// lstsq.F90:15-20: the least-squares filter.  The reference writes its
// decimal constants as default-kind (single precision) literals converted
// to FP64, so the FP64 coefficients are the FP32 roundings of 0.6744132 and
// -0.1744132, halved.
static const double LSQ_C[5] = {0.5, (double)0.6744132f / 2.0, 0.0, (double)-0.1744132f / 2.0, 0.0};
// rows 1-4 of a non-periodic pencil (gaussian.F90:24-43, lstsq.F90:24-43):
// centre weight, then the weights of the pairs at distance 1, 2, 3
static const double B_ROWS[4][4] = {{5.0 / 6.0, 1.0 / 6.0, 0.0, 0.0},
                                    {2.0 / 3.0, 1.0 / 6.0, 0.0, 0.0},
                                    {31.0 / 64.0, 7.0 / 32.0, 5.0 / 128.0, 0.0},
                                    {17.0 / 48.0, 15.0 / 64.0, 7.0 / 96.0, 1.0 / 64.0}};

struct XFilterObject {
    int n = 0;
    bool periodic = true;
    bool gaussian = true;  // else least squares (no bc options)
};

// closure rows at the left end for boundary option bc (gaussian.F90:183-251):
// 0 one-sided rows; +1 / -1 the full stencil on the even / odd reflection
// f(-m) = bc * f(m) about point 0.
static void left_rows(const double *c, int bc, double (&w)[EXPL_NB * EXPL_WB])
{
    std::memset(w, 0, sizeof(w));
    for (int j = 0; j < EXPL_NB; ++j) {
        double *row = w + j * EXPL_WB;
        if (bc == 0) {
            row[j] += B_ROWS[j][0];
            for (int q = 1; q <= 3; ++q) {
                if (B_ROWS[j][q] == 0.0) continue;
                row[j + q] += B_ROWS[j][q];
                if (j - q >= 0) row[j - q] += B_ROWS[j][q];
            }
        } else {
            row[j] += c[0];
            for (int q = 1; q <= EXPL_R; ++q) {
                row[j + q] += c[q];
                const int m = j - q;
                if (m >= 0) row[m] += c[q];
                else row[-m] += bc * c[q];
            }
        }
    }
}

}  // namespace pdo

struct pdo_xfilter_s : public pdo::XFilterObject {};

using namespace pdo;

static int make_xfilter(pdo_xfilter *h, int n, int periodic, bool gaussian)
{
    if (!h) { set_error("null handle pointer"); return PDO_ERR_ARG; }
    *h = nullptr;
    if (n < 1 || (n > 1 && n < 8)) {
        set_error("explicit filters need n == 1 or n >= 8 (the 9-point stencil and its closure rows)");
        return PDO_ERR_ARG;
    }
    auto *o = new pdo_xfilter_s();
    o->n = n;
    o->periodic = periodic != 0;
    o->gaussian = gaussian;
    *h = o;
    return 0;
}

static void fill_args(const XFilterObject *h, int bc1, int bcn, ExplicitArgs &E)
{
    std::memset(&E, 0, sizeof(E));
    E.n = h->n;
    E.periodic = h->periodic ? 1 : 0;
    const double *c = h->gaussian ? GAUSS_C : LSQ_C;
    std::memcpy(E.c, c, sizeof(E.c));
    if (!h->periodic) {
        double l[EXPL_NB * EXPL_WB], r[EXPL_NB * EXPL_WB];
        left_rows(c, bc1, l);
        left_rows(c, bcn, r);
        std::memcpy(E.left, l, sizeof(l));
        // the right end is the mirror image: row n-1-j against point n-1-k
        // takes the left table's row j, point k
        for (int j = 0; j < EXPL_NB; ++j)
            for (int k = 0; k < EXPL_WB; ++k)
                E.right[(EXPL_NB - 1 - j) * EXPL_WB + (EXPL_WB - 1 - k)] = r[j * EXPL_WB + k];
    }
}

static int xfilter_apply(pdo_xfilter h, int dir, int prec, const void *f, void *fil, long na, long nb, int bc1,
                         int bcn, void *stream)
{
    if (!h || !f || !fil || na < 0 || nb < 0) { set_error("invalid argument"); return PDO_ERR_ARG; }
    const int mode = prec_mode(prec);
    if (mode < 0) { set_error("unknown precision mode"); return PDO_ERR_ARG; }
    if (h->gaussian) {
        // gaussian.F90:121-135: GracefulExit(..., 324)
        if (bc1 < -1 || bc1 > 1 || bcn < -1 || bcn > 1) {
            set_error("Incorrect boundary specification for bc1/bcn (should be 0, 1 or -1)");
            return 324;
        }
    } else {
        bc1 = bcn = 0;  // the least-squares filter has no boundary options
    }
    const bool dbl = mode == MODE_F64;
    hipStream_t st = (hipStream_t)stream;
    const long count = (long)h->n * na * nb;
    if (count == 0) return 0;
    if (h->n == 1) {
        hipError_t e = hipMemcpyAsync(fil, f, count * (dbl ? 8 : 4), hipMemcpyDeviceToDevice, st);
        return e == hipSuccess ? 0 : hip_fail(e, "hipMemcpyAsync");
    }
    if (na * nb * (long)h->n >= (1L << 31)) { set_error("field too large"); return PDO_ERR_ARG; }
    ExplicitArgs E;
    fill_args(h, bc1, bcn, E);
    E.na = (int)na;
    E.nb = (int)nb;
    E.dir = dir;
    hipError_t e = launch_explicit(dbl, E, f, fil, st);
    return e == hipSuccess ? 0 : hip_fail(e, "explicit filter launch");
}

extern "C" {

int pdo_gaussian_init(pdo_xfilter *h, int n, int periodic) { return make_xfilter(h, n, periodic, true); }

int pdo_lstsq_init(pdo_xfilter *h, int n, int periodic) { return make_xfilter(h, n, periodic, false); }

int pdo_xfilter_filter1(pdo_xfilter h, int prec, const void *f, void *fil, int na, int nb, int bc1, int bcn,
                        void *stream)
{
    return xfilter_apply(h, 1, prec, f, fil, na, nb, bc1, bcn, stream);
}

int pdo_xfilter_filter2(pdo_xfilter h, int prec, const void *f, void *fil, int na, int nb, int bc1, int bcn,
                        void *stream)
{
    return xfilter_apply(h, 2, prec, f, fil, na, nb, bc1, bcn, stream);
}

int pdo_xfilter_filter3(pdo_xfilter h, int prec, const void *f, void *fil, int na, int nb, int bc1, int bcn,
                        void *stream)
{
    return xfilter_apply(h, 3, prec, f, fil, na, nb, bc1, bcn, stream);
}

int pdo_xfilter_filter3_cmplx(pdo_xfilter h, int prec, const void *f, void *fil, int na, int nb, int bc1, int bcn,
                              void *stream)
{
    // interleaved (re, im): a real z filter of (2 na, nb, n)
    return xfilter_apply(h, 3, prec, f, fil, 2L * na, nb, bc1, bcn, stream);
}

int pdo_xfilter_dense(pdo_xfilter h, int bc1, int bcn, double *W)
{
    if (!h || !W) { set_error("invalid argument"); return PDO_ERR_ARG; }
    if (!h->gaussian) bc1 = bcn = 0;
    if (bc1 < -1 || bc1 > 1 || bcn < -1 || bcn > 1) { set_error("Incorrect boundary specification"); return 324; }
    const int n = h->n;
    std::memset(W, 0, sizeof(double) * n * n);
    if (n == 1) { W[0] = 1.0; return 0; }
    ExplicitArgs E;
    fill_args(h, bc1, bcn, E);
    // the kernel's row rules (pdo_explicit.hip), row-major W[r*n + col]
    for (int r = 0; r < n; ++r) {
        double *row = W + (size_t)r * n;
        if (E.periodic || (r >= EXPL_NB && r < n - EXPL_NB)) {
            row[r] += E.c[0];
            for (int q = 1; q <= EXPL_R; ++q) {
                row[((r + q) % n + n) % n] += E.c[q];
                row[((r - q) % n + n) % n] += E.c[q];
            }
        } else if (r < EXPL_NB) {
            for (int q = 0; q < EXPL_WB; ++q) row[q] += E.left[r * EXPL_WB + q];
        } else {
            const int j = r - (n - EXPL_NB);
            for (int q = 0; q < EXPL_WB; ++q) row[n - EXPL_WB + q] += E.right[j * EXPL_WB + q];
        }
    }
    return 0;
}

int pdo_xfilter_destroy(pdo_xfilter h)
{
    delete h;
    return 0;
}

}  // extern "C"
