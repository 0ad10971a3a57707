// The per-direction dispatch types of the reference and the operators built
// on them:
//   derivatives (src/derivatives/derivatives.F90): cd10 / cd06 / Fourier
//     collocation ("four", src/derivatives/ffts.F90) per direction;
//   filters (src/filters/filters.F90): cf90 / gaussian / lstsq per direction;
//   operators (src/utilities/operators.F90): gradient, curl, divergence,
//     filter3D.
// On one GPU the reference's pencil sizes xsz = ysz = zsz = (nx, ny, nz) and
// its transposes are identities: every direction works on the same array.
#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/padeops_hip.h"
#include "pdo_explicit.hpp"
#include "pdo_fft.hpp"
#include "pdo_kernels.hpp"
#include "pdo_ops.hpp"
#include "pdo_poisson.hpp"

namespace pdo {

// ffts (ffts.F90:67-183): one direction's Fourier-collocation derivative
struct FourierAxis {
    int dir = 0, n = 0;
    PoissonPlanDev pl{};
    std::vector<void *> allocs;
    double *k1 = nullptr, *k2 = nullptr;  // i k and -k^2 multipliers, the n/2 mode zero
    void *work = nullptr;                  // x: half spectrum of every line
    size_t work_bytes = 0;
    ~FourierAxis()
    {
        for (void *p : allocs) (void)hipFree(p);
        if (work) (void)hipFree(work);
    }
    int init(int dir_, int n_, double dx)
    {
        dir = dir_;
        n = n_;
        if (n == 1) return 0;
        const bool ok = dir == 1 ? fft_pow2(n, 16, 1024) : fft_pow2(n, 8, 1024);
        if (!ok) {
            set_error("Fourier collocation here needs a power-of-two length (x: 16..1024, y/z: 8..1024)");
            return PDO_ERR_UNSUP;
        }
        std::memset(&pl, 0, sizeof(pl));
        pl.nx = dir == 1 ? n : 2;
        pl.ny = dir == 2 ? n : 1;
        pl.nz = dir == 3 ? n : 1;
        pl.nxp = ((pl.nx / 2 + 1) + 15) / 16 * 16;
        int st = fft_tables(pl, pl.nx, dir == 2 ? n : 0, dir == 3 ? n : 0, allocs);
        if (st) return st;
        // GetWaveNums (ffts.F90:370-388): m 2 pi / (n dx) for m < n/2, the
        // n/2 mode removed (ffts.F90:199,262); the full-length tables of y
        // and z also hold the negative modes m - n
        const int len = dir == 1 ? n / 2 + 1 : n;
        std::vector<double> a(len), b(len);
        for (int m = 0; m < len; ++m) {
            const int s = (m <= n / 2) ? m : m - n;
            const double k = (m == n / 2) ? 0.0 : 2.0 * M_PI * s / (n * dx);
            a[m] = k;
            b[m] = -k * k;
        }
        if ((st = upload_vec(a, &k1))) return st;
        allocs.push_back(k1);
        if ((st = upload_vec(b, &k2))) return st;
        allocs.push_back(k2);
        return 0;
    }
    int apply(int order, int prec, const void *f, void *df, long na, long nb, hipStream_t st)
    {
        const int mode = fft_mode(prec);
        if (mode < 0) { set_error("unknown precision mode"); return PDO_ERR_ARG; }
        const size_t W = mode == 0 ? 8 : 4;
        const long count = (long)n * na * nb;
        if (count == 0) return 0;
        if (n == 1) {
            // ffts.F90:190-193: df = 0
            hipError_t e = launch_zero_words(df, count * (long)W / 4, st);
            return e == hipSuccess ? 0 : hip_fail(e, "zero fill");
        }
        if (dir != 1 && ((dir == 2 ? na : na * nb) % 2) != 0) {
            set_error("Fourier derivative along y/z needs an even count of lines along the contiguous index");
            return PDO_ERR_UNSUP;
        }
        if (dir == 1) {
            const size_t need = (size_t)pl.nxp * na * nb * 2 * W;
            if (work_bytes < need) {
                if (work) (void)hipFree(work);
                work = nullptr;
                work_bytes = 0;
                hipError_t e = hipMalloc(&work, need);
                if (e != hipSuccess) return hip_fail(e, "hipMalloc(Fourier work)");
                work_bytes = need;
            }
        }
        hipError_t e = run_fourier_deriv(pl, mode, dir, order == 1 ? 1 : 0, f, df, work, na, nb,
                                         order == 1 ? k1 : k2, st);
        return e == hipSuccess ? 0 : hip_fail(e, "Fourier derivative");
    }
};

enum { METHOD_CD10 = 1, METHOD_CD06 = 2, METHOD_FOUR = 3, METHOD_CHEB = 4 };
enum { FIL_CF90 = 1, FIL_GAUSSIAN = 2, FIL_LSTSQ = 3 };

struct DerivativesObject {
    int nx = 0, ny = 0, nz = 0;
    int method[3] = {0, 0, 0};
    pdo_op op[3] = {nullptr, nullptr, nullptr};
    FourierAxis *four[3] = {nullptr, nullptr, nullptr};
    void *tmp = nullptr;  // one field of work for the operators
    size_t tmp_bytes = 0;
    ~DerivativesObject()
    {
        for (int d = 0; d < 3; ++d) {
            if (op[d]) pdo_op_destroy(op[d]);
            delete four[d];
        }
        if (tmp) (void)hipFree(tmp);
    }
    int n_of(int d) const { return d == 0 ? nx : (d == 1 ? ny : nz); }
    // na, nb of the direction's pencils in the (nx, ny, nz) array
    void pencils(int d, int &na, int &nb) const
    {
        if (d == 0) { na = ny; nb = nz; }
        else if (d == 1) { na = nx; nb = nz; }
        else { na = nx; nb = ny; }
    }
    int scratch(size_t bytes)
    {
        if (tmp_bytes >= bytes) return 0;
        if (tmp) (void)hipFree(tmp);
        tmp = nullptr;
        tmp_bytes = 0;
        hipError_t e = hipMalloc(&tmp, bytes);
        if (e != hipSuccess) return hip_fail(e, "hipMalloc(operator work)");
        tmp_bytes = bytes;
        return 0;
    }
    // derivatives%ddx..ddz (order 1) and d2dx2..d2dz2 (order 2), derivatives.F90:447-569
    int apply(int d, int order, int prec, const void *f, void *df, int bc1, int bcn, hipStream_t st)
    {
        int na, nb;
        pencils(d, na, nb);
        switch (method[d]) {
        case METHOD_CD10: {
            using Fn = int (*)(pdo_op, int, const void *, void *, int, int, int, int, void *);
            static const Fn f1[3] = {pdo_cd10_dd1, pdo_cd10_dd2, pdo_cd10_dd3};
            static const Fn f2[3] = {pdo_cd10_d2d1, pdo_cd10_d2d2, pdo_cd10_d2d3};
            return (order == 1 ? f1 : f2)[d](op[d], prec, f, df, na, nb, bc1, bcn, st);
        }
        case METHOD_CD06: {
            if (order == 2) {
                // derivatives.F90:524-525,543-544,562-563: GracefulExit("CD06 is incomplete right now", 21)
                set_error("CD06 is incomplete right now");
                return PDO_ERR_UNSUP;
            }
            using Fn = int (*)(pdo_op, int, const void *, void *, int, int, void *);
            static const Fn f1[3] = {pdo_cd06_dd1, pdo_cd06_dd2, pdo_cd06_dd3};
            return f1[d](op[d], prec, f, df, na, nb, st);
        }
        case METHOD_FOUR:
            return four[d]->apply(order, prec, f, df, na, nb, st);
        }
        set_error("derivatives object not initialised");
        return PDO_ERR_ARG;
    }
};

struct FiltersObject {
    int nx = 0, ny = 0, nz = 0;
    int method[3] = {0, 0, 0};
    pdo_op cf[3] = {nullptr, nullptr, nullptr};
    pdo_xfilter xf[3] = {nullptr, nullptr, nullptr};
    void *tmp[2] = {nullptr, nullptr};
    size_t tmp_bytes = 0;
    ~FiltersObject()
    {
        for (int d = 0; d < 3; ++d) {
            if (cf[d]) pdo_op_destroy(cf[d]);
            if (xf[d]) pdo_xfilter_destroy(xf[d]);
        }
        for (void *p : tmp) if (p) (void)hipFree(p);
    }
    void pencils(int d, int &na, int &nb) const
    {
        if (d == 0) { na = ny; nb = nz; }
        else if (d == 1) { na = nx; nb = nz; }
        else { na = nx; nb = ny; }
    }
    // filters%filterx/y/z (filters.F90:220-269)
    int apply(int d, int prec, const void *f, void *ff, int bc1, int bcn, hipStream_t st)
    {
        int na, nb;
        pencils(d, na, nb);
        if (method[d] == FIL_CF90) {
            using Fn = int (*)(pdo_op, int, const void *, void *, int, int, int, int, void *);
            static const Fn fn[3] = {pdo_cf90_filter1, pdo_cf90_filter2, pdo_cf90_filter3};
            return fn[d](cf[d], prec, f, ff, na, nb, bc1, bcn, st);
        }
        if (method[d] == FIL_GAUSSIAN || method[d] == FIL_LSTSQ) {
            using Fn = int (*)(pdo_xfilter, int, const void *, void *, int, int, int, int, void *);
            static const Fn fn[3] = {pdo_xfilter_filter1, pdo_xfilter_filter2, pdo_xfilter_filter3};
            return fn[d](xf[d], prec, f, ff, na, nb, bc1, bcn, st);
        }
        set_error("filters object not initialised");
        return PDO_ERR_ARG;
    }
    int scratch(size_t bytes)
    {
        if (tmp_bytes >= bytes) return 0;
        for (void *&p : tmp) {
            if (p) (void)hipFree(p);
            p = nullptr;
        }
        tmp_bytes = 0;
        for (void *&p : tmp) {
            hipError_t e = hipMalloc(&p, bytes);
            if (e != hipSuccess) return hip_fail(e, "hipMalloc(filter3D work)");
        }
        tmp_bytes = bytes;
        return 0;
    }
};

}  // namespace pdo

struct pdo_derivatives_s : public pdo::DerivativesObject {};
struct pdo_filters_s : public pdo::FiltersObject {};

using namespace pdo;

static size_t word_of(int prec) { return prec == PDO_F64 ? 8 : 4; }

extern "C" {

int pdo_derivatives_init(pdo_derivatives *h, int nx, int ny, int nz, double dx, double dy, double dz,
                         const int *periodic, const int *method)
{
    if (!h || !periodic || !method || nx < 1 || ny < 1 || nz < 1) { set_error("invalid argument"); return PDO_ERR_ARG; }
    *h = nullptr;
    auto *o = new pdo_derivatives_s();
    o->nx = nx; o->ny = ny; o->nz = nz;
    const double dxs[3] = {dx, dy, dz};
    for (int d = 0; d < 3; ++d) {
        const int n = o->n_of(d);
        int st = 0;
        switch (method[d]) {
        case METHOD_CD10:
            st = pdo_cd10_init(&o->op[d], n, dxs[d], periodic[d], 0, 0);
            break;
        case METHOD_CD06:
            st = pdo_cd06_init(&o->op[d], n, dxs[d], periodic[d], 0, 0);
            break;
        case METHOD_FOUR:
            o->four[d] = new FourierAxis();
            st = o->four[d]->init(d + 1, n, dxs[d]);
            break;
        case METHOD_CHEB:
            set_error("Chebychev is incomplete right now");  // derivatives.F90:465
            st = PDO_ERR_UNSUP;
            break;
        default:
            set_error("Invalid method selected");  // derivatives.F90:317 (GracefulExit 01)
            st = PDO_ERR_ARG;
        }
        o->method[d] = method[d];
        if (st) {
            delete o;
            return st;
        }
    }
    *h = o;
    return 0;
}

int pdo_derivatives_method(pdo_derivatives h, int dir, int *method)
{
    if (!h || !method || dir < 1 || dir > 3) { set_error("invalid argument"); return PDO_ERR_ARG; }
    *method = h->method[dir - 1];
    return 0;
}

#define PDO_DER(NAME, D, ORDER)                                                                                  \
    int NAME(pdo_derivatives h, int prec, const void *f, void *df, int bc1, int bcn, void *stream)               \
    {                                                                                                            \
        if (!h || !f || !df) { set_error("invalid argument"); return PDO_ERR_ARG; }                              \
        return h->apply(D, ORDER, prec, f, df, bc1, bcn, (hipStream_t)stream);                                   \
    }
PDO_DER(pdo_derivatives_ddx, 0, 1)
PDO_DER(pdo_derivatives_ddy, 1, 1)
PDO_DER(pdo_derivatives_ddz, 2, 1)
PDO_DER(pdo_derivatives_d2dx2, 0, 2)
PDO_DER(pdo_derivatives_d2dy2, 1, 2)
PDO_DER(pdo_derivatives_d2dz2, 2, 2)
#undef PDO_DER

int pdo_derivatives_destroy(pdo_derivatives h)
{
    delete h;
    return 0;
}

int pdo_filters_init(pdo_filters *h, int nx, int ny, int nz, const int *periodic, const int *method)
{
    if (!h || !periodic || !method || nx < 1 || ny < 1 || nz < 1) { set_error("invalid argument"); return PDO_ERR_ARG; }
    *h = nullptr;
    auto *o = new pdo_filters_s();
    o->nx = nx; o->ny = ny; o->nz = nz;
    const int ns[3] = {nx, ny, nz};
    for (int d = 0; d < 3; ++d) {
        int st = 0;
        switch (method[d]) {
        case FIL_CF90: st = pdo_cf90_init(&o->cf[d], ns[d], periodic[d]); break;
        case FIL_GAUSSIAN: st = pdo_gaussian_init(&o->xf[d], ns[d], periodic[d]); break;
        case FIL_LSTSQ: st = pdo_lstsq_init(&o->xf[d], ns[d], periodic[d]); break;
        default:
            set_error("Incorrect method select");  // filters.F90:138 (GracefulExit 52)
            st = PDO_ERR_ARG;
        }
        o->method[d] = method[d];
        if (st) {
            delete o;
            return st;
        }
    }
    *h = o;
    return 0;
}

int pdo_filters_method(pdo_filters h, int dir, int *method)
{
    if (!h || !method || dir < 1 || dir > 3) { set_error("invalid argument"); return PDO_ERR_ARG; }
    *method = h->method[dir - 1];
    return 0;
}

#define PDO_FIL(NAME, D)                                                                                         \
    int NAME(pdo_filters h, int prec, const void *f, void *ff, int bc1, int bcn, void *stream)                   \
    {                                                                                                            \
        if (!h || !f || !ff) { set_error("invalid argument"); return PDO_ERR_ARG; }                              \
        return h->apply(D, prec, f, ff, bc1, bcn, (hipStream_t)stream);                                          \
    }
PDO_FIL(pdo_filters_filterx, 0)
PDO_FIL(pdo_filters_filtery, 1)
PDO_FIL(pdo_filters_filterz, 2)
#undef PDO_FIL

int pdo_filters_destroy(pdo_filters h)
{
    delete h;
    return 0;
}

// operators.F90:18-53
int pdo_operators_gradient(pdo_derivatives der, int prec, const void *f, void *dfdx, void *dfdy, void *dfdz,
                           const int *bc, void *stream)
{
    if (!der || !f || !dfdx || !dfdy || !dfdz) { set_error("invalid argument"); return PDO_ERR_ARG; }
    static const int zero[6] = {0, 0, 0, 0, 0, 0};
    const int *b = bc ? bc : zero;
    hipStream_t st = (hipStream_t)stream;
    int r;
    if ((r = der->apply(1, 1, prec, f, dfdy, b[2], b[3], st))) return r;
    if ((r = der->apply(0, 1, prec, f, dfdx, b[0], b[1], st))) return r;
    return der->apply(2, 1, prec, f, dfdz, b[4], b[5], st);
}

// operators.F90:55-116: curl = (dw/dy - dv/dz, du/dz - dw/dx, dv/dx - du/dy),
// three consecutive fields
int pdo_operators_curl(pdo_derivatives der, int prec, const void *u, const void *v, const void *w, void *curl,
                       const int *bc, void *stream)
{
    if (!der || !u || !v || !w || !curl) { set_error("invalid argument"); return PDO_ERR_ARG; }
    static const int zero[6] = {0, 0, 0, 0, 0, 0};
    const int *b = bc ? bc : zero;
    hipStream_t st = (hipStream_t)stream;
    const long N = (long)der->nx * der->ny * der->nz;
    const size_t W = word_of(prec);
    int r;
    if ((r = der->scratch(N * W))) return r;
    char *c = (char *)curl;
    void *c1 = c, *c2 = c + N * W, *c3 = c + 2 * N * W;
    const bool dbl = prec == PDO_F64;
    hipError_t e;
    if ((r = der->apply(1, 1, prec, w, c1, b[2], b[3], st))) return r;
    if ((r = der->apply(2, 1, prec, v, der->tmp, b[4], b[5], st))) return r;
    if ((e = launch_axpy(dbl, c1, c1, der->tmp, -1.0, N, st)) != hipSuccess) return hip_fail(e, "curl");
    if ((r = der->apply(2, 1, prec, u, c2, b[4], b[5], st))) return r;
    if ((r = der->apply(0, 1, prec, w, der->tmp, b[0], b[1], st))) return r;
    if ((e = launch_axpy(dbl, c2, c2, der->tmp, -1.0, N, st)) != hipSuccess) return hip_fail(e, "curl");
    if ((r = der->apply(0, 1, prec, v, c3, b[0], b[1], st))) return r;
    if ((r = der->apply(1, 1, prec, u, der->tmp, b[2], b[3], st))) return r;
    if ((e = launch_axpy(dbl, c3, c3, der->tmp, -1.0, N, st)) != hipSuccess) return hip_fail(e, "curl");
    return 0;
}

// operators.F90:118-156: div = dv/dy + du/dx + dw/dz with the bc signs
// reversed (a velocity component that is even across a boundary has an odd
// normal derivative)
int pdo_operators_divergence(pdo_derivatives der, int prec, const void *u, const void *v, const void *w, void *div,
                             const int *bc, void *stream)
{
    if (!der || !u || !v || !w || !div) { set_error("invalid argument"); return PDO_ERR_ARG; }
    static const int zero[6] = {0, 0, 0, 0, 0, 0};
    const int *b = bc ? bc : zero;
    hipStream_t st = (hipStream_t)stream;
    const long N = (long)der->nx * der->ny * der->nz;
    const bool dbl = prec == PDO_F64;
    int r;
    if ((r = der->scratch(N * word_of(prec)))) return r;
    hipError_t e;
    if ((r = der->apply(1, 1, prec, v, div, -b[2], -b[3], st))) return r;
    if ((r = der->apply(0, 1, prec, u, der->tmp, -b[0], -b[1], st))) return r;
    if ((e = launch_axpy(dbl, div, div, der->tmp, 1.0, N, st)) != hipSuccess) return hip_fail(e, "divergence");
    if ((r = der->apply(2, 1, prec, w, der->tmp, -b[4], -b[5], st))) return r;
    if ((e = launch_axpy(dbl, div, div, der->tmp, 1.0, N, st)) != hipSuccess) return hip_fail(e, "divergence");
    return 0;
}

// operators.F90:158-223: filter in y, then x, then z, numtimes each, in place
int pdo_operators_filter3d(pdo_filters fil, int prec, void *arr, int numtimes, const int *bc, void *stream)
{
    if (!fil || !arr) { set_error("invalid argument"); return PDO_ERR_ARG; }
    static const int zero[6] = {0, 0, 0, 0, 0, 0};
    const int *b = bc ? bc : zero;
    const int times = numtimes >= 1 ? numtimes : 1;  // operators.F90:176-180
    hipStream_t st = (hipStream_t)stream;
    const long N = (long)fil->nx * fil->ny * fil->nz;
    const size_t W = word_of(prec);
    int r;
    if ((r = fil->scratch(N * W))) return r;
    // passes alternate between two work fields (a pass never filters in
    // place); the result is copied back into the caller's array
    const void *cur = arr;
    int nb = 0;
    const int order[3] = {1, 0, 2};
    for (int d : order) {
        for (int t = 0; t < times; ++t) {
            void *dst = fil->tmp[nb];
            if ((r = fil->apply(d, prec, cur, dst, b[2 * d], b[2 * d + 1], st))) return r;
            cur = dst;
            nb ^= 1;
        }
    }
    hipError_t e = hipMemcpyAsync(arr, cur, N * W, hipMemcpyDeviceToDevice, st);
    return e == hipSuccess ? 0 : hip_fail(e, "filter3D copy");
}

}  // extern "C"
