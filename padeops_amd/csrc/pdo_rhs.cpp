// Full RHS evaluation on a triply periodic n^3 grid: the composite the
// benchmark times.  Built from the same operator objects the C ABI exposes
// (cd06 derivatives, cf90 filters, PoissonPeriodic), in the order of the
// reference call sites: derivatives%ddx/ddy/ddz (derivatives.F90:447-510),
// filters%filterx/y/z (filters.F90:220-269), PoissonPeriodic%poisson_solve
// with modified-wavenumber callbacks (PoissonPeriodic.F90:62-87,188-204).
#include <algorithm>
#include <cstring>
#include <vector>

#include "../../include/padeops_hip.h"
#include "pdo_ops.hpp"
#include "pdo_poisson.hpp"

// run the nine gradient passes on a second stream (1) or in order on the
// caller's stream (0); tuning constant
#ifndef PDO_RHS_SIDE_STREAM
#define PDO_RHS_SIDE_STREAM 1
#endif
// gradient components (of 3) launched beside the filter chain; the others
// run beside the divergence, the Poisson solve and the projection; tuning
// constant
#ifndef PDO_RHS_GRAD_EARLY
#define PDO_RHS_GRAD_EARLY 3
#endif
// form the x terms of the divergence and of the projection inside the
// Poisson solve's x transforms (1) or as compact x passes (0)
#ifndef PDO_RHS_SPECTRAL_X
#define PDO_RHS_SPECTRAL_X 1
#endif

namespace pdo {

struct RhsObject {
    int n = 0;
    double L = 0;
    int prec = 0;
    size_t word = 8;
    // d1g serves the gradient passes on the side stream, d1 those on the
    // caller's stream: one object per stream, so that no variant state (the
    // general kernel's work area) is shared by concurrent launches
    OpObject d1, d1g, fil;
    PoissonObject poisson;
    void *ut[3] = {nullptr, nullptr, nullptr};
    void *tmp = nullptr, *div = nullptr;
    double *partial = nullptr;
    int nblocks = 1024;
    // the nine gradient passes run on a second stream, concurrently with the
    // filter chain; fork and join are events on the caller's stream
    hipStream_t side = nullptr;
    hipEvent_t ev_fork = nullptr, ev_fork2 = nullptr, ev_join = nullptr;
    ~RhsObject()
    {
        if (ev_fork) (void)hipEventDestroy(ev_fork);
        if (ev_fork2) (void)hipEventDestroy(ev_fork2);
        if (ev_join) (void)hipEventDestroy(ev_join);
        if (side) (void)hipStreamDestroy(side);
        for (void *p : ut) if (p) (void)hipFree(p);
        if (tmp) (void)hipFree(tmp);
        if (div) (void)hipFree(div);
        if (partial) (void)hipFree(partial);
    }
    void *field(void *base, int q) const { return (char *)base + (size_t)q * n * n * n * word; }
};

}  // namespace pdo

struct pdo_rhs_s : public pdo::RhsObject {};

using namespace pdo;

extern "C" {

int pdo_rhs_init(pdo_rhs *h, int n, double L, int prec)
{
    if (!h) return PDO_ERR_ARG;
    *h = nullptr;
    if (prec_mode(prec) < 0) { set_error("unknown precision mode"); return PDO_ERR_ARG; }
    auto *o = new pdo_rhs_s();
    o->n = n; o->L = L; o->prec = prec; o->word = (prec == PDO_F64) ? 8 : 4;
    const double dx = L / n;
    for (OpObject *d : {&o->d1, &o->d1g}) {
        d->scheme = SCHEME_CD06; d->n = n; d->dx = dx; d->periodic = true;
    }
    o->fil.scheme = SCHEME_CF90; o->fil.n = n; o->fil.dx = 1.0; o->fil.periodic = true;
    int st = o->poisson.init(n, n, n, dx, dx, dx, PDO_MODK_CD06, PDO_MODK_CD06, PDO_MODK_CD06, 1);
    if (st) { delete o; return st; }
    const size_t bytes = (size_t)n * n * n * o->word;
    hipError_t e = hipSuccess;
    for (int i = 0; i < 3 && e == hipSuccess; i++) e = hipMalloc(&o->ut[i], bytes);
    if (e == hipSuccess) e = hipMalloc(&o->tmp, bytes);
    if (e == hipSuccess) e = hipMalloc(&o->div, bytes);
    if (e == hipSuccess) e = hipMalloc(&o->partial, o->nblocks * sizeof(double));
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&o->side, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&o->ev_fork, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&o->ev_fork2, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&o->ev_join, hipEventDisableTiming);
    if (e != hipSuccess) { delete o; return hip_fail(e, "hipMalloc(rhs work buffers)"); }
    *h = o;
    return 0;
}

int pdo_rhs_eval(pdo_rhs h, const void *u, const void *v, const void *w, void *grad, void *uo, void *vo, void *wo,
                 void *stream)
{
    if (!h || !u || !v || !w || !grad || !uo || !vo || !wo) { set_error("null argument"); return PDO_ERR_ARG; }
    hipStream_t st = (hipStream_t)stream;
    const int n = h->n;
    const long nn = (long)n * n;
    const void *ui[3] = {u, v, w};
    void *uoi[3] = {uo, vo, wo};
    int r;
    // derivative / filter along direction d of a field (n, n, n)
    auto op_on = [&](hipStream_t s, OpObject *o, int kind, int d, const void *f, void *out, int epi,
                     const void *aux) {
        return apply(o, kind, 0, 0, d, h->prec, f, out, d == 3 ? nn : n, d == 3 ? 1 : n, s, epi, aux);
    };
    auto op = [&](OpObject *o, int kind, int d, const void *f, void *out, int epi, const void *aux) {
        return op_on(st, o, kind, d, f, out, epi, aux);
    };
    // the 9 velocity-gradient components run on the side stream (they only
    // read u_i): those of the first PDO_RHS_GRAD_EARLY components beside the
    // filter chain, the others beside the divergence / Poisson / projection
    hipError_t e = hipSuccess;
    auto gradients = [&](int i0, int i1, hipEvent_t fork) -> int {
        if (i0 >= i1) return 0;
        hipError_t ee = hipEventRecord(fork, st);
        if (ee == hipSuccess) ee = hipStreamWaitEvent(h->side, fork, 0);
        if (ee != hipSuccess) return hip_fail(ee, "stream fork");
        for (int i = i0; i < i1; i++)
            for (int j = 0; j < 3; j++) {
                const int rr = op_on(PDO_RHS_SIDE_STREAM ? h->side : st, &h->d1g, KIND_D1, j + 1, ui[i],
                                     h->field(grad, 3 * i + j), 0, nullptr);
                if (rr) return rr;
            }
        return 0;
    };
    if ((r = gradients(0, PDO_RHS_GRAD_EARLY, h->ev_fork))) return r;
    // filtered velocity ut_i = F_z F_y F_x u_i (uo_i is the x-pass scratch)
    for (int i = 0; i < 3; i++) {
        if ((r = op(&h->fil, KIND_FILTER, 1, ui[i], uoi[i], 0, nullptr))) return r;
        if ((r = op(&h->fil, KIND_FILTER, 2, uoi[i], h->tmp, 0, nullptr))) return r;
        if ((r = op(&h->fil, KIND_FILTER, 3, h->tmp, h->ut[i], 0, nullptr))) return r;
    }
    if ((r = gradients(PDO_RHS_GRAD_EARLY, 3, h->ev_fork2))) return r;
    if ((e = hipEventRecord(h->ev_join, h->side)) != hipSuccess) return hip_fail(e, "stream join");
    // divergence of the filtered velocity, accumulated in place
    if (PDO_RHS_SPECTRAL_X) {
        // the z and y terms as compact passes; the x term, and the x
        // derivative of the pressure, are taken by the Poisson solve's x
        // transforms with the cd06 eigenvalues i k'(kx) (on the periodic
        // grid the same operator as the compact x solve)
        if ((r = op(&h->d1, KIND_D1, 3, h->ut[2], h->div, 0, nullptr))) return r;
        if ((r = op(&h->d1, KIND_D1, 2, h->ut[1], h->div, 1, h->div))) return r;
        // p = Poisson(div + d ut/dx) (in place) and u = ut - dp/dx
        if ((r = h->poisson.solve_dx(h->prec, h->div, h->ut[0], h->div, uoi[0], st))) return r;
        for (int i = 1; i < 3; i++)
            if ((r = op(&h->d1, KIND_D1, i + 1, h->div, uoi[i], 2, h->ut[i]))) return r;
    } else {
        if ((r = op(&h->d1, KIND_D1, 1, h->ut[0], h->div, 0, nullptr))) return r;
        if ((r = op(&h->d1, KIND_D1, 2, h->ut[1], h->div, 1, h->div))) return r;
        if ((r = op(&h->d1, KIND_D1, 3, h->ut[2], h->div, 1, h->div))) return r;
        // pressure (in place) and projection u = ut - grad p
        if ((r = h->poisson.solve(h->prec, h->div, h->div, st))) return r;
        for (int i = 0; i < 3; i++)
            if ((r = op(&h->d1, KIND_D1, i + 1, h->div, uoi[i], 2, h->ut[i]))) return r;
    }
    // everything after this call on the caller's stream also follows the gradients
    if ((e = hipStreamWaitEvent(st, h->ev_join, 0)) != hipSuccess) return hip_fail(e, "stream join");
    return 0;
}

int pdo_rhs_pressure(pdo_rhs h, const void **p)
{
    if (!h || !p) return PDO_ERR_ARG;
    *p = h->div;
    return 0;
}

int pdo_rhs_max_divergence(pdo_rhs h, const void *u, const void *v, const void *w, double *maxdiv, void *stream)
{
    if (!h || !u || !v || !w || !maxdiv) return PDO_ERR_ARG;
    hipStream_t st = (hipStream_t)stream;
    const int n = h->n;
    const long nn = (long)n * n;
    int r;
    if ((r = apply(&h->d1, KIND_D1, 0, 0, 1, h->prec, u, h->tmp, n, n, st))) return r;
    if ((r = apply(&h->d1, KIND_D1, 0, 0, 2, h->prec, v, h->tmp, n, n, st, 1, h->tmp))) return r;
    if ((r = apply(&h->d1, KIND_D1, 0, 0, 3, h->prec, w, h->tmp, nn, 1, st, 1, h->tmp))) return r;
    hipError_t e = launch_absmax(h->tmp, h->prec == PDO_F64, nn * n, h->partial, h->nblocks, st);
    if (e != hipSuccess) return hip_fail(e, "absmax");
    std::vector<double> part(h->nblocks);
    e = hipMemcpyAsync(part.data(), h->partial, part.size() * sizeof(double), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return hip_fail(e, "divergence readback");
    *maxdiv = *std::max_element(part.begin(), part.end());
    return 0;
}

int pdo_rhs_destroy(pdo_rhs h)
{
    delete h;
    return 0;
}

}  // extern "C"
