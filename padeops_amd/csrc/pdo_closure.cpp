// Learned SGS closure object behind the C ABI: the network (pdo_mlp), the
// feature and output buffers of one field, and the reductions; kernels in
// pdo_closure.hip.
#include <cmath>
#include <vector>

#include "../../include/padeops_hip.h"
#include "pdo_closure.hpp"
#include "pdo_ops.hpp"

namespace pdo {

struct ClosureObject {
    pdo_mlp net = nullptr;
    long max_points = 0;
    float *X = nullptr, *G = nullptr;
    double *partial = nullptr, *sumsq = nullptr, *loss = nullptr;
    unsigned *nfall = nullptr;
    long nnorm = 0;       // points behind the current scale (sumsq)
    int fall_blocks = 0;  // entries of nfall written by the last stress call
    std::vector<void *> allocs;
    ~ClosureObject()
    {
        for (void *p : allocs) (void)hipFree(p);
    }
    template <typename T> int alloc(T **p, size_t n)
    {
        hipError_t e = hipMalloc((void **)p, n * sizeof(T));
        if (e != hipSuccess) return hip_fail(e, "hipMalloc(closure)");
        allocs.push_back(*p);
        return 0;
    }
};

}  // namespace pdo

struct pdo_closure_s : public pdo::ClosureObject {};

long pdo::closure_max_points(pdo_closure c) { return c ? c->max_points : 0; }

using namespace pdo;

static int ok(hipError_t e, const char *w) { return e == hipSuccess ? 0 : hip_fail(e, w); }

static bool prec_ok(int prec) { return prec == PDO_F64 || prec == PDO_F32_REFINE || prec == PDO_F32_STORE || prec == PDO_F32; }

extern "C" {

int pdo_closure_create(pdo_closure *h, pdo_mlp net, long max_points)
{
    if (!h || max_points <= 0) { set_error("invalid argument"); return PDO_ERR_ARG; }
    *h = nullptr;
    auto *o = new pdo_closure_s();
    o->net = net;
    o->max_points = max_points;
    int st = 0;
    if (!st) st = o->alloc(&o->X, (size_t)max_points * 10);
    if (!st) st = o->alloc(&o->G, (size_t)max_points * 6);
    if (!st) st = o->alloc(&o->partial, 8192);
    if (!st) st = o->alloc(&o->sumsq, 1);
    if (!st) st = o->alloc(&o->loss, 1);
    if (!st) st = o->alloc(&o->nfall, 8192);
    if (st) { delete o; return st; }
    *h = o;
    return 0;
}

int pdo_closure_features(pdo_closure h, int prec, const void *duidxj, long npts, double dratio, float *X, float *A9,
                         void *stream)
{
    if (!h || !duidxj || !X || npts <= 0 || !prec_ok(prec)) { set_error("invalid argument"); return PDO_ERR_ARG; }
    hipStream_t st = (hipStream_t)stream;
    const bool dbl = prec == PDO_F64;
    int rc;
    if ((rc = ok(launch_grad_sumsq(dbl, duidxj, npts, npts, h->partial, h->sumsq, st), "closure scale"))) return rc;
    h->nnorm = npts;
    return ok(launch_closure_features(dbl, duidxj, npts, npts, h->sumsq, npts, dratio, X, A9, st), "closure features");
}

int pdo_closure_stress(pdo_closure h, int prec, const void *duidxj, long npts, double delta, double dratio, int which,
                       int fallback, int rescale, void *tau, void *stream)
{
    if (!h || !duidxj || !tau || npts <= 0 || !prec_ok(prec) || which < 0 || which > 2) {
        set_error("invalid argument");
        return PDO_ERR_ARG;
    }
    if (!h->net) { set_error("closure object has no network"); return PDO_ERR_ARG; }
    if (npts > h->max_points) { set_error("closure: more points than the object was created for"); return PDO_ERR_ARG; }
    if (!rescale && h->nnorm <= 0) { set_error("closure: no scale to reuse (call with rescale first)"); return PDO_ERR_ARG; }
    hipStream_t st = (hipStream_t)stream;
    const bool dbl = prec == PDO_F64;
    int rc;
    if (rescale) {
        if ((rc = ok(launch_grad_sumsq(dbl, duidxj, npts, npts, h->partial, h->sumsq, st), "closure scale"))) return rc;
        h->nnorm = npts;
    }
    if ((rc = ok(launch_closure_features(dbl, duidxj, npts, npts, h->sumsq, h->nnorm, dratio, h->X, nullptr, st),
                 "closure features")))
        return rc;
    if ((rc = pdo_mlp_forward(h->net, h->X, h->G, npts, stream))) return rc;
    h->fall_blocks = closure_blocks(npts);
    return ok(launch_closure_stress(dbl, duidxj, npts, npts, h->sumsq, h->nnorm, delta * delta, h->G, which, fallback,
                                    tau, npts, h->nfall, st),
              "closure stress");
}

int pdo_closure_fallback_count(pdo_closure h, long *count, void *stream)
{
    if (!h || !count) { set_error("invalid argument"); return PDO_ERR_ARG; }
    *count = 0;
    if (h->fall_blocks <= 0) return 0;
    std::vector<unsigned> c(h->fall_blocks);
    hipStream_t st = (hipStream_t)stream;
    hipError_t e = hipMemcpyAsync(c.data(), h->nfall, c.size() * sizeof(unsigned), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return hip_fail(e, "closure fallback count");
    long s = 0;
    for (unsigned v : c) s += v;
    *count = s;
    return 0;
}

int pdo_closure_scale(pdo_closure h, double *rms, void *stream)
{
    if (!h || !rms) { set_error("invalid argument"); return PDO_ERR_ARG; }
    double s = 0;
    hipStream_t st = (hipStream_t)stream;
    hipError_t e = hipMemcpyAsync(&s, h->sumsq, sizeof(double), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return hip_fail(e, "closure scale readback");
    *rms = h->nnorm > 0 && s > 0 ? std::sqrt(s / (double)h->nnorm) : 0.0;
    return 0;
}

int pdo_closure_loss(pdo_closure h, const float *G, const float *A9, const float *T, long B, float *dG, double *loss,
                     void *stream)
{
    if (!h || !G || !A9 || !T || !dG || B <= 0) { set_error("invalid argument"); return PDO_ERR_ARG; }
    hipStream_t st = (hipStream_t)stream;
    int rc;
    if ((rc = ok(launch_closure_loss(G, A9, T, B, dG, h->partial, h->loss, st), "closure loss"))) return rc;
    if (!loss) return 0;
    hipError_t e = hipMemcpyAsync(loss, h->loss, sizeof(double), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    return ok(e, "closure loss readback");
}

int pdo_closure_destroy(pdo_closure h)
{
    delete h;
    return 0;
}

}  // extern "C"
