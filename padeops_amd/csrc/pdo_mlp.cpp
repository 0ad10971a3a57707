// Learned SGS closure network behind the C ABI: parameters, Adam state and
// training buffers live on the device; see pdo_mlp.hip for the kernels.
#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/padeops_hip.h"
#include "pdo_mlp.hpp"
#include "pdo_ops.hpp"

namespace pdo {

struct MlpObject {
    long max_batch = 0;
    float *P = nullptr, *M = nullptr, *V = nullptr, *grads = nullptr, *partial = nullptr, *Y = nullptr;
    unsigned short *Wb = nullptr, *act = nullptr;
    MlpCache cache{};
    long t = 0;  // Adam step count
    long cached_batch = 0;  // batch of the activations in cache (pdo_mlp_forward_train)
    std::vector<void *> allocs;
    ~MlpObject()
    {
        for (void *p : allocs) (void)hipFree(p);
    }
    template <typename T> int alloc(T **p, size_t n)
    {
        hipError_t e = hipMalloc((void **)p, n * sizeof(T));
        if (e != hipSuccess) return hip_fail(e, "hipMalloc(mlp)");
        allocs.push_back(*p);
        e = hipMemset(*p, 0, n * sizeof(T));
        return e == hipSuccess ? 0 : hip_fail(e, "hipMemset(mlp)");
    }
};

}  // namespace pdo

struct pdo_mlp_s : public pdo::MlpObject {};

using namespace pdo;

static int ok(hipError_t e, const char *w) { return e == hipSuccess ? 0 : hip_fail(e, w); }

extern "C" {

int pdo_mlp_nparam(void) { return MLP_NPARAM; }

int pdo_mlp_create(pdo_mlp *h, long max_batch)
{
    if (!h || max_batch <= 0) { set_error("invalid argument"); return PDO_ERR_ARG; }
    *h = nullptr;
    auto *o = new pdo_mlp_s();
    o->max_batch = max_batch;
    const long nb = (max_batch + 63) / 64;
    int st = 0;
    if (!st) st = o->alloc(&o->P, MLP_NPARAM);
    if (!st) st = o->alloc(&o->M, MLP_NPARAM);
    if (!st) st = o->alloc(&o->V, MLP_NPARAM);
    if (!st) st = o->alloc(&o->grads, MLP_NPARAM);
    if (!st) st = o->alloc(&o->Wb, MLP_NPARAM);
    if (!st) st = o->alloc(&o->partial, (size_t)nb * MLP_NPARAM);
    if (!st) st = o->alloc(&o->Y, (size_t)max_batch * MLP_OUT);
    if (!st) st = o->alloc(&o->act, (size_t)max_batch * (MLP_IN_PAD + 6 * MLP_HID));
    if (st) { delete o; return st; }
    unsigned short *a = o->act;
    o->cache.H[0] = a;
    a += (size_t)max_batch * MLP_IN_PAD;
    for (int l = 0; l < 3; l++) {
        o->cache.Z[l] = a;
        a += (size_t)max_batch * MLP_HID;
        o->cache.H[l + 1] = a;
        a += (size_t)max_batch * MLP_HID;
    }
    *h = o;
    return 0;
}

int pdo_mlp_set_params(pdo_mlp h, const float *params)
{
    if (!h || !params) { set_error("invalid argument"); return PDO_ERR_ARG; }
    int st = ok(hipMemcpy(h->P, params, MLP_NPARAM * sizeof(float), hipMemcpyHostToDevice), "hipMemcpy");
    if (st) return st;
    if ((st = ok(hipMemset(h->M, 0, MLP_NPARAM * sizeof(float)), "hipMemset"))) return st;
    if ((st = ok(hipMemset(h->V, 0, MLP_NPARAM * sizeof(float)), "hipMemset"))) return st;
    h->t = 0;
    if ((st = ok(launch_mlp_to_bf16(h->P, h->Wb, nullptr), "mlp"))) return st;
    return ok(hipDeviceSynchronize(), "hipDeviceSynchronize");
}

int pdo_mlp_get_params(pdo_mlp h, float *params)
{
    if (!h || !params) { set_error("invalid argument"); return PDO_ERR_ARG; }
    return ok(hipMemcpy(params, h->P, MLP_NPARAM * sizeof(float), hipMemcpyDeviceToHost), "hipMemcpy");
}

int pdo_mlp_get_grads(pdo_mlp h, float *grads)
{
    if (!h || !grads) { set_error("invalid argument"); return PDO_ERR_ARG; }
    return ok(hipMemcpy(grads, h->grads, MLP_NPARAM * sizeof(float), hipMemcpyDeviceToHost), "hipMemcpy");
}

int pdo_mlp_forward(pdo_mlp h, const float *X, float *Y, long B, void *stream)
{
    if (!h || !X || !Y || B < 0) { set_error("invalid argument"); return PDO_ERR_ARG; }
    const MlpCache none{};
    return ok(launch_mlp_forward(X, Y, h->Wb, h->P, B, none, (hipStream_t)stream), "mlp forward");
}

int pdo_mlp_gradients(pdo_mlp h, const float *X, const float *T, long B, float *Y, void *stream)
{
    if (!h || !X || !T || B <= 0 || B > h->max_batch) { set_error("invalid argument or batch too large"); return PDO_ERR_ARG; }
    float *Yo = Y ? Y : h->Y;
    hipStream_t st = (hipStream_t)stream;
    int rc;
    if ((rc = ok(launch_mlp_forward(X, Yo, h->Wb, h->P, B, h->cache, st), "mlp forward"))) return rc;
    h->cached_batch = B;
    return ok(launch_mlp_backward(Yo, T, nullptr, h->Wb, B, h->cache, h->partial, h->grads, st), "mlp backward");
}

int pdo_mlp_forward_train(pdo_mlp h, const float *X, float *Y, long B, void *stream)
{
    if (!h || !X || !Y || B <= 0 || B > h->max_batch) { set_error("invalid argument or batch too large"); return PDO_ERR_ARG; }
    h->cached_batch = B;
    return ok(launch_mlp_forward(X, Y, h->Wb, h->P, B, h->cache, (hipStream_t)stream), "mlp forward");
}

int pdo_mlp_backward(pdo_mlp h, const float *dY, long B, void *stream)
{
    if (!h || !dY || B <= 0) { set_error("invalid argument"); return PDO_ERR_ARG; }
    if (B != h->cached_batch) {
        set_error("pdo_mlp_backward: batch differs from the last pdo_mlp_forward_train");
        return PDO_ERR_ARG;
    }
    return ok(launch_mlp_backward(nullptr, nullptr, dY, h->Wb, B, h->cache, h->partial, h->grads,
                                  (hipStream_t)stream), "mlp backward");
}

int pdo_mlp_adam_step(pdo_mlp h, float lr, float b1, float b2, float eps, void *stream)
{
    if (!h) { set_error("invalid argument"); return PDO_ERR_ARG; }
    h->t++;
    const float c1 = 1.0f - (float)std::pow((double)b1, (double)h->t);
    const float c2 = 1.0f - (float)std::pow((double)b2, (double)h->t);
    return ok(launch_mlp_adam(h->P, h->M, h->V, h->grads, h->Wb, lr, b1, b2, eps, c1, c2, (hipStream_t)stream),
              "mlp adam");
}

int pdo_mlp_train_step(pdo_mlp h, const float *X, const float *T, long B, float lr, float b1, float b2, float eps,
                       float *Y, void *stream)
{
    int rc = pdo_mlp_gradients(h, X, T, B, Y, stream);
    if (rc) return rc;
    return pdo_mlp_adam_step(h, lr, b1, b2, eps, stream);
}

int pdo_mlp_destroy(pdo_mlp h)
{
    delete h;
    return 0;
}

}  // extern "C"
