// Allocation-free compact-operator entry point, pdo_op_apply_noalloc.
//
// The one-lane-per-pencil kernel (sizes outside the register path) keeps an
// FP64 work area of n_sys values per pencil.  The ordinary entry points take
// it from stream-ordered memory per call; this one takes it from `scratch`, a
// fixed array in device memory of which every device holds its own copy for
// the life of the process.  Nothing here calls hipMalloc or hipFree, so
// callers that may not allocate (code under stream capture, serving loops)
// can apply any operator variant that an earlier, ordinary call has built.
#include <hip/hip_runtime.h>

#include <mutex>

#include "../../include/padeops_hip.h"
#include "pdo_ops.hpp"

namespace pdo {

constexpr size_t NOALLOC_WORDS = (size_t)PDO_NOALLOC_SCRATCH_BYTES / sizeof(double);

// The work area.  In host code this name is a host-side stand-in, not a device
// address: kernels receive the per-device address from hipGetSymbolAddress.
static __device__ __align__(256) double scratch[NOALLOC_WORDS];

namespace {
constexpr int MAX_DEVICES = 64;
std::mutex g_mu;
double *g_addr[MAX_DEVICES] = {};
}  // namespace

// device address of this device's copy of `scratch`, looked up once per
// device (with that device current)
static int scratch_address(double **out)
{
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
    if (dev < 0 || dev >= MAX_DEVICES) {
        set_error("device index out of range for the allocation-free work area");
        return PDO_ERR_ARG;
    }
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_addr[dev]) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)");
        g_addr[dev] = (double *)p;
    }
    *out = g_addr[dev];
    return 0;
}

}  // namespace pdo

using namespace pdo;

extern "C" {

int pdo_op_apply_noalloc(pdo_op h, int kind, int dir, int prec, const void *f, void *out, int na, int nb, int bc1,
                         int bcn, void *stream)
{
    if (!h) { set_error("null handle"); return PDO_ERR_ARG; }
    double *work = nullptr;
    if (int st = scratch_address(&work)) return st;
    return apply_noalloc(h, kind, bc1, bcn, dir, prec, f, out, na, nb, (hipStream_t)stream, work, NOALLOC_WORDS);
}

}  // extern "C"
