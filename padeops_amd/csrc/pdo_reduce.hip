// Max-abs reduction used by the divergence check.
#include <hip/hip_runtime.h>

#include "pdo_kernels.hpp"

namespace pdo {

template <typename T>
__global__ void __launch_bounds__(256) k_absmax(const T *__restrict__ x, long n, double *__restrict__ partial)
{
    __shared__ double sm[256];
    double m = 0.0;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const double v = fabs((double)x[i]);
        m = v > m ? v : m;
    }
    sm[threadIdx.x] = m;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) sm[threadIdx.x] = fmax(sm[threadIdx.x], sm[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = sm[0];
}

hipError_t launch_absmax(const void *x, bool is_double, long n, double *partial, int nblocks, hipStream_t st)
{
    if (is_double)
        hipLaunchKernelGGL((k_absmax<double>), dim3(nblocks), dim3(256), 0, st, (const double *)x, n, partial);
    else
        hipLaunchKernelGGL((k_absmax<float>), dim3(nblocks), dim3(256), 0, st, (const float *)x, n, partial);
    return hipGetLastError();
}

}  // namespace pdo
