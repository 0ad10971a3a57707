// Pointwise kernels on half spectra (nxh, ny, nplanes) of interleaved complex
// values: multiplication by i k (spectral.F90:235-312), the 2/3-rule mask
// (spectral.F90:314-341) and the two pointwise steps of the periodic pressure
// projection and divergence check (PadePoisson.F90:386-432, 1165-1244).
// One thread per complex value, grid-stride; every kernel is HBM-bound.
#include <hip/hip_runtime.h>

#include "pdo_spectral.hpp"

namespace pdo {

template <typename TS> struct CV;
template <> struct CV<double> {
    using V = double2;
};
template <> struct CV<float> {
    using V = float2;
};

static unsigned grid_for(long n)
{
    long b = (n + 255) / 256;
    return (unsigned)(b < 8192 ? (b > 0 ? b : 1) : 8192);
}

// out = i k f with k = kv[index along dim] (dim 1: i, dim 2: j)
template <typename TS, typename TC>
__global__ void __launch_bounds__(256) k_mtimes_ik(const TS *in, TS *out, const double *__restrict__ kv, int nxh,
                                                  int ny, long n, int dim)
{
    using V = typename CV<TS>::V;
    const V *src = (const V *)in;
    V *dst = (V *)out;
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) {
        const long ij = dim == 1 ? t % nxh : (t / nxh) % ny;
        const TC k = (TC)kv[ij];
        const V f = src[t];
        dst[t] = V{(TS)(-k * (TC)f.y), (TS)(k * (TC)f.x)};
    }
}

// f *= m1[i] m2[j]
template <typename TS>
__global__ void __launch_bounds__(256) k_mask2(TS *f, const double *__restrict__ m1, const double *__restrict__ m2,
                                              int nxh, int ny, long n)
{
    using V = typename CV<TS>::V;
    V *p = (V *)f;
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) {
        const long i = t % nxh, j = (t / nxh) % ny;
        const TS g = (TS)(m1[i] * m2[j]);
        const V x = p[t];
        p[t] = V{x.x * g, x.y * g};
    }
}

// f = d + i (k1 u + k2 v)
template <typename TS, typename TC>
__global__ void __launch_bounds__(256) k_div_spec(const TS *u, const TS *v, const TS *d, TS *f,
                                                 const double *__restrict__ k1, const double *__restrict__ k2, int nxh,
                                                 int ny, long n)
{
    using V = typename CV<TS>::V;
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) {
        const long i = t % nxh, j = (t / nxh) % ny;
        const TC a = (TC)k1[i], b = (TC)k2[j];
        const V uu = ((const V *)u)[t], vv = ((const V *)v)[t], dd = ((const V *)d)[t];
        const TC sr = a * (TC)uu.x + b * (TC)vv.x, si = a * (TC)uu.y + b * (TC)vv.y;
        ((V *)f)[t] = V{(TS)((TC)dd.x - si), (TS)((TC)dd.y + sr)};
    }
}

// u -= i k1 p, v -= i k2 p on the nz cell planes; w -= dp on the nz+1 edge planes
template <typename TS, typename TC>
__global__ void __launch_bounds__(256) k_proj_update(TS *u, TS *v, TS *w, const TS *p, const TS *dp,
                                                    const double *__restrict__ k1, const double *__restrict__ k2,
                                                    int nxh, int ny, long ncell, long nedge)
{
    using V = typename CV<TS>::V;
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < nedge; t += (long)gridDim.x * blockDim.x) {
        if (t < ncell) {
            const long i = t % nxh, j = (t / nxh) % ny;
            const TC a = (TC)k1[i], b = (TC)k2[j];
            const V pp = ((const V *)p)[t];
            V uu = ((V *)u)[t], vv = ((V *)v)[t];
            uu.x = (TS)((TC)uu.x + a * (TC)pp.y);
            uu.y = (TS)((TC)uu.y - a * (TC)pp.x);
            vv.x = (TS)((TC)vv.x + b * (TC)pp.y);
            vv.y = (TS)((TC)vv.y - b * (TC)pp.x);
            ((V *)u)[t] = uu;
            ((V *)v)[t] = vv;
        }
        const V dd = ((const V *)dp)[t];
        V ww = ((V *)w)[t];
        ww.x = (TS)((TC)ww.x - (TC)dd.x);
        ww.y = (TS)((TC)ww.y - (TC)dd.y);
        ((V *)w)[t] = ww;
    }
}

template <typename TS, typename TC>
static hipError_t mtimes_ik_t(const void *in, void *out, const double *kv, int nxh, int ny, long n, int dim,
                              hipStream_t st)
{
    hipLaunchKernelGGL((k_mtimes_ik<TS, TC>), dim3(grid_for(n)), dim3(256), 0, st, (const TS *)in, (TS *)out, kv,
                       nxh, ny, n, dim);
    return hipGetLastError();
}

hipError_t launch_mtimes_ik(int mode, const void *in, void *out, const double *kv, int nxh, int ny, long n, int dim,
                            hipStream_t st)
{
    switch (mode) {
    case 0: return mtimes_ik_t<double, double>(in, out, kv, nxh, ny, n, dim, st);
    case 1: return mtimes_ik_t<float, float>(in, out, kv, nxh, ny, n, dim, st);
    case 2: return mtimes_ik_t<float, double>(in, out, kv, nxh, ny, n, dim, st);
    }
    return hipErrorInvalidValue;
}

hipError_t launch_mask2(int mode, void *f, const double *m1, const double *m2, int nxh, int ny, long n,
                        hipStream_t st)
{
    if (mode == 0)
        hipLaunchKernelGGL((k_mask2<double>), dim3(grid_for(n)), dim3(256), 0, st, (double *)f, m1, m2, nxh, ny, n);
    else
        hipLaunchKernelGGL((k_mask2<float>), dim3(grid_for(n)), dim3(256), 0, st, (float *)f, m1, m2, nxh, ny, n);
    return hipGetLastError();
}

template <typename TS, typename TC>
static hipError_t div_spec_t(const void *u, const void *v, const void *d, void *f, const double *k1,
                             const double *k2, int nxh, int ny, long n, hipStream_t st)
{
    hipLaunchKernelGGL((k_div_spec<TS, TC>), dim3(grid_for(n)), dim3(256), 0, st, (const TS *)u, (const TS *)v,
                       (const TS *)d, (TS *)f, k1, k2, nxh, ny, n);
    return hipGetLastError();
}

hipError_t launch_div_spec(int mode, const void *u, const void *v, const void *d, void *f, const double *k1,
                           const double *k2, int nxh, int ny, long n, hipStream_t st)
{
    switch (mode) {
    case 0: return div_spec_t<double, double>(u, v, d, f, k1, k2, nxh, ny, n, st);
    case 1: return div_spec_t<float, float>(u, v, d, f, k1, k2, nxh, ny, n, st);
    case 2: return div_spec_t<float, double>(u, v, d, f, k1, k2, nxh, ny, n, st);
    }
    return hipErrorInvalidValue;
}

template <typename TS, typename TC>
static hipError_t proj_update_t(void *u, void *v, void *w, const void *p, const void *dp, const double *k1,
                                const double *k2, int nxh, int ny, long ncell, long nedge, hipStream_t st)
{
    hipLaunchKernelGGL((k_proj_update<TS, TC>), dim3(grid_for(nedge)), dim3(256), 0, st, (TS *)u, (TS *)v, (TS *)w,
                       (const TS *)p, (const TS *)dp, k1, k2, nxh, ny, ncell, nedge);
    return hipGetLastError();
}

hipError_t launch_proj_update(int mode, void *u, void *v, void *w, const void *p, const void *dp, const double *k1,
                              const double *k2, int nxh, int ny, long ncell, long nedge, hipStream_t st)
{
    switch (mode) {
    case 0: return proj_update_t<double, double>(u, v, w, p, dp, k1, k2, nxh, ny, ncell, nedge, st);
    case 1: return proj_update_t<float, float>(u, v, w, p, dp, k1, k2, nxh, ny, ncell, nedge, st);
    case 2: return proj_update_t<float, double>(u, v, w, p, dp, k1, k2, nxh, ny, ncell, nedge, st);
    }
    return hipErrorInvalidValue;
}

}  // namespace pdo
