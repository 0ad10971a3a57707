// Pointwise kernels on half spectra (nxh, ny, nplanes) of interleaved complex
// values: multiplication by i k (spectral.F90:235-312), the 2/3-rule mask
// (spectral.F90:314-341), the two pointwise steps of the pressure projection
// and divergence check (PadePoisson.F90:386-432, 1165-1244) and the Stokes
// pressure of the projection with walls (:236-384, 641-714).
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

// f = src m1[i] m2[j] (src may be f)
template <typename TS>
__global__ void __launch_bounds__(256) k_mask2(const TS *src, TS *f, const double *__restrict__ m1,
                                              const double *__restrict__ m2, int nxh, int ny, long n)
{
    using V = typename CV<TS>::V;
    const V *q = (const V *)src;
    V *p = (V *)f;
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) {
        const long i = t % nxh, j = (t / nxh) % ny;
        const TS g = (TS)(m1[i] * m2[j]);
        const V x = q[t];
        p[t] = V{x.x * g, x.y * g};
    }
}

// f = d + i (k1 u + k2 v) (d null: zero)
template <typename TS, typename TC>
__global__ void __launch_bounds__(256) k_div_spec(const TS *u, const TS *v, const TS *d, TS *f,
                                                 const double *__restrict__ k1, const double *__restrict__ k2, int nxh,
                                                 int ny, long n)
{
    using V = typename CV<TS>::V;
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) {
        const long i = t % nxh, j = (t / nxh) % ny;
        const TC a = (TC)k1[i], b = (TC)k2[j];
        const V uu = ((const V *)u)[t], vv = ((const V *)v)[t];
        const V dd = d ? ((const V *)d)[t] : V{(TS)0, (TS)0};
        const TC sr = a * (TC)uu.x + b * (TC)vv.x, si = a * (TC)uu.y + b * (TC)vv.y;
        ((V *)f)[t] = V{(TS)((TC)dd.x - si), (TS)((TC)dd.y + sr)};
    }
}

// u -= i k1 p, v -= i k2 p on the nz cell planes; w -= dp on the nz+1 edge
// planes (dp null: w untouched)
template <typename TS, typename TC>
__global__ void __launch_bounds__(256) k_proj_update(TS *u, TS *v, TS *w, const TS *p, const TS *dp,
                                                    const double *__restrict__ k1, const double *__restrict__ k2,
                                                    int nxh, int ny, long ncell, long nedge)
{
    using V = typename CV<TS>::V;
    const long n = dp ? nedge : ncell;
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) {
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
        if (!dp) continue;
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
    return launch_mask2_oop(mode, f, f, m1, m2, nxh, ny, n, st);
}

hipError_t launch_mask2_oop(int mode, const void *src, void *f, const double *m1, const double *m2, int nxh, int ny,
                            long n, hipStream_t st)
{
    if (mode == 0)
        hipLaunchKernelGGL((k_mask2<double>), dim3(grid_for(n)), dim3(256), 0, st, (const double *)src, (double *)f,
                           m1, m2, nxh, ny, n);
    else
        hipLaunchKernelGGL((k_mask2<float>), dim3(grid_for(n)), dim3(256), 0, st, (const float *)src, (float *)f, m1,
                           m2, nxh, ny, n);
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

// ---------------------------------------------------------------------------
// Stokes pressure with walls (PadePoisson.F90:236-307 tables, 320-384 and
// 641-714 the step), one thread per (kx, ky) line.  With lam = |(k1, k2)|
// (plain GetWaveNums values) and denFact d of the line:
//   chat_b = -w(bottom) d; chat_t = (w(top) - chat_b sinh_bot(Lz)) d
//   u -= k1 (i chat_b cosh_bot + i chat_t cosh_top) on the cells (v with k2)
//   w -= chat_b sinh_bot + chat_t sinh_top on the inner edges, 0 on the walls
// cosh_bot(z) = cosh(lam (Lz - z)), cosh_top(z) = cosh(lam z) (caps 4e13,
// 1e13 from an argument of 32 on), sinh_bot(z) = -lam sinh(lam (Lz - z)),
// sinh_top(z) = lam sinh(lam z) (caps -+4e13): the reference's tables,
// evaluated on the fly.  Arithmetic in FP64 for every storage type.  coef
// (4 doubles per line) keeps chat_b, chat_t for the pressure.
// ---------------------------------------------------------------------------
__device__ inline double st_cosh_bot(double l, double Lz, double z)
{
    const double t = l * (Lz - z);
    return t < 32.0 ? cosh(t) : 4.0e13;
}
__device__ inline double st_cosh_top(double l, double z)
{
    const double t = l * z;
    return t < 32.0 ? cosh(t) : 1.0e13;
}
__device__ inline double st_sinh_bot(double l, double Lz, double z)
{
    const double t = l * (Lz - z);
    return t < 32.0 ? -l * sinh(t) : -4.0e13;
}
__device__ inline double st_sinh_top(double l, double z)
{
    const double t = l * z;
    return t < 32.0 ? l * sinh(t) : 4.0e13;
}

template <typename TS>
__global__ void __launch_bounds__(256) k_stokes(TS *u, TS *v, TS *w, double *coef, StokesTables t)
{
    using V = typename CV<TS>::V;
    const long np = (long)t.nxh * t.ny;
    const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= np) return;
    const double a = t.k1[q % t.nxh], b = t.k2[q / t.nxh], l = t.lam[q], d = t.den[q];
    V *wq = (V *)w + q;
    const V w0 = wq[0], wn = wq[(long)t.nz * np];
    const double cbr = -(double)w0.x * d, cbi = -(double)w0.y * d;
    const double sbn = st_sinh_bot(l, t.Lz, t.ze[t.nz]);
    const double ctr = ((double)wn.x - cbr * sbn) * d, cti = ((double)wn.y - cbi * sbn) * d;
    coef[4 * q] = cbr;
    coef[4 * q + 1] = cbi;
    coef[4 * q + 2] = ctr;
    coef[4 * q + 3] = cti;
    V *uq = (V *)u + q, *vq = (V *)v + q;
    for (int k = 0; k < t.nz; ++k) {
        const double z = t.zc[k], hb = st_cosh_bot(l, t.Lz, z), ht = st_cosh_top(l, z);
        // i chat cosh
        const double pbr = -cbi * hb, pbi = cbr * hb, ptr = -cti * ht, pti = ctr * ht;
        V uu = uq[k * np], vv = vq[k * np];
        uu.x = (TS)(((double)uu.x - a * pbr) - a * ptr);
        uu.y = (TS)(((double)uu.y - a * pbi) - a * pti);
        vv.x = (TS)(((double)vv.x - b * pbr) - b * ptr);
        vv.y = (TS)(((double)vv.y - b * pbi) - b * pti);
        uq[k * np] = uu;
        vq[k * np] = vv;
    }
    {
        const double s0 = st_sinh_top(l, t.ze[0]);
        wq[0] = V{(TS)(0.0 - ctr * s0), (TS)(0.0 - cti * s0)};
    }
    for (int k = 1; k < t.nz; ++k) {
        const double z = t.ze[k], sb = st_sinh_bot(l, t.Lz, z), stp = st_sinh_top(l, z);
        V ww = wq[k * np];
        ww.x = (TS)(((double)ww.x - cbr * sb) - ctr * stp);
        ww.y = (TS)(((double)ww.y - cbi * sb) - cti * stp);
        wq[k * np] = ww;
    }
    wq[(long)t.nz * np] = V{(TS)0, (TS)0};
}

// p += i chat_b cosh_bot + i chat_t cosh_top on the nz cell planes (the
// reference's pressure adds phat_z1 + phat_z2, :882-890)
template <typename TS>
__global__ void __launch_bounds__(256) k_stokes_pressure(TS *p, const double *coef, StokesTables t)
{
    using V = typename CV<TS>::V;
    const long np = (long)t.nxh * t.ny, n = np * t.nz;
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) {
        const long q = e % np;
        const int k = (int)(e / np);
        const double l = t.lam[q], z = t.zc[k], hb = st_cosh_bot(l, t.Lz, z), ht = st_cosh_top(l, z);
        const double *c = coef + 4 * q;
        V pp = ((V *)p)[e];
        pp.x = (TS)(((double)pp.x + (-c[1] * hb)) + (-c[3] * ht));
        pp.y = (TS)(((double)pp.y + c[0] * hb) + c[2] * ht);
        ((V *)p)[e] = pp;
    }
}

hipError_t launch_stokes(int mode, void *u, void *v, void *w, double *coef, const StokesTables &t, hipStream_t st)
{
    const long np = (long)t.nxh * t.ny;
    const dim3 grid((unsigned)((np + 255) / 256));
    if (mode == 0)
        hipLaunchKernelGGL((k_stokes<double>), grid, dim3(256), 0, st, (double *)u, (double *)v, (double *)w, coef,
                           t);
    else
        hipLaunchKernelGGL((k_stokes<float>), grid, dim3(256), 0, st, (float *)u, (float *)v, (float *)w, coef, t);
    return hipGetLastError();
}

hipError_t launch_stokes_pressure(int mode, void *p, const double *coef, const StokesTables &t, hipStream_t st)
{
    const long n = (long)t.nxh * t.ny * t.nz;
    if (mode == 0)
        hipLaunchKernelGGL((k_stokes_pressure<double>), dim3(grid_for(n)), dim3(256), 0, st, (double *)p, coef, t);
    else
        hipLaunchKernelGGL((k_stokes_pressure<float>), dim3(grid_for(n)), dim3(256), 0, st, (float *)p, coef, t);
    return hipGetLastError();
}

}  // namespace pdo
