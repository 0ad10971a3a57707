// Subgrid-scale eddy-viscosity kernels of igrid's sgs_igrid
// (src/incompressible/sgs_models): the strain rate of the velocity gradient
// (eddyViscosity.F90:1-21), the Smagorinsky kernel |S| = sqrt(2 S_ij S_ij)
// (smagorinsky.F90:45-68), the sigma kernel (sigma.F90:30-100), the AMD
// kernel (AMD.F90:23-73), the model
// constant (eddyViscosity.F90:74-103), and the pointwise steps and planar
// averages of the standard dynamic procedure (standardDynamicProcedure.F90:
// 74-250).  Fields are (npts) points per component, components `stride`
// apart (the reference's (nx, ny, nz, 9) / (.., 6) arrays).  All kernels are
// pointwise and HBM-bound except the planar sums.
#include <hip/hip_runtime.h>

#include "pdo_sgs.hpp"

namespace pdo {

static unsigned grid_of(long n)
{
    long b = (n + 255) / 256;
    return (unsigned)(b < 16384 ? (b > 0 ? b : 1) : 16384);
}

// S_ij from duidxj (eddyViscosity.F90:8-16), then the model kernel:
// model 0 Smagorinsky |S|, 1 sigma, 2 AMD (camd_x, y, z); nu may be null
template <typename TS, typename TC>
__global__ void __launch_bounds__(256) k_sgs_kernel(const TS *__restrict__ g, TS *__restrict__ S,
                                                   TS *__restrict__ nu, long n, long stride, int model, TC cx, TC cy,
                                                   TC cz)
{
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) {
        TC d[9];
#pragma unroll
        for (int c = 0; c < 9; ++c) d[c] = (TC)g[t + c * stride];
        TC s[6];
        s[0] = d[0];
        s[3] = d[4];
        s[5] = d[8];
        s[1] = (TC)0.5 * (d[1] + d[3]);
        s[2] = (TC)0.5 * (d[2] + d[6]);
        s[4] = (TC)0.5 * (d[5] + d[7]);
        if (S) {
#pragma unroll
            for (int c = 0; c < 6; ++c) S[t + c * stride] = (TS)s[c];
        }
        if (!nu) continue;
        if (model == 0) {
            TC q = s[0] * s[0];
            q += (TC)2 * (s[1] * s[1]);
            q += (TC)2 * (s[2] * s[2]);
            q += s[3] * s[3];
            q += (TC)2 * (s[4] * s[4]);
            q += s[5] * s[5];
            nu[t] = (TS)sqrt((TC)2 * q);
        } else if (model == 1) {
            // sigma.F90:46-96: singular values of the gradient from the
            // invariants of G = A^T A, nu = s3 (s1 - s2) (s2 - s3) / s1^2,
            // in FP64 for every storage type: the closed-form cubic roots
            // lose about half their digits to cancellation
            using TD = double;
            TD e[9];
#pragma unroll
            for (int c = 0; c < 9; ++c) e[c] = (TD)d[c];
            const TD G11 = e[0] * e[0] + e[3] * e[3] + e[6] * e[6];
            const TD G12 = e[0] * e[1] + e[3] * e[4] + e[6] * e[7];
            const TD G13 = e[0] * e[2] + e[3] * e[5] + e[6] * e[8];
            const TD G22 = e[1] * e[1] + e[4] * e[4] + e[7] * e[7];
            const TD G23 = e[1] * e[2] + e[4] * e[5] + e[7] * e[8];
            const TD G33 = e[2] * e[2] + e[5] * e[5] + e[8] * e[8];
            const TD I1 = G11 + G22 + G33, I1sq = I1 * I1, I1cu = I1sq * I1;
            TD I2 = -G11 * G11 - G22 * G22 - G33 * G33;
            I2 = I2 - (TD)2 * G12 * G12 - (TD)2 * G13 * G13;
            I2 = I2 - (TD)2 * G23 * G23;
            I2 = (TD)0.5 * (I2 + I1sq);
            TD I3 = G11 * (G22 * G33 - G23 * G23);
            I3 = I3 + G12 * (G13 * G23 - G12 * G33);
            I3 = I3 + G13 * (G12 * G23 - G22 * G13);
            TD a1 = I1sq / (TD)9 - I2 / (TD)3;
            a1 = a1 > (TD)0 ? a1 : (TD)0;
            const TD a2 = I1cu / (TD)27 - I1 * I2 / (TD)6 + I3 / (TD)2;
            const TD a1s = sqrt(a1);
            TD ca = a2 / (a1 * a1s + (TD)1e-13);
            ca = ca < (TD)1 ? ca : (TD)1;
            ca = ca > (TD)-1 ? ca : (TD)-1;
            const TD a3 = acos(ca) / (TD)3;
            constexpr TD pi3 = 1.04719755119659774615;
            TD s1sq = I1 / (TD)3 + (TD)2 * a1s * cos(a3);
            s1sq = s1sq > (TD)0 ? s1sq : (TD)0;
            const TD s1 = sqrt(s1sq);
            TD s2 = (TD)-2 * a1s * cos(pi3 + a3) + I1 / (TD)3;
            s2 = sqrt(s2 > (TD)0 ? s2 : (TD)0);
            TD s3 = (TD)-2 * a1s * cos(pi3 - a3) + I1 / (TD)3;
            s3 = sqrt(s3 > (TD)0 ? s3 : (TD)0);
            nu[t] = (TS)(s3 * (s1 - s2) * (s2 - s3) / (s1sq + (TD)1e-15));
        } else {
            // AMD.F90:43-70: -(C grad u_i . C grad u_j) S_ij / (grad u : grad u)
            auto row = [&](int i, int j) {
                return (d[3 * i] * cx) * (d[3 * j] * cx) + (d[3 * i + 1] * cy) * (d[3 * j + 1] * cy) +
                       (d[3 * i + 2] * cz) * (d[3 * j + 2] * cz);
            };
            TC num = row(0, 0) * s[0];
            num += row(1, 1) * s[3];
            num += row(2, 2) * s[5];
            num += (TC)2 * row(0, 1) * s[1];
            num += (TC)2 * row(0, 2) * s[2];
            num += (TC)2 * row(1, 2) * s[4];
            TC den = 0;
#pragma unroll
            for (int c = 0; c < 9; ++c) den += d[c] * d[c];
            const TC v = -num / (den + (TC)1e-32);
            nu[t] = (TS)(v > (TC)0 ? v : (TC)0);
        }
    }
}

// Smagorinsky kernel of a strain-rate field (smagorinsky.F90:45-68)
template <typename TS>
__global__ void __launch_bounds__(256) k_smag_S(const TS *S, TS *nu, long n, long stride)
{
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) {
        double s[6];
        for (int c = 0; c < 6; ++c) s[c] = (double)S[t + c * stride];
        double q = s[0] * s[0];
        q += 2.0 * (s[1] * s[1]);
        q += 2.0 * (s[2] * s[2]);
        q += s[3] * s[3];
        q += 2.0 * (s[4] * s[4]);
        q += s[5] * s[5];
        nu[t] = (TS)sqrt(2.0 * q);
    }
}

// nu *= c (global) or c_k (plane k of nxy points)
template <typename TS>
__global__ void __launch_bounds__(256) k_sgs_scale(TS *nu, long nxy, long n, const double *__restrict__ cplane,
                                                  double cglobal)
{
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x)
        nu[t] = (TS)((double)nu[t] * (cplane ? cplane[t / nxy] : cglobal));
}

// out_c = a * b_c (c < ncomp, components `stride` apart); a may be shared
template <typename TS>
__global__ void __launch_bounds__(256) k_mul(const TS *a, const TS *b, TS *out, long n, long stride, int ncomp)
{
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) {
        const double av = (double)a[t];
        for (int c = 0; c < ncomp; ++c) out[t + c * stride] = (TS)(av * (double)b[t + c * stride]);
    }
}

// M_c = F(D S_c) - r2 * Df * Sf_c (standardDynamicProcedure.F90:103-107),
// written over fDS
template <typename TS>
__global__ void __launch_bounds__(256) k_germano_M(TS *fDS, const TS *Df, const TS *Sf, long n, long stride, double r2)
{
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) {
        const double df = (double)Df[t];
        for (int c = 0; c < 6; ++c)
            fDS[t + c * stride] = (TS)((double)fDS[t + c * stride] - r2 * df * (double)Sf[t + c * stride]);
    }
}

// the six products u_i u_j (ordering 11 12 13 22 23 33) of u, v, w
template <typename TS>
__global__ void __launch_bounds__(256) k_uiuj(const TS *u, const TS *v, const TS *w, TS *out, long n, long stride)
{
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) {
        const double a = (double)u[t], b = (double)v[t], c = (double)w[t];
        out[t] = (TS)(a * a);
        out[t + stride] = (TS)(a * b);
        out[t + 2 * stride] = (TS)(a * c);
        out[t + 3 * stride] = (TS)(b * b);
        out[t + 4 * stride] = (TS)(b * c);
        out[t + 5 * stride] = (TS)(c * c);
    }
}

// L_c = F(u_i u_j) - uf_i uf_j, then num = sum L M (off-diagonals twice),
// den = 2 sum M M (off-diagonals twice)  (standardDynamicProcedure.F90:113-163)
template <typename TS>
__global__ void __launch_bounds__(256) k_germano_ratio(const TS *fuu, const TS *uf, const TS *M, double *num,
                                                      double *den, long n, long stride)
{
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) {
        const double a = (double)uf[t], b = (double)uf[t + stride], c = (double)uf[t + 2 * stride];
        const double prod[6] = {a * a, a * b, a * c, b * b, b * c, c * c};
        const double w[6] = {1, 2, 2, 1, 2, 1};
        double nm = 0, dn = 0;
        for (int q = 0; q < 6; ++q) {
            const double L = (double)fuu[t + q * stride] - prod[q];
            const double m = (double)M[t + q * stride];
            nm += w[q] * L * m;
            dn += w[q] * m * m;
        }
        num[t] = nm;
        den[t] = 2.0 * dn;
    }
}

// per-plane sums of two FP64 fields: block (x, k) sums a slice of plane k,
// one atomic-free partial per block; the host-side launcher finishes with
// k_plane_ratio over the partials
__global__ void __launch_bounds__(256) k_plane_partial(const double *a, const double *b, long nxy, int nblk,
                                                      double *pa, double *pb)
{
    __shared__ double sa[256], sb[256];
    const int k = blockIdx.y;
    double ra = 0, rb = 0;
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < nxy; t += (long)nblk * blockDim.x) {
        ra += a[(long)k * nxy + t];
        rb += b[(long)k * nxy + t];
    }
    sa[threadIdx.x] = ra;
    sb[threadIdx.x] = rb;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            sa[threadIdx.x] += sa[threadIdx.x + s];
            sb[threadIdx.x] += sb[threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        pa[(long)k * nblk + blockIdx.x] = sa[0];
        pb[(long)k * nblk + blockIdx.x] = sb[0];
    }
}

// lambda_k = max(<num>_k, 0) / (<den>_k + 1e-18), <.> = plane sum * meanfact
// (planarAverage_and_TakeRatio, standardDynamicProcedure.F90:222-250)
__global__ void k_plane_ratio(const double *pa, const double *pb, int nblk, int nz, double meanfact, double *lam)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nz) return;
    double sa = 0, sb = 0;
    for (int i = 0; i < nblk; ++i) {
        sa += pa[(long)k * nblk + i];
        sb += pb[(long)k * nblk + i];
    }
    const double m = sa * meanfact;
    lam[k] = (m > 0 ? m : 0.0) / (sb * meanfact + 1.0e-18);
}

template <typename TS, typename TC>
static hipError_t sgs_kernel_t(const void *g, void *S, void *nu, long n, long stride, int model, double cx,
                               double cy, double cz, hipStream_t st)
{
    hipLaunchKernelGGL((k_sgs_kernel<TS, TC>), dim3(grid_of(n)), dim3(256), 0, st, (const TS *)g, (TS *)S, (TS *)nu,
                       n, stride, model, (TC)cx, (TC)cy, (TC)cz);
    return hipGetLastError();
}

hipError_t launch_sgs_kernel(int mode, const void *g, void *S, void *nu, long n, long stride, int model, double cx,
                             double cy, double cz, hipStream_t st)
{
    switch (mode) {
    case 0: return sgs_kernel_t<double, double>(g, S, nu, n, stride, model, cx, cy, cz, st);
    case 1: return sgs_kernel_t<float, float>(g, S, nu, n, stride, model, cx, cy, cz, st);
    case 2: return sgs_kernel_t<float, double>(g, S, nu, n, stride, model, cx, cy, cz, st);
    }
    return hipErrorInvalidValue;
}

hipError_t launch_smag_from_S(bool dbl, const void *S, void *nu, long n, long stride, hipStream_t st)
{
    if (dbl)
        hipLaunchKernelGGL((k_smag_S<double>), dim3(grid_of(n)), dim3(256), 0, st, (const double *)S, (double *)nu, n,
                           stride);
    else
        hipLaunchKernelGGL((k_smag_S<float>), dim3(grid_of(n)), dim3(256), 0, st, (const float *)S, (float *)nu, n,
                           stride);
    return hipGetLastError();
}

hipError_t launch_sgs_scale(bool dbl, void *nu, long nxy, long n, const double *cplane, double cglobal,
                            hipStream_t st)
{
    if (dbl)
        hipLaunchKernelGGL((k_sgs_scale<double>), dim3(grid_of(n)), dim3(256), 0, st, (double *)nu, nxy, n, cplane,
                           cglobal);
    else
        hipLaunchKernelGGL((k_sgs_scale<float>), dim3(grid_of(n)), dim3(256), 0, st, (float *)nu, nxy, n, cplane,
                           cglobal);
    return hipGetLastError();
}

hipError_t launch_mul(bool dbl, const void *a, const void *b, void *out, long n, long stride, int ncomp,
                      hipStream_t st)
{
    if (dbl)
        hipLaunchKernelGGL((k_mul<double>), dim3(grid_of(n)), dim3(256), 0, st, (const double *)a, (const double *)b,
                           (double *)out, n, stride, ncomp);
    else
        hipLaunchKernelGGL((k_mul<float>), dim3(grid_of(n)), dim3(256), 0, st, (const float *)a, (const float *)b,
                           (float *)out, n, stride, ncomp);
    return hipGetLastError();
}

hipError_t launch_germano_M(bool dbl, void *fDS, const void *Df, const void *Sf, long n, long stride, double r2,
                            hipStream_t st)
{
    if (dbl)
        hipLaunchKernelGGL((k_germano_M<double>), dim3(grid_of(n)), dim3(256), 0, st, (double *)fDS,
                           (const double *)Df, (const double *)Sf, n, stride, r2);
    else
        hipLaunchKernelGGL((k_germano_M<float>), dim3(grid_of(n)), dim3(256), 0, st, (float *)fDS, (const float *)Df,
                           (const float *)Sf, n, stride, r2);
    return hipGetLastError();
}

hipError_t launch_uiuj(bool dbl, const void *u, const void *v, const void *w, void *out, long n, long stride,
                       hipStream_t st)
{
    if (dbl)
        hipLaunchKernelGGL((k_uiuj<double>), dim3(grid_of(n)), dim3(256), 0, st, (const double *)u, (const double *)v,
                           (const double *)w, (double *)out, n, stride);
    else
        hipLaunchKernelGGL((k_uiuj<float>), dim3(grid_of(n)), dim3(256), 0, st, (const float *)u, (const float *)v,
                           (const float *)w, (float *)out, n, stride);
    return hipGetLastError();
}

hipError_t launch_germano_ratio(bool dbl, const void *fuu, const void *uf, const void *M, double *num, double *den,
                                long n, long stride, hipStream_t st)
{
    if (dbl)
        hipLaunchKernelGGL((k_germano_ratio<double>), dim3(grid_of(n)), dim3(256), 0, st, (const double *)fuu,
                           (const double *)uf, (const double *)M, num, den, n, stride);
    else
        hipLaunchKernelGGL((k_germano_ratio<float>), dim3(grid_of(n)), dim3(256), 0, st, (const float *)fuu,
                           (const float *)uf, (const float *)M, num, den, n, stride);
    return hipGetLastError();
}

hipError_t launch_plane_ratio(const double *num, const double *den, long nxy, int nz, double meanfact,
                              double *partial, double *lam, hipStream_t st)
{
    const int nblk = (int)((nxy + 255) / 256 < 64 ? (nxy + 255) / 256 : 64);
    double *pa = partial, *pb = partial + (size_t)nz * nblk;
    hipLaunchKernelGGL(k_plane_partial, dim3(nblk, nz), dim3(256), 0, st, num, den, nxy, nblk, pa, pb);
    hipLaunchKernelGGL(k_plane_ratio, dim3((nz + 63) / 64), dim3(64), 0, st, pa, pb, nblk, nz, meanfact, lam);
    return hipGetLastError();
}

}  // namespace pdo
