// Learned SGS closure: pointwise features, stress and training loss of the
// tensor-basis MLP (no counterpart in the reference; its baseline is the
// eddy-viscosity family of src/incompressible/sgs_models).
//
// From the resolved velocity gradient A_ij = du_i/dx_j (igrid's duidxj
// order, component 3 i + j, components `stride` apart) and the scale
// g = sqrt(mean over the field of |A|^2) the kernels form the normalised
// gradient a = A / g, its strain s = sym(a) - tr(a)/3 I and rotation
// r = skew(a), and
//   features x (10): sqrt tr s^2, sqrt -tr r^2, cbrt tr s^3, cbrt tr r^2 s,
//     (-tr r^2 s^2)^(1/4), signed (tr r^2 s r s^2)^(1/6) (Pope's integrity
//     basis of (s, r), each root taken to the degree of its invariant), the
//     three principal strains of s (descending), and Delta / Delta_0;
//   basis (6 symmetric traceless tensors, Pope 1975):
//     T1 = s, T2 = s r - r s, T3 = s^2 - tr(s^2)/3 I, T4 = r^2 - tr(r^2)/3 I,
//     T5 = r s^2 - s^2 r, T6 = r^2 s + s r^2 - 2/3 tr(s r^2) I;
//   stress tau = Delta^2 g^2 sum_n G_n T_n, G the network's 6 outputs.
// The stress is frame-indifferent: rotating A rotates tau and leaves x and G
// unchanged.  Arithmetic is FP64 per point (full rate on gfx950), except the
// FP32 roots and principal-strain trigonometry of the features.
#include <hip/hip_runtime.h>

#include "pdo_closure.hpp"

namespace pdo {

namespace {

struct M3 {
    double a[3][3];
};

__device__ __forceinline__ M3 mul(const M3 &x, const M3 &y)
{
    M3 z;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) z.a[i][j] = x.a[i][0] * y.a[0][j] + x.a[i][1] * y.a[1][j] + x.a[i][2] * y.a[2][j];
    return z;
}

__device__ __forceinline__ double tr(const M3 &x) { return x.a[0][0] + x.a[1][1] + x.a[2][2]; }

// the six independent components (11 12 13 22 23 33) of a symmetric tensor
__device__ __forceinline__ void sym6(const M3 &x, double (&o)[6])
{
    o[0] = x.a[0][0];
    o[1] = 0.5 * (x.a[0][1] + x.a[1][0]);
    o[2] = 0.5 * (x.a[0][2] + x.a[2][0]);
    o[3] = x.a[1][1];
    o[4] = 0.5 * (x.a[1][2] + x.a[2][1]);
    o[5] = x.a[2][2];
}

template <typename TS>
__device__ __forceinline__ void strain_rotation(const TS *g, long t, long stride, double inv, M3 &s, M3 &r)
{
    double A[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) A[i][j] = inv * (double)g[t + (3 * i + j) * stride];
    const double t3 = (A[0][0] + A[1][1] + A[2][2]) / 3.0;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            s.a[i][j] = 0.5 * (A[i][j] + A[j][i]) - (i == j ? t3 : 0.0);
            r.a[i][j] = 0.5 * (A[i][j] - A[j][i]);
        }
}

// The invariants (traces, where cancellation lives) are FP64; their roots and
// the principal-strain trigonometry are FP32, whose ~1e-7 relative error is
// far below the BF16 rounding the network applies to its inputs.
__device__ void features_of(const M3 &s, const M3 &r, double dratio, float *x)
{
    const M3 s2 = mul(s, s), r2 = mul(r, r);
    const double i1 = tr(s2), i2 = -tr(r2), i3 = tr(mul(s2, s)), i4 = tr(mul(r2, s)), i5 = tr(mul(r2, s2));
    const double i6 = tr(mul(mul(r2, s), mul(r, s2)));
    x[0] = sqrtf(i1 > 0 ? (float)i1 : 0.0f);
    x[1] = sqrtf(i2 > 0 ? (float)i2 : 0.0f);
    x[2] = cbrtf((float)i3);
    x[3] = cbrtf((float)i4);
    x[4] = sqrtf(sqrtf(-i5 > 0 ? (float)-i5 : 0.0f));
    x[5] = copysignf(cbrtf(sqrtf(fabsf((float)i6))), (float)i6);
    // principal strains: roots of l^3 - p l - q = 0, p = tr(s^2)/2, q = det s = tr(s^3)/3
    const double p = 0.5 * i1, q = i3 / 3.0;
    float l0 = 0, l1 = 0, l2 = 0;
    if (p > 1e-300) {
        double c = 1.5 * q / p * sqrt(3.0 / p);
        c = c > 1.0 ? 1.0 : (c < -1.0 ? -1.0 : c);
        const float ph = acosf((float)c) / 3.0f, m = 2.0f * sqrtf((float)(p / 3.0));
        l0 = m * cosf(ph);
        l1 = m * cosf(ph - 2.0943951023931955f);   // 2 pi / 3
        l2 = m * cosf(ph - 4.1887902047863905f);   // 4 pi / 3
    }
    // descending order: l0 >= l1 >= l2 for ph in [0, pi/3]
    x[6] = l0;
    x[7] = l1;
    x[8] = l2;
    x[9] = (float)dratio;
}

// the six basis tensors, components (11 12 13 22 23 33)
__device__ void basis_of(const M3 &s, const M3 &r, double (&T)[6][6])
{
    const M3 s2 = mul(s, s), r2 = mul(r, r), sr = mul(s, r), rs = mul(r, s);
    const M3 rs2 = mul(r, s2), s2r = mul(s2, r), r2s = mul(r2, s), sr2 = mul(s, r2);
    const double ts2 = tr(s2) / 3.0, tr2 = tr(r2) / 3.0, tsr2 = 2.0 * tr(sr2) / 3.0;
    M3 b[6];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const double d = i == j ? 1.0 : 0.0;
            b[0].a[i][j] = s.a[i][j];
            b[1].a[i][j] = sr.a[i][j] - rs.a[i][j];
            b[2].a[i][j] = s2.a[i][j] - ts2 * d;
            b[3].a[i][j] = r2.a[i][j] - tr2 * d;
            b[4].a[i][j] = rs2.a[i][j] - s2r.a[i][j];
            b[5].a[i][j] = r2s.a[i][j] + sr2.a[i][j] - tsr2 * d;
        }
#pragma unroll
    for (int n = 0; n < 6; ++n) sym6(b[n], T[n]);
}

__device__ __forceinline__ double inv_scale(const double *sumsq, long npts)
{
    const double ms = sumsq[0] / (double)npts;
    return ms > 0 ? 1.0 / sqrt(ms) : 0.0;
}

unsigned grid_of(long n)
{
    long b = (n + 255) / 256;
    return (unsigned)(b < 8192 ? (b > 0 ? b : 1) : 8192);
}

}  // namespace

// ---------------------------------------------------------------------------
// sum over the field of |A|^2 (partials per block, then one block in order)
// ---------------------------------------------------------------------------
template <typename TS>
__global__ void __launch_bounds__(256) k_grad_sumsq(const TS *__restrict__ g, long n, long stride,
                                                   double *__restrict__ partial)
{
    __shared__ double red[256];
    double acc = 0;
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) {
#pragma unroll
        for (int c = 0; c < 9; ++c) {
            const double v = (double)g[t + c * stride];
            acc = fma(v, v, acc);
        }
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

__global__ void __launch_bounds__(256) k_sum_partials(const double *__restrict__ partial, int nb,
                                                     double *__restrict__ out)
{
    __shared__ double red[256];
    double acc = 0;
    for (int i = threadIdx.x; i < nb; i += blockDim.x) acc += partial[i];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = red[0];
}

// ---------------------------------------------------------------------------
// features X (n, 10) and, optionally, the normalised gradient a (n, 9)
// ---------------------------------------------------------------------------
template <typename TS>
__global__ void __launch_bounds__(256) k_closure_features(const TS *__restrict__ g, long n, long stride,
                                                         const double *__restrict__ sumsq, long nnorm, double dratio,
                                                         float *__restrict__ X, float *__restrict__ A9)
{
    const double inv = inv_scale(sumsq, nnorm);
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) {
        M3 s, r;
        strain_rotation(g, t, stride, inv, s, r);
        float x[10];
        features_of(s, r, dratio, x);
#pragma unroll
        for (int k = 0; k < 10; ++k) X[t * 10 + k] = x[k];
        if (A9) {
#pragma unroll
            for (int c = 0; c < 9; ++c) A9[t * 9 + c] = (float)(inv * (double)g[t + c * stride]);
        }
    }
}

// ---------------------------------------------------------------------------
// stress of the network outputs G (n, 6).  which = 0: write tau 11 12 22 33
// to tau (4 components, stride apart, cells); 1: tau 13 23 (2, edges);
// 2: all six (11 12 13 22 23 33).  fallback: keep the value already in tau
// wherever the network's dissipation -tau_ij S_ij is negative; the number
// of such points of each block goes to nfall[blockIdx.x].
// ---------------------------------------------------------------------------
template <typename TS>
__global__ void __launch_bounds__(256) k_closure_stress(const TS *__restrict__ g, long n, long stride,
                                                       const double *__restrict__ sumsq, long nnorm, double delta2,
                                                       const float *__restrict__ G, int which, int fallback,
                                                       TS *__restrict__ tau, long tstride,
                                                       unsigned *__restrict__ nfall)
{
    __shared__ unsigned cnt[256];
    const double inv = inv_scale(sumsq, nnorm);
    // tau = Delta^2 g^2 sum G_n T_n(a), a = A / g
    const double scale = inv > 0 ? delta2 / (inv * inv) : 0.0;
    unsigned mine = 0;
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) {
        M3 s, r;
        strain_rotation(g, t, stride, inv, s, r);
        double T[6][6];
        basis_of(s, r, T);
        double tv[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll
        for (int k = 0; k < 6; ++k) {
            const double gk = (double)G[t * 6 + k];
#pragma unroll
            for (int c = 0; c < 6; ++c) tv[c] = fma(gk, T[k][c], tv[c]);
        }
        // dissipation sign of the normalised stress against the strain
        const double pi = -(tv[0] * s.a[0][0] + tv[3] * s.a[1][1] + tv[5] * s.a[2][2] +
                            2.0 * (tv[1] * s.a[0][1] + tv[2] * s.a[0][2] + tv[4] * s.a[1][2]));
        if (fallback && pi < 0) {
            ++mine;
            continue;
        }
        if (which == 0) {
            tau[t] = (TS)(scale * tv[0]);
            tau[t + tstride] = (TS)(scale * tv[1]);
            tau[t + 2 * tstride] = (TS)(scale * tv[3]);
            tau[t + 3 * tstride] = (TS)(scale * tv[5]);
        } else if (which == 1) {
            tau[t] = (TS)(scale * tv[2]);
            tau[t + tstride] = (TS)(scale * tv[4]);
        } else {
#pragma unroll
            for (int c = 0; c < 6; ++c) tau[t + c * tstride] = (TS)(scale * tv[c]);
        }
    }
    if (!nfall) return;
    cnt[threadIdx.x] = mine;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) cnt[threadIdx.x] += cnt[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) nfall[blockIdx.x] = cnt[0];
}

// ---------------------------------------------------------------------------
// training loss: L = sum over the batch and the six components of
// w_c (p_c - t_c)^2 / (9 B), w = (1 2 2 1 2 1) (the tensor's nine entries),
// p = sum_n G_n T_n(a); dG (B, 6) = dL/dG; per-block loss partials
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_closure_loss(const float *__restrict__ G, const float *__restrict__ A9,
                                                     const float *__restrict__ Tt, long B, float *__restrict__ dG,
                                                     double *__restrict__ partial)
{
    __shared__ double red[256];
    const double wc[6] = {1, 2, 2, 1, 2, 1};
    const double nrm = 1.0 / (9.0 * (double)B);
    double acc = 0;
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < B; t += (long)gridDim.x * blockDim.x) {
        M3 s, r;
        strain_rotation(A9, t * 9, 1, 1.0, s, r);
        double T[6][6];
        basis_of(s, r, T);
        double p[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll
        for (int k = 0; k < 6; ++k) {
            const double gk = (double)G[t * 6 + k];
#pragma unroll
            for (int c = 0; c < 6; ++c) p[c] = fma(gk, T[k][c], p[c]);
        }
        double e[6];
#pragma unroll
        for (int c = 0; c < 6; ++c) {
            e[c] = p[c] - (double)Tt[t * 6 + c];
            acc += wc[c] * e[c] * e[c];
        }
#pragma unroll
        for (int k = 0; k < 6; ++k) {
            double d = 0;
#pragma unroll
            for (int c = 0; c < 6; ++c) d = fma(2.0 * wc[c] * e[c], T[k][c], d);
            dG[t * 6 + k] = (float)(d * nrm);
        }
    }
    red[threadIdx.x] = acc * nrm;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
int closure_blocks(long n) { return (int)grid_of(n); }

hipError_t launch_grad_sumsq(bool dbl, const void *g, long n, long stride, double *partial, double *out,
                             hipStream_t st)
{
    const unsigned nb = grid_of(n);
    if (dbl)
        hipLaunchKernelGGL(k_grad_sumsq<double>, dim3(nb), dim3(256), 0, st, (const double *)g, n, stride, partial);
    else
        hipLaunchKernelGGL(k_grad_sumsq<float>, dim3(nb), dim3(256), 0, st, (const float *)g, n, stride, partial);
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(256), 0, st, partial, (int)nb, out);
    return hipGetLastError();
}

hipError_t launch_closure_features(bool dbl, const void *g, long n, long stride, const double *sumsq, long nnorm,
                                   double dratio, float *X, float *A9, hipStream_t st)
{
    const unsigned nb = grid_of(n);
    if (dbl)
        hipLaunchKernelGGL(k_closure_features<double>, dim3(nb), dim3(256), 0, st, (const double *)g, n, stride,
                           sumsq, nnorm, dratio, X, A9);
    else
        hipLaunchKernelGGL(k_closure_features<float>, dim3(nb), dim3(256), 0, st, (const float *)g, n, stride,
                           sumsq, nnorm, dratio, X, A9);
    return hipGetLastError();
}

hipError_t launch_closure_stress(bool dbl, const void *g, long n, long stride, const double *sumsq, long nnorm,
                                 double delta2, const float *G, int which, int fallback, void *tau, long tstride,
                                 unsigned *nfall, hipStream_t st)
{
    const unsigned nb = grid_of(n);
    if (dbl)
        hipLaunchKernelGGL(k_closure_stress<double>, dim3(nb), dim3(256), 0, st, (const double *)g, n, stride, sumsq,
                           nnorm, delta2, G, which, fallback, (double *)tau, tstride, nfall);
    else
        hipLaunchKernelGGL(k_closure_stress<float>, dim3(nb), dim3(256), 0, st, (const float *)g, n, stride, sumsq,
                           nnorm, delta2, G, which, fallback, (float *)tau, tstride, nfall);
    return hipGetLastError();
}

hipError_t launch_closure_loss(const float *G, const float *A9, const float *T, long B, float *dG, double *partial,
                               double *loss, hipStream_t st)
{
    const unsigned nb = grid_of(B);
    hipLaunchKernelGGL(k_closure_loss, dim3(nb), dim3(256), 0, st, G, A9, T, B, dG, partial);
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(256), 0, st, partial, (int)nb, loss);
    return hipGetLastError();
}

}  // namespace pdo
