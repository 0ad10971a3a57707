// Explicit symmetric 9-point filters (gaussianstuff, lstsqstuff: the
// Gaussian filter of width 4 dx and the least-squares filter) along x, y or
// z of a field f(n, na, nb) / (na, n, nb) / (na, nb, n).
//
// One thread per output point, threads running along the contiguous index
// of the array, so the nine loads of every row are coalesced in each
// direction and the overlapping windows of neighbouring rows hit the cache.
// Interior rows use the symmetric stencil c0 f_j + sum_q c_q (f_{j+q} +
// f_{j-q}); periodic rows wrap; the first and last NB rows of a
// non-periodic pencil take their weights from a closure table (NB rows of
// WB weights against the first / last WB points).  HBM-bound: one read and
// one write per point.
#include "pdo_explicit.hpp"

namespace pdo {

template <typename TS, typename TA>
__global__ void __launch_bounds__(256) k_explicit(ExplicitArgs E, const TS *__restrict__ f, TS *__restrict__ out)
{
    const long total = (long)E.n * E.na * E.nb;
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (long)gridDim.x * blockDim.x) {
        // row r of the pencil through element e, and the element stride of rows
        long r, st;
        if (E.dir == 1) { r = e % E.n; st = 1; }
        else if (E.dir == 2) { r = (e / E.na) % E.n; st = E.na; }
        else { r = e / ((long)E.na * E.nb); st = (long)E.na * E.nb; }
        const TS *p = f + (e - r * st);  // row 0 of this pencil
        TA acc;
        if (E.periodic) {
            acc = (TA)E.c[0] * (TA)p[r * st];
#pragma unroll
            for (int q = 1; q <= EXPL_R; ++q) {
                long a = r + q, b = r - q;
                a -= (a >= E.n) ? E.n : 0;
                b += (b < 0) ? E.n : 0;
                acc += (TA)E.c[q] * ((TA)p[a * st] + (TA)p[b * st]);
            }
        } else if (r < EXPL_NB) {
            acc = 0;
#pragma unroll
            for (int q = 0; q < EXPL_WB; ++q) acc += (TA)E.left[r * EXPL_WB + q] * (TA)p[q * st];
        } else if (r >= E.n - EXPL_NB) {
            const long j = r - (E.n - EXPL_NB);
            const TS *pe = p + (E.n - EXPL_WB) * st;
            acc = 0;
#pragma unroll
            for (int q = 0; q < EXPL_WB; ++q) acc += (TA)E.right[j * EXPL_WB + q] * (TA)pe[q * st];
        } else {
            acc = (TA)E.c[0] * (TA)p[r * st];
#pragma unroll
            for (int q = 1; q <= EXPL_R; ++q) acc += (TA)E.c[q] * ((TA)p[(r + q) * st] + (TA)p[(r - q) * st]);
        }
        out[e] = (TS)acc;
    }
}

hipError_t launch_explicit(bool dbl, const ExplicitArgs &E, const void *f, void *out, hipStream_t st)
{
    const long total = (long)E.n * E.na * E.nb;
    if (total <= 0) return hipSuccess;
    long blocks = (total + 255) / 256;
    if (blocks > 1 << 20) blocks = 1 << 20;
    // FP32 storage is filtered in FP64 arithmetic: the stencil is cheap next
    // to the memory traffic, and every FP32 precision mode gives one result
    if (dbl)
        hipLaunchKernelGGL((k_explicit<double, double>), dim3((unsigned)blocks), dim3(256), 0, st, E,
                           (const double *)f, (double *)out);
    else
        hipLaunchKernelGGL((k_explicit<float, double>), dim3((unsigned)blocks), dim3(256), 0, st, E,
                           (const float *)f, (float *)out);
    return hipGetLastError();
}

// out = a + s * b (the sums of the operators module)
template <typename T>
__global__ void __launch_bounds__(256) k_axpy(T *out, const T *a, const T *b, double s, long n)
{
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        out[i] = (T)((double)a[i] + s * (double)b[i]);
}

hipError_t launch_axpy(bool dbl, void *out, const void *a, const void *b, double s, long n, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    long blocks = (n + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    if (dbl)
        hipLaunchKernelGGL((k_axpy<double>), dim3((unsigned)blocks), dim3(256), 0, st, (double *)out,
                           (const double *)a, (const double *)b, s, n);
    else
        hipLaunchKernelGGL((k_axpy<float>), dim3((unsigned)blocks), dim3(256), 0, st, (float *)out,
                           (const float *)a, (const float *)b, s, n);
    return hipGetLastError();
}

}  // namespace pdo
