// Pointwise kernels of the igrid pieces (interp_PrimitiveVars, getRHS_SGS,
// the nonlinear and viscous terms, the Runge-Kutta stage updates).  All of
// them stream their operands once (HBM-bound, grid-stride loops).  mode: 0
// FP64, 1 FP32, 2 FP32 storage with FP64 arithmetic.
#include <hip/hip_runtime.h>

#include "pdo_igrid.hpp"

namespace pdo {

namespace {

template <typename TS> struct C2;
template <> struct C2<double> { using V = double2; };
template <> struct C2<float> { using V = float2; };

unsigned blocks_for(long n)
{
    long b = (n + 255) / 256;
    return (unsigned)(b < 8192 ? (b > 0 ? b : 1) : 8192);
}

#define PDO_GRID_STRIDE(t, n) \
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < (n); t += (long)gridDim.x * blockDim.x)

// plane 0 of a (nxh, ny, nz+1) spectrum set to zero, its mean mode to v
template <typename TS>
__global__ void __launch_bounds__(256) k_surface_plane(TS *f, long nplane, double v)
{
    using V = typename C2<TS>::V;
    V *z = (V *)f;
    PDO_GRID_STRIDE(t, nplane) z[t] = V{(TS)(t == 0 ? v : 0.0), (TS)0};
}

// tau_ij = -2 nu S_ij: cells 11, 12, 22, 33 from (nuC, SC), edges 13, 23
// from (nuE, SE); S holds (11, 12, 13, 22, 23, 33) as in get_Sij_from_duidxj
template <typename TS, typename TC>
__global__ void __launch_bounds__(256) k_tau(const TS *nuC, const TS *SC, const TS *nuE, const TS *SE, TS *tauC,
                                           TS *tauE, long nC, long nE)
{
    PDO_GRID_STRIDE(t, nC)
    {
        const TC m = (TC)-2 * (TC)nuC[t];
        tauC[t] = (TS)(m * (TC)SC[t]);
        tauC[t + nC] = (TS)(m * (TC)SC[t + nC]);
        tauC[t + 2 * nC] = (TS)(m * (TC)SC[t + 3 * nC]);
        tauC[t + 3 * nC] = (TS)(m * (TC)SC[t + 5 * nC]);
    }
    PDO_GRID_STRIDE(t, nE)
    {
        const TC m = (TC)-2 * (TC)nuE[t];
        tauE[t] = (TS)(m * (TC)SE[t + 2 * nE]);
        tauE[t + nE] = (TS)(m * (TC)SE[t + 4 * nE]);
    }
}

template <typename TS>
__global__ void __launch_bounds__(256) k_clip_negative(TS *a, long n)
{
    PDO_GRID_STRIDE(t, n) if (a[t] < (TS)0) a[t] = (TS)0;
}

// rhs -= i k f (k = kv[index along dim]) or, kv == NULL, rhs -= f; complex
// arrays of (nxh, ny, planes)
template <typename TS, typename TC>
__global__ void __launch_bounds__(256) k_sub(TS *rhs, const TS *f, const double *__restrict__ kv, int nxh, int ny,
                                           long n, int dim)
{
    using V = typename C2<TS>::V;
    V *r = (V *)rhs;
    const V *s = (const V *)f;
    PDO_GRID_STRIDE(t, n)
    {
        const V a = s[t];
        V o = r[t];
        if (kv) {
            const long ij = dim == 1 ? t % nxh : (t / nxh) % ny;
            const TC k = (TC)kv[ij];
            // i k (a.x + i a.y) = -k a.y + i k a.x
            o.x = (TS)((TC)o.x - (-k * (TC)a.y));
            o.y = (TS)((TC)o.y - k * (TC)a.x);
        } else {
            o.x = (TS)((TC)o.x - (TC)a.x);
            o.y = (TS)((TC)o.y - (TC)a.y);
        }
        r[t] = o;
    }
}

// out = (a1 - b1) c1 [+ (a2 - b2) c2]: the rotational-form products of
// AddNonLinearTerm_Rot (igrid.F90:1497-1566)
template <typename TS, typename TC>
__global__ void __launch_bounds__(256) k_cross(const TS *a1, const TS *b1, const TS *c1, const TS *a2, const TS *b2,
                                             const TS *c2, TS *out, long n)
{
    PDO_GRID_STRIDE(t, n)
    {
        TC v = ((TC)a1[t] - (TC)b1[t]) * (TC)c1[t];
        if (a2) v += ((TC)a2[t] - (TC)b2[t]) * (TC)c2[t];
        out[t] = (TS)v;
    }
}

// r += s (real words)
template <typename TS, typename TC>
__global__ void __launch_bounds__(256) k_add(TS *r, const TS *s, long n)
{
    PDO_GRID_STRIDE(t, n) r[t] = (TS)((TC)r[t] + (TC)s[t]);
}

// addViscousTerm (igrid.F90:1914-1959): rhs += (1/Re) (-(k1^2 + k2^2) f + d2f)
template <typename TS, typename TC>
__global__ void __launch_bounds__(256) k_visc(TS *rhs, const TS *f, const TS *d2, const double *__restrict__ k1,
                                            const double *__restrict__ k2, int nxh, int ny, long n, double invRe)
{
    using V = typename C2<TS>::V;
    V *r = (V *)rhs;
    const V *a = (const V *)f;
    const V *b = (const V *)d2;
    PDO_GRID_STRIDE(t, n)
    {
        const long i = t % nxh, j = (t / nxh) % ny;
        const TC ks = (TC)(k1[i] * k1[i] + k2[j] * k2[j]);
        const V x = a[t], y = b[t];
        V o = r[t];
        o.x = (TS)((TC)o.x + (TC)invRe * (-ks * (TC)x.x + (TC)y.x));
        o.y = (TS)((TC)o.y + (TC)invRe * (-ks * (TC)x.y + (TC)y.y));
        r[t] = o;
    }
}

// out = sum_q c[q] x[q], terms added left to right (real words); NULL
// operands end the list
struct LinComb {
    const void *x[5];
    double c[5];
};
template <typename TS, typename TC>
__global__ void __launch_bounds__(256) k_lincomb(TS *out, LinComb L, long n)
{
    PDO_GRID_STRIDE(t, n)
    {
        TC v = (TC)L.c[0] * (TC)((const TS *)L.x[0])[t];
#pragma unroll
        for (int q = 1; q < 5; ++q)
            if (L.x[q]) v += (TC)L.c[q] * (TC)((const TS *)L.x[q])[t];
        out[t] = (TS)v;
    }
}

// partial maxima of |u|/dx + |v|/dy + |w|/dz over the blocks of the grid
template <typename TS>
__global__ void __launch_bounds__(256) k_cfl(const TS *u, const TS *v, const TS *w, long n, double idx, double idy,
                                           double idz, double *partial)
{
    __shared__ double red[256];
    double m = 0.0;
    PDO_GRID_STRIDE(t, n)
    {
        const double s = fabs((double)u[t] * idx) + fabs((double)v[t] * idy) + fabs((double)w[t] * idz);
        m = s > m ? s : m;
    }
    red[threadIdx.x] = m;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + o]);
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// ---------------------------------------------------------------------------
// HIT_shell_forcing (forcingIsotropic.F90:123-314) on a periodic-z grid: for
// each of nwaves random wave vectors (kx, ky, kz) >= 0 the 3D Fourier
// coefficients of (uhat, vhat, what) at that mode (a z transform of the
// 2D spectra at one frequency; w shifted from edges to cells), the forcing
// fac conj(u_k) with fac = normfact eps / (|u_k|^2 + |v_k|^2 + |w_k|^2 +
// 1e-14) / nwaves, and its inverse z transform (1/nz; w shifted back to the
// edges, plane nz repeating plane 0) added to the right-hand sides.  One
// workgroup takes the modes in order, so repeated modes add up as in the
// reference.  The random draws: splitmix64 of (seed, draw counter, wave,
// variable), uniform in [0, 1) from the top 53 bits.
// ---------------------------------------------------------------------------
__device__ __forceinline__ unsigned long long splitmix64(unsigned long long x)
{
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

__device__ __forceinline__ double uniform01(unsigned long long seed, unsigned long long draw, int ik, int q)
{
    const unsigned long long key = splitmix64(splitmix64(seed + draw) + 3ull * (unsigned long long)ik + (unsigned)q);
    return (double)(key >> 11) * 0x1.0p-53;
}

template <typename TS>
__global__ void __launch_bounds__(256) k_hit_forcing(const TS *uhat, const TS *vhat, const TS *what, TS *urhs,
                                                   TS *vrhs, TS *wrhs, HitForcingArgs F)
{
    using V = typename C2<TS>::V;
    __shared__ double red[6][256];
    __shared__ double sfac[7];
    const int tid = threadIdx.x;
    const long plane = (long)F.nxh * F.ny;
    const int nz = F.nz;
    for (int ik = 0; ik < F.nwaves; ++ik) {
        // pick_random_wavenumbers (:133-156)
        const double kabs = F.kmin + (F.kmax - F.kmin) * uniform01(F.seed, F.draw, ik, 0);
        const double zeta = -1.0 + 2.0 * uniform01(F.seed, F.draw, ik, 1);
        const double theta = 2.0 * M_PI * uniform01(F.seed, F.draw, ik, 2);
        const double sz = sqrt(1.0 - zeta * zeta);
        const int wx = (int)ceil(fabs(kabs * sz * cos(theta)));
        const int wy = (int)ceil(fabs(kabs * sz * sin(theta)));
        const int wz = (int)ceil(fabs(kabs * zeta));
        // the host keeps ceil(kmax) below nx/2, ny and nz
        const long col = (long)wx + (long)F.nxh * wy;
        double acc[6] = {0, 0, 0, 0, 0, 0};
        for (int k = tid; k < nz; k += 256) {
            double sn, cs;
            sincospi(-2.0 * (double)((long)wz * k % nz) / (double)nz, &sn, &cs);
            const V a = ((const V *)uhat)[col + (long)k * plane];
            const V b = ((const V *)vhat)[col + (long)k * plane];
            const V c = ((const V *)what)[col + (long)k * plane];
            acc[0] += (double)a.x * cs - (double)a.y * sn;
            acc[1] += (double)a.x * sn + (double)a.y * cs;
            acc[2] += (double)b.x * cs - (double)b.y * sn;
            acc[3] += (double)b.x * sn + (double)b.y * cs;
            acc[4] += (double)c.x * cs - (double)c.y * sn;
            acc[5] += (double)c.x * sn + (double)c.y * cs;
        }
        for (int q = 0; q < 6; ++q) red[q][tid] = acc[q];
        __syncthreads();
        for (int o = 128; o > 0; o >>= 1) {
            if (tid < o)
                for (int q = 0; q < 6; ++q) red[q][tid] += red[q][tid + o];
            __syncthreads();
        }
        if (tid == 0) {
            // w: shiftz_E2C, times exp(i k3 dz / 2)
            const double k3 = F.k3[wz];
            double es, ec;
            sincos(0.5 * k3 * F.dz, &es, &ec);
            const double wr = red[4][0] * ec - red[5][0] * es, wi = red[4][0] * es + red[5][0] * ec;
            const double den = red[0][0] * red[0][0] + red[1][0] * red[1][0] + red[2][0] * red[2][0] +
                               red[3][0] * red[3][0] + wr * wr + wi * wi + 1.0e-14;
            const double fac = F.normfact * F.eps / den / (double)F.nwaves;
            // forcing coefficients fac conj(.) ; w shifted back by exp(-i k3 dz / 2)
            sfac[0] = fac * red[0][0];
            sfac[1] = -fac * red[1][0];
            sfac[2] = fac * red[2][0];
            sfac[3] = -fac * red[3][0];
            const double fr = fac * wr, fi = -fac * wi;
            sfac[4] = fr * ec + fi * es;
            sfac[5] = -fr * es + fi * ec;
        }
        __syncthreads();
        // inverse z transform (1/nz) of the single mode into every plane
        const double inz = 1.0 / (double)nz;
        for (int k = tid; k <= nz; k += 256) {
            const int kk = k == nz ? 0 : k;
            double sn, cs;
            sincospi(2.0 * (double)((long)wz * kk % nz) / (double)nz, &sn, &cs);
            const long idx = col + (long)k * plane;
            if (k < nz) {
                V *ru = (V *)urhs + idx;
                V *rv = (V *)vrhs + idx;
                V o = *ru;
                o.x = (TS)((double)o.x + (sfac[0] * cs - sfac[1] * sn) * inz);
                o.y = (TS)((double)o.y + (sfac[0] * sn + sfac[1] * cs) * inz);
                *ru = o;
                o = *rv;
                o.x = (TS)((double)o.x + (sfac[2] * cs - sfac[3] * sn) * inz);
                o.y = (TS)((double)o.y + (sfac[2] * sn + sfac[3] * cs) * inz);
                *rv = o;
            }
            V *rw = (V *)wrhs + idx;
            V o = *rw;
            o.x = (TS)((double)o.x + (sfac[4] * cs - sfac[5] * sn) * inz);
            o.y = (TS)((double)o.y + (sfac[4] * sn + sfac[5] * cs) * inz);
            *rw = o;
        }
        __syncthreads();
    }
}

template <typename TS, typename TC>
hipError_t surface_t(void *f, long nplane, double v, hipStream_t st)
{
    hipLaunchKernelGGL((k_surface_plane<TS>), dim3(blocks_for(nplane)), dim3(256), 0, st, (TS *)f, nplane, v);
    return hipGetLastError();
}

}  // namespace

hipError_t launch_surface_plane(int mode, void *f, long nplane, double v, hipStream_t st)
{
    if (mode == 0) return surface_t<double, double>(f, nplane, v, st);
    return surface_t<float, float>(f, nplane, v, st);
}

hipError_t launch_tau(int mode, const void *nuC, const void *SC, const void *nuE, const void *SE, void *tauC,
                      void *tauE, long nC, long nE, hipStream_t st)
{
    const dim3 g(blocks_for(nC > nE ? nC : nE)), b(256);
    switch (mode) {
    case 0:
        hipLaunchKernelGGL((k_tau<double, double>), g, b, 0, st, (const double *)nuC, (const double *)SC,
                           (const double *)nuE, (const double *)SE, (double *)tauC, (double *)tauE, nC, nE);
        break;
    case 1:
        hipLaunchKernelGGL((k_tau<float, float>), g, b, 0, st, (const float *)nuC, (const float *)SC,
                           (const float *)nuE, (const float *)SE, (float *)tauC, (float *)tauE, nC, nE);
        break;
    default:
        hipLaunchKernelGGL((k_tau<float, double>), g, b, 0, st, (const float *)nuC, (const float *)SC,
                           (const float *)nuE, (const float *)SE, (float *)tauC, (float *)tauE, nC, nE);
    }
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Moeng wall model: first-plane sums, the surface quantities, the embedded
// wall stresses
// ---------------------------------------------------------------------------
namespace {

template <typename TS>
__global__ void __launch_bounds__(256) k_wall_sums(const TS *u, const TS *v, long nxy, double *part)
{
    __shared__ double ss[256], su[256];
    double a = 0, b = 0;
    for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < nxy; i += (long)gridDim.x * 256) {
        const double x = (double)u[i], y = (double)v[i];
        a += sqrt(x * x + y * y);
        b += x;
    }
    ss[threadIdx.x] = a;
    su[threadIdx.x] = b;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            ss[threadIdx.x] += ss[threadIdx.x + w];
            su[threadIdx.x] += su[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = ss[0];
        part[2 * blockIdx.x + 1] = su[0];
    }
}

__global__ void __launch_bounds__(256) k_wall_finalize(const double *part, int nb, double meanfact, double kappa,
                                                      double lg, double *wm)
{
    __shared__ double ss[256], su[256];
    double a = 0, b = 0;
    for (int i = threadIdx.x; i < nb; i += 256) {
        a += part[2 * i];
        b += part[2 * i + 1];
    }
    ss[threadIdx.x] = a;
    su[threadIdx.x] = b;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            ss[threadIdx.x] += ss[threadIdx.x + w];
            su[threadIdx.x] += su[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const double Uspmn = ss[0] * meanfact;
        const double ustar = Uspmn * kappa / lg;
        wm[0] = Uspmn;
        wm[1] = ustar;
        wm[2] = -ustar * ustar / (Uspmn + 1.0e-13);
        wm[3] = su[0] * meanfact;
    }
}

template <typename TS>
__global__ void __launch_bounds__(256) k_wall_embed(const TS *u, const TS *v, TS *tauE, long nxy, long nE,
                                                   const double *wm)
{
    const double M = wm[2];
    for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < nxy; i += (long)gridDim.x * 256) {
        tauE[i] = (TS)(M * (double)u[i]);
        tauE[nE + i] = (TS)(M * (double)v[i]);
    }
}

template <typename TS>
hipError_t wall_t(const void *u, const void *v, void *tauE, long nxy, long nE, double *part, int nb, double kappa,
                  double lg, double *wm, hipStream_t st)
{
    hipLaunchKernelGGL((k_wall_sums<TS>), dim3(nb), dim3(256), 0, st, (const TS *)u, (const TS *)v, nxy, part);
    hipLaunchKernelGGL(k_wall_finalize, dim3(1), dim3(256), 0, st, (const double *)part, nb, 1.0 / (double)nxy,
                       kappa, lg, wm);
    hipLaunchKernelGGL((k_wall_embed<TS>), dim3(blocks_for(nxy)), dim3(256), 0, st, (const TS *)u, (const TS *)v,
                       (TS *)tauE, nxy, nE, (const double *)wm);
    return hipGetLastError();
}

}  // namespace

hipError_t launch_wall_model(int mode, const void *u, const void *v, void *tauE, long nxy, long nE, double *part,
                             int nb, double kappa, double lg, double *wm, hipStream_t st)
{
    if (mode == 0) return wall_t<double>(u, v, tauE, nxy, nE, part, nb, kappa, lg, wm, st);
    return wall_t<float>(u, v, tauE, nxy, nE, part, nb, kappa, lg, wm, st);
}

namespace {
template <typename TS>
__global__ void __launch_bounds__(256) k_mean_add(TS *f, long nplane, int nz, double v)
{
    for (int k = blockIdx.x * 256 + threadIdx.x; k < nz; k += gridDim.x * 256) {
        TS *m = f + 2 * (long)k * nplane;
        *m = (TS)((double)*m + v);
    }
}
}  // namespace

hipError_t launch_mean_add(int mode, void *f, long nplane, int nz, double v, hipStream_t st)
{
    const dim3 g((unsigned)((nz + 255) / 256));
    if (mode == 0) hipLaunchKernelGGL((k_mean_add<double>), g, dim3(256), 0, st, (double *)f, nplane, nz, v);
    else hipLaunchKernelGGL((k_mean_add<float>), g, dim3(256), 0, st, (float *)f, nplane, nz, v);
    return hipGetLastError();
}

hipError_t launch_clip_negative(int mode, void *a, long n, hipStream_t st)
{
    if (mode == 0)
        hipLaunchKernelGGL((k_clip_negative<double>), dim3(blocks_for(n)), dim3(256), 0, st, (double *)a, n);
    else
        hipLaunchKernelGGL((k_clip_negative<float>), dim3(blocks_for(n)), dim3(256), 0, st, (float *)a, n);
    return hipGetLastError();
}

hipError_t launch_sub(int mode, void *rhs, const void *f, const double *kv, int nxh, int ny, long n, int dim,
                      hipStream_t st)
{
    const dim3 g(blocks_for(n)), b(256);
    switch (mode) {
    case 0:
        hipLaunchKernelGGL((k_sub<double, double>), g, b, 0, st, (double *)rhs, (const double *)f, kv, nxh, ny, n,
                           dim);
        break;
    case 1:
        hipLaunchKernelGGL((k_sub<float, float>), g, b, 0, st, (float *)rhs, (const float *)f, kv, nxh, ny, n, dim);
        break;
    default:
        hipLaunchKernelGGL((k_sub<float, double>), g, b, 0, st, (float *)rhs, (const float *)f, kv, nxh, ny, n, dim);
    }
    return hipGetLastError();
}

hipError_t launch_cross(int mode, const void *a1, const void *b1, const void *c1, const void *a2, const void *b2,
                        const void *c2, void *out, long n, hipStream_t st)
{
    const dim3 g(blocks_for(n)), b(256);
    switch (mode) {
    case 0:
        hipLaunchKernelGGL((k_cross<double, double>), g, b, 0, st, (const double *)a1, (const double *)b1,
                           (const double *)c1, (const double *)a2, (const double *)b2, (const double *)c2,
                           (double *)out, n);
        break;
    case 1:
        hipLaunchKernelGGL((k_cross<float, float>), g, b, 0, st, (const float *)a1, (const float *)b1,
                           (const float *)c1, (const float *)a2, (const float *)b2, (const float *)c2, (float *)out,
                           n);
        break;
    default:
        hipLaunchKernelGGL((k_cross<float, double>), g, b, 0, st, (const float *)a1, (const float *)b1,
                           (const float *)c1, (const float *)a2, (const float *)b2, (const float *)c2, (float *)out,
                           n);
    }
    return hipGetLastError();
}

hipError_t launch_add(int mode, void *r, const void *s, long n, hipStream_t st)
{
    const dim3 g(blocks_for(n)), b(256);
    if (mode == 0)
        hipLaunchKernelGGL((k_add<double, double>), g, b, 0, st, (double *)r, (const double *)s, n);
    else
        hipLaunchKernelGGL((k_add<float, float>), g, b, 0, st, (float *)r, (const float *)s, n);
    return hipGetLastError();
}

hipError_t launch_visc(int mode, void *rhs, const void *f, const void *d2, const double *k1, const double *k2, int nxh,
                       int ny, long n, double invRe, hipStream_t st)
{
    const dim3 g(blocks_for(n)), b(256);
    switch (mode) {
    case 0:
        hipLaunchKernelGGL((k_visc<double, double>), g, b, 0, st, (double *)rhs, (const double *)f,
                           (const double *)d2, k1, k2, nxh, ny, n, invRe);
        break;
    case 1:
        hipLaunchKernelGGL((k_visc<float, float>), g, b, 0, st, (float *)rhs, (const float *)f, (const float *)d2, k1,
                           k2, nxh, ny, n, invRe);
        break;
    default:
        hipLaunchKernelGGL((k_visc<float, double>), g, b, 0, st, (float *)rhs, (const float *)f, (const float *)d2, k1,
                           k2, nxh, ny, n, invRe);
    }
    return hipGetLastError();
}

hipError_t launch_lincomb(int mode, void *out, const void *const *x, const double *c, int nterms, long n,
                          hipStream_t st)
{
    if (nterms < 1 || nterms > 5) return hipErrorInvalidValue;
    LinComb L{};
    for (int q = 0; q < nterms; ++q) { L.x[q] = x[q]; L.c[q] = c[q]; }
    const dim3 g(blocks_for(n)), b(256);
    switch (mode) {
    case 0: hipLaunchKernelGGL((k_lincomb<double, double>), g, b, 0, st, (double *)out, L, n); break;
    case 1: hipLaunchKernelGGL((k_lincomb<float, float>), g, b, 0, st, (float *)out, L, n); break;
    default: hipLaunchKernelGGL((k_lincomb<float, double>), g, b, 0, st, (float *)out, L, n);
    }
    return hipGetLastError();
}

hipError_t launch_cfl(int mode, const void *u, const void *v, const void *w, long n, double idx, double idy,
                      double idz, double *partial, int nblocks, hipStream_t st)
{
    if (mode == 0)
        hipLaunchKernelGGL((k_cfl<double>), dim3(nblocks), dim3(256), 0, st, (const double *)u, (const double *)v,
                           (const double *)w, n, idx, idy, idz, partial);
    else
        hipLaunchKernelGGL((k_cfl<float>), dim3(nblocks), dim3(256), 0, st, (const float *)u, (const float *)v,
                           (const float *)w, n, idx, idy, idz, partial);
    return hipGetLastError();
}

hipError_t launch_hit_forcing(int mode, const void *uhat, const void *vhat, const void *what, void *urhs, void *vrhs,
                              void *wrhs, const HitForcingArgs &F, hipStream_t st)
{
    if (mode == 0)
        hipLaunchKernelGGL((k_hit_forcing<double>), dim3(1), dim3(256), 0, st, (const double *)uhat,
                           (const double *)vhat, (const double *)what, (double *)urhs, (double *)vrhs,
                           (double *)wrhs, F);
    else
        hipLaunchKernelGGL((k_hit_forcing<float>), dim3(1), dim3(256), 0, st, (const float *)uhat, (const float *)vhat,
                           (const float *)what, (float *)urhs, (float *)vrhs, (float *)wrhs, F);
    return hipGetLastError();
}

}  // namespace pdo
