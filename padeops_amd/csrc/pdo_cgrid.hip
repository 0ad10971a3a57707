// Pointwise kernels of the cgrid compressible right-hand side for one species
// of ideal gas (src/compressible/cgrid.F90, IdealGasEOS.F90,
// PowerLawViscosity.F90).  Every kernel is elementwise over n points, FP64
// arithmetic for both storage types, grid-stride, HBM-bound.
#include <hip/hip_runtime.h>

#include "pdo_cgrid.hpp"

namespace pdo {

namespace {

unsigned grid_of(long n)
{
    long b = (n + 255) / 256;
    return (unsigned)(b < 16384 ? (b > 0 ? b : 1) : 16384);
}

#define PDO_CG_LOOP(n) for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < (n); t += (long)gridDim.x * blockDim.x)

template <typename TS> __device__ __forceinline__ double ld(const TS *p, long i) { return (double)p[i]; }

// get_primitive (cgrid.F90:763-797) with the ideal gas (IdealGasEOS.F90:45-70):
// W = (rho, rho u, rho v, rho w, TE), n apart; prim = (rho, u, v, w, e, p, T)
template <typename TS>
__global__ void __launch_bounds__(256) k_primitive(const TS *W, TS *prim, long n, double gm1, double onebyR)
{
    PDO_CG_LOOP(n)
    {
        const double rho = ld(W, t), onebyrho = 1.0 / rho;
        const double u = ld(W, t + n) * onebyrho, v = ld(W, t + 2 * n) * onebyrho, w = ld(W, t + 3 * n) * onebyrho;
        const double e = (ld(W, t + 4 * n) * onebyrho) - 0.5 * (u * u + v * v + w * w);
        prim[t] = (TS)rho;
        prim[t + n] = (TS)u;
        prim[t + 2 * n] = (TS)v;
        prim[t + 3 * n] = (TS)w;
        prim[t + 4 * n] = (TS)e;
        prim[t + 5 * n] = (TS)(gm1 * rho * e);
        prim[t + 6 * n] = (TS)(gm1 * e * onebyR);
    }
}

// get_transport_properties, one species (MixtureEOS.F90:175-203) with the
// power law (PowerLawViscosity.F90:46-76): mu, bulk = 0, kappa = Cp mu / Pr;
// all zero when inviscid
template <typename TS>
__global__ void __launch_bounds__(256) k_transport(const TS *T, TS *mu, TS *bulk, TS *kap, long n, int inviscid,
                                                  double mu_ref, double T_ref, double vn, double Cp, double Pr)
{
    PDO_CG_LOOP(n)
    {
        double m = 0, k = 0;
        if (!inviscid) {
            m = mu_ref * pow(ld(T, t) / T_ref, vn);
            k = Cp * m / Pr;
        }
        mu[t] = (TS)m;
        bulk[t] = (TS)0;
        kap[t] = (TS)k;
    }
}

// getLAD (cgrid.F90:1081-1084): the strain-rate magnitude of the gradients g
// (9 fields n apart, du_i/dx_j at 3 i + j)
template <typename TS>
__global__ void __launch_bounds__(256) k_shear_func(const TS *g, TS *func, long n)
{
    PDO_CG_LOOP(n)
    {
        double d[9];
#pragma unroll
        for (int c = 0; c < 9; ++c) d[c] = ld(g, t + c * n);
        const double a = d[3] + d[1], b = d[6] + d[2], c = d[7] + d[5];
        func[t] = (TS)sqrt(d[0] * d[0] + 0.5 * a * a + 0.5 * b * b + d[4] * d[4] + 0.5 * c * c + d[8] * d[8]);
    }
}

// dilatation dudx + dvdy + dwdz (:1116)
template <typename TS>
__global__ void __launch_bounds__(256) k_dilatation(const TS *g, TS *func, long n)
{
    PDO_CG_LOOP(n) func[t] = (TS)(ld(g, t) + ld(g, t + 4 * n) + ld(g, t + 8 * n));
}

// mustar = Cmu rho |X dx^6 + Z dz^6 + Y dy^6| (:1086-1110), X, Y, Z the
// fourth derivatives along x, y, z
template <typename TS>
__global__ void __launch_bounds__(256) k_mustar(const TS *X, const TS *Y, const TS *Z, const TS *rho, TS *out, long n,
                                               double Cmu, double dx6, double dy6, double dz6)
{
    PDO_CG_LOOP(n)
    {
        double m = ld(X, t) * dx6;
        m = m + ld(Z, t) * dz6;
        m = m + ld(Y, t) * dy6;
        out[t] = (TS)(Cmu * ld(rho, t) * fabs(m));
    }
}

// bulkstar (:1118-1158): the fourth derivatives of the dilatation weighted by
// (d_i (drho/dx_i)^2 / |grad rho|^2)^2, the switch dil^2 / (dil^2 + |curl u|^2)
// where the dilatation is negative, Cbeta rho |.|
template <typename TS>
__global__ void __launch_bounds__(256) k_bulkstar(const TS *X, const TS *Y, const TS *Z, const TS *R, const TS *g,
                                                 const TS *dil, const TS *rho, TS *out, long n, double Cbeta, double dx,
                                                 double dy, double dz, double dx4, double dy4, double dz4)
{
    PDO_CG_LOOP(n)
    {
        double r1 = ld(R, t), r2 = ld(R, t + n), r3 = ld(R, t + 2 * n);
        r1 = r1 * r1;
        r2 = r2 * r2;
        r3 = r3 * r3;
        const double s = r1 + r2 + r3 + 1.0e-32;
        double wx = dx * r1 / s, wy = dy * r2 / s, wz = dz * r3 / s;
        double b = (ld(X, t) * dx4) * (wx * wx);
        b = b + (ld(Z, t) * dz4) * (wz * wz);
        b = b + (ld(Y, t) * dy4) * (wy * wy);
        const double c1 = ld(g, t + 7 * n) - ld(g, t + 5 * n), c2 = ld(g, t + 2 * n) - ld(g, t + 6 * n),
                     c3 = ld(g, t + 3 * n) - ld(g, t + n);
        const double curl2 = c1 * c1 + c2 * c2 + c3 * c3;
        const double f = ld(dil, t), d2 = f * f;
        const double fsw = f >= 0.0 ? 0.0 : d2 / (d2 + curl2 + 1.0e-32);
        out[t] = (TS)(Cbeta * ld(rho, t) * fsw * fabs(b));
    }
}

// kapstar (:1163-1196): the fourth derivatives of e weighted by
// d_i (de/dx_i)^2 / |grad e|^2, Ckap rho c |.| / T with c = sqrt(gam p / rho)
template <typename TS>
__global__ void __launch_bounds__(256) k_kapstar(const TS *X, const TS *Y, const TS *Z, const TS *R, const TS *prim,
                                                TS *out, long n, double Ckap, double gam, double dx, double dy,
                                                double dz, double dx4, double dy4, double dz4)
{
    PDO_CG_LOOP(n)
    {
        double e1 = ld(R, t), e2 = ld(R, t + n), e3 = ld(R, t + 2 * n);
        e1 = e1 * e1;
        e2 = e2 * e2;
        e3 = e3 * e3;
        const double s = e1 + e2 + e3 + 1.0e-32;
        double k = (ld(X, t) * dx4) * (dx * e1 / s);
        k = k + (ld(Z, t) * dz4) * (dz * e3 / s);
        k = k + (ld(Y, t) * dy4) * (dy * e2 / s);
        const double rho = ld(prim, t), p = ld(prim, t + 5 * n), T = ld(prim, t + 6 * n);
        const double sos = sqrt(gam * p / rho);
        out[t] = (TS)(Ckap * rho * sos * fabs(k) / T);
    }
}

// mu += mustar, bulk += bulkstar, kap += kapstar (:1251-1255)
template <typename TS>
__global__ void __launch_bounds__(256) k_add3(TS *mu, TS *bulk, TS *kap, const TS *ms, const TS *bs, const TS *ks,
                                             long n)
{
    PDO_CG_LOOP(n)
    {
        mu[t] = (TS)(ld(mu, t) + ld(ms, t));
        bulk[t] = (TS)(ld(bulk, t) + ld(bs, t));
        kap[t] = (TS)(ld(kap, t) + ld(ks, t));
    }
}

// get_tau (:1360-1404): tau = (xx, xy, xz, yy, yz, zz), n apart
template <typename TS>
__global__ void __launch_bounds__(256) k_tau(const TS *g, const TS *mu, const TS *bulk, TS *tau, long n)
{
    PDO_CG_LOOP(n)
    {
        double d[9];
#pragma unroll
        for (int c = 0; c < 9; ++c) d[c] = ld(g, t + c * n);
        const double m = ld(mu, t), bk = ld(bulk, t);
        const double bambda = (4.0 / 3.0) * m + bk, lam = bk - (2.0 / 3.0) * m;
        tau[t] = (TS)(bambda * d[0] + lam * (d[4] + d[8]));
        tau[t + n] = (TS)(m * (d[1] + d[3]));
        tau[t + 2 * n] = (TS)(m * (d[2] + d[6]));
        tau[t + 3 * n] = (TS)(bambda * d[4] + lam * (d[0] + d[8]));
        tau[t + 4 * n] = (TS)(m * (d[5] + d[7]));
        tau[t + 5 * n] = (TS)(bambda * d[8] + lam * (d[0] + d[4]));
    }
}

// q = -kappa dT (get_q :1428-1444)
template <typename TS>
__global__ void __launch_bounds__(256) k_neg_mul(const TS *a, const TS *b, TS *out, long n)
{
    PDO_CG_LOOP(n) out[t] = (TS)(-ld(a, t) * ld(b, t));
}

// the five fluxes along direction dir (getRHS_x/y/z :896-1056):
//   mass rho u_d; momentum rho u_d u_i + p delta_id - tau_id;
//   energy (TE + p - tau_dd) u_d - sum_{i != d} u_i tau_id + q_d
template <typename TS>
__global__ void __launch_bounds__(256) k_fluxes(const TS *W, const TS *prim, const TS *tau, const TS *q, TS *flux,
                                               long n, int dir)
{
    // tau component of (i, d), components xx xy xz yy yz zz
    constexpr int TI[3][3] = {{0, 1, 2}, {1, 3, 4}, {2, 4, 5}};
    PDO_CG_LOOP(n)
    {
        const double m = ld(W, t + (1 + dir) * n);  // rho u_d
        const double vel[3] = {ld(prim, t + n), ld(prim, t + 2 * n), ld(prim, t + 3 * n)};
        const double p = ld(prim, t + 5 * n);
        double ti[3];
#pragma unroll
        for (int i = 0; i < 3; ++i) ti[i] = ld(tau, t + TI[i][dir] * n);
        flux[t] = (TS)m;
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            double f = m * vel[i];
            if (i == dir) f = f + p;
            flux[t + (1 + i) * n] = (TS)(f - ti[i]);
        }
        double E = (ld(W, t + 4 * n) + p - ti[dir]) * vel[dir];
#pragma unroll
        for (int i = 0; i < 3; ++i)
            if (i != dir) E = E - vel[i] * ti[i];
        flux[t + 4 * n] = (TS)(E + ld(q, t + dir * n));
    }
}

// rhs -= d (five fields n apart), or rhs = -d when first
template <typename TS>
__global__ void __launch_bounds__(256) k_sub5(TS *rhs, const TS *d, long n5, int first)
{
    PDO_CG_LOOP(n5) rhs[t] = (TS)((first ? 0.0 : ld(rhs, t)) - ld(d, t));
}

template <typename TS>
__global__ void __launch_bounds__(256) k_copy(const TS *a, TS *b, long n)
{
    PDO_CG_LOOP(n) b[t] = a[t];
}

#undef PDO_CG_LOOP

template <typename TS> TS *P(void *p) { return (TS *)p; }
template <typename TS> const TS *C(const void *p) { return (const TS *)p; }

template <typename TS>
hipError_t pointwise_t(CgOp op, const CgArgs &a, hipStream_t st)
{
    const dim3 g(grid_of(op == CG_SUB5 ? 5 * a.n : a.n)), b(256);
    switch (op) {
    case CG_PRIMITIVE:
        hipLaunchKernelGGL((k_primitive<TS>), g, b, 0, st, C<TS>(a.in[0]), P<TS>(a.out[0]), a.n, a.c[0], a.c[1]);
        break;
    case CG_TRANSPORT:
        hipLaunchKernelGGL((k_transport<TS>), g, b, 0, st, C<TS>(a.in[0]), P<TS>(a.out[0]), P<TS>(a.out[1]),
                           P<TS>(a.out[2]), a.n, a.i0, a.c[0], a.c[1], a.c[2], a.c[3], a.c[4]);
        break;
    case CG_SHEAR:
        hipLaunchKernelGGL((k_shear_func<TS>), g, b, 0, st, C<TS>(a.in[0]), P<TS>(a.out[0]), a.n);
        break;
    case CG_DILATATION:
        hipLaunchKernelGGL((k_dilatation<TS>), g, b, 0, st, C<TS>(a.in[0]), P<TS>(a.out[0]), a.n);
        break;
    case CG_MUSTAR:
        hipLaunchKernelGGL((k_mustar<TS>), g, b, 0, st, C<TS>(a.in[0]), C<TS>(a.in[1]), C<TS>(a.in[2]),
                           C<TS>(a.in[3]), P<TS>(a.out[0]), a.n, a.c[0], a.c[1], a.c[2], a.c[3]);
        break;
    case CG_BULKSTAR:
        hipLaunchKernelGGL((k_bulkstar<TS>), g, b, 0, st, C<TS>(a.in[0]), C<TS>(a.in[1]), C<TS>(a.in[2]),
                           C<TS>(a.in[3]), C<TS>(a.in[4]), C<TS>(a.in[5]), C<TS>(a.in[6]), P<TS>(a.out[0]), a.n,
                           a.c[0], a.c[1], a.c[2], a.c[3], a.c[4], a.c[5], a.c[6]);
        break;
    case CG_KAPSTAR:
        hipLaunchKernelGGL((k_kapstar<TS>), g, b, 0, st, C<TS>(a.in[0]), C<TS>(a.in[1]), C<TS>(a.in[2]),
                           C<TS>(a.in[3]), C<TS>(a.in[4]), P<TS>(a.out[0]), a.n, a.c[0], a.c[1], a.c[2], a.c[3],
                           a.c[4], a.c[5], a.c[6], a.c[7]);
        break;
    case CG_ADD3:
        hipLaunchKernelGGL((k_add3<TS>), g, b, 0, st, P<TS>(a.out[0]), P<TS>(a.out[1]), P<TS>(a.out[2]),
                           C<TS>(a.in[0]), C<TS>(a.in[1]), C<TS>(a.in[2]), a.n);
        break;
    case CG_TAU:
        hipLaunchKernelGGL((k_tau<TS>), g, b, 0, st, C<TS>(a.in[0]), C<TS>(a.in[1]), C<TS>(a.in[2]),
                           P<TS>(a.out[0]), a.n);
        break;
    case CG_NEGMUL:
        hipLaunchKernelGGL((k_neg_mul<TS>), g, b, 0, st, C<TS>(a.in[0]), C<TS>(a.in[1]), P<TS>(a.out[0]), a.n);
        break;
    case CG_FLUXES:
        hipLaunchKernelGGL((k_fluxes<TS>), g, b, 0, st, C<TS>(a.in[0]), C<TS>(a.in[1]), C<TS>(a.in[2]),
                           C<TS>(a.in[3]), P<TS>(a.out[0]), a.n, a.i0);
        break;
    case CG_SUB5:
        hipLaunchKernelGGL((k_sub5<TS>), g, b, 0, st, P<TS>(a.out[0]), C<TS>(a.in[0]), 5 * a.n, a.i0);
        break;
    case CG_COPY:
        hipLaunchKernelGGL((k_copy<TS>), g, b, 0, st, C<TS>(a.in[0]), P<TS>(a.out[0]), a.n);
        break;
    }
    return hipGetLastError();
}

}  // namespace

hipError_t launch_cgrid(bool dbl, CgOp op, const CgArgs &a, hipStream_t st)
{
    return dbl ? pointwise_t<double>(op, a, st) : pointwise_t<float>(op, a, st);
}

}  // namespace pdo
