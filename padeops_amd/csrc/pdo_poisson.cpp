// PoissonPeriodic object (src/utilities/PoissonPeriodic.F90) over the FFT
// kernels of pdo_fft.hip.
#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/padeops_hip.h"
#include "pdo_fft.hpp"
#include "pdo_ops.hpp"
#include "pdo_poisson.hpp"

namespace pdo {

// wavenumbers of PoissonPeriodic.F90:226-261 (GetWaveNums + ifftshift):
// k = (0, 1, .., n/2-1, -n/2, .., -1) * 2 pi / (n dx)
std::vector<double> wavenums(int n, double dx)
{
    const int dummy = n - (n % 2);
    std::vector<double> t(n), k(n);
    for (int i = 1; i <= n; i++) t[i - 1] = (-M_PI + (i - 1) * 2.0 * M_PI / (double)dummy) / dx;
    if (n % 2 == 0) {
        for (int i = 0; i < n / 2; i++) { k[i] = t[n / 2 + i]; k[n / 2 + i] = t[i]; }
    } else {
        for (int i = 0; i < (n + 1) / 2; i++) k[i] = t[(n + 1) / 2 - 1 + i];
        for (int i = 0; i < (n - 1) / 2; i++) k[(n + 1) / 2 + i] = t[i];
    }
    return k;
}

// modified wavenumber k'(k) of the derivative scheme (omega = k dx)
double modified_k(int kind, double k, double dx)
{
    const double w = k * dx;
    switch (kind) {
    case PDO_MODK_CD06: {
        // collocated cd06 first derivative, cd06.F90:14-16
        const double al = 1.0 / 3.0, a = (14.0 / 9.0) / 2.0, b = (1.0 / 9.0) / 4.0;
        return (2 * a * std::sin(w) + 2 * b * std::sin(2 * w)) / (1 + 2 * al * std::cos(w)) / dx;
    }
    case PDO_MODK_CD06STAGG: {
        // PadeDerOps.F90:1034-1053
        const double al = 9.0 / 62.0, a = 63.0 / 62.0, b = 17.0 / 62.0;
        return (2 * a * std::sin(w / 2) + (2.0 / 3.0) * b * std::sin(3 * w / 2)) / (1 + 2 * al * std::cos(w)) / dx;
    }
    case PDO_MODK_CD10: {
        // collocated cd10 first derivative, cd10.F90:15-19
        const double al = 0.5, be = 1.0 / 20.0, a = (17.0 / 12.0) / 2.0, b = (101.0 / 150.0) / 4.0,
                     c = (1.0 / 100.0) / 6.0;
        return (2 * a * std::sin(w) + 2 * b * std::sin(2 * w) + 2 * c * std::sin(3 * w)) /
               (1 + 2 * al * std::cos(w) + 2 * be * std::cos(2 * w)) / dx;
    }
    default: return k;
    }
}

template <typename T> int upload_vec(const std::vector<T> &h, T **d)
{
    hipError_t e = hipMalloc(d, h.size() * sizeof(T));
    if (e != hipSuccess) return hip_fail(e, "hipMalloc");
    e = hipMemcpy(*d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice);
    if (e != hipSuccess) return hip_fail(e, "hipMemcpy");
    return 0;
}
template int upload_vec<double>(const std::vector<double> &, double **);
template int upload_vec<float>(const std::vector<float> &, float **);

// exp(-2 pi i m / L), m < L, in FP64 and FP32 (computed in long double)
static void twiddles(int L, std::vector<double> &d, std::vector<float> &f)
{
    d.resize(2 * L);
    f.resize(2 * L);
    for (int m = 0; m < L; m++) {
        const long double a = -2.0L * 3.14159265358979323846264338327950288L * m / L;
        d[2 * m] = (double)cosl(a);
        d[2 * m + 1] = (double)sinl(a);
        f[2 * m] = (float)cosl(a);
        f[2 * m + 1] = (float)sinl(a);
    }
}

bool fft_pow2(int n, int lo, int hi) { return n >= lo && n <= hi && (n & (n - 1)) == 0; }

int fft_tables(PoissonPlanDev &dev, int nx, int ny, int nz, std::vector<void *> &allocs)
{
    int st;
    const int nxh = nx / 2 + 1;
    const int Ls[3] = {nx / 2, ny, nz};
    for (int d = 0; d < 3; d++) {
        if (Ls[d] <= 0) continue;
        std::vector<double> td; std::vector<float> tf;
        twiddles(Ls[d], td, tf);
        double *pd; float *pf;
        if ((st = upload_vec(td, &pd))) return st;
        allocs.push_back(pd);
        if ((st = upload_vec(tf, &pf))) return st;
        allocs.push_back(pf);
        dev.tw_d[d] = pd; dev.tw_f[d] = pf;
    }
    // exp(-2 pi i k / nx), k = 0..nx/2: the real-transform split
    std::vector<double> td(2 * nxh); std::vector<float> tf(2 * nxh);
    for (int k = 0; k < nxh; k++) {
        const long double a = -2.0L * 3.14159265358979323846264338327950288L * k / nx;
        td[2 * k] = (double)cosl(a); td[2 * k + 1] = (double)sinl(a);
        tf[2 * k] = (float)cosl(a); tf[2 * k + 1] = (float)sinl(a);
    }
    double *pd; float *pf;
    if ((st = upload_vec(td, &pd))) return st;
    allocs.push_back(pd);
    if ((st = upload_vec(tf, &pf))) return st;
    allocs.push_back(pf);
    dev.twn_d = pd; dev.twn_f = pf;
    return 0;
}

int fft_mode(int prec)
{
    switch (prec) {
    case PDO_F64: return 0;
    case PDO_F32_REFINE: case PDO_F32: return 1;
    case PDO_F32_STORE: return 2;
    }
    return -1;
}

PoissonObject::~PoissonObject()
{
    for (void *p : allocs) (void)hipFree(p);
    if (spec) (void)hipFree(spec);
}

int PoissonObject::init(int nx_, int ny_, int nz_, double dx, double dy, double dz, int mx, int my, int mz, int zp)
{
    auto kp = [](int n, double d, int kind, int count) {
        std::vector<double> k = wavenums(n, d), out(count);
        for (int i = 0; i < count; i++) out[i] = modified_k(kind, k[i], d);
        return out;
    };
    const std::vector<double> kx = kp(nx_, dx, mx, nx_ / 2 + 1), ky = kp(ny_, dy, my, ny_), kz = kp(nz_, dz, mz, nz_);
    return init_k(nx_, ny_, nz_, kx.data(), ky.data(), kz.data(), zp);
}

int PoissonObject::init_k(int nx_, int ny_, int nz_, const double *kx, const double *ky, const double *kz, int zp)
{
    if (!fft_pow2(nx_, 16, 1024) || !fft_pow2(ny_, 8, 1024) || !fft_pow2(nz_, 8, 1024)) {
        set_error("Poisson solver sizes must be powers of two in [16,1024] (x) / [8,1024] (y, z)");
        return PDO_ERR_ARG;
    }
    if (!kx || !ky || !kz) { set_error("null wavenumber table"); return PDO_ERR_ARG; }
    if (zp != 0 && zp != 1) { set_error("zero_policy must be 0 or 1"); return PDO_ERR_ARG; }
    nx = nx_; ny = ny_; nz = nz_;
    std::memset(&dev, 0, sizeof(dev));
    dev.nx = nx; dev.ny = ny; dev.nz = nz; dev.zero_policy = zp;
    const int nxh = nx / 2 + 1;
    dev.nxp = (nxh + 15) / 16 * 16;
    auto sq = [](const double *k, int count) {
        std::vector<double> out(count);
        for (int i = 0; i < count; i++) out[i] = k[i] * k[i];
        return out;
    };
    double *p;
    int st;
    if ((st = upload_vec(sq(kx, nxh), &p))) return st;
    allocs.push_back(p); dev.kx2 = p;
    if ((st = upload_vec(sq(ky, ny), &p))) return st;
    allocs.push_back(p); dev.ky2 = p;
    if ((st = upload_vec(sq(kz, nz), &p))) return st;
    allocs.push_back(p); dev.kz2 = p;
    {
        std::vector<double> kp(nxh);
        for (int i = 0; i < nxh; i++) kp[i] = (i == 0 || i == nx / 2) ? 0.0 : kx[i];
        if ((st = upload_vec(kp, &p))) return st;
        allocs.push_back(p); kpx = p;
    }
    return fft_tables(dev, nx, ny, nz, allocs);
}

int PoissonObject::solve_dx(int prec, const void *div, const void *ut, void *p, void *ux, hipStream_t st)
{
    if (!div || !ut || !p || !ux) { set_error("null field"); return PDO_ERR_ARG; }
    const int mode = fft_mode(prec);
    if (mode < 0) { set_error("unknown precision mode"); return PDO_ERR_ARG; }
    const size_t need = (size_t)dev.nxp * ny * nz * 2 * (mode == 0 ? 8 : 4);
    if (spec_bytes < need) {
        if (spec) (void)hipFree(spec);
        spec = nullptr; spec_bytes = 0;
        hipError_t e = hipMalloc(&spec, need);
        if (e != hipSuccess) return hip_fail(e, "hipMalloc(spectrum)");
        spec_bytes = need;
    }
    hipError_t e = run_poisson_dx(dev, mode, div, ut, p, ux, spec, kpx, st);
    return e == hipSuccess ? 0 : hip_fail(e, "poisson kernels");
}

int PoissonObject::solve(int prec, const void *rhs, void *f, hipStream_t st)
{
    if (!rhs || !f) { set_error("null field"); return PDO_ERR_ARG; }
    const int mode = fft_mode(prec);
    if (mode < 0) { set_error("unknown precision mode"); return PDO_ERR_ARG; }
    const size_t need = (size_t)dev.nxp * ny * nz * 2 * (mode == 0 ? 8 : 4);
    if (spec_bytes < need) {
        if (spec) (void)hipFree(spec);
        spec = nullptr; spec_bytes = 0;
        hipError_t e = hipMalloc(&spec, need);
        if (e != hipSuccess) return hip_fail(e, "hipMalloc(spectrum)");
        spec_bytes = need;
    }
    hipError_t e = run_poisson(dev, mode, rhs, f, spec, st);
    return e == hipSuccess ? 0 : hip_fail(e, "poisson kernels");
}

}  // namespace pdo

using namespace pdo;

extern "C" {

int pdo_poisson_periodic_init(pdo_poisson *h, int nx, int ny, int nz, double dx, double dy, double dz, int modk_x,
                              int modk_y, int modk_z, int zero_policy)
{
    if (!h) return PDO_ERR_ARG;
    *h = nullptr;
    auto *o = new pdo_poisson_s();
    int st = o->init(nx, ny, nz, dx, dy, dz, modk_x, modk_y, modk_z, zero_policy);
    if (st) { delete o; return st; }
    *h = o;
    return 0;
}

int pdo_poisson_periodic_init_k(pdo_poisson *h, int nx, int ny, int nz, const double *kx, const double *ky,
                                const double *kz, int zero_policy)
{
    if (!h) return PDO_ERR_ARG;
    *h = nullptr;
    auto *o = new pdo_poisson_s();
    int st = o->init_k(nx, ny, nz, kx, ky, kz, zero_policy);
    if (st) { delete o; return st; }
    *h = o;
    return 0;
}

int pdo_poisson_periodic_solve(pdo_poisson h, int prec, const void *rhs, void *f, void *stream)
{
    if (!h) { set_error("null handle"); return PDO_ERR_ARG; }
    return h->solve(prec, rhs, f, (hipStream_t)stream);
}

int pdo_poisson_destroy(pdo_poisson h)
{
    delete h;
    return 0;
}

}  // extern "C"
