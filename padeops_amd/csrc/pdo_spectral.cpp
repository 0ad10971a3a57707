// spectral (2D transforms, i k multiplication, dealiasing) and the periodic
// PadePoisson projection, over the FFT kernels of pdo_fft.hip and the compact
// staggered z operators.
//   spectral:    src/incompressible/spectral.F90 (init :867-931,
//                initializeEverything :933-1260, init_periodic_inZ_procedures
//                :755-865, mTimes_ik :235-312, dealias :314-363, fft/ifft
//                :1413-1453 -> fft_3d.F90:616-663)
//   PadePoisson: src/incompressible/PadePoisson.F90 (InitPeriodicPoissonSolver
//                :76-128, PeriodicProjection :386-432, DivergenceCheck
//                :1165-1244)
#include <cmath>
#include <cstring>

#include "../../include/padeops_hip.h"
#include "pdo_poisson.hpp"
#include "pdo_spectral.hpp"
#include "pdo_system.hpp"

namespace pdo {

SpectralObject::~SpectralObject()
{
    for (void *p : allocs) (void)hipFree(p);
    if (work) (void)hipFree(work);
}

int SpectralObject::init(int nx_, int ny_, int nz_, double dx_, double dy_, double dz_, int fix, double fact,
                         int pz)
{
    if (!fft_pow2(nx_, 16, 1024) || !fft_pow2(ny_, 8, 1024) || nz_ < 1) {
        set_error("spectral: nx must be a power of two in [16,1024], ny in [8,1024], nz >= 1");
        return PDO_ERR_ARG;
    }
    if (pz && !fft_pow2(nz_, 8, 1024)) {
        set_error("spectral: periodic z needs nz a power of two in [8,1024]");
        return PDO_ERR_ARG;
    }
    nx = nx_; ny = ny_; nz = nz_; nxh = nx / 2 + 1;
    dx = dx_; dy = dy_; dz = dz_;
    fix_oddball = fix != 0; periodic_z = pz != 0; dealias_fact = fact;
    // 1-D wavenumbers with the sign of the oddball (Nyquist) flipped to +,
    // initializeEverything STEP 3
    std::vector<double> k1d = wavenums(nx, dx), k2d = wavenums(ny, dy), k3d = wavenums(nz, dz);
    k1d[nx / 2] = -k1d[nx / 2];
    k2d[ny / 2] = -k2d[ny / 2];
    if (nz > 1) k3d[nz / 2] = -k3d[nz / 2];
    k1.assign(k1d.begin(), k1d.begin() + nxh);
    k2 = k2d;
    k3 = k3d;
    // STEP 7: two-thirds rule on (k1, k2), taken before the oddball fix
    const double kdx = ((2.0 / 3.0) * M_PI) / dx, kdy = ((2.0 / 3.0) * M_PI) / dy;
    g1.resize(nxh); g2.resize(ny); g3.assign(nz, 1.0);
    for (int i = 0; i < nxh; i++) g1[i] = std::fabs(k1[i]) < kdx ? 1.0 : 0.0;
    for (int j = 0; j < ny; j++) g2[j] = std::fabs(k2[j]) < kdy ? 1.0 : 0.0;
    // STEP 11 (two periodic directions): zero the x and y oddballs
    if (fix_oddball) {
        k2[ny / 2] = 0.0;
        k1[nx / 2] = 0.0;
    }
    if (periodic_z) {
        // init_periodic_inZ_procedures: the mask of the stored k1, k2, k3
        const double cx = (dealias_fact * M_PI) / dx, cy = (dealias_fact * M_PI) / dy, cz = (dealias_fact * M_PI) / dz;
        for (int i = 0; i < nxh; i++) g1[i] = std::fabs(k1[i]) >= cx ? 0.0 : 1.0;
        for (int j = 0; j < ny; j++) g2[j] = std::fabs(k2[j]) >= cy ? 0.0 : 1.0;
        for (int k = 0; k < nz; k++) g3[k] = std::fabs(k3[k]) >= cz ? 0.0 : 1.0;
    }
    int st;
    double *p;
    const std::vector<double> *src[5] = {&k1, &k2, &g1, &g2, &g3};
    const double **dst[5] = {&d_k1, &d_k2, &d_g1, &d_g2, &d_g3};
    for (int q = 0; q < 5; q++) {
        if ((st = upload_vec(*src[q], &p))) return st;
        allocs.push_back(p);
        *dst[q] = p;
    }
    std::memset(&plan, 0, sizeof(plan));
    plan.nx = nx; plan.ny = ny; plan.nz = nz; plan.nxp = nxh;
    return fft_tables(plan, nx, ny, periodic_z ? nz : 0, allocs);
}

int SpectralObject::ensure_work(size_t bytes)
{
    if (work_bytes >= bytes) return 0;
    if (work) (void)hipFree(work);
    work = nullptr; work_bytes = 0;
    hipError_t e = hipMalloc(&work, bytes);
    if (e != hipSuccess) return hip_fail(e, "hipMalloc(spectral work)");
    work_bytes = bytes;
    return 0;
}

PadePoissonObject::~PadePoissonObject()
{
    for (void *p : allocs) (void)hipFree(p);
    if (f) (void)hipFree(f);
    if (dp) (void)hipFree(dp);
    if (derz) pdo_op_destroy(derz);
}

int PadePoissonObject::init(SpectralObject *s, double dz)
{
    sp = s;
    nz = s->nz;
    if (!fft_pow2(nz, 8, 1024)) {
        set_error("PadePoisson: periodic z needs nz a power of two in [8,1024]");
        return PDO_ERR_ARG;
    }
    int st = pdo_cd06stagg_init_periodic(&derz, nz, dz);
    if (st) return st;
    // |k|^2 = k1^2 + k2^2 + k3'^2 with GetWaveNums k1, k2 and the cd06stagg
    // modified k3 (InitPeriodicPoissonSolver)
    std::vector<double> k1 = wavenums(s->nx, s->dx), k2 = wavenums(s->ny, s->dy), k3 = wavenums(nz, dz);
    std::vector<double> ax(s->nxh), ay(s->ny), az(nz);
    for (int i = 0; i < s->nxh; i++) ax[i] = k1[i] * k1[i];
    for (int j = 0; j < s->ny; j++) ay[j] = k2[j] * k2[j];
    for (int k = 0; k < nz; k++) {
        const double km = modified_k(PDO_MODK_CD06STAGG, k3[k], dz);
        az[k] = km * km;
    }
    double *p;
    if ((st = upload_vec(ax, &p))) return st;
    allocs.push_back(p); d_ax = p;
    if ((st = upload_vec(ay, &p))) return st;
    allocs.push_back(p); d_ay = p;
    if ((st = upload_vec(az, &p))) return st;
    allocs.push_back(p); d_az = p;
    std::memset(&zplan, 0, sizeof(zplan));
    zplan.nx = s->nx; zplan.ny = s->ny; zplan.nz = nz; zplan.nxp = s->nxh;
    return fft_tables(zplan, s->nx, s->ny, nz, allocs);
}

int PadePoissonObject::ensure(size_t fb, size_t db)
{
    if (f_bytes < fb) {
        if (f) (void)hipFree(f);
        f = nullptr; f_bytes = 0;
        hipError_t e = hipMalloc(&f, fb);
        if (e != hipSuccess) return hip_fail(e, "hipMalloc(projection work)");
        f_bytes = fb;
    }
    if (dp_bytes < db) {
        if (dp) (void)hipFree(dp);
        dp = nullptr; dp_bytes = 0;
        hipError_t e = hipMalloc(&dp, db);
        if (e != hipSuccess) return hip_fail(e, "hipMalloc(projection work)");
        dp_bytes = db;
    }
    return 0;
}

static int hip_status(hipError_t e, const char *what) { return e == hipSuccess ? 0 : hip_fail(e, what); }

}  // namespace pdo

using namespace pdo;

extern "C" {

int pdo_spectral_init(pdo_spectral *h, int nx, int ny, int nz, double dx, double dy, double dz, int fix_oddball,
                      double dealias_fact, int periodic_z)
{
    if (!h) return PDO_ERR_ARG;
    *h = nullptr;
    auto *o = new pdo_spectral_s();
    int st = o->init(nx, ny, nz, dx, dy, dz, fix_oddball, dealias_fact, periodic_z);
    if (st) { delete o; return st; }
    *h = o;
    return 0;
}

int pdo_spectral_fft(pdo_spectral h, int prec, const void *in, void *out, void *stream)
{
    if (!h || !in || !out) { set_error("invalid argument"); return PDO_ERR_ARG; }
    const int mode = fft_mode(prec);
    if (mode < 0) { set_error("unknown precision mode"); return PDO_ERR_ARG; }
    return hip_status(run_fft2(h->plan, mode, in, out, h->nz, (hipStream_t)stream), "spectral fft");
}

int pdo_spectral_ifft(pdo_spectral h, int prec, const void *in, void *out, int set_oddball, void *stream)
{
    if (!h || !in || !out) { set_error("invalid argument"); return PDO_ERR_ARG; }
    const int mode = fft_mode(prec);
    if (mode < 0) { set_error("unknown precision mode"); return PDO_ERR_ARG; }
    int st = h->ensure_work((size_t)h->nxh * h->ny * h->nz * 2 * (mode == 0 ? 8 : 4));
    if (st) return st;
    return hip_status(run_ifft2(h->plan, mode, in, h->work, out, h->nz, set_oddball, (hipStream_t)stream),
                      "spectral ifft");
}

int pdo_spectral_mtimes_ik(pdo_spectral h, int prec, int dir, const void *in, void *out, void *stream)
{
    if (!h || !in || !out || (dir != 1 && dir != 2)) { set_error("invalid argument"); return PDO_ERR_ARG; }
    const int mode = fft_mode(prec);
    if (mode < 0) { set_error("unknown precision mode"); return PDO_ERR_ARG; }
    const long n = (long)h->nxh * h->ny * h->nz;
    return hip_status(launch_mtimes_ik(mode, in, out, dir == 1 ? h->d_k1 : h->d_k2, h->nxh, h->ny, n, dir,
                                       (hipStream_t)stream),
                      "mTimes_ik");
}

int pdo_spectral_dealias(pdo_spectral h, int prec, void *fhat, void *stream)
{
    if (!h || !fhat) { set_error("invalid argument"); return PDO_ERR_ARG; }
    const int mode = fft_mode(prec);
    if (mode < 0) { set_error("unknown precision mode"); return PDO_ERR_ARG; }
    if (!h->periodic_z) {
        const long n = (long)h->nxh * h->ny * h->nz;
        return hip_status(launch_mask2(mode, fhat, h->d_g1, h->d_g2, h->nxh, h->ny, n, (hipStream_t)stream),
                          "dealias");
    }
    const SpecMul m{h->d_g1, h->d_g2, h->d_g3, 2, 1.0 / h->nz};
    return hip_status(run_zmul(h->plan, mode, fhat, fhat, m, (hipStream_t)stream), "dealias");
}

int pdo_spectral_dealias_edge(pdo_spectral h, int prec, void *fhat, void *stream)
{
    if (!h || !fhat) { set_error("invalid argument"); return PDO_ERR_ARG; }
    const int mode = fft_mode(prec);
    if (mode < 0) { set_error("unknown precision mode"); return PDO_ERR_ARG; }
    if (!h->periodic_z) return 0;  // dealias_edgeField acts only with periodic z
    const SpecMul m{h->d_g1, h->d_g2, h->d_g3, 2, 1.0 / h->nz};
    hipError_t e = run_zmul(h->plan, mode, fhat, fhat, m, (hipStream_t)stream);
    if (e != hipSuccess) return hip_fail(e, "dealias_edgeField");
    // edge nz+1 repeats edge 1
    const size_t plane = (size_t)h->nxh * h->ny * 2 * (mode == 0 ? 8 : 4);
    e = hipMemcpyAsync((char *)fhat + plane * h->nz, fhat, plane, hipMemcpyDeviceToDevice, (hipStream_t)stream);
    return hip_status(e, "dealias_edgeField copy");
}

int pdo_spectral_wavenumbers(pdo_spectral h, double *k1, double *k2, double *k3, double *g1, double *g2, double *g3)
{
    if (!h) { set_error("null handle"); return PDO_ERR_ARG; }
    if (k1) std::memcpy(k1, h->k1.data(), h->k1.size() * sizeof(double));
    if (k2) std::memcpy(k2, h->k2.data(), h->k2.size() * sizeof(double));
    if (k3) std::memcpy(k3, h->k3.data(), h->k3.size() * sizeof(double));
    if (g1) std::memcpy(g1, h->g1.data(), h->g1.size() * sizeof(double));
    if (g2) std::memcpy(g2, h->g2.data(), h->g2.size() * sizeof(double));
    if (g3) std::memcpy(g3, h->g3.data(), h->g3.size() * sizeof(double));
    return 0;
}

int pdo_spectral_set_factors(pdo_spectral h, const double *g1, const double *g2, const double *g3)
{
    if (!h) { set_error("null handle"); return PDO_ERR_ARG; }
    struct F { const double *src; std::vector<double> *host; const double *dev; size_t n; };
    const F f[3] = {{g1, &h->g1, h->d_g1, (size_t)h->nxh}, {g2, &h->g2, h->d_g2, (size_t)h->ny},
                    {g3, &h->g3, h->d_g3, (size_t)h->nz}};
    for (const F &q : f) {
        if (!q.src) continue;
        q.host->assign(q.src, q.src + q.n);
        hipError_t e = hipMemcpy((void *)q.dev, q.src, q.n * sizeof(double), hipMemcpyHostToDevice);
        if (e != hipSuccess) return hip_fail(e, "hipMemcpy(filter factors)");
    }
    return 0;
}

int pdo_spectral_destroy(pdo_spectral h)
{
    delete h;
    return 0;
}

int pdo_padepoisson_init_periodic(pdo_padepoisson *h, pdo_spectral sp, double dz)
{
    if (!h || !sp) { set_error("invalid argument"); return PDO_ERR_ARG; }
    *h = nullptr;
    auto *o = new pdo_padepoisson_s();
    int st = o->init(sp, dz);
    if (st) { delete o; return st; }
    *h = o;
    return 0;
}

int pdo_padepoisson_projection(pdo_padepoisson h, int prec, void *uhat, void *vhat, void *what, void *stream)
{
    if (!h || !uhat || !vhat || !what) { set_error("invalid argument"); return PDO_ERR_ARG; }
    const int mode = fft_mode(prec);
    if (mode < 0) { set_error("unknown precision mode"); return PDO_ERR_ARG; }
    SpectralObject *s = h->sp;
    const size_t W = mode == 0 ? 8 : 4;
    const long ncell = (long)s->nxh * s->ny * h->nz, nedge = (long)s->nxh * s->ny * (h->nz + 1);
    int st = h->ensure((size_t)ncell * 2 * W, (size_t)nedge * 2 * W);
    if (st) return st;
    hipStream_t str = (hipStream_t)stream;
    // f = d(w)/dz (E2C) + i (k1 u + k2 v)
    if ((st = apply(h->derz, KIND_STG_D1_E2C, 0, 0, 3, prec, what, h->f, 2L * s->nxh, s->ny, str))) return st;
    hipError_t e = launch_div_spec(mode, uhat, vhat, h->f, h->f, s->d_k1, s->d_k2, s->nxh, s->ny, ncell, str);
    if (e != hipSuccess) return hip_fail(e, "projection divergence");
    // p = z-inverse( -1/|k|^2 z-forward(f) ) / nz
    const SpecMul m{h->d_ax, h->d_ay, h->d_az, 1, 1.0 / h->nz};
    if ((e = run_zmul(h->zplan, mode, h->f, h->f, m, str)) != hipSuccess) return hip_fail(e, "projection solve");
    // w -= dp/dz (C2E); u -= i k1 p; v -= i k2 p
    if ((st = apply(h->derz, KIND_STG_D1_C2E, 0, 0, 3, prec, h->f, h->dp, 2L * s->nxh, s->ny, str))) return st;
    e = launch_proj_update(mode, uhat, vhat, what, h->f, h->dp, s->d_k1, s->d_k2, s->nxh, s->ny, ncell, nedge, str);
    return hip_status(e, "projection update");
}

int pdo_padepoisson_divergence(pdo_padepoisson h, int prec, const void *uhat, const void *vhat, const void *what,
                               void *divergence, void *stream)
{
    if (!h || !uhat || !vhat || !what || !divergence) { set_error("invalid argument"); return PDO_ERR_ARG; }
    const int mode = fft_mode(prec);
    if (mode < 0) { set_error("unknown precision mode"); return PDO_ERR_ARG; }
    SpectralObject *s = h->sp;
    const size_t W = mode == 0 ? 8 : 4;
    const long ncell = (long)s->nxh * s->ny * h->nz, nedge = (long)s->nxh * s->ny * (h->nz + 1);
    int st = h->ensure((size_t)ncell * 2 * W, (size_t)nedge * 2 * W);
    if (st) return st;
    hipStream_t str = (hipStream_t)stream;
    if ((st = apply(h->derz, KIND_STG_D1_E2C, 0, 0, 3, prec, what, h->f, 2L * s->nxh, s->ny, str))) return st;
    hipError_t e = launch_div_spec(mode, uhat, vhat, h->f, h->f, s->d_k1, s->d_k2, s->nxh, s->ny, ncell, str);
    if (e != hipSuccess) return hip_fail(e, "divergence");
    return pdo_spectral_ifft(static_cast<pdo_spectral>(s), prec, h->f, divergence, 0, stream);
}

int pdo_padepoisson_destroy(pdo_padepoisson h)
{
    delete h;
    return 0;
}

}  // extern "C"
