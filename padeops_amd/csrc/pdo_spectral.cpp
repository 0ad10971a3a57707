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
#include "pdo_kernels.hpp"
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
    if (dz_odd) pdo_op_destroy(dz_odd);
    if (coef) (void)hipFree(coef);
    if (cp) (void)hipFree(cp);
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

// PadePoisson.F90:181-232 (PeriodicInZ false, no Stokes pressure): the
// wavenumbers of the 2 nz-point extensions, k3' the cd06stagg modified
// wavenumber (getmodCD06stagg), the half-cell phases exp(-+ i k3 dz / 2),
// |k|^2 from GetWaveNums of x and y; derZ of DivergenceCheck (odd ends, :150)
int PadePoissonObject::init_nonperiodic(SpectralObject *s, double dz)
{
    sp = s;
    nz = s->nz;
    periodic = false;
    if (!fft_pow2(2 * nz, 8, 1024)) {
        set_error("PadePoisson: walls in z need nz a power of two in [4,512]");
        return PDO_ERR_ARG;
    }
    int st = pdo_cd06stagg_init(&dz_odd, nz, dz, 0, 0, 0, 0);
    if (st) return st;
    const int L = 2 * nz;
    std::vector<double> k1 = wavenums(s->nx, s->dx), k2 = wavenums(s->ny, s->dy), k3 = wavenums(L, dz);
    std::vector<double> ax(s->nxh), ay(s->ny), az(L), cm(2 * L), cp(2 * L);
    for (int i = 0; i < s->nxh; i++) ax[i] = k1[i] * k1[i];
    for (int j = 0; j < s->ny; j++) ay[j] = k2[j] * k2[j];
    for (int k = 0; k < L; k++) {
        const double km = modified_k(PDO_MODK_CD06STAGG, k3[k], dz), ph = 0.5 * dz * k3[k];
        az[k] = km * km;
        cm[2 * k] = km * std::cos(ph);
        cm[2 * k + 1] = -km * std::sin(ph);
        cp[2 * k] = km * std::cos(ph);
        cp[2 * k + 1] = km * std::sin(ph);
    }
    const std::vector<double> *src[5] = {&ax, &ay, &az, &cm, &cp};
    const double **dst[5] = {&wt.ax, &wt.ay, &wt.az, &wt.kcm, &wt.kcp};
    for (int q = 0; q < 5; q++) {
        double *p;
        if ((st = upload_vec(*src[q], &p))) return st;
        allocs.push_back(p);
        *dst[q] = p;
    }
    std::memset(&zplan2, 0, sizeof(zplan2));
    zplan2.nx = s->nx; zplan2.ny = s->ny; zplan2.nz = L; zplan2.nxp = s->nxh;
    return fft_tables(zplan2, s->nx, s->ny, L, allocs);
}

// PadePoisson.F90:236-307 (computeStokesPressure): per (kx, ky) line
// lam = |(k1, k2)| of the plain GetWaveNums values and
// denFact = 1 / (lam sinh(lam Lz) + 1e-13) where lam Lz < 500, else 0, then 0
// below 1e-16 and at the mean line; the edges linspace(0, Lz, nz + 1)
// (running sums of the step, gridtools.F90:51-69) and the cells between them
int PadePoissonObject::set_stokes(double Lz)
{
    if (periodic) {
        set_error("PadePoisson: the Stokes pressure needs walls in z");
        return PDO_ERR_ARG;
    }
    if (!(Lz > 0)) {
        set_error("PadePoisson: Lz must be positive");
        return PDO_ERR_ARG;
    }
    if (stokes) {
        set_error("PadePoisson: the Stokes pressure is set already");
        return PDO_ERR_ARG;
    }
    SpectralObject *s = sp;
    std::vector<double> k1 = wavenums(s->nx, s->dx), k2 = wavenums(s->ny, s->dy);
    k1.resize(s->nxh);
    const long np = (long)s->nxh * s->ny;
    std::vector<double> lam(np), den(np), ze(nz + 1), zc(nz);
    for (int j = 0; j < s->ny; j++)
        for (int i = 0; i < s->nxh; i++) {
            const long q = i + (long)s->nxh * j;
            const double l = std::sqrt(k1[i] * k1[i] + k2[j] * k2[j]);
            double d = (l * Lz < 500.0) ? 1.0 / (l * std::sinh(l * Lz) + 1.0e-13) : 0.0;
            if (d < 1.0e-16) d = 0.0;
            lam[q] = l;
            den[q] = d;
        }
    den[0] = 0.0;
    const double step = (Lz - 0.0) / ((double)(nz + 1) - 1.0);
    ze[0] = 0.0;
    for (int k = 1; k <= nz; k++) ze[k] = ze[k - 1] + step;
    for (int k = 0; k < nz; k++) zc[k] = 0.5 * (ze[k] + ze[k + 1]);
    const std::vector<double> *src[6] = {&k1, &k2, &lam, &den, &zc, &ze};
    const double **dst[6] = {&stt.k1, &stt.k2, &stt.lam, &stt.den, &stt.zc, &stt.ze};
    for (int q = 0; q < 6; q++) {
        double *p;
        int st = upload_vec(*src[q], &p);
        if (st) return st;
        allocs.push_back(p);
        *dst[q] = p;
    }
    hipError_t e = hipMalloc((void **)&coef, (size_t)np * 4 * sizeof(double));
    if (e != hipSuccess) return hip_fail(e, "hipMalloc(Stokes coefficients)");
    stt.Lz = Lz;
    stt.nxh = s->nxh;
    stt.ny = s->ny;
    stt.nz = nz;
    stokes = true;
    return 0;
}

int PadePoissonObject::ensure_copies(size_t bytes)
{
    if (cp_bytes >= bytes) return 0;
    if (cp) (void)hipFree(cp);
    cp = nullptr;
    cp_bytes = 0;
    hipError_t e = hipMalloc(&cp, bytes);
    if (e != hipSuccess) return hip_fail(e, "hipMalloc(getPressure copies)");
    cp_bytes = bytes;
    return 0;
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

int pdo_padepoisson_init_nonperiodic(pdo_padepoisson *h, pdo_spectral sp, double dz)
{
    if (!h || !sp) { set_error("invalid argument"); return PDO_ERR_ARG; }
    *h = nullptr;
    if (sp->periodic_z) {
        set_error("PadePoisson with walls: the spectral object must not be periodic in z");
        return PDO_ERR_ARG;
    }
    auto *o = new pdo_padepoisson_s();
    int st = o->init_nonperiodic(sp, dz);
    if (st) { delete o; return st; }
    *h = o;
    return 0;
}

// the walls solve of PressureProjection (PadePoisson.F90:434-624) and
// getPressure (:752-898): Step 0 the Stokes step in place on (u, v, w), from
// where on k1, k2 are the plain GetWaveNums values (k1inZ, k2inZ); Step 1
// f = i (k1 u + k2 v); Steps 3-8 the extensions, transforms, pressure,
// (upd) w update and walls; Step 9 (upd) u -= i k1 p, v -= i k2 p.  The
// pressure stays in h->f (with the Stokes part when pressure is set).
static int walls_solve(pdo_padepoisson h, int mode, void *uhat, void *vhat, void *what, bool upd, bool pressure,
                       hipStream_t str)
{
    SpectralObject *s = h->sp;
    const size_t W = mode == 0 ? 8 : 4;
    const long ncell = (long)s->nxh * s->ny * h->nz;
    int st = h->ensure((size_t)ncell * 2 * W, 0);
    if (st) return st;
    const double *k1 = h->stokes ? h->stt.k1 : s->d_k1, *k2 = h->stokes ? h->stt.k2 : s->d_k2;
    hipError_t e;
    if (h->stokes && (e = launch_stokes(mode, uhat, vhat, what, h->coef, h->stt, str)) != hipSuccess)
        return hip_fail(e, "Stokes pressure");
    if ((e = launch_div_spec(mode, uhat, vhat, nullptr, h->f, k1, k2, s->nxh, s->ny, ncell, str)) != hipSuccess)
        return hip_fail(e, "projection divergence");
    if ((e = run_pp_walls(h->zplan2, mode, h->f, what, h->wt, upd, str)) != hipSuccess)
        return hip_fail(e, "projection solve");
    if (upd && (e = launch_proj_update(mode, uhat, vhat, what, h->f, nullptr, k1, k2, s->nxh, s->ny, ncell, ncell,
                                       str)) != hipSuccess)
        return hip_fail(e, "projection update");
    if (pressure && h->stokes && (e = launch_stokes_pressure(mode, h->f, h->coef, h->stt, str)) != hipSuccess)
        return hip_fail(e, "Stokes pressure");
    return 0;
}

static int projection_walls(pdo_padepoisson h, int mode, void *uhat, void *vhat, void *what, hipStream_t str)
{
    return walls_solve(h, mode, uhat, vhat, what, true, false, str);
}

// PeriodicProjection (PadePoisson.F90:386-432) and Periodic_getPressure
// (:716-750): f = ddz_E2C(w) + i (k1 u + k2 v), the z solve
// p = z-inverse(-1/|k|^2 z-forward(f)) / nz into h->f; upd: w -= ddz_C2E(p),
// u -= i k1 p, v -= i k2 p
static int periodic_solve(pdo_padepoisson h, int mode, int prec, void *uhat, void *vhat, void *what, bool upd,
                          hipStream_t str)
{
    SpectralObject *s = h->sp;
    const size_t W = mode == 0 ? 8 : 4;
    const long ncell = (long)s->nxh * s->ny * h->nz, nedge = (long)s->nxh * s->ny * (h->nz + 1);
    int st = h->ensure((size_t)ncell * 2 * W, (size_t)nedge * 2 * W);
    if (st) return st;
    if ((st = apply(h->derz, KIND_STG_D1_E2C, 0, 0, 3, prec, what, h->f, 2L * s->nxh, s->ny, str))) return st;
    hipError_t e = launch_div_spec(mode, uhat, vhat, h->f, h->f, s->d_k1, s->d_k2, s->nxh, s->ny, ncell, str);
    if (e != hipSuccess) return hip_fail(e, "projection divergence");
    const SpecMul m{h->d_ax, h->d_ay, h->d_az, 1, 1.0 / h->nz};
    if ((e = run_zmul(h->zplan, mode, h->f, h->f, m, str)) != hipSuccess) return hip_fail(e, "projection solve");
    if (!upd) return 0;
    if ((st = apply(h->derz, KIND_STG_D1_C2E, 0, 0, 3, prec, h->f, h->dp, 2L * s->nxh, s->ny, str))) return st;
    e = launch_proj_update(mode, uhat, vhat, what, h->f, h->dp, s->d_k1, s->d_k2, s->nxh, s->ny, ncell, nedge, str);
    return hip_status(e, "projection update");
}

int pdo_padepoisson_projection(pdo_padepoisson h, int prec, void *uhat, void *vhat, void *what, void *stream)
{
    if (!h || !uhat || !vhat || !what) { set_error("invalid argument"); return PDO_ERR_ARG; }
    const int mode = fft_mode(prec);
    if (mode < 0) { set_error("unknown precision mode"); return PDO_ERR_ARG; }
    if (!h->periodic) return projection_walls(h, mode, uhat, vhat, what, (hipStream_t)stream);
    return periodic_solve(h, mode, prec, uhat, vhat, what, true, (hipStream_t)stream);
}

int pdo_padepoisson_set_stokes(pdo_padepoisson h, double Lz)
{
    if (!h) { set_error("invalid argument"); return PDO_ERR_ARG; }
    return h->set_stokes(Lz);
}

int pdo_padepoisson_pressure(pdo_padepoisson h, int prec, void *uhat, void *vhat, void *what, void *pressure,
                             int update, void *stream)
{
    if (!h || !uhat || !vhat || !what || !pressure || (update != 0 && update != 1)) {
        set_error("invalid argument");
        return PDO_ERR_ARG;
    }
    const int mode = fft_mode(prec);
    if (mode < 0) { set_error("unknown precision mode"); return PDO_ERR_ARG; }
    hipStream_t str = (hipStream_t)stream;
    int st;
    if (h->periodic) {
        st = periodic_solve(h, mode, prec, uhat, vhat, what, update == 1, str);
    } else if (update || !h->stokes) {
        st = walls_solve(h, mode, uhat, vhat, what, update == 1, true, str);
    } else {
        // getPressure with the Stokes pressure: its step works on copies
        SpectralObject *s = h->sp;
        const size_t W = mode == 0 ? 8 : 4;
        const size_t bc = (size_t)s->nxh * s->ny * h->nz * 2 * W, be = (size_t)s->nxh * s->ny * (h->nz + 1) * 2 * W;
        if ((st = h->ensure_copies(2 * bc + be))) return st;
        char *c = (char *)h->cp;
        const void *srcs[3] = {uhat, vhat, what};
        const size_t sizes[3] = {bc, bc, be};
        size_t off = 0;
        for (int q = 0; q < 3; q++) {
            hipError_t e = hipMemcpyAsync(c + off, srcs[q], sizes[q], hipMemcpyDeviceToDevice, str);
            if (e != hipSuccess) return hip_fail(e, "getPressure copies");
            off += sizes[q];
        }
        st = walls_solve(h, mode, c, c + bc, c + 2 * bc, false, true, str);
    }
    if (st) return st;
    return pdo_spectral_ifft(static_cast<pdo_spectral>(h->sp), prec, h->f, pressure, 0, stream);
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
    // the z term: periodic cd06stagg, or with walls the odd-ended one of the
    // reference's derZ (PadePoisson.F90:150-151)
    if ((st = apply(h->periodic ? h->derz : h->dz_odd, KIND_STG_D1_E2C, 0, 0, 3, prec, what, h->f, 2L * s->nxh,
                    s->ny, str)))
        return st;
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
