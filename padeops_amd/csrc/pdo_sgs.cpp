// sgs_igrid eddy-viscosity models and the standard dynamic procedure behind
// the C ABI (src/incompressible/sgs_models/*.F90, spectral.F90:1333-1367).
#include <cmath>
#include <vector>

#include "../../include/padeops_hip.h"
#include "pdo_fft.hpp"
#include "pdo_poisson.hpp"
#include "pdo_sgs.hpp"
#include "pdo_spectral.hpp"

namespace pdo {

struct DynProcObject {
    SpectralObject *sp = nullptr;
    double deltaRat = 2.0;
    const double *d_t1 = nullptr, *d_t2 = nullptr;  // test-filter factors (spectral.F90:1343-1362)
    std::vector<void *> allocs;
    void *buf = nullptr;
    size_t buf_bytes = 0;
    ~DynProcObject()
    {
        for (void *p : allocs) (void)hipFree(p);
        if (buf) (void)hipFree(buf);
    }
};

}  // namespace pdo

struct pdo_dynproc_s : public pdo::DynProcObject {};

using namespace pdo;

static int hip_ok(hipError_t e, const char *what) { return e == hipSuccess ? 0 : hip_fail(e, what); }

extern "C" {

int pdo_sgs_kernel(int prec, int model, const void *duidxj, void *Sij, void *nu_sgs, long npts, double camd_x,
                   double camd_y, double camd_z, void *stream)
{
    if (!duidxj || npts < 0 || model < 0 || model > 2 || (!Sij && !nu_sgs)) {
        set_error("invalid argument");
        return PDO_ERR_ARG;
    }
    const int mode = fft_mode(prec);
    if (mode < 0) { set_error("unknown precision mode"); return PDO_ERR_ARG; }
    if (npts == 0) return 0;
    return hip_ok(launch_sgs_kernel(mode, duidxj, Sij, nu_sgs, npts, npts, model, camd_x, camd_y, camd_z,
                                    (hipStream_t)stream),
                  "sgs kernel");
}

int pdo_sgs_scale(int prec, void *nu_sgs, long nxy, int nz, const double *cplane, double cglobal, void *stream)
{
    if (!nu_sgs || nxy < 0 || nz < 0) { set_error("invalid argument"); return PDO_ERR_ARG; }
    const int mode = fft_mode(prec);
    if (mode < 0) { set_error("unknown precision mode"); return PDO_ERR_ARG; }
    if (nxy == 0 || nz == 0) return 0;
    return hip_ok(launch_sgs_scale(mode == 0, nu_sgs, nxy, nxy * nz, cplane, cglobal, (hipStream_t)stream),
                  "sgs model constant");
}

int pdo_dynproc_init(pdo_dynproc *h, pdo_spectral spC, double deltaRat)
{
    if (!h || !spC || !(deltaRat > 0)) { set_error("invalid argument"); return PDO_ERR_ARG; }
    *h = nullptr;
    if (spC->periodic_z) {
        // spectral.F90:1339-1342
        set_error("Test filtering currently not supported for problems with periodic BC in z");
        return PDO_ERR_UNSUP;
    }
    auto *o = new pdo_dynproc_s();
    o->sp = spC;
    o->deltaRat = deltaRat;
    const double kfx = (((2.0 / 3.0) * M_PI) / spC->dx) / deltaRat, kfy = (((2.0 / 3.0) * M_PI) / spC->dy) / deltaRat;
    std::vector<double> t1(spC->nxh), t2(spC->ny);
    for (int i = 0; i < spC->nxh; i++) t1[i] = std::fabs(spC->k1[i]) < kfx ? 1.0 : 0.0;
    for (int j = 0; j < spC->ny; j++) t2[j] = std::fabs(spC->k2[j]) < kfy ? 1.0 : 0.0;
    double *p;
    int st;
    if ((st = upload_vec(t1, &p))) { delete o; return st; }
    o->allocs.push_back(p); o->d_t1 = p;
    if ((st = upload_vec(t2, &p))) { delete o; return st; }
    o->allocs.push_back(p); o->d_t2 = p;
    *h = o;
    return 0;
}

int pdo_dynproc_compute(pdo_dynproc h, int prec, const void *u, const void *v, const void *w, const void *uhat,
                        const void *vhat, const void *what, const void *Sij, const void *Dsgs, double *lambda_1d,
                        void *stream)
{
    if (!h || !u || !v || !w || !uhat || !vhat || !what || !Sij || !Dsgs || !lambda_1d) {
        set_error("invalid argument");
        return PDO_ERR_ARG;
    }
    const int mode = fft_mode(prec);
    if (mode < 0) { set_error("unknown precision mode"); return PDO_ERR_ARG; }
    SpectralObject &sp = *h->sp;
    const bool dbl = mode == 0;
    const size_t W = dbl ? 8 : 4;
    const long n = (long)sp.nx * sp.ny * sp.nz, nc = (long)sp.nxh * sp.ny * sp.nz, nxy = (long)sp.nx * sp.ny;
    const int nblk = (int)((nxy + 255) / 256 < 64 ? (nxy + 255) / 256 : 64);
    // work: spectrum | Sf(6) | Df | DS(6) | uf(3) | uu(6) | num, den (FP64) | partials
    const size_t spec = (size_t)nc * 2 * W, fld = (size_t)n * W;
    const size_t need = spec + fld * 22 + (size_t)n * 16 + (size_t)sp.nz * nblk * 16 + 256;
    if (h->buf_bytes < need) {
        if (h->buf) (void)hipFree(h->buf);
        h->buf = nullptr; h->buf_bytes = 0;
        hipError_t e = hipMalloc(&h->buf, need);
        if (e != hipSuccess) return hip_fail(e, "hipMalloc(dynamic procedure work)");
        h->buf_bytes = need;
    }
    char *b = (char *)h->buf;
    void *T = b;
    char *Sf = b + spec, *Df = Sf + 6 * fld, *DS = Df + fld, *uf = DS + 6 * fld, *uu = uf + 3 * fld;
    double *num = (double *)(uu + 6 * fld), *den = num + n, *part = den + n;
    hipStream_t st = (hipStream_t)stream;
    hipError_t e;
    int rc;
    // TestFilter_Real_to_Real (standardDynamicProcedure.F90:165-179): fft,
    // the cutoff, ifft, with the y transforms and the cutoff in one pass
    auto tf = [&](const void *in, void *out) -> int {
        if ((e = run_filter2(sp.plan, mode, in, T, out, sp.nz, h->d_t1, h->d_t2, st)) != hipSuccess)
            return hip_fail(e, "test filter");
        return 0;
    };
    // TestFilter_Cmplx_to_Real (:195-207): the cutoff of a copy of fhat, ifft
    auto tfc = [&](const void *fhat, void *out) -> int {
        if ((e = launch_mask2_oop(mode, fhat, T, h->d_t1, h->d_t2, sp.nxh, sp.ny, nc, st)) != hipSuccess)
            return hip_fail(e, "test filter");
        return pdo_spectral_ifft(static_cast<pdo_spectral>(h->sp), prec, T, out, 0, stream);
    };
    const char *S = (const char *)Sij;
    for (int c = 0; c < 6; c++)
        if ((rc = tf(S + c * fld, Sf + c * fld))) return rc;
    if ((e = launch_smag_from_S(dbl, Sf, Df, n, n, st)) != hipSuccess) return hip_fail(e, "dynamic procedure");
    if ((e = launch_mul(dbl, Dsgs, Sij, DS, n, n, 6, st)) != hipSuccess) return hip_fail(e, "dynamic procedure");
    for (int c = 0; c < 6; c++)
        if ((rc = tf(DS + c * fld, DS + c * fld))) return rc;
    if ((e = launch_germano_M(dbl, DS, Df, Sf, n, n, h->deltaRat * h->deltaRat, st)) != hipSuccess)
        return hip_fail(e, "dynamic procedure");
    const void *hats[3] = {uhat, vhat, what};
    for (int c = 0; c < 3; c++)
        if ((rc = tfc(hats[c], uf + c * fld))) return rc;
    if ((e = launch_uiuj(dbl, u, v, w, uu, n, n, st)) != hipSuccess) return hip_fail(e, "dynamic procedure");
    for (int c = 0; c < 6; c++)
        if ((rc = tf(uu + c * fld, uu + c * fld))) return rc;
    if ((e = launch_germano_ratio(dbl, uu, uf, DS, num, den, n, n, st)) != hipSuccess)
        return hip_fail(e, "dynamic procedure");
    e = launch_plane_ratio(num, den, nxy, sp.nz, 1.0 / ((double)sp.nx * sp.ny), part, lambda_1d, st);
    return hip_ok(e, "planar averages");
}

int pdo_dynproc_destroy(pdo_dynproc h)
{
    delete h;
    return 0;
}

}  // extern "C"
