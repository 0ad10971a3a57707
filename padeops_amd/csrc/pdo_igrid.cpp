// igrid%compute_duidxj (src/incompressible/igrid.F90:2553-2683): the nine
// velocity-gradient components on cells and on edges, in physical and in
// spectral space, from the spectral velocity of igrid, over the spectral
// objects (x, y: i k and 2D inverse transforms) and Pade6stagg (z).  On one
// GPU the reference's y <-> z transposes are identities, so the z operators
// work on the spectral arrays in place of transposed copies.
#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/padeops_hip.h"
#include "pdo_closure.hpp"
#include "pdo_igrid.hpp"
#include "pdo_ops.hpp"
#include "pdo_poisson.hpp"
#include "pdo_spectral.hpp"

namespace pdo {

struct DuidxjObject {
    pdo_spectral spC = nullptr, spE = nullptr;
    pdo_pade6stagg dz = nullptr;
    int bc[12] = {};
    bool inviscid = false;
};

}  // namespace pdo

struct pdo_duidxj_s : public pdo::DuidxjObject {};

using namespace pdo;

extern "C" {

int pdo_duidxj_init(pdo_duidxj *h, pdo_spectral spC, pdo_spectral spE, pdo_pade6stagg dzop, const int *bc,
                    int inviscid)
{
    if (!h || !spC || !spE || !dzop || !bc) { set_error("invalid argument"); return PDO_ERR_ARG; }
    *h = nullptr;
    if (spE->nx != spC->nx || spE->ny != spC->ny || spE->nz != spC->nz + 1) {
        set_error("duidxj: the edge spectral object must have nz+1 planes of the cell object's grid");
        return PDO_ERR_ARG;
    }
    auto *o = new pdo_duidxj_s();
    o->spC = spC; o->spE = spE; o->dz = dzop;
    std::memcpy(o->bc, bc, sizeof(o->bc));
    o->inviscid = inviscid != 0;
    *h = o;
    return 0;
}

int pdo_duidxj_compute(pdo_duidxj h, int prec, const pdo_duidxj_fields *F, void *stream)
{
    if (!h || !F || !F->uhat || !F->vhat || !F->whatC || !F->what || !F->uEhat || !F->vEhat || !F->duidxjC ||
        !F->duidxjE || !F->duidxjChat || !F->duidxjEhat) {
        set_error("invalid argument");
        return PDO_ERR_ARG;
    }
    if (!h->inviscid && (!F->d2wdz2hatE || !F->d2udz2hatC || !F->d2vdz2hatC)) {
        set_error("viscous duidxj needs the second-derivative outputs");
        return PDO_ERR_ARG;
    }
    const int mode = prec_mode(prec);
    if (mode < 0) { set_error("unknown precision mode"); return PDO_ERR_ARG; }
    const size_t W = (mode == MODE_F64) ? 8 : 4;
    SpectralObject &C = *h->spC;
    const int nxh = C.nxh, ny = C.ny, nz = C.nz, nx = C.nx;
    const size_t rC = (size_t)nx * ny * nz * W, rE = (size_t)nx * ny * (nz + 1) * W;
    const size_t cC = (size_t)nxh * ny * nz * 2 * W, cE = (size_t)nxh * ny * (nz + 1) * 2 * W;
    auto slot = [](void *base, size_t bytes, int k) { return (void *)((char *)base + bytes * (k - 1)); };
    // igrid.F90:2575-2589 slot order: 1 dudx 2 dudy 3 dudz 4 dvdx 5 dvdy 6 dvdz 7 dwdx 8 dwdy 9 dwdz
    void *RC = F->duidxjC, *RE = F->duidxjE, *HC = F->duidxjChat, *HE = F->duidxjEhat;
    const int *bc = h->bc;
    const int uB = bc[0], uT = bc[1], vB = bc[2], vT = bc[3], wB = bc[4], wT = bc[5];
    const int dUB = bc[6], dUT = bc[7], dVB = bc[8], dVT = bc[9], dWB = bc[10], dWT = bc[11];
    int st;
#define PDO_TRY(x)            \
    do {                      \
        if ((st = (x))) return st; \
    } while (0)
    // x and y derivatives: i k then the 2D inverse transform, on cells and edges
    struct XY { const void *c, *e; int kC; };
    const XY xy[6] = {{F->uhat, F->uEhat, 1}, {F->uhat, F->uEhat, 2}, {F->vhat, F->vEhat, 4},
                      {F->vhat, F->vEhat, 5}, {F->whatC, F->what, 7}, {F->whatC, F->what, 8}};
    for (const XY &q : xy) {
        const int dir = (q.kC == 1 || q.kC == 4 || q.kC == 7) ? 1 : 2;
        PDO_TRY(pdo_spectral_mtimes_ik(h->spC, prec, dir, q.c, slot(HC, cC, q.kC), stream));
        PDO_TRY(pdo_spectral_ifft(h->spC, prec, slot(HC, cC, q.kC), slot(RC, rC, q.kC), 0, stream));
        PDO_TRY(pdo_spectral_mtimes_ik(h->spE, prec, dir, q.e, slot(HE, cE, q.kC), stream));
        PDO_TRY(pdo_spectral_ifft(h->spE, prec, slot(HE, cE, q.kC), slot(RE, rE, q.kC), 0, stream));
    }
    // dwdz (:2621-2627): E2C on the edge w, C2E interpolation of it
    PDO_TRY(pdo_pade6stagg_apply(h->dz, 10, wB, wT, prec, F->what, slot(HC, cC, 9), nxh, ny, 1, stream));
    PDO_TRY(pdo_spectral_ifft(h->spC, prec, slot(HC, cC, 9), slot(RC, rC, 9), 0, stream));
    PDO_TRY(pdo_pade6stagg_apply(h->dz, 15, dWB, dWT, prec, slot(HC, cC, 9), slot(HE, cE, 9), nxh, ny, 1, stream));
    PDO_TRY(pdo_spectral_ifft(h->spE, prec, slot(HE, cE, 9), slot(RE, rE, 9), 0, stream));
    // d2wdz2 (:2630-2633)
    if (!h->inviscid) PDO_TRY(pdo_pade6stagg_apply(h->dz, 13, wB, wT, prec, F->what, F->d2wdz2hatE, nxh, ny, 1, stream));
    // dudz, d2udz2 and dvdz, d2vdz2 (:2636-2679)
    struct Z { const void *c; int k; int b, t, db, dt; void *d2; };
    const Z zz[2] = {{F->uhat, 3, uB, uT, dUB, dUT, F->d2udz2hatC}, {F->vhat, 6, vB, vT, dVB, dVT, F->d2vdz2hatC}};
    for (const Z &q : zz) {
        void *eh = slot(HE, cE, q.k);
        PDO_TRY(pdo_pade6stagg_apply(h->dz, 11, q.b, q.t, prec, q.c, eh, nxh, ny, 1, stream));
        PDO_TRY(pdo_spectral_ifft(h->spE, prec, eh, slot(RE, rE, q.k), 0, stream));
        if (!h->inviscid) {
            // a one-sided boundary takes the second derivative as E2C(C2E(.));
            // the C2E result is the edge derivative just computed
            if (q.t == 0 || q.b == 0) {
                PDO_TRY(pdo_pade6stagg_apply(h->dz, 10, q.db, q.dt, prec, eh, q.d2, nxh, ny, 1, stream));
            } else {
                PDO_TRY(pdo_pade6stagg_apply(h->dz, 12, q.b, q.t, prec, q.c, q.d2, nxh, ny, 1, stream));
            }
        }
        PDO_TRY(pdo_pade6stagg_apply(h->dz, 14, q.db, q.dt, prec, eh, slot(HC, cC, q.k), nxh, ny, 1, stream));
        PDO_TRY(pdo_spectral_ifft(h->spC, prec, slot(HC, cC, q.k), slot(RC, rC, q.k), 0, stream));
    }
#undef PDO_TRY
    return 0;
}

int pdo_duidxj_destroy(pdo_duidxj h)
{
    delete h;
    return 0;
}

// igrid%interp_PrimitiveVars (igrid.F90:1423-1463): w to cells, u and v (and
// T) to edges through Pade6stagg, each followed by the 2D inverse transform
int pdo_igrid_interp_primitive(pdo_duidxj h, int prec, const pdo_interp_fields *F, void *stream)
{
    if (!h || !F || !F->uhat || !F->vhat || !F->what || !F->whatC || !F->wC || !F->uEhat || !F->uE || !F->vEhat ||
        !F->vE || (F->That && (!F->TEhat || !F->TE))) {
        set_error("invalid argument");
        return PDO_ERR_ARG;
    }
    const int mode = fft_mode(prec);
    if (mode < 0) { set_error("unknown precision mode"); return PDO_ERR_ARG; }
    SpectralObject &C = *h->spC;
    const int nxh = C.nxh, ny = C.ny;
    const int *bc = h->bc;
    int st;
    // Step 1: w -> wC (wBC_bottom, wBC_top)
    if ((st = pdo_pade6stagg_apply(h->dz, 14, bc[4], bc[5], prec, F->what, F->whatC, nxh, ny, 1, stream))) return st;
    if ((st = pdo_spectral_ifft(h->spC, prec, F->whatC, F->wC, 0, stream))) return st;
    // Steps 2, 3: u -> uE, v -> vE
    if ((st = pdo_pade6stagg_apply(h->dz, 15, bc[0], bc[1], prec, F->uhat, F->uEhat, nxh, ny, 1, stream))) return st;
    if ((st = pdo_spectral_ifft(h->spE, prec, F->uEhat, F->uE, 0, stream))) return st;
    if ((st = pdo_pade6stagg_apply(h->dz, 15, bc[2], bc[3], prec, F->vhat, F->vEhat, nxh, ny, 1, stream))) return st;
    if ((st = pdo_spectral_ifft(h->spE, prec, F->vEhat, F->vE, 0, stream))) return st;
    // Step 4: T -> TE; a prescribed surface temperature (botBC_Temp == 0)
    // replaces the bottom edge plane by its mean Tsurf (:1449-1454)
    if (F->That) {
        if ((st = pdo_pade6stagg_apply(h->dz, 15, F->TBC_bottom, F->TBC_top, prec, F->That, F->TEhat, nxh, ny, 1,
                                       stream)))
            return st;
        if (F->botBC_Temp == 0) {
            hipError_t e = launch_surface_plane(mode, F->TEhat, (long)nxh * ny, F->Tsurf * (double)C.nx * C.ny,
                                                (hipStream_t)stream);
            if (e != hipSuccess) return hip_fail(e, "surface temperature");
        }
        if ((st = pdo_spectral_ifft(h->spE, prec, F->TEhat, F->TE, 0, stream))) return st;
    }
    return 0;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// sgs_igrid%getRHS_SGS (sgsmod_igrid.F90:204-266) with getTauSGS (:156-201)
// for the eddy-viscosity models
// ---------------------------------------------------------------------------
namespace pdo {

struct SgsRhsObject {
    pdo_spectral spC = nullptr, spE = nullptr;
    pdo_pade6stagg dz = nullptr;
    pdo_sgs_params p{};
    double *d_cC = nullptr, *d_cE = nullptr;  // planar model constants (cmodelC, cmodelE), device
    size_t nC = 0, nE = 0, sC = 0, sE = 0;     // points (cells, edges), spectral values (cells, edges)
    int wprec = -1;                            // precision the work buffers are sized for
    std::vector<void *> bufs;
    void *SC = nullptr, *SE = nullptr, *nuC = nullptr, *nuE = nullptr, *tauC = nullptr, *tauE = nullptr;
    void *cb1 = nullptr, *cb2 = nullptr;
    // learned closure (pdo_sgsrhs_set_closure): its stress replaces the
    // eddy-viscosity stress, which stays where the closure's dissipation is
    // negative when `fallback` is set
    pdo_closure closure = nullptr;
    double cl_delta = 0, cl_dratio = 1;
    int cl_fallback = 0;
    // Moeng wall model: block sums and {Uspmn, ustar, WallMFactor, Umn}
    double *d_wpart = nullptr, *d_wm = nullptr;
    static constexpr int WALL_BLOCKS = 256;
    void free_bufs()
    {
        for (void *b : bufs) (void)hipFree(b);
        bufs.clear();
    }
    ~SgsRhsObject()
    {
        free_bufs();
        if (d_cC) (void)hipFree(d_cC);
        if (d_cE) (void)hipFree(d_cE);
        if (d_wpart) (void)hipFree(d_wpart);
        if (d_wm) (void)hipFree(d_wm);
    }
    int alloc(int prec)
    {
        if (wprec == prec) return 0;
        free_bufs();
        const size_t W = prec == PDO_F64 ? 8 : 4;
        const size_t sizes[8] = {6 * nC * W, 6 * nE * W, nC * W, nE * W, 4 * nC * W, 2 * nE * W, 2 * sC * W, 2 * sE * W};
        void **dst[8] = {&SC, &SE, &nuC, &nuE, &tauC, &tauE, &cb1, &cb2};
        for (int i = 0; i < 8; i++) {
            hipError_t e = hipMalloc(dst[i], sizes[i]);
            if (e != hipSuccess) return hip_fail(e, "hipMalloc(sgs work)");
            bufs.push_back(*dst[i]);
        }
        wprec = prec;
        return 0;
    }
};

}  // namespace pdo

struct pdo_sgsrhs_s : public pdo::SgsRhsObject {};

extern "C" {

int pdo_sgsrhs_init(pdo_sgsrhs *h, pdo_spectral spC, pdo_spectral spE, pdo_pade6stagg dzop, const pdo_sgs_params *p,
                    int prec)
{
    if (!h || !spC || !spE || !dzop || !p) { set_error("invalid argument"); return PDO_ERR_ARG; }
    *h = nullptr;
    if (p->model < 0 || p->model > 2) {
        set_error("eddy-viscosity model: 0 Smagorinsky, 1 sigma or 2 AMD");
        return PDO_ERR_ARG;
    }
    if (p->wall_model != 0) {
        // the Bou-Zeid model (2) filters with spectral%GsurfaceFilter, which
        // the reference never sets up, so only the Moeng model is offered
        if (p->wall_model != 1) { set_error("wall model: 0 none or 1 Moeng"); return PDO_ERR_ARG; }
        if (pade6stagg_periodic(dzop)) {
            set_error("a wall model needs walls in z (a non-periodic Pade6stagg)");
            return PDO_ERR_ARG;
        }
        if (!(p->z0 > 0) || !(spC->dz > 2 * p->z0)) {
            set_error("wall model: need 0 < z0 < dz / 2");
            return PDO_ERR_ARG;
        }
    }
    if (spE->nx != spC->nx || spE->ny != spC->ny || spE->nz != spC->nz + 1) {
        set_error("sgs: the edge spectral object must have nz+1 planes of the cell object's grid");
        return PDO_ERR_ARG;
    }
    if (prec_mode(prec) < 0) { set_error("unknown precision mode"); return PDO_ERR_ARG; }
    auto *o = new pdo_sgsrhs_s();
    o->spC = spC; o->spE = spE; o->dz = dzop; o->p = *p;
    o->p.cmodelC = nullptr; o->p.cmodelE = nullptr;
    const int nx = spC->nx, ny = spC->ny, nz = spC->nz;
    o->nC = (size_t)nx * ny * nz; o->nE = (size_t)nx * ny * (nz + 1);
    o->sC = (size_t)spC->nxh * ny * nz; o->sE = (size_t)spC->nxh * ny * (nz + 1);
    hipError_t e = hipSuccess;
    if (p->cmodelC) {
        e = hipMalloc(&o->d_cC, nz * sizeof(double));
        if (e == hipSuccess) e = hipMemcpy(o->d_cC, p->cmodelC, nz * sizeof(double), hipMemcpyHostToDevice);
    }
    if (e == hipSuccess && p->cmodelE) {
        e = hipMalloc(&o->d_cE, (nz + 1) * sizeof(double));
        if (e == hipSuccess) e = hipMemcpy(o->d_cE, p->cmodelE, (nz + 1) * sizeof(double), hipMemcpyHostToDevice);
    }
    if (e == hipSuccess && p->wall_model) {
        e = hipMalloc(&o->d_wpart, 2 * SgsRhsObject::WALL_BLOCKS * sizeof(double));
        if (e == hipSuccess) e = hipMalloc(&o->d_wm, 4 * sizeof(double));
    }
    if (e != hipSuccess) { delete o; return hip_fail(e, "sgs model constants"); }
    if (int st = o->alloc(prec)) { delete o; return st; }
    *h = o;
    return 0;
}

static int sgsrhs_compute(pdo_sgsrhs h, int prec, const void *duidxjC, const void *duidxjE, const double *lambda,
                          const void *u, const void *v, void *urhs, void *vrhs, void *wrhs, void *stream)
{
    if (!h || !duidxjC || !duidxjE || !urhs || !vrhs || !wrhs) { set_error("invalid argument"); return PDO_ERR_ARG; }
    if (h->p.wall_model && (!u || !v)) {
        set_error("sgs: the wall model needs the first-level velocity (pdo_sgsrhs_compute_wall)");
        return PDO_ERR_ARG;
    }
    const int mode = fft_mode(prec);
    if (mode < 0) { set_error("unknown precision mode"); return PDO_ERR_ARG; }
    if (h->wprec != prec) {
        set_error("sgs: the object's work buffers were sized for another precision (pass it to init)");
        return PDO_ERR_ARG;
    }
    hipStream_t st = (hipStream_t)stream;
    SpectralObject &C = *h->spC, &E = *h->spE;
    const int nx = C.nx, ny = C.ny, nz = C.nz, nxh = C.nxh;
    const long nxy = (long)nx * ny;
    const pdo_sgs_params &p = h->p;
    int r;
#define PDO_TRY(x)                \
    do {                          \
        if ((r = (x))) return r;  \
    } while (0)
#define PDO_HIP(x, what)                                        \
    do {                                                        \
        hipError_t ee = (x);                                    \
        if (ee != hipSuccess) return hip_fail(ee, what);        \
    } while (0)
    // Step 0/1: S_ij and the model kernel on cells; on edges S_ij and, with
    // explicitCalcEdgeEddyViscosity, the kernel too
    const bool edge_kernel = p.explicit_edge && !lambda;
    PDO_TRY(pdo_sgs_kernel(prec, p.model, duidxjC, h->SC, h->nuC, (long)h->nC, p.camd_x, p.camd_y, p.camd_z, stream));
    PDO_TRY(pdo_sgs_kernel(prec, p.model, duidxjE, h->SE, edge_kernel ? h->nuE : nullptr, (long)h->nE, p.camd_x,
                           p.camd_y, p.camd_z, stream));
    // Step 3: multiply_by_model_constant (eddyViscosity.F90:74-103)
    if (lambda) {
        // the dynamic procedure's planar coefficient, then C2E interpolation
        PDO_TRY(pdo_sgs_scale(prec, h->nuC, nxy, nz, lambda, 0.0, stream));
        PDO_TRY(pdo_pade6stagg_apply(h->dz, 15, 0, 0, prec, h->nuC, h->nuE, nx, ny, 0, stream));
    } else {
        PDO_TRY(pdo_sgs_scale(prec, h->nuC, nxy, nz, h->d_cC, p.cglobal, stream));
        if (p.explicit_edge) {
            PDO_TRY(pdo_sgs_scale(prec, h->nuE, nxy, nz + 1, h->d_cE, p.cglobal, stream));
        } else {
            // interpolate_eddy_viscosity(checknegative = .true.)
            PDO_TRY(pdo_pade6stagg_apply(h->dz, 15, 0, 0, prec, h->nuC, h->nuE, nx, ny, 0, stream));
            PDO_HIP(launch_clip_negative(mode, h->nuE, (long)h->nE, st), "clip");
        }
    }
    // tau_ij = -2 nu S_ij
    PDO_HIP(launch_tau(mode, h->nuC, h->SC, h->nuE, h->SE, h->tauC, h->tauE, (long)h->nC, (long)h->nE, st), "tau");
    if (h->closure) {
        // the learned stress on cells (scale from the cell gradients) and on
        // edges (the same scale)
        PDO_TRY(pdo_closure_stress(h->closure, prec, duidxjC, (long)h->nC, h->cl_delta, h->cl_dratio, 0,
                                   h->cl_fallback, 1, h->tauC, stream));
        PDO_TRY(pdo_closure_stress(h->closure, prec, duidxjE, (long)h->nE, h->cl_delta, h->cl_dratio, 1,
                                   h->cl_fallback, 0, h->tauE, stream));
    }
    // computeWallStress and embed_WM_stress (sgsmod_igrid.F90:164, 197): the
    // wall stresses on the bottom edge plane of tau13, tau23
    if (p.wall_model)
        PDO_HIP(launch_wall_model(mode, u, v, h->tauE, nxy, (long)h->nE, h->d_wpart, SgsRhsObject::WALL_BLOCKS, 0.40,
                                  std::log(C.dz / 2.0 / p.z0), h->d_wm, st),
                "wall model");
    const size_t W = prec == PDO_F64 ? 8 : 4;
    auto tc = [&](int q) { return (const void *)((const char *)h->tauC + (size_t)q * h->nC * W); };
    auto te = [&](int q) { return (const void *)((const char *)h->tauE + (size_t)q * h->nE * W); };
    const long nsC = (long)h->sC, nsE = (long)h->sE;
    // ddx(tau11) -> urhs; ddy(tau22) -> vrhs
    PDO_TRY(pdo_spectral_fft(h->spC, prec, tc(0), h->cb1, stream));
    PDO_HIP(launch_sub(mode, urhs, h->cb1, C.d_k1, nxh, ny, nsC, 1, st), "rhs");
    PDO_TRY(pdo_spectral_fft(h->spC, prec, tc(2), h->cb1, stream));
    PDO_HIP(launch_sub(mode, vrhs, h->cb1, C.d_k2, nxh, ny, nsC, 2, st), "rhs");
    // ddz(tau33) -> wrhs (ddz_C2E with (0, 0) ends)
    PDO_TRY(pdo_spectral_fft(h->spC, prec, tc(3), h->cb1, stream));
    PDO_TRY(pdo_pade6stagg_apply(h->dz, 11, 0, 0, prec, h->cb1, h->cb2, nxh, ny, 1, stream));
    PDO_HIP(launch_sub(mode, wrhs, h->cb2, nullptr, nxh, ny, nsE, 0, st), "rhs");
    // ddx(tau12) -> vrhs, ddy(tau12) -> urhs
    PDO_TRY(pdo_spectral_fft(h->spC, prec, tc(1), h->cb1, stream));
    PDO_HIP(launch_sub(mode, vrhs, h->cb1, C.d_k1, nxh, ny, nsC, 1, st), "rhs");
    PDO_HIP(launch_sub(mode, urhs, h->cb1, C.d_k2, nxh, ny, nsC, 2, st), "rhs");
    // ddz(tau13) -> urhs, ddx(tau13) -> wrhs; ddz(tau23) -> vrhs, ddy(tau23) -> wrhs
    for (int q = 0; q < 2; q++) {
        PDO_TRY(pdo_spectral_fft(h->spE, prec, te(q), h->cb2, stream));
        PDO_TRY(pdo_pade6stagg_apply(h->dz, 10, 0, 0, prec, h->cb2, h->cb1, nxh, ny, 1, stream));
        PDO_HIP(launch_sub(mode, q == 0 ? urhs : vrhs, h->cb1, nullptr, nxh, ny, nsC, 0, st), "rhs");
        PDO_HIP(launch_sub(mode, wrhs, h->cb2, q == 0 ? E.d_k1 : E.d_k2, nxh, ny, nsE, q + 1, st), "rhs");
    }
#undef PDO_TRY
#undef PDO_HIP
    return 0;
}

int pdo_sgsrhs_compute(pdo_sgsrhs h, int prec, const void *duidxjC, const void *duidxjE, const double *lambda,
                       void *urhs, void *vrhs, void *wrhs, void *stream)
{
    return sgsrhs_compute(h, prec, duidxjC, duidxjE, lambda, nullptr, nullptr, urhs, vrhs, wrhs, stream);
}

int pdo_sgsrhs_compute_wall(pdo_sgsrhs h, int prec, const void *duidxjC, const void *duidxjE, const double *lambda,
                            const void *u, const void *v, void *urhs, void *vrhs, void *wrhs, void *stream)
{
    return sgsrhs_compute(h, prec, duidxjC, duidxjE, lambda, u, v, urhs, vrhs, wrhs, stream);
}

int pdo_sgsrhs_wall_state(pdo_sgsrhs h, double *state)
{
    if (!h || !state) { set_error("invalid argument"); return PDO_ERR_ARG; }
    if (!h->d_wm) { set_error("sgs: no wall model"); return PDO_ERR_ARG; }
    hipError_t e = hipMemcpy(state, h->d_wm, 4 * sizeof(double), hipMemcpyDeviceToHost);
    return e == hipSuccess ? 0 : hip_fail(e, "wall state");
}

int pdo_sgsrhs_set_closure(pdo_sgsrhs h, pdo_closure c, double delta, double dratio, int fallback)
{
    if (!h) { set_error("invalid argument"); return PDO_ERR_ARG; }
    if (c && (long)h->nE > closure_max_points(c)) {
        set_error("sgs: the closure object holds fewer points than the edge field");
        return PDO_ERR_ARG;
    }
    h->closure = c;
    h->cl_delta = delta;
    h->cl_dratio = dratio;
    h->cl_fallback = fallback ? 1 : 0;
    return 0;
}

int pdo_sgsrhs_fields(pdo_sgsrhs h, const void **tauC, const void **tauE, const void **nuC, const void **nuE)
{
    if (!h) return PDO_ERR_ARG;
    if (tauC) *tauC = h->tauC;
    if (tauE) *tauE = h->tauE;
    if (nuC) *nuC = h->nuC;
    if (nuE) *nuE = h->nuE;
    return 0;
}

int pdo_sgsrhs_destroy(pdo_sgsrhs h)
{
    delete h;
    return 0;
}

}  // extern "C"
