// Time advancement of type(igrid) on one GPU, for domains periodic in z (the
// forced isotropic turbulence setup) or with walls in z (the channel and
// boundary-layer setups): src/incompressible/igrid.F90
//   timeAdvance :1056-1076 -> TVD_RK3 :1105-1170 / SSP_RK45 :1172-1299
//   populate_rhs :1793-1912: AddNonLinearTerm_Rot :1497-1566, addViscousTerm
//     :1914-1959, sgs_igrid%getRHS_SGS, HIT_shell_forcing%getRHS_HITforcing
//   project_and_prep :1961-1992: dealiasFields :1020-1037, PadePoisson
//     PressureProjection, the inverse 2D transforms, interp_PrimitiveVars,
//     compute_duidxj
//   compute_deltaT :1374-1400, wrapup_timestep :2067- (time and step count)
// The state lives in device memory for the life of the object; a step
// launches only kernels (and, with a CFL number, reads one maximum back).
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/padeops_hip.h"
#include "pdo_igrid.hpp"
#include "pdo_ops.hpp"
#include "pdo_poisson.hpp"
#include "pdo_spectral.hpp"

namespace pdo {

struct IgridObject {
    pdo_igrid_params p{};
    int prec = PDO_F64, mode = 0;
    size_t W = 8;
    int nx = 0, ny = 0, nz = 0, nxh = 0;
    double dx = 1, dy = 1, dz = 1;
    pdo_spectral spC = nullptr, spE = nullptr;
    pdo_pade6stagg d6 = nullptr;
    pdo_padepoisson pp = nullptr;
    pdo_duidxj dj = nullptr;
    pdo_sgsrhs sgs = nullptr;
    // SGS closure kind (pdo_igrid_set_closure): 0 eddy viscosity with the
    // init constants, 1 dynamic Smagorinsky, 2 learned + dynamic fallback,
    // 3 learned; the dynamic procedure's strain, kernel and coefficient
    int ckind = 0;
    pdo_dynproc dyn = nullptr;
    pdo_spectral spD = nullptr;  // the dynamic procedure's (x-y) spectral object
    void *dS = nullptr, *dD = nullptr;
    double *dlam = nullptr;
    std::vector<void *> bufs;
    // sizes in real words: physical cells / edges, spectra (complex values
    // counted twice) of cells / edges
    size_t nC = 0, nE = 0, sC = 0, sE = 0;
    // state and stages (spectral): index 0 the state, 1..3 the extra stages
    void *uh[4] = {}, *vh[4] = {}, *wh[4] = {};
    void *ur = nullptr, *vr = nullptr, *wr = nullptr;     // rhs
    void *ur5 = nullptr, *vr5 = nullptr, *wr5 = nullptr;  // SSP-RK45 stage-5 rhs
    // prepared fields of the current stage
    void *u = nullptr, *v = nullptr, *wC = nullptr, *w = nullptr, *uE = nullptr, *vE = nullptr;
    void *whC = nullptr, *uEh = nullptr, *vEh = nullptr;
    void *gC = nullptr, *gE = nullptr, *gCh = nullptr, *gEh = nullptr;
    void *d2w = nullptr, *d2u = nullptr, *d2v = nullptr;
    void *tC1 = nullptr, *tE1 = nullptr, *fC1 = nullptr, *fE1 = nullptr;  // nonlinear-term work
    double *k3 = nullptr;
    double *partial = nullptr;
    int nblocks = 1024;
    // the stage fields prepared by project_and_prep belong to these spectra
    void *cur_u = nullptr, *cur_v = nullptr, *cur_w = nullptr;
    double tsim = 0, dt = 0;
    long step = 0;
    bool newTimeStep = true;
    unsigned long long draws = 0;

    ~IgridObject()
    {
        for (void *b : bufs) (void)hipFree(b);
        if (sgs) pdo_sgsrhs_destroy(sgs);
        if (dyn) pdo_dynproc_destroy(dyn);
        if (spD) pdo_spectral_destroy(spD);
        if (dj) pdo_duidxj_destroy(dj);
        if (pp) pdo_padepoisson_destroy(pp);
        if (d6) pdo_pade6stagg_destroy(d6);
        if (spE) pdo_spectral_destroy(spE);
        if (spC) pdo_spectral_destroy(spC);
    }
    int alloc(void **p, size_t bytes)
    {
        hipError_t e = hipMalloc(p, bytes);
        if (e != hipSuccess) return hip_fail(e, "hipMalloc(igrid state)");
        bufs.push_back(*p);
        return 0;
    }
};

}  // namespace pdo

struct pdo_igrid_s : public pdo::IgridObject {};

using namespace pdo;

namespace {

#define PDO_TRY(x)                  \
    do {                            \
        if (int r_ = (x)) return r_; \
    } while (0)
#define PDO_HIP(x, what)                                 \
    do {                                                 \
        hipError_t e_ = (x);                             \
        if (e_ != hipSuccess) return hip_fail(e_, what); \
    } while (0)

// project_and_prep (igrid.F90:1961-1992) for the spectra (uh, vh, wh)
int project_and_prep(pdo_igrid h, void *uh, void *vh, void *wh, hipStream_t st)
{
    void *s = (void *)st;
    // Step 1: dealiasFields (:1020-1037; periodic in z: the 3D mask on cells,
    // the edge variant on w; with walls the x-y mask of each object)
    PDO_TRY(pdo_spectral_dealias(h->spC, h->prec, uh, s));
    PDO_TRY(pdo_spectral_dealias(h->spC, h->prec, vh, s));
    if (h->p.walls) PDO_TRY(pdo_spectral_dealias(h->spE, h->prec, wh, s));
    else PDO_TRY(pdo_spectral_dealias_edge(h->spC, h->prec, wh, s));
    // Step 2: pressure projection
    PDO_TRY(pdo_padepoisson_projection(h->pp, h->prec, uh, vh, wh, s));
    // Step 3: physical fields
    PDO_TRY(pdo_spectral_ifft(h->spC, h->prec, uh, h->u, 0, s));
    PDO_TRY(pdo_spectral_ifft(h->spC, h->prec, vh, h->v, 0, s));
    PDO_TRY(pdo_spectral_ifft(h->spE, h->prec, wh, h->w, 0, s));
    // Step 4: interp_PrimitiveVars
    pdo_interp_fields I{};
    I.uhat = uh; I.vhat = vh; I.what = wh;
    I.whatC = h->whC; I.wC = h->wC; I.uEhat = h->uEh; I.uE = h->uE; I.vEhat = h->vEh; I.vE = h->vE;
    PDO_TRY(pdo_igrid_interp_primitive(h->dj, h->prec, &I, s));
    // Step 5: compute_duidxj
    pdo_duidxj_fields D{};
    D.uhat = uh; D.vhat = vh; D.whatC = h->whC; D.what = wh; D.uEhat = h->uEh; D.vEhat = h->vEh;
    D.duidxjC = h->gC; D.duidxjE = h->gE; D.duidxjChat = h->gCh; D.duidxjEhat = h->gEh;
    D.d2wdz2hatE = h->d2w; D.d2udz2hatC = h->d2u; D.d2vdz2hatC = h->d2v;
    PDO_TRY(pdo_duidxj_compute(h->dj, h->prec, &D, s));
    h->cur_u = uh; h->cur_v = vh; h->cur_w = wh;
    return 0;
}

// populate_rhs (igrid.F90:1793-1912) of the prepared stage into (ur, vr, wr)
int populate_rhs(pdo_igrid h, void *ur, void *vr, void *wr, hipStream_t st)
{
    void *s = (void *)st;
    const int mode = h->mode;
    const size_t W = h->W;
    auto g = [&](void *base, size_t n, int slot) { return (const void *)((const char *)base + (slot - 1) * n * W); };
    const size_t nC = h->nC, nE = h->nE;
    const long nsC = (long)(h->sC / 2), nsE = (long)(h->sE / 2);
    // Step 1: AddNonLinearTerm_Rot (:1497-1566)
    //   u: T1C = (dvdx - dudy) v, T2E = (dwdx - dudz) w; urhs = E2C(fft T2E) + fft T1C
    PDO_HIP(launch_cross(mode, g(h->gC, nC, 4), g(h->gC, nC, 2), h->v, nullptr, nullptr, nullptr, h->tC1, (long)nC, st),
            "nonlinear");
    PDO_TRY(pdo_spectral_fft(h->spC, h->prec, h->tC1, h->fC1, s));
    PDO_HIP(launch_cross(mode, g(h->gE, nE, 7), g(h->gE, nE, 3), h->w, nullptr, nullptr, nullptr, h->tE1, (long)nE, st),
            "nonlinear");
    PDO_TRY(pdo_spectral_fft(h->spE, h->prec, h->tE1, h->fE1, s));
    PDO_TRY(pdo_pade6stagg_apply(h->d6, 14, 0, 0, h->prec, h->fE1, ur, h->nxh, h->ny, 1, s));
    PDO_HIP(launch_add(mode, ur, h->fC1, (long)h->sC, st), "nonlinear");
    //   v: T1C = (dudy - dvdx) u, T2E = (dwdy - dvdz) w
    PDO_HIP(launch_cross(mode, g(h->gC, nC, 2), g(h->gC, nC, 4), h->u, nullptr, nullptr, nullptr, h->tC1, (long)nC, st),
            "nonlinear");
    PDO_TRY(pdo_spectral_fft(h->spC, h->prec, h->tC1, h->fC1, s));
    PDO_HIP(launch_cross(mode, g(h->gE, nE, 8), g(h->gE, nE, 6), h->w, nullptr, nullptr, nullptr, h->tE1, (long)nE, st),
            "nonlinear");
    PDO_TRY(pdo_spectral_fft(h->spE, h->prec, h->tE1, h->fE1, s));
    PDO_TRY(pdo_pade6stagg_apply(h->d6, 14, 0, 0, h->prec, h->fE1, vr, h->nxh, h->ny, 1, s));
    PDO_HIP(launch_add(mode, vr, h->fC1, (long)h->sC, st), "nonlinear");
    //   w: T1E = (dudz - dwdx) uE + (dvdz - dwdy) vE; wrhs = fft T1E
    PDO_HIP(launch_cross(mode, g(h->gE, nE, 3), g(h->gE, nE, 7), h->uE, g(h->gE, nE, 6), g(h->gE, nE, 8), h->vE, h->tE1,
                         (long)nE, st),
            "nonlinear");
    PDO_TRY(pdo_spectral_fft(h->spE, h->prec, h->tE1, wr, s));
    // Step 5: addViscousTerm (:1914-1959)
    if (h->p.Re > 0) {
        const double invRe = 1.0 / h->p.Re;
        PDO_HIP(launch_visc(mode, ur, h->cur_u, h->d2u, h->spC->d_k1, h->spC->d_k2, h->nxh, h->ny, nsC, invRe, st),
                "viscous");
        PDO_HIP(launch_visc(mode, vr, h->cur_v, h->d2v, h->spC->d_k1, h->spC->d_k2, h->nxh, h->ny, nsC, invRe, st),
                "viscous");
        PDO_HIP(launch_visc(mode, wr, h->cur_w, h->d2w, h->spE->d_k1, h->spE->d_k2, h->nxh, h->ny, nsE, invRe, st),
                "viscous");
    }
    // Step 6: SGS stress divergence (with the dynamic procedure's planar
    // coefficient of this stage for closure kinds 1 and 2)
    if (h->sgs) {
        const double *lam = nullptr;
        if (h->ckind == 1 || h->ckind == 2) {
            PDO_TRY(pdo_sgs_kernel(h->prec, 0, h->gC, h->dS, h->dD, (long)nC, 0.0, 0.0, 0.0, s));
            PDO_TRY(pdo_dynproc_compute(h->dyn, h->prec, h->u, h->v, h->wC, h->cur_u, h->cur_v, h->whC, h->dS, h->dD,
                                        h->dlam, s));
            lam = h->dlam;
        }
        if (h->p.sgs.wall_model)
            PDO_TRY(pdo_sgsrhs_compute_wall(h->sgs, h->prec, h->gC, h->gE, lam, h->u, h->v, ur, vr, wr, s));
        else
            PDO_TRY(pdo_sgsrhs_compute(h->sgs, h->prec, h->gC, h->gE, lam, ur, vr, wr, s));
    }
    // Step 3a: addExtraForcingTerm (:1759-1765): the transform of the
    // uniform dpFdx, i.e. dpFdx nx ny on the mean mode of every plane of u
    if (h->p.dpFdx != 0.0)
        PDO_HIP(launch_mean_add(mode, ur, (long)h->nxh * h->ny, h->nz, h->p.dpFdx * h->nx * h->ny, st), "forcing");
    // Step 8: HIT forcing; new wave vectors at the first stage of a step
    if (h->p.use_forcing) {
        if (h->newTimeStep) h->draws++;
        HitForcingArgs F{};
        F.nxh = h->nxh; F.ny = h->ny; F.nz = h->nz; F.nwaves = h->p.nwaves;
        F.kmin = h->p.kmin; F.kmax = h->p.kmax; F.eps = h->p.eps_amplitude;
        const double N = (double)h->nx * h->ny * h->nz;
        F.normfact = N * N;
        F.dz = h->dz; F.seed = h->p.seed; F.draw = h->draws; F.k3 = h->k3;
        PDO_HIP(launch_hit_forcing(mode, h->cur_u, h->cur_v, h->cur_w, ur, vr, wr, F, st), "HIT forcing");
    }
    return 0;
}

// compute_deltaT (:1374-1400)
int compute_dt(pdo_igrid h, hipStream_t st)
{
    if (!(h->p.CFL > 0)) {
        h->dt = h->p.dt;
        return 0;
    }
    PDO_HIP(launch_cfl(h->mode, h->u, h->v, h->wC, (long)h->nC, 1.0 / h->dx, 1.0 / h->dy, 1.0 / h->dz, h->partial,
                       h->nblocks, st),
            "CFL");
    std::vector<double> part(h->nblocks);
    PDO_HIP(hipMemcpyAsync(part.data(), h->partial, part.size() * sizeof(double), hipMemcpyDeviceToHost, st),
            "CFL readback");
    PDO_HIP(hipStreamSynchronize(st), "CFL readback");
    double m = 0;
    for (double x : part) m = x > m ? x : m;
    h->dt = h->p.CFL / m;
    if (h->p.Re > 0) {
        const double dmin = std::fmin(h->dx, std::fmin(h->dy, h->dz));
        h->dt = std::fmin(h->dt, h->p.CFL * h->p.Re * dmin * dmin);
    }
    return 0;
}

// one linear combination per velocity component; x[q] and c[q] of the
// cell components, the edge component's operands follow the same pattern
int combine(pdo_igrid h, void *const out[3], const std::vector<void *const *> &terms, const std::vector<double> &c,
            hipStream_t st)
{
    for (int comp = 0; comp < 3; comp++) {
        const void *x[5];
        for (size_t q = 0; q < terms.size(); q++) x[q] = terms[q][comp];
        const long n = (long)(comp == 2 ? h->sE : h->sC);
        PDO_HIP(launch_lincomb(h->mode, out[comp], x, c.data(), (int)terms.size(), n, st), "stage update");
    }
    return 0;
}

int ssp_rk45(pdo_igrid h, hipStream_t st)
{
    // igrid.F90:1176-1183
    const double b01 = 0.39175222657189, b12 = 0.368410593050371, b23 = 0.25189177427169, b34 = 0.54497475022852;
    const double b35 = 0.06369246866629, b45 = 0.22600748323690;
    const double a20 = 0.444370493651235, a21 = 0.555629506348765;
    const double a30 = 0.620101851488403, a32 = 0.379898148511597;
    const double a40 = 0.17807995439313, a43 = 0.821920045606868;
    const double a52 = 0.517231671970585, a53 = 0.096059710526147, a54 = 0.386708617503269;
    PDO_TRY(compute_dt(h, st));
    const double dt = h->dt;
    void *S0[3] = {h->uh[0], h->vh[0], h->wh[0]}, *S1[3] = {h->uh[1], h->vh[1], h->wh[1]};
    void *S2[3] = {h->uh[2], h->vh[2], h->wh[2]}, *S3[3] = {h->uh[3], h->vh[3], h->wh[3]};
    void *R[3] = {h->ur, h->vr, h->wr}, *R5[3] = {h->ur5, h->vr5, h->wr5};
    // stage 1
    PDO_TRY(populate_rhs(h, h->ur, h->vr, h->wr, st));
    h->newTimeStep = false;
    PDO_TRY(combine(h, S1, {S0, R}, {1.0, b01 * dt}, st));
    PDO_TRY(project_and_prep(h, S1[0], S1[1], S1[2], st));
    // stage 2
    PDO_TRY(populate_rhs(h, h->ur, h->vr, h->wr, st));
    PDO_TRY(combine(h, S2, {S0, S1, R}, {a20, a21, b12 * dt}, st));
    PDO_TRY(project_and_prep(h, S2[0], S2[1], S2[2], st));
    // stage 3
    PDO_TRY(populate_rhs(h, h->ur, h->vr, h->wr, st));
    PDO_TRY(combine(h, S3, {S0, S2, R}, {a30, a32, b23 * dt}, st));
    PDO_TRY(project_and_prep(h, S3[0], S3[1], S3[2], st));
    // stage 4 (u4 overwrites the state in place)
    PDO_TRY(populate_rhs(h, h->ur, h->vr, h->wr, st));
    PDO_TRY(combine(h, S0, {S0, S3, R}, {a40, a43, b34 * dt}, st));
    PDO_TRY(project_and_prep(h, S0[0], S0[1], S0[2], st));
    // stage 5: rhs of u4 into the spare buffers; R still holds L(u3)
    PDO_TRY(populate_rhs(h, h->ur5, h->vr5, h->wr5, st));
    PDO_TRY(combine(h, S0, {S2, S3, R, S0, R5}, {a52, a53, b35 * dt, a54, b45 * dt}, st));
    PDO_TRY(project_and_prep(h, S0[0], S0[1], S0[2], st));
    return 0;
}

int tvd_rk3(pdo_igrid h, hipStream_t st)
{
    PDO_TRY(compute_dt(h, st));
    const double dt = h->dt;
    void *S0[3] = {h->uh[0], h->vh[0], h->wh[0]}, *S1[3] = {h->uh[1], h->vh[1], h->wh[1]};
    void *R[3] = {h->ur, h->vr, h->wr};
    PDO_TRY(populate_rhs(h, h->ur, h->vr, h->wr, st));
    h->newTimeStep = false;
    PDO_TRY(combine(h, S1, {S0, R}, {1.0, dt}, st));
    PDO_TRY(project_and_prep(h, S1[0], S1[1], S1[2], st));
    PDO_TRY(populate_rhs(h, h->ur, h->vr, h->wr, st));
    PDO_TRY(combine(h, S1, {S0, S1, R}, {3.0 / 4.0, 1.0 / 4.0, (1.0 / 4.0) * dt}, st));
    PDO_TRY(project_and_prep(h, S1[0], S1[1], S1[2], st));
    PDO_TRY(populate_rhs(h, h->ur, h->vr, h->wr, st));
    PDO_TRY(combine(h, S0, {S0, S1, R}, {1.0 / 3.0, 2.0 / 3.0, (2.0 / 3.0) * dt}, st));
    PDO_TRY(project_and_prep(h, S0[0], S0[1], S0[2], st));
    return 0;
}

}  // namespace

extern "C" {

int pdo_igrid_init(pdo_igrid *out, const pdo_igrid_params *p, int prec)
{
    if (!out || !p) { set_error("invalid argument"); return PDO_ERR_ARG; }
    *out = nullptr;
    const int mode = fft_mode(prec);
    if (mode < 0) { set_error("unknown precision mode"); return PDO_ERR_ARG; }
    if (p->scheme != 1 && p->scheme != 2) { set_error("timeSteppingScheme: 1 TVD-RK3 or 2 SSP-RK45"); return PDO_ERR_ARG; }
    if (!(p->CFL > 0) && !(p->dt > 0)) { set_error("a CFL number or a fixed time step is needed"); return PDO_ERR_ARG; }
    if (p->walls) {
        if (p->botWall < 1 || p->botWall > 3 || p->topWall < 1 || p->topWall > 2) {
            set_error("walls: botWall 1 no-slip, 2 slip or 3 wall model; topWall 1 or 2");
            return PDO_ERR_ARG;
        }
        if ((p->botWall == 3) != (p->use_sgs && p->sgs.wall_model == 1)) {
            set_error("walls: a wall-model bottom wall (botWall 3) goes with the SGS wall model, and only then");
            return PDO_ERR_ARG;
        }
        if (p->use_forcing) { set_error("walls: HIT forcing needs periodic z"); return PDO_ERR_ARG; }
    }
    if (p->use_forcing) {
        const int kc = (int)std::ceil(p->kmax);
        if (p->nwaves < 1 || !(p->kmin >= 0) || !(p->kmax >= p->kmin) || kc >= p->nx / 2 || kc >= p->ny ||
            kc >= p->nz / 2) {
            set_error("HIT forcing: nwaves >= 1 and 0 <= kmin <= kmax with ceil(kmax) below nx/2, ny and nz/2");
            return PDO_ERR_ARG;
        }
    }
    auto *h = new pdo_igrid_s();
    h->p = *p;
    h->p.sgs.cmodelC = nullptr;
    h->p.sgs.cmodelE = nullptr;
    h->prec = prec; h->mode = mode; h->W = (prec == PDO_F64) ? 8 : 4;
    h->nx = p->nx; h->ny = p->ny; h->nz = p->nz; h->nxh = p->nx / 2 + 1;
    h->dx = p->Lx / p->nx; h->dy = p->Ly / p->ny; h->dz = p->Lz / p->nz;
    int st = 0;
    // igrid.F90:489-493 (periodic in z: the cell object carries the 3D dealias
    // mask), Pade6stagg and PadePoisson periodic or with walls
    const int per = p->walls ? 0 : 1;
    if (!st) st = pdo_spectral_init(&h->spC, h->nx, h->ny, h->nz, h->dx, h->dy, h->dz, 0, 2.0 / 3.0, per);
    if (!st) st = pdo_spectral_init(&h->spE, h->nx, h->ny, h->nz + 1, h->dx, h->dy, h->dz, 0, 2.0 / 3.0, 0);
    if (!st) st = pdo_pade6stagg_init(&h->d6, h->nz, h->dz, per);
    if (!st) {
        if (per) {
            st = pdo_padepoisson_init_periodic(&h->pp, h->spC, h->dz);
        } else {
            st = pdo_padepoisson_init_nonperiodic(&h->pp, h->spC, h->dz);
            if (!st && p->stokes_pressure) st = pdo_padepoisson_set_stokes(h->pp, p->Lz);
        }
    }
    if (!st) {
        // boundary kinds of get_boundary_conditions_stencil (:5148-5220):
        // {uBC, vBC, wBC, dUdzBC, dVdzBC, dWdzBC} x {bottom, top}; slip walls
        // for periodic z (the setting of the periodic isotropic runs, where
        // they only select d2dz2_C2C for the second z derivatives of u, v)
        int bc[12] = {1, 1, 1, 1, -1, -1, -1, -1, -1, -1, 0, 0};
        if (p->walls) {
            for (int side = 0; side < 2; side++) {
                const int wall = side == 0 ? p->botWall : p->topWall;
                int uB, wB, dUB, dWB;
                if (wall == 1) { uB = -1; wB = 1; dUB = 0; dWB = -1; }      // no-slip
                else if (wall == 2) { uB = 1; wB = -1; dUB = -1; dWB = 0; }  // slip
                else { uB = 0; wB = -1; dUB = 0; dWB = 0; }                  // wall model
                bc[0 + side] = uB; bc[2 + side] = uB; bc[4 + side] = wB;
                bc[6 + side] = dUB; bc[8 + side] = dUB; bc[10 + side] = dWB;
            }
        }
        st = pdo_duidxj_init(&h->dj, h->spC, h->spE, h->d6, bc, p->Re > 0 ? 0 : 1);
    }
    if (!st && p->use_sgs) st = pdo_sgsrhs_init(&h->sgs, h->spC, h->spE, h->d6, &p->sgs, prec);
    if (st) { delete h; return st; }
    const size_t W = h->W;
    h->nC = (size_t)h->nx * h->ny * h->nz;
    h->nE = (size_t)h->nx * h->ny * (h->nz + 1);
    h->sC = 2 * (size_t)h->nxh * h->ny * h->nz;
    h->sE = 2 * (size_t)h->nxh * h->ny * (h->nz + 1);
    for (int i = 0; i < 4 && !st; i++) {
        if (!st) st = h->alloc(&h->uh[i], h->sC * W);
        if (!st) st = h->alloc(&h->vh[i], h->sC * W);
        if (!st) st = h->alloc(&h->wh[i], h->sE * W);
    }
    void **cbufs[] = {&h->ur, &h->vr, &h->ur5, &h->vr5, &h->whC, &h->d2u, &h->d2v, &h->fC1};
    for (void **b : cbufs)
        if (!st) st = h->alloc(b, h->sC * W);
    void **ebufs[] = {&h->wr, &h->wr5, &h->uEh, &h->vEh, &h->d2w, &h->fE1};
    for (void **b : ebufs)
        if (!st) st = h->alloc(b, h->sE * W);
    void **rC[] = {&h->u, &h->v, &h->wC, &h->tC1};
    for (void **b : rC)
        if (!st) st = h->alloc(b, h->nC * W);
    void **rE[] = {&h->w, &h->uE, &h->vE, &h->tE1};
    for (void **b : rE)
        if (!st) st = h->alloc(b, h->nE * W);
    if (!st) st = h->alloc(&h->gC, 9 * h->nC * W);
    if (!st) st = h->alloc(&h->gE, 9 * h->nE * W);
    if (!st) st = h->alloc(&h->gCh, 9 * h->sC * W);
    if (!st) st = h->alloc(&h->gEh, 9 * h->sE * W);
    if (!st) st = h->alloc((void **)&h->partial, h->nblocks * sizeof(double));
    if (!st) st = h->alloc((void **)&h->k3, h->nz * sizeof(double));
    if (!st) {
        // GetWaveNums (spectral.F90:1748-1763) of z
        std::vector<double> k(h->nz);
        for (int i = 0; i < h->nz; i++) {
            const int f = i < h->nz / 2 ? i : i - h->nz;
            k[i] = 2.0 * M_PI * f / (h->nz * h->dz);
        }
        hipError_t e = hipMemcpy(h->k3, k.data(), k.size() * sizeof(double), hipMemcpyHostToDevice);
        if (e != hipSuccess) st = hip_fail(e, "wavenumbers");
    }
    if (st) { delete h; return st; }
    *out = h;
    return 0;
}

int pdo_igrid_set_velocity(pdo_igrid h, const void *u, const void *v, const void *w, void *stream)
{
    if (!h || !u || !v || !w) { set_error("invalid argument"); return PDO_ERR_ARG; }
    hipStream_t st = (hipStream_t)stream;
    PDO_TRY(pdo_spectral_fft(h->spC, h->prec, u, h->uh[0], stream));
    PDO_TRY(pdo_spectral_fft(h->spC, h->prec, v, h->vh[0], stream));
    PDO_TRY(pdo_spectral_fft(h->spE, h->prec, w, h->wh[0], stream));
    PDO_TRY(project_and_prep(h, h->uh[0], h->vh[0], h->wh[0], st));
    h->tsim = 0; h->step = 0; h->newTimeStep = true; h->draws = 0;
    return 0;
}

int pdo_igrid_step(pdo_igrid h, int nsteps, void *stream)
{
    if (!h || nsteps < 0) { set_error("invalid argument"); return PDO_ERR_ARG; }
    if (h->cur_u != h->uh[0]) { set_error("no velocity set (pdo_igrid_set_velocity)"); return PDO_ERR_ARG; }
    hipStream_t st = (hipStream_t)stream;
    for (int i = 0; i < nsteps; i++) {
        PDO_TRY(h->p.scheme == 2 ? ssp_rk45(h, st) : tvd_rk3(h, st));
        // wrapup_timestep (:2067-2072)
        h->step++;
        h->tsim += h->dt;
        h->newTimeStep = true;
    }
    return 0;
}

int pdo_igrid_field(pdo_igrid h, int which, const void **ptr)
{
    if (!h || !ptr) { set_error("invalid argument"); return PDO_ERR_ARG; }
    const void *f[] = {h->u, h->v, h->wC, h->w, h->uh[0], h->vh[0], h->wh[0], h->gC, h->gE, h->ur, h->vr, h->wr,
                       h->uE, h->vE, h->whC};
    if (which < 0 || which >= (int)(sizeof(f) / sizeof(f[0]))) { set_error("unknown field"); return PDO_ERR_ARG; }
    *ptr = f[which];
    return 0;
}

int pdo_igrid_set_closure(pdo_igrid h, int kind, pdo_closure c, double delta, double dratio, double deltaRat)
{
    if (!h || kind < 0 || kind > 3) { set_error("invalid argument"); return PDO_ERR_ARG; }
    if (kind > 0 && (!h->sgs || h->p.sgs.model != 0)) {
        set_error("igrid closure: init with use_sgs and the Smagorinsky model");
        return PDO_ERR_ARG;
    }
    if (kind >= 2 && !c) { set_error("igrid closure: kinds 2 and 3 need a closure object"); return PDO_ERR_ARG; }
    if ((kind == 1 || kind == 2) && !h->dyn) {
        // the test filter acts in x and y only (standardDynamicProcedure.F90:
        // 165-207); InitTestFilter refuses spectral objects with periodic z
        // (spectral.F90:1339-1342), so the procedure gets an object of this
        // grid without the periodic-z masks: the same 2D transforms
        int st = pdo_spectral_init(&h->spD, h->nx, h->ny, h->nz, h->dx, h->dy, h->dz, 0, 2.0 / 3.0, 0);
        if (!st) st = pdo_dynproc_init(&h->dyn, h->spD, deltaRat);
        if (!st) st = h->alloc(&h->dS, 6 * h->nC * h->W);
        if (!st) st = h->alloc(&h->dD, h->nC * h->W);
        if (!st) st = h->alloc((void **)&h->dlam, h->nz * sizeof(double));
        if (st) return st;
    }
    PDO_TRY(pdo_sgsrhs_set_closure(h->sgs, kind >= 2 ? c : nullptr, delta, dratio, kind == 2));
    h->ckind = kind;
    return 0;
}

int pdo_igrid_time(pdo_igrid h, double *tsim, long *step, double *dt)
{
    if (!h) { set_error("invalid argument"); return PDO_ERR_ARG; }
    if (tsim) *tsim = h->tsim;
    if (step) *step = h->step;
    if (dt) *dt = h->dt;
    return 0;
}

// ---- restart files (igrid.F90:2719-2800) -----------------------------------
// RESTART_RunRR_u.TTTTTT, _v, _w: the global physical fields as raw FP64 in
// Fortran order (what decomp_2d_write_one writes: u, v on cells, w on the
// nz+1 edges); RESTART_RunRR_info.TTTTTT: tsim as one "(100g15.5)" record.
static std::string restart_name(const char *dir, int rid, const char *var, long tid)
{
    char buf[64];
    std::snprintf(buf, sizeof(buf), "RESTART_Run%02d_%s.%06ld", rid, var, tid);
    return std::string(dir) + "/" + buf;
}

// one G15.5 field: F(11).(5-k) plus four blanks for 0.1 <= |x| < 1e5 (k
// digits before the point), E15.5 otherwise, as the Fortran G edit
// descriptor writes it
static std::string g15_5(double x)
{
    char buf[64];
    const double m = std::fabs(x);
    if (m == 0.0) {
        std::snprintf(buf, sizeof(buf), "%11.4f    ", 0.0);
        return buf;
    }
    if (m >= 0.1 - 0.5e-6 && m < 1e5 - 0.5) {
        // digits before the point after rounding to 5 significant digits
        char t[32];
        std::snprintf(t, sizeof(t), "%.4e", m);
        const int e = std::atoi(std::strchr(t, 'e') + 1);
        const int k = e + 1;
        std::snprintf(buf, sizeof(buf), "%11.*f    ", 5 - k, x);
        return buf;
    }
    char t[32];
    std::snprintf(t, sizeof(t), "%.4e", m);  // d.dddde+XX
    const int e = std::atoi(std::strchr(t, 'e') + 1) + 1;
    char mant[16] = {'0', '.', t[0], t[2], t[3], t[4], t[5], 0};
    std::snprintf(buf, sizeof(buf), "%15s", (std::string(x < 0 ? "-" : "") + mant + "E" + (e < 0 ? "-" : "+") +
                                            (std::abs(e) < 10 ? "0" : "") + std::to_string(std::abs(e))).c_str());
    return buf;
}

int pdo_igrid_dump_restart(pdo_igrid h, const char *dir, int rid, void *stream)
{
    if (!h || !dir) { set_error("invalid argument"); return PDO_ERR_ARG; }
    hipStream_t st = (hipStream_t)stream;
    PDO_HIP(hipStreamSynchronize(st), "restart");
    const struct { const char *name; const void *dev; size_t n; } F[3] = {
        {"u", h->u, h->nC}, {"v", h->v, h->nC}, {"w", h->w, h->nE}};
    for (const auto &f : F) {
        std::vector<double> host(f.n);
        if (h->W == 8) {
            PDO_HIP(hipMemcpy(host.data(), f.dev, f.n * 8, hipMemcpyDeviceToHost), "restart");
        } else {
            std::vector<float> tmp(f.n);
            PDO_HIP(hipMemcpy(tmp.data(), f.dev, f.n * 4, hipMemcpyDeviceToHost), "restart");
            for (size_t i = 0; i < f.n; i++) host[i] = tmp[i];
        }
        const std::string name = restart_name(dir, rid, f.name, h->step);
        FILE *fp = std::fopen(name.c_str(), "wb");
        if (!fp) { set_error("cannot write " + name); return PDO_ERR_ARG; }
        const size_t wr = std::fwrite(host.data(), sizeof(double), f.n, fp);
        std::fclose(fp);
        if (wr != f.n) { set_error("short write " + name); return PDO_ERR_ARG; }
    }
    const std::string info = restart_name(dir, rid, "info", h->step);
    FILE *fp = std::fopen(info.c_str(), "w");
    if (!fp) { set_error("cannot write " + info); return PDO_ERR_ARG; }
    std::fprintf(fp, "%s\n", g15_5(h->tsim).c_str());
    std::fclose(fp);
    return 0;
}

int pdo_igrid_read_restart(pdo_igrid h, const char *dir, int rid, long tid, void *stream)
{
    if (!h || !dir) { set_error("invalid argument"); return PDO_ERR_ARG; }
    hipStream_t st = (hipStream_t)stream;
    void *dst[3] = {h->u, h->v, h->w};
    const char *names[3] = {"u", "v", "w"};
    const size_t n[3] = {h->nC, h->nC, h->nE};
    for (int q = 0; q < 3; q++) {
        const std::string name = restart_name(dir, rid, names[q], tid);
        FILE *fp = std::fopen(name.c_str(), "rb");
        if (!fp) { set_error("cannot read " + name); return PDO_ERR_ARG; }
        std::vector<double> host(n[q]);
        const size_t rd = std::fread(host.data(), sizeof(double), n[q], fp);
        std::fclose(fp);
        if (rd != n[q]) { set_error("short read " + name); return PDO_ERR_ARG; }
        PDO_HIP(hipStreamSynchronize(st), "restart");
        if (h->W == 8) {
            PDO_HIP(hipMemcpy(dst[q], host.data(), n[q] * 8, hipMemcpyHostToDevice), "restart");
        } else {
            std::vector<float> tmp(host.begin(), host.end());
            PDO_HIP(hipMemcpy(dst[q], tmp.data(), n[q] * 4, hipMemcpyHostToDevice), "restart");
        }
    }
    const std::string info = restart_name(dir, rid, "info", tid);
    FILE *fp = std::fopen(info.c_str(), "r");
    if (!fp) { set_error("cannot read " + info); return PDO_ERR_ARG; }
    double t = 0;
    const int got = std::fscanf(fp, "%lf", &t);
    std::fclose(fp);
    if (got != 1) { set_error("bad restart info " + info); return PDO_ERR_ARG; }
    // the physical fields are the state: transform them and prepare the stage
    PDO_TRY(pdo_igrid_set_velocity(h, h->u, h->v, h->w, stream));
    h->tsim = t;
    h->step = tid;
    // the forcing's draws continue from the restart's time index, as the
    // reference seeds them from it (forcingIsotropic.F90:92)
    h->draws = (unsigned long long)tid;
    return 0;
}

int pdo_igrid_destroy(pdo_igrid h)
{
    delete h;
    return 0;
}

}  // extern "C"
