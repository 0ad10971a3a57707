// igrid%compute_duidxj (src/incompressible/igrid.F90:2553-2683): the nine
// velocity-gradient components on cells and on edges, in physical and in
// spectral space, from the spectral velocity of igrid, over the spectral
// objects (x, y: i k and 2D inverse transforms) and Pade6stagg (z).  On one
// GPU the reference's y <-> z transposes are identities, so the z operators
// work on the spectral arrays in place of transposed copies.
#include <cstring>

#include "../../include/padeops_hip.h"
#include "pdo_ops.hpp"
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

}  // extern "C"
