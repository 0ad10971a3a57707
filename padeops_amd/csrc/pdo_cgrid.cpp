// The cgrid compressible right-hand side for one species of ideal gas on one
// GPU (src/compressible/cgrid.F90): get_primitive :763-797, getRHS :822-894
// with getRHS_x/y/z :896-1056, getLAD :1058-1257 (artificial shear and bulk
// viscosity and conductivity), filter :1259-1336, get_tau :1360-1404, get_q
// :1406-1452.  Derivatives and filters are the library's derivatives /
// filters objects (derivatives%ddx.., filters%filterx..), so the compact
// solves of the collocated drop-in carry the work; the rest are pointwise
// kernels (pdo_cgrid.hip).  All fields live in device memory for the life of
// the object.
#include <cmath>
#include <vector>

#include "../../include/padeops_hip.h"
#include "pdo_cgrid.hpp"
#include "pdo_ops.hpp"

namespace pdo {

// work field indices (each n words of the storage type)
enum {
    F_PRIM = 0,            // rho u v w e p T
    F_G = F_PRIM + 7,      // du_i/dx_j at 3 i + j
    F_MU = F_G + 9,        // mu, bulk, kap (physical + artificial)
    F_STAR = F_MU + 3,     // mustar, bulkstar, kapstar
    F_TAU = F_STAR + 3,    // xx xy xz yy yz zz
    F_Q = F_TAU + 6,       // qx qy qz
    F_FUNC = F_Q + 3,
    F_T1 = F_FUNC + 1,
    F_D4 = F_T1 + 1,       // fourth derivatives along x, y, z
    F_R = F_D4 + 3,        // gradient of rho or e
    F_FIL = F_R + 3,       // two filter buffers
    F_FLUX = F_FIL + 2,    // five fluxes
    F_DFLUX = F_FLUX + 5,  // their derivatives
    F_COUNT = F_DFLUX + 5
};

struct CgridObject {
    pdo_cgrid_params p{};
    int prec = PDO_F64;
    size_t W = 8;
    long n = 0;
    pdo_derivatives der = nullptr;
    pdo_filters fil = nullptr, gfil = nullptr;
    void *buf = nullptr;
    ~CgridObject()
    {
        if (der) pdo_derivatives_destroy(der);
        if (fil) pdo_filters_destroy(fil);
        if (gfil) pdo_filters_destroy(gfil);
        if (buf) (void)hipFree(buf);
    }
    void *f(int i) { return (char *)buf + (size_t)i * n * W; }
};

}  // namespace pdo

struct pdo_cgrid_s : public pdo::CgridObject {};

using namespace pdo;

namespace {

#define PDO_TRY(x)                   \
    do {                             \
        if (int r_ = (x)) return r_; \
    } while (0)

int cg(pdo_cgrid h, CgOp op, const CgArgs &a, hipStream_t st)
{
    hipError_t e = launch_cgrid(h->W == 8, op, a, st);
    return e == hipSuccess ? 0 : hip_fail(e, "cgrid kernel");
}

// derivatives%dd{x,y,z} / d2d{x,y,z}2 along direction d (0, 1, 2)
int dd(pdo_cgrid h, int d, const void *f, void *df, int bc1, int bcn, hipStream_t st)
{
    void *s = (void *)st;
    if (d == 0) return pdo_derivatives_ddx(h->der, h->prec, f, df, bc1, bcn, s);
    if (d == 1) return pdo_derivatives_ddy(h->der, h->prec, f, df, bc1, bcn, s);
    return pdo_derivatives_ddz(h->der, h->prec, f, df, bc1, bcn, s);
}
int d2(pdo_cgrid h, int d, const void *f, void *df, int bc1, int bcn, hipStream_t st)
{
    void *s = (void *)st;
    if (d == 0) return pdo_derivatives_d2dx2(h->der, h->prec, f, df, bc1, bcn, s);
    if (d == 1) return pdo_derivatives_d2dy2(h->der, h->prec, f, df, bc1, bcn, s);
    return pdo_derivatives_d2dz2(h->der, h->prec, f, df, bc1, bcn, s);
}

const int *bc_of(pdo_cgrid h, int d) { return d == 0 ? h->p.x_bc : (d == 1 ? h->p.y_bc : h->p.z_bc); }

// gradient (cgrid.F90:446-480); flip[d] negates the bc of direction d
int gradient(pdo_cgrid h, const void *f, void *out3, int flipdir, hipStream_t st)
{
    for (int d = 0; d < 3; d++) {
        const int *b = bc_of(h, d);
        const int sg = d == flipdir ? -1 : 1;
        PDO_TRY(dd(h, d, f, (char *)out3 + (size_t)d * h->n * h->W, sg * b[0], sg * b[1], st));
    }
    return 0;
}

// the fourth derivative along each direction (two second-derivative passes)
// into F_D4 + d
int fourth(pdo_cgrid h, const void *f, hipStream_t st)
{
    for (int d = 0; d < 3; d++) {
        const int *b = bc_of(h, d);
        PDO_TRY(d2(h, d, f, h->f(F_T1), b[0], b[1], st));
        PDO_TRY(d2(h, d, h->f(F_T1), h->f(F_D4 + d), b[0], b[1], st));
    }
    return 0;
}

// filter (cgrid.F90:1259-1336): numtimes passes along y, then x, then z, in
// place on arr, through the two filter buffers
int filter3(pdo_cgrid h, pdo_filters fl, void *arr, int numtimes, hipStream_t st)
{
    void *s = (void *)st;
    void *bufs[2] = {h->f(F_FIL), h->f(F_FIL + 1)};
    const int order[3] = {1, 0, 2};
    const int total = 3 * numtimes;
    void *cur = arr;
    int k = 0, pass = 0;
    for (int q = 0; q < 3; q++) {
        const int d = order[q];
        const int *b = bc_of(h, d);
        for (int r = 0; r < numtimes; r++, pass++) {
            // out of place: the last pass writes arr unless it also reads it
            void *dst = (pass == total - 1 && cur != arr) ? arr : bufs[k];
            const int e = d == 0 ? pdo_filters_filterx(fl, h->prec, cur, dst, b[0], b[1], s)
                        : d == 1 ? pdo_filters_filtery(fl, h->prec, cur, dst, b[0], b[1], s)
                                 : pdo_filters_filterz(fl, h->prec, cur, dst, b[0], b[1], s);
            if (e) return e;
            if (dst != arr) k ^= 1;
            cur = dst;
        }
    }
    if (cur != arr) {
        CgArgs a;
        a.in[0] = cur;
        a.out[0] = arr;
        a.n = h->n;
        PDO_TRY(cg(h, CG_COPY, a, st));
    }
    return 0;
}

// getLAD (cgrid.F90:1058-1257), one species: adds the filtered artificial
// shear and bulk viscosity and conductivity to mu, bulk, kap
int lad(pdo_cgrid h, hipStream_t st)
{
    const pdo_cgrid_params &p = h->p;
    const double dx = p.dx, dy = p.dy, dz = p.dz;
    const double dx2 = dx * dx, dy2 = dy * dy, dz2 = dz * dz;
    const double dx4 = dx2 * dx2, dy4 = dy2 * dy2, dz4 = dz2 * dz2;
    const long n = h->n;
    // artificial shear viscosity
    {
        CgArgs a;
        a.in[0] = h->f(F_G);
        a.out[0] = h->f(F_FUNC);
        a.n = n;
        PDO_TRY(cg(h, CG_SHEAR, a, st));
    }
    PDO_TRY(fourth(h, h->f(F_FUNC), st));
    {
        CgArgs a;
        a.in[0] = h->f(F_D4);
        a.in[1] = h->f(F_D4 + 1);
        a.in[2] = h->f(F_D4 + 2);
        a.in[3] = h->f(F_PRIM);
        a.out[0] = h->f(F_STAR);
        a.n = n;
        a.c[0] = p.Cmu;
        a.c[1] = dx4 * dx2;
        a.c[2] = dy4 * dy2;
        a.c[3] = dz4 * dz2;
        PDO_TRY(cg(h, CG_MUSTAR, a, st));
    }
    PDO_TRY(filter3(h, h->gfil, h->f(F_STAR), 2, st));
    // artificial bulk viscosity
    {
        CgArgs a;
        a.in[0] = h->f(F_G);
        a.out[0] = h->f(F_FUNC);
        a.n = n;
        PDO_TRY(cg(h, CG_DILATATION, a, st));
    }
    PDO_TRY(gradient(h, h->f(F_PRIM), h->f(F_R), -1, st));
    PDO_TRY(fourth(h, h->f(F_FUNC), st));
    {
        CgArgs a;
        a.in[0] = h->f(F_D4);
        a.in[1] = h->f(F_D4 + 1);
        a.in[2] = h->f(F_D4 + 2);
        a.in[3] = h->f(F_R);
        a.in[4] = h->f(F_G);
        a.in[5] = h->f(F_FUNC);
        a.in[6] = h->f(F_PRIM);
        a.out[0] = h->f(F_STAR + 1);
        a.n = n;
        const double c[7] = {p.Cbeta, dx, dy, dz, dx4, dy4, dz4};
        for (int i = 0; i < 7; i++) a.c[i] = c[i];
        PDO_TRY(cg(h, CG_BULKSTAR, a, st));
    }
    PDO_TRY(filter3(h, h->gfil, h->f(F_STAR + 1), 2, st));
    // artificial conductivity
    PDO_TRY(gradient(h, h->f(F_PRIM + 4), h->f(F_R), -1, st));
    PDO_TRY(fourth(h, h->f(F_PRIM + 4), st));
    {
        CgArgs a;
        a.in[0] = h->f(F_D4);
        a.in[1] = h->f(F_D4 + 1);
        a.in[2] = h->f(F_D4 + 2);
        a.in[3] = h->f(F_R);
        a.in[4] = h->f(F_PRIM);
        a.out[0] = h->f(F_STAR + 2);
        a.n = n;
        const double c[8] = {p.Ckap, p.gam, dx, dy, dz, dx4, dy4, dz4};
        for (int i = 0; i < 8; i++) a.c[i] = c[i];
        PDO_TRY(cg(h, CG_KAPSTAR, a, st));
    }
    PDO_TRY(filter3(h, h->gfil, h->f(F_STAR + 2), 2, st));
    CgArgs a;
    a.out[0] = h->f(F_MU);
    a.out[1] = h->f(F_MU + 1);
    a.out[2] = h->f(F_MU + 2);
    a.in[0] = h->f(F_STAR);
    a.in[1] = h->f(F_STAR + 1);
    a.in[2] = h->f(F_STAR + 2);
    a.n = n;
    return cg(h, CG_ADD3, a, st);
}

}  // namespace

extern "C" {

int pdo_cgrid_init(pdo_cgrid *out, const pdo_cgrid_params *p, int prec)
{
    if (!out || !p) { set_error("invalid argument"); return PDO_ERR_ARG; }
    *out = nullptr;
    if (prec_mode(prec) < 0) { set_error("unknown precision mode"); return PDO_ERR_ARG; }
    if (p->nx < 1 || p->ny < 1 || p->nz < 1 || !(p->dx > 0) || !(p->dy > 0) || !(p->dz > 0)) {
        set_error("cgrid: sizes and spacings must be positive");
        return PDO_ERR_ARG;
    }
    if (!(p->gam > 1) || !(p->Rgas > 0) || (!p->inviscid && (!(p->T_ref > 0) || !(p->Pr > 0)))) {
        set_error("cgrid: need gam > 1, Rgas > 0 and, when viscous, T_ref > 0 and Pr > 0");
        return PDO_ERR_ARG;
    }
    auto *h = new pdo_cgrid_s();
    h->p = *p;
    h->prec = prec;
    h->W = prec == PDO_F64 ? 8 : 4;
    h->n = (long)p->nx * p->ny * p->nz;
    int st = pdo_derivatives_init(&h->der, p->nx, p->ny, p->nz, p->dx, p->dy, p->dz, p->periodic, p->derivative);
    if (!st) st = pdo_filters_init(&h->fil, p->nx, p->ny, p->nz, p->periodic, p->filter);
    const int gauss[3] = {2, 2, 2};
    if (!st) st = pdo_filters_init(&h->gfil, p->nx, p->ny, p->nz, p->periodic, gauss);
    if (st) { delete h; return st; }
    hipError_t e = hipMalloc(&h->buf, (size_t)F_COUNT * h->n * h->W);
    if (e != hipSuccess) { delete h; return hip_fail(e, "hipMalloc(cgrid work)"); }
    *out = h;
    return 0;
}

int pdo_cgrid_rhs(pdo_cgrid h, int prec, const void *Wc, void *rhs, void *stream)
{
    if (!h || !Wc || !rhs) { set_error("invalid argument"); return PDO_ERR_ARG; }
    if (prec != h->prec) { set_error("cgrid: the object's work buffers were sized for another precision"); return PDO_ERR_ARG; }
    hipStream_t st = (hipStream_t)stream;
    const pdo_cgrid_params &p = h->p;
    const long n = h->n;
    // get_primitive
    {
        CgArgs a;
        a.in[0] = Wc;
        a.out[0] = h->f(F_PRIM);
        a.n = n;
        a.c[0] = p.gam - 1.0;
        a.c[1] = 1.0 / p.Rgas;
        PDO_TRY(cg(h, CG_PRIMITIVE, a, st));
    }
    // velocity gradients: u with -x_bc, v with -y_bc, w with -z_bc
    for (int c = 0; c < 3; c++) PDO_TRY(gradient(h, h->f(F_PRIM + 1 + c), h->f(F_G + 3 * c), c, st));
    // transport properties, then the artificial ones
    {
        CgArgs a;
        a.in[0] = h->f(F_PRIM + 6);
        a.out[0] = h->f(F_MU);
        a.out[1] = h->f(F_MU + 1);
        a.out[2] = h->f(F_MU + 2);
        a.n = n;
        a.i0 = p.inviscid;
        a.c[0] = p.mu_ref;
        a.c[1] = p.T_ref;
        a.c[2] = p.visc_n;
        a.c[3] = p.gam * p.Rgas * (1.0 / (p.gam - 1.0));
        a.c[4] = p.Pr;
        PDO_TRY(cg(h, CG_TRANSPORT, a, st));
    }
    PDO_TRY(lad(h, st));
    // tau and q
    {
        CgArgs a;
        a.in[0] = h->f(F_G);
        a.in[1] = h->f(F_MU);
        a.in[2] = h->f(F_MU + 1);
        a.out[0] = h->f(F_TAU);
        a.n = n;
        PDO_TRY(cg(h, CG_TAU, a, st));
    }
    for (int d = 0; d < 3; d++) {
        const int *b = bc_of(h, d);
        PDO_TRY(dd(h, d, h->f(F_PRIM + 6), h->f(F_T1), b[0], b[1], st));
        CgArgs a;
        a.in[0] = h->f(F_MU + 2);
        a.in[1] = h->f(F_T1);
        a.out[0] = h->f(F_Q + d);
        a.n = n;
        PDO_TRY(cg(h, CG_NEGMUL, a, st));
    }
    // rhs = -sum_d d/dx_d (fluxes along d): symmetric ends for the momentum
    // along d, antisymmetric for the others
    for (int d = 0; d < 3; d++) {
        CgArgs a;
        a.in[0] = Wc;
        a.in[1] = h->f(F_PRIM);
        a.in[2] = h->f(F_TAU);
        a.in[3] = h->f(F_Q);
        a.out[0] = h->f(F_FLUX);
        a.n = n;
        a.i0 = d;
        PDO_TRY(cg(h, CG_FLUXES, a, st));
        const int *b = bc_of(h, d);
        for (int c = 0; c < 5; c++) {
            const int sg = c == 1 + d ? 1 : -1;
            PDO_TRY(dd(h, d, h->f(F_FLUX + c), h->f(F_DFLUX + c), sg * b[0], sg * b[1], st));
        }
        CgArgs s;
        s.in[0] = h->f(F_DFLUX);
        s.out[0] = rhs;
        s.n = n;
        s.i0 = d == 0;
        PDO_TRY(cg(h, CG_SUB5, s, st));
    }
    return 0;
}

int pdo_cgrid_filter(pdo_cgrid h, int prec, void *arr, int numtimes, int gaussian, void *stream)
{
    if (!h || !arr || numtimes < 1) { set_error("invalid argument"); return PDO_ERR_ARG; }
    if (prec != h->prec) { set_error("cgrid: the object's work buffers were sized for another precision"); return PDO_ERR_ARG; }
    return filter3(h, gaussian ? h->gfil : h->fil, arr, numtimes, (hipStream_t)stream);
}

int pdo_cgrid_field(pdo_cgrid h, int which, const void **ptr)
{
    if (!h || !ptr || which < 0 || which > 12) { set_error("invalid argument"); return PDO_ERR_ARG; }
    static const int base[13] = {F_PRIM,     F_PRIM + 1, F_PRIM + 2, F_PRIM + 3, F_PRIM + 4, F_PRIM + 5, F_PRIM + 6,
                                 F_MU,       F_MU + 1,   F_MU + 2,   F_G,        F_TAU,      F_Q};
    *ptr = h->f(base[which]);
    return 0;
}

int pdo_cgrid_destroy(pdo_cgrid h)
{
    delete h;
    return 0;
}

}  // extern "C"
