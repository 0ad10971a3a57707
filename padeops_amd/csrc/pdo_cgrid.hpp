// cgrid compressible right-hand side (pdo_cgrid.cpp) and its pointwise
// kernels (pdo_cgrid.hip).
#pragma once

#include <hip/hip_runtime.h>

namespace pdo {

enum CgOp {
    CG_PRIMITIVE,   // in W; out prim (7 fields); c: gam - 1, 1 / Rgas
    CG_TRANSPORT,   // in T; out mu, bulk, kap; i0 inviscid; c: mu_ref, T_ref, n, Cp, Pr
    CG_SHEAR,       // in g (9 fields); out func
    CG_DILATATION,  // in g; out func
    CG_MUSTAR,      // in X, Y, Z, rho; out; c: Cmu, dx^6, dy^6, dz^6
    CG_BULKSTAR,    // in X, Y, Z, grad rho (3), g, dil, rho; out; c: Cbeta, dx, dy, dz, dx^4, dy^4, dz^4
    CG_KAPSTAR,     // in X, Y, Z, grad e (3), prim; out; c: Ckap, gam, dx, dy, dz, dx^4, dy^4, dz^4
    CG_ADD3,        // out mu, bulk, kap += in mustar, bulkstar, kapstar
    CG_TAU,         // in g, mu, bulk; out tau (6 fields)
    CG_NEGMUL,      // out = -in0 in1
    CG_FLUXES,      // in W, prim, tau, q (3 fields); out 5 fluxes; i0 direction
    CG_SUB5,        // out (5 fields) -= in (5 fields); i0: first (out = -in)
    CG_COPY,        // out = in
};

struct CgArgs {
    const void *in[8] = {};
    void *out[3] = {};
    long n = 0;  // points per field
    int i0 = 0;
    double c[8] = {};
};

hipError_t launch_cgrid(bool dbl, CgOp op, const CgArgs &a, hipStream_t st);

}  // namespace pdo
