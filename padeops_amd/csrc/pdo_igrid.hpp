// Pointwise kernels of the igrid pieces (pdo_igrid.hip).  mode: 0 FP64,
// 1 FP32, 2 FP32 storage with FP64 arithmetic (as the spectral objects).
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/padeops_hip.h"

namespace pdo {

// plane 0 of a spectrum of nplane complex values per plane: zero, except the
// mean mode, which is set to v
hipError_t launch_surface_plane(int mode, void *f, long nplane, double v, hipStream_t st);
// tau = -2 nu S: cells (nC points) 11, 12, 22, 33; edges (nE points) 13, 23
hipError_t launch_tau(int mode, const void *nuC, const void *SC, const void *nuE, const void *SE, void *tauC,
                      void *tauE, long nC, long nE, hipStream_t st);
// f[k nplane] += v (the real part of the mean mode of each of nz planes of
// a spectrum with nplane complex values per plane)
hipError_t launch_mean_add(int mode, void *f, long nplane, int nz, double v, hipStream_t st);
// a = max(a, 0) (real)
hipError_t launch_clip_negative(int mode, void *a, long n, hipStream_t st);
// rhs -= i k f (kv per index along dim 1 or 2) or rhs -= f (kv NULL); complex
hipError_t launch_sub(int mode, void *rhs, const void *f, const double *kv, int nxh, int ny, long n, int dim,
                      hipStream_t st);

// out = (a1 - b1) c1 + (a2 - b2) c2 (second term when a2 != NULL), real
hipError_t launch_cross(int mode, const void *a1, const void *b1, const void *c1, const void *a2, const void *b2,
                        const void *c2, void *out, long n, hipStream_t st);
// r += s (n real words)
hipError_t launch_add(int mode, void *r, const void *s, long n, hipStream_t st);
// rhs += invRe (-(k1^2 + k2^2) f + d2) on complex (nxh, ny, planes)
hipError_t launch_visc(int mode, void *rhs, const void *f, const void *d2, const double *k1, const double *k2, int nxh,
                       int ny, long n, double invRe, hipStream_t st);
// out = c[0] x[0] + ... + c[nterms-1] x[nterms-1] (n real words, left to right)
hipError_t launch_lincomb(int mode, void *out, const void *const *x, const double *c, int nterms, long n,
                          hipStream_t st);
// partial[b] = max over block b of |u|/dx + |v|/dy + |w|/dz (nblocks blocks)
hipError_t launch_cfl(int mode, const void *u, const void *v, const void *w, long n, double idx, double idy,
                      double idz, double *partial, int nblocks, hipStream_t st);

// Moeng wall model (wallmodel.F90:33-86, 106-134, 162-199 unstratified):
// sums of sqrt(u^2 + v^2) and u over the first cell plane (nxy points of u,
// v) in nb blocks, then wm = {Uspmn, ustar, WallMFactor, Umn} with
// Uspmn = sum / nxy, ustar = Uspmn kappa / lg (lg = log(dz / 2 / z0)),
// WallMFactor = -ustar^2 / (Uspmn + 1e-13), and on the bottom edge plane of
// tauE (13 at 0, 23 at nE) tau13 = WallMFactor u, tau23 = WallMFactor v.
// part holds 2 nb doubles; the sums run in a fixed order.
hipError_t launch_wall_model(int mode, const void *u, const void *v, void *tauE, long nxy, long nE, double *part,
                             int nb, double kappa, double lg, double *wm, hipStream_t st);
// whether a Pade6stagg object is periodic (pdo_pade6.cpp)
bool pade6stagg_periodic(pdo_pade6stagg h);

struct HitForcingArgs {
    int nxh, ny, nz, nwaves;
    double kmin, kmax, eps, normfact, dz;
    unsigned long long seed, draw;
    const double *k3;  // device: GetWaveNums(nz, dz)
};
// HIT_shell_forcing on the spectra (uhat, vhat: nz planes; what: nz+1)
hipError_t launch_hit_forcing(int mode, const void *uhat, const void *vhat, const void *what, void *urhs, void *vrhs,
                              void *wrhs, const HitForcingArgs &F, hipStream_t st);

}  // namespace pdo
