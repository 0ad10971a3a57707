// Pointwise kernels of the igrid pieces (pdo_igrid.hip).  mode: 0 FP64,
// 1 FP32, 2 FP32 storage with FP64 arithmetic (as the spectral objects).
#pragma once

#include <hip/hip_runtime.h>

namespace pdo {

// plane 0 of a spectrum of nplane complex values per plane: zero, except the
// mean mode, which is set to v
hipError_t launch_surface_plane(int mode, void *f, long nplane, double v, hipStream_t st);
// tau = -2 nu S: cells (nC points) 11, 12, 22, 33; edges (nE points) 13, 23
hipError_t launch_tau(int mode, const void *nuC, const void *SC, const void *nuE, const void *SE, void *tauC,
                      void *tauE, long nC, long nE, hipStream_t st);
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
