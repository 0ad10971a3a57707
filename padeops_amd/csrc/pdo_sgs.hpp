// SGS eddy-viscosity kernels and the standard dynamic procedure (pdo_sgs.hip).
#pragma once

#include <hip/hip_runtime.h>

namespace pdo {

// mode as the FFT kernels: 0 FP64, 1 FP32, 2 FP32 storage / FP64 arithmetic
hipError_t launch_sgs_kernel(int mode, const void *g, void *S, void *nu, long n, long stride, int model, double cx,
                             double cy, double cz, hipStream_t st);
hipError_t launch_smag_from_S(bool dbl, const void *S, void *nu, long n, long stride, hipStream_t st);
hipError_t launch_sgs_scale(bool dbl, void *nu, long nxy, long n, const double *cplane, double cglobal,
                            hipStream_t st);
hipError_t launch_mul(bool dbl, const void *a, const void *b, void *out, long n, long stride, int ncomp,
                      hipStream_t st);
hipError_t launch_germano_M(bool dbl, void *fDS, const void *Df, const void *Sf, long n, long stride, double r2,
                            hipStream_t st);
hipError_t launch_uiuj(bool dbl, const void *u, const void *v, const void *w, void *out, long n, long stride,
                       hipStream_t st);
hipError_t launch_germano_ratio(bool dbl, const void *fuu, const void *uf, const void *M, double *num, double *den,
                                long n, long stride, hipStream_t st);
hipError_t launch_plane_ratio(const double *num, const double *den, long nxy, int nz, double meanfact,
                              double *partial, double *lam, hipStream_t st);

}  // namespace pdo
