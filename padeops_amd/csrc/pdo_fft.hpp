// Periodic FFT Poisson solver: device plan and launcher.
#pragma once

#include <hip/hip_runtime.h>

namespace pdo {

// Factor applied between the forward and inverse transforms of a z line
// (kx, ky, kz):
//   kind 0: -1/(ax[kx]+ay[ky]+az[kz]+1e-20), zero at (0,0,0)    (PoissonPeriodic)
//   kind 1: -1/(ax+ay+az), zero where ax+ay+az <= 1e-14          (PadePoisson)
//   kind 2: ax[kx]*ay[ky]*az[kz]                                   (dealias mask)
//   kind 3: i*az[kz]       (Fourier first derivative; ax, ay unused)
//   kind 4: az[kz]         (Fourier second derivative; ax, ay unused)
// then times `scale`.
struct SpecMul {
    const double *ax, *ay, *az;
    int kind;
    double scale;
};

struct PoissonPlanDev {
    int nx, ny, nz;
    int nxp;                             // row pitch of the half spectrum (>= nx/2+1, multiple of 16)
    int zero_policy;
    const double *kx2, *ky2, *kz2;       // device, sizes nx/2+1, ny, nz
    const double *tw_d[3];               // exp(-2 pi i m / L) for x (L = nx/2), y, z
    const float *tw_f[3];
    const double *twn_d;                 // exp(-2 pi i k / nx), k = 0..nx/2
    const float *twn_f;
};

// mode 0: FP64, 1: FP32 storage and arithmetic, 2: FP32 storage, FP64 arithmetic.
// spec: work buffer of nxp*ny*nz complex values of the storage type.
hipError_t run_poisson(const PoissonPlanDev &pl, int mode, const void *rhs, void *f, void *spec, hipStream_t st);

// The Poisson solve of the full RHS with its x terms in spectral space:
// p = Poisson(div + D_x ut), ux = ut - D_x p, where D_x is the operator with
// eigenvalue i kp[kx] on x mode kx (kp: nx/2+1 values; for the cd06 first
// derivative, its modified wavenumber, zero at kx = 0 and nx/2).
hipError_t run_poisson_dx(const PoissonPlanDev &pl, int mode, const void *div, const void *ut, void *p, void *ux,
                          void *spec, const double *kp, hipStream_t st);

// 2D (x r2c, y c2c) forward transform of each of `nplanes` (nx, ny) planes into
// an unpadded (nx/2+1, ny, nplanes) spectrum; unnormalized.
hipError_t run_fft2(const PoissonPlanDev &pl, int mode, const void *in, void *out, int nplanes, hipStream_t st);
// filtering of real (nx, ny, nplanes) data by t1(kx) t2(ky) (device arrays of
// nx/2+1 and ny): r2c into spec (a half spectrum of scratch), one fused
// forward-multiply-inverse y pass in place, c2r into out (out may be in)
hipError_t run_filter2(const PoissonPlanDev &pl, int mode, const void *in, void *spec, void *out, int nplanes,
                       const double *t1, const double *t2, hipStream_t st);
// inverse of run_fft2 scaled by 1/(nx ny); `in` is not modified (work holds
// the y-inverse planes); oddball != 0 drops the kx = nx/2 modes.
hipError_t run_ifft2(const PoissonPlanDev &pl, int mode, const void *in, void *work, void *out, int nplanes,
                     int oddball, hipStream_t st);
// z transform, multiply by m, inverse z transform of an unpadded
// (nx/2+1, ny, nz) spectrum (in may equal out).
hipError_t run_zmul(const PoissonPlanDev &pl, int mode, const void *in, void *out, const SpecMul &m, hipStream_t st);
// tables of the PadePoisson solve with walls (device, FP64): ax (nx/2+1),
// ay (ny), az (2 nz) the squared wavenumbers, kcm and kcp (2 nz complex,
// interleaved) k3' exp(-+ i k3 dz / 2)
struct PadeWallTables {
    const double *ax, *ay, *az, *kcm, *kcp;
};
// the extended solve and (upd) w update of PadePoisson with walls, in place
// on f (nx/2+1, ny, nz) and w (nx/2+1, ny, nz+1); upd false leaves w as it
// is; the plan's transforms have length 2 nz (its nz)
hipError_t run_pp_walls(const PoissonPlanDev &pl, int mode, void *f, void *w, const PadeWallTables &t, bool upd,
                        hipStream_t st);

// Fourier-collocation derivative along dir of real f (n along dir; x:
// (nx, na, nb), y: (na, ny, nb), z: (na, nb, nz)): the spectrum is multiplied
// by i kv (imag != 0) or kv, kv holding nx/2+1 (x) or n (y, z) values with
// the n/2 mode zero.  x needs work of nxp * na * nb complex values; y and z
// need na (y) or na*nb (z) even.
hipError_t run_fourier_deriv(const PoissonPlanDev &pl, int mode, int dir, int imag, const void *in, void *out,
                             void *work, long na, long nb, const double *kv, hipStream_t st);

}  // namespace pdo
