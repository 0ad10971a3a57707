// Periodic FFT Poisson solver: device plan and launcher.
#pragma once

#include <hip/hip_runtime.h>

namespace pdo {

// squared modified wavenumbers (FP64) of the half spectrum (nx/2+1, ny, nz)
template <typename TC> struct PoissonMul {
    const double *kx2, *ky2, *kz2;
    int nxh;
    int zero_policy;
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

}  // namespace pdo
