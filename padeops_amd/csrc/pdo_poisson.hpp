// PoissonPeriodic object behind the C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "pdo_fft.hpp"

namespace pdo {

double modified_k(int kind, double k, double dx);
// GetWaveNums + ifftshift: (0, 1, .., n/2-1, -n/2, .., -1) * 2 pi / (n dx)
std::vector<double> wavenums(int n, double dx);
template <typename T> int upload_vec(const std::vector<T> &h, T **d);
bool fft_pow2(int n, int lo, int hi);
// twiddle tables of the x (length nx/2 plus the split), y and z (nz > 0) transforms
int fft_tables(PoissonPlanDev &dev, int nx, int ny, int nz, std::vector<void *> &allocs);
// storage/arithmetic mode of the FFT kernels for a precision mode (-1: unknown)
int fft_mode(int prec);

struct PoissonObject {
    int nx = 0, ny = 0, nz = 0;
    PoissonPlanDev dev;
    std::vector<void *> allocs;
    void *spec = nullptr;
    size_t spec_bytes = 0;
    double *kpx = nullptr;  // x modified wavenumber k'(kx), zero at kx = 0 and nx/2 (device, nx/2+1)
    int init(int nx, int ny, int nz, double dx, double dy, double dz, int mx, int my, int mz, int zero_policy);
    // kx (nx/2+1), ky (ny), kz (nz): the wavenumbers the solver divides by
    int init_k(int nx, int ny, int nz, const double *kx, const double *ky, const double *kz, int zero_policy);
    int solve(int prec, const void *rhs, void *f, hipStream_t st);
    // p = Poisson(div + D_x ut) and ux = ut - D_x p with D_x the x derivative
    // whose modified wavenumber the solver uses (p may alias div)
    int solve_dx(int prec, const void *div, const void *ut, void *p, void *ux, hipStream_t st);
    ~PoissonObject();
};

}  // namespace pdo

struct pdo_poisson_s : public pdo::PoissonObject {};
