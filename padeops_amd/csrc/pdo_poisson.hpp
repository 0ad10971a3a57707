// PoissonPeriodic object behind the C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "pdo_fft.hpp"

namespace pdo {

double modified_k(int kind, double k, double dx);

struct PoissonObject {
    int nx = 0, ny = 0, nz = 0;
    PoissonPlanDev dev;
    std::vector<void *> allocs;
    void *spec = nullptr;
    size_t spec_bytes = 0;
    int init(int nx, int ny, int nz, double dx, double dy, double dz, int mx, int my, int mz, int zero_policy);
    int solve(int prec, const void *rhs, void *f, hipStream_t st);
    ~PoissonObject();
};

}  // namespace pdo

struct pdo_poisson_s : public pdo::PoissonObject {};
