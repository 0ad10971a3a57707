// Explicit symmetric filters (Gaussian, least squares): launch arguments;
// the elementwise sums of the operators module.
#pragma once

#include <hip/hip_runtime.h>

namespace pdo {

constexpr int EXPL_R = 4;   // stencil half-width
constexpr int EXPL_NB = 4;  // closure rows per end (non-periodic)
constexpr int EXPL_WB = 8;  // closure weights per row

// Passed by value: the coefficients travel in the kernel arguments.
struct ExplicitArgs {
    int n, na, nb, dir, periodic;
    double c[EXPL_R + 1];                 // c0, c1..c4 of the symmetric stencil
    double left[EXPL_NB * EXPL_WB];       // rows 0..3 against points 0..7
    double right[EXPL_NB * EXPL_WB];      // rows n-4..n-1 against points n-8..n-1
};

hipError_t launch_explicit(bool dbl, const ExplicitArgs &E, const void *f, void *out, hipStream_t st);
// out = a + s b, elementwise over n values (FP64 arithmetic)
hipError_t launch_axpy(bool dbl, void *out, const void *a, const void *b, double s, long n, hipStream_t st);

}  // namespace pdo
