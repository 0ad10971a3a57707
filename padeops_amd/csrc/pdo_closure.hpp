// Learned SGS closure kernels (pdo_closure.hip): features, tensor-basis
// stress, training loss.
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/padeops_hip.h"

namespace pdo {

// points a closure object was created for
long closure_max_points(pdo_closure c);

// blocks of the closure kernels for n points (at most 8192): the length of
// the partial arrays they fill
int closure_blocks(long n);
// out[0] = sum over n points of the 9 gradient components squared
hipError_t launch_grad_sumsq(bool dbl, const void *g, long n, long stride, double *partial, double *out,
                             hipStream_t st);
// features X (n, 10) (and the normalised gradient A9 (n, 9) unless null) with
// the scale sqrt(sumsq[0] / nnorm)
hipError_t launch_closure_features(bool dbl, const void *g, long n, long stride, const double *sumsq, long nnorm,
                                   double dratio, float *X, float *A9, hipStream_t st);
// stress of the outputs G (n, 6); which 0: cells (11 12 22 33), 1: edges
// (13 23), 2: all six; fallback keeps tau where the dissipation is negative
hipError_t launch_closure_stress(bool dbl, const void *g, long n, long stride, const double *sumsq, long nnorm,
                                 double delta2, const float *G, int which, int fallback, void *tau, long tstride,
                                 unsigned *nfall, hipStream_t st);
// training loss (to loss[0]) and its gradient dG (B, 6)
hipError_t launch_closure_loss(const float *G, const float *A9, const float *T, long B, float *dG, double *partial,
                               double *loss, hipStream_t st);

}  // namespace pdo
