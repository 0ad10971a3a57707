// Shared declarations of the compact-operator kernels and their launchers.
#pragma once

#include <hip/hip_runtime.h>

namespace pdo {

constexpr int MAXM = 32;   // rows per segment of the register-resident kernels
constexpr int PMAX = 16;   // most segments per pencil (n <= 512 on that path)
constexpr int NBMAX = 4;   // closure rows per end of a non-periodic system
constexpr int WBMAX = 8;   // closure columns per end

enum KMode : int { MODE_F64 = 0, MODE_F32R = 1, MODE_F32S = 2, MODE_F32 = 3 };
enum KShape : int { SHAPE_STRIDED = 0, SHAPE_X = 1, SHAPE_GENERAL = 2 };

// Arguments of one apply launch (passed by value).  A pencil is addressed by
// (a, b): input element r of pencil (a,b) is f[a*sa_in + b*sb_in + r*se_in].
struct ApplyArgs {
    const void *f;
    void *out;
    const void *aux;   // epilogue operand (epi 1: out = aux + y, 2: out = aux - y)
    int epi;
    int periodic;
    long na, nb;
    long sa_in, sb_in, se_in;
    long sa_out, sb_out, se_out;
    int n_sys, n_in, n_read, n_out;
    int m, P;
    int xr;            // rows the last segment holds beyond m (0 or 1)
    int lgn;           // log2(n_in) (x kernel; n_in is a power of two there)
    int D32, D64;      // circulant coupling offsets kept by FP32 / FP64 solves
    int D32R, D32C;    // ... by the refined mode's first solve / its correction
    int nbl, nbr, wb;
    const double *tab_d;
    const float *tab_f;
    int o_fac, o_spk, o_lhs, o_G, o_sp;
    double c[5];
    double band[5];    // interior LHS band (offsets -2..2), used by periodic systems
};

hipError_t launch_absmax(const void *x, bool is_double, long n, double *partial, int nblocks, hipStream_t st);
hipError_t launch_zero_words(void *p, long words, hipStream_t st);
hipError_t launch_compact(int family, int shape, int mode, const ApplyArgs &A, double *scratch, hipStream_t st);

}  // namespace pdo
