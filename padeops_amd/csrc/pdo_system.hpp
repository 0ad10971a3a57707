// Host-side description of one compact operator as the banded linear system
//     A y = B f
// that it solves along a pencil, plus the partitioned ("spike") solve tables
// the GPU kernels consume.  Everything here is FP64 host code; it runs once per
// operator variant at init time.
#pragma once

#include <array>
#include <cstdint>
#include <map>
#include <string>
#include <utility>
#include <vector>

#include "pdo_kernels.hpp"

namespace pdo {

// operator families: fix the shape of the interior stencil the kernels
// evaluate (number of pairs R, symmetric or antisymmetric, center term,
// half-point shift of staggered operators) and the LHS half-bandwidth K.
enum Family : int {
    FAM_CD06_D1 = 0,  // K1 R2 antisym
    FAM_SYM2C = 1,    // K1 R2 sym + center  (cd06 d2, staggered d2)
    FAM_CD10_D1 = 2,  // K2 R3 antisym
    FAM_CD10_D2 = 3,  // K2 R3 sym + center
    FAM_CF90 = 4,     // K2 R4 sym + center
    FAM_STG_E2C_A = 5,  // K1 R2 antisym, shifted +1/2 (ddz_E2C)
    FAM_STG_C2E_A = 6,  // K1 R2 antisym, shifted -1/2 (ddz_C2E)
    FAM_STG_E2C_S = 7,  // K1 R2 sym, shifted +1/2 (InterpZ_E2C)
    FAM_STG_C2E_S = 8,  // K1 R2 sym, shifted -1/2 (InterpZ_C2E)
    FAM_COUNT = 9
};

// operator kinds of pdo_op_dense()
enum OpKind : int {
    KIND_D1 = 0, KIND_D2 = 1, KIND_FILTER = 2,
    KIND_STG_D1_E2C = 10, KIND_STG_D1_C2E = 11, KIND_STG_D2_C2C = 12, KIND_STG_D2_E2E = 13,
    KIND_STG_IN_E2C = 14, KIND_STG_IN_C2E = 15, KIND_STG_D1_C2C = 16, KIND_STG_D1_E2E = 17
};

enum Scheme : int { SCHEME_CD06 = 0, SCHEME_CD10 = 1, SCHEME_CF90 = 2, SCHEME_CD06STAGG = 3 };

struct SystemDesc {
    int family = 0;
    int n_sys = 0;      // rows of the solved system
    int n_in = 0;       // input points per pencil
    int n_out = 0;      // output points per pencil (n_sys, or n_sys+1 for a copied edge)
    bool periodic = true;
    int K = 1;          // LHS half bandwidth (1 tri, 2 penta)
    int RB = 2;         // window half width needed by every row of the RHS
    double c[5] = {0, 0, 0, 0, 0};  // interior stencil: c[0] center, c[1..R] pair weights
    // LHS band per row, offsets -2..2 (index 2 = diagonal)
    std::vector<std::array<double, 5>> lhs;
    // explicit RHS rows (non-periodic closures): row -> (input column, coefficient)
    std::map<int, std::vector<std::pair<int, double>>> special;
};

// Family shape constants shared with the kernels.
struct FamShape { int K, R, sigma, shift, center; };
FamShape family_shape(int fam);

// Builders.  Return 0 or the reference ierr / status code.
int build_cd06(SystemDesc &s, int kind, int n, double dx, bool periodic, int bc1, int bcn);
int build_cd10(SystemDesc &s, int kind, int n, double dx, bool periodic, int bc1, int bcn);
int build_cf90(SystemDesc &s, int n, bool periodic, int bc1, int bcn);
int build_stagg_periodic(SystemDesc &s, int kind, int n, double dz);
// flags: bit 0 isTopEven, 1 isBotEven, 2 isTopSided, 3 isBotSided
int build_stagg_nonperiodic(SystemDesc &s, int kind, int n, double dz, int flags);

// Evaluate the RHS stencil row i on input f (host, for dense export/tests).
double rhs_row(const SystemDesc &s, int i, const double *f);
void dense(const SystemDesc &s, double *A, double *B);

// ---- partitioned solve tables ---------------------------------------------
// Segment types: 0 first (non-periodic left end), 1 interior / periodic,
// 2 last (non-periodic right end), 3 single segment with both ends.

struct SpikeTables {
    int m = 0, P = 0;        // segment length and count (P*m + xr == n_sys)
    int xr = 0;              // extra rows of the last segment (0, or 1 for 32 P + 1 non-periodic rows)
    bool general = false;    // single-pencil kernel (any n): P == 1, m == n_sys
    int K = 1, nbl = 0, nbr = 0, wb = 0;
    // per segment type t, row j (0..m+xr-1; types 0 and 1 use the first m):
    std::vector<double> fac;   // [4][m+xr][5]: l1, l2, inv, u1*inv, u2*inv
    std::vector<double> spk;   // [4][m+xr][4]: V0, V1, W0, W1
    std::vector<double> lhs;   // [4][m+xr][5]: band of A (offsets -2..2)
    std::vector<double> G;     // [P][2K][2K*P]: rows of R^-1 giving yL[0..K), yR[0..K)
    std::vector<double> sp;    // [2][NBMAX][WBMAX]: closure rows (left: cols 0..WBMAX-1, right: cols n_in-WBMAX..n_in-1)
    int D32 = 0, D64 = 0;      // periodic: segment offsets +-D whose coupling exceeds 1e-9 / 1e-18 of the largest
    // the refined FP32 mode: its first solve keeps couplings above 1e-6 (its
    // error is removed by the correction), the correction those above 1e-3
    // (the correction is ~1e-5 of the solution, so its own relative error
    // of 1e-3 leaves ~1e-8)
    int D32R = 0, D32C = 0;
};

int build_spike(const SystemDesc &s, SpikeTables &t, bool force_general = false);

// Choose (m, P, xr) for a system of n rows; returns false when only the
// general kernel applies.
bool choose_partition(int n_sys, int K, int nb, bool periodic, int &m, int &P, int &xr);

}  // namespace pdo
