// Fused compact-operator kernels for gfx950.
//
// One launch applies  y = A^-1 (B f)  along every pencil of a field, where B
// is the explicit RHS stencil and A the tri/pentadiagonal (cyclic when
// periodic) left-hand side of a PadeOps compact scheme.  The solve is a
// partitioned ("spike") factorization: each pencil of n = P*m rows is cut into
// P segments of m = 32 rows held in the registers of one lane; each lane
//   1. evaluates the stencil for its rows from a register window of the input,
//   2. solves its local banded block (no pivoting, factors precomputed in FP64),
//   3. publishes the first and last K values of the local solution in LDS,
//   4. reads the interface values of its neighbours from the precomputed
//      inverse of the 2KP-unknown reduced system, and
//   5. corrects its rows with the precomputed spike vectors.
// Input is read once from HBM and the result written once: no intermediate
// global traffic.  The FP32 refinement mode repeats steps 2-5 on the residual
// b - A y formed in FP64 from the register window, so it costs no extra bytes.
//
// Layout of the two kernel shapes:
//   strided (y and z pencils): a workgroup owns 64 consecutive pencils along
//     the contiguous index (one lane each) and all P segments (one wave each),
//     so every load/store instruction of a wave touches 64 consecutive values.
//   x (contiguous pencils): a workgroup stages TP whole pencils in LDS with
//     coalesced loads, lanes are (pencil, segment) pairs reading their window
//     from LDS (segment rows padded by one element against bank conflicts),
//     and results go back through LDS to coalesced stores.
// Sizes outside the register path (n not a multiple of 32 or above 512) use a
// one-lane-per-pencil kernel with global scratch, computing in FP64.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "pdo_kernels.hpp"

namespace pdo {

template <int FAM> struct Fam;
//                                  K  R  sigma shift center
#define PDO_FAM(ID, K_, R_, S_, H_, C_)                                                   \
    template <> struct Fam<ID> {                                                          \
        static constexpr int K = K_, R = R_, SIGMA = S_, SHIFT = H_;                      \
        static constexpr bool CENTER = C_;                                                \
    };
PDO_FAM(0, 1, 2, -1, 0, false)   // cd06 d1
PDO_FAM(1, 1, 2, +1, 0, true)    // cd06 d2, staggered d2
PDO_FAM(2, 2, 3, -1, 0, false)   // cd10 d1
PDO_FAM(3, 2, 3, +1, 0, true)    // cd10 d2
PDO_FAM(4, 2, 4, +1, 0, true)    // cf90
PDO_FAM(5, 1, 2, -1, +1, false)  // staggered ddz E2C
PDO_FAM(6, 1, 2, -1, -1, false)  // staggered ddz C2E
PDO_FAM(7, 1, 2, +1, +1, false)  // staggered interp E2C
PDO_FAM(8, 1, 2, +1, -1, false)  // staggered interp C2E
#undef PDO_FAM

template <int MODE> struct Mode;
template <> struct Mode<MODE_F64> { using TS = double; using TC = double; static constexpr bool REFINE = false; };
template <> struct Mode<MODE_F32R> { using TS = float; using TC = float; static constexpr bool REFINE = true; };
template <> struct Mode<MODE_F32S> { using TS = float; using TC = double; static constexpr bool REFINE = false; };
template <> struct Mode<MODE_F32> { using TS = float; using TC = float; static constexpr bool REFINE = false; };


template <typename T> __device__ __forceinline__ const T *tab(const ApplyArgs &A);
template <> __device__ __forceinline__ const double *tab<double>(const ApplyArgs &A) { return A.tab_d; }
template <> __device__ __forceinline__ const float *tab<float>(const ApplyArgs &A) { return A.tab_f; }


// interior stencil of output row j; window index of input row (row0 + j + o)
// is j + R + o.
template <int FAM, typename T, typename TW, int W>
__device__ __forceinline__ T stencil_row(const TW (&win)[W], int j, const T *c)
{
    using F = Fam<FAM>;
    T acc = F::CENTER ? c[0] * (T)win[j + F::R] : T(0);
#pragma unroll
    for (int q = 1; q <= F::R; ++q) {
        const int hi = j + F::R + q + (F::SHIFT < 0 ? -1 : 0);
        const int lo = j + F::R - q + (F::SHIFT > 0 ? 1 : 0);
        const T pr = F::SIGMA > 0 ? (T)win[hi] + (T)win[lo] : (T)win[hi] - (T)win[lo];
        acc = fma(c[q], pr, acc);
    }
    return acc;
}

// right closure row against input columns [n_in - WBMAX, n_in): window index
// MSEG + R - WBMAX + q + DN with DN = n_in - n_sys (MSEG: rows of the last
// segment)
template <int FAM, int DN, int MSEG, typename T, typename TW, int W>
__device__ __forceinline__ T right_closure(const TW (&win)[W], const double *sp)
{
    constexpr int base = MSEG + Fam<FAM>::R - WBMAX + DN;
    static_assert(base >= 0 && base + WBMAX <= W, "closure window");
    T acc = 0;
#pragma unroll
    for (int q = 0; q < WBMAX; ++q) acc = fma((T)sp[q], (T)win[base + q], acc);
    return acc;
}

// RHS row j of segment `seg` with the closure rows of non-periodic systems.
// Closure rows are stored WBMAX wide: left rows against input columns
// [0, WBMAX), right rows against [n_in - WBMAX, n_in).  Staggered systems
// read one input point more (E2C: n_in = n_sys + 1) or less (C2E: n_in =
// n_sys - 1) than they solve; the uniform branch on the difference keeps
// every window index a compile-time constant.
template <int FAM, int MSEG, typename T, typename TW, int W>
__device__ __forceinline__ T right_closure_dn(const TW (&win)[W], const double *sp, int dn)
{
    if (dn == 0) return right_closure<FAM, 0, MSEG, T>(win, sp);
    if (dn > 0) return right_closure<FAM, 1, MSEG, T>(win, sp);
    return right_closure<FAM, -1, MSEG, T>(win, sp);
}

// (closure rows are at most NBMAX per end, so rows farther in are pruned at
// compile time; `ext`: this is the last segment of a system of 32 P + 1 rows,
// which holds row MAXM too)
template <int FAM, typename T, int XR = 0, typename TW, int W>
__device__ __forceinline__ T rhs_row(const TW (&win)[W], int j, const T *c, const ApplyArgs &A, int ty, int m,
                                     bool ext = false)
{
    using F = Fam<FAM>;
    static_assert(MAXM + F::R - WBMAX >= 1, "closure window");
    if (!A.periodic) {
        if ((ty == 0 || ty == 3) && j < NBMAX && j < A.nbl) {
            const double *sp = A.tab_d + A.o_sp + j * WBMAX;
            T acc = 0;
#pragma unroll
            for (int q = 0; q < WBMAX; ++q) acc = fma((T)sp[q], (T)win[q + F::R], acc);
            return acc;
        }
        const int dn = A.n_in - A.n_sys;
        if constexpr (XR > 0) {
            if ((ty == 2 || ty == 3) && ext) {
                if (j >= MAXM + 1 - NBMAX && j >= MAXM + 1 - A.nbr)
                    return right_closure_dn<FAM, MAXM + 1, T>(
                        win, A.tab_d + A.o_sp + (NBMAX + j - (MAXM + 1 - A.nbr)) * WBMAX, dn);
                return stencil_row<FAM, T>(win, j, c);
            }
        }
        if ((ty == 2 || ty == 3) && j >= m - NBMAX && j >= m - A.nbr)
            return right_closure_dn<FAM, MAXM, T>(win, A.tab_d + A.o_sp + (NBMAX + j - (m - A.nbr)) * WBMAX, dn);
    }
    return stencil_row<FAM, T>(win, j, c);
}

// refined modes: form the first solve's RHS in FP64 (shared with the
// residual) for the wide stencils of periodic systems, in FP32 otherwise.
// 512^3 periodic passes: cf90 x/y/z 0.246/0.266/0.270 ms with FP64,
// 0.273/0.288/0.300 with FP32; cd06 x/y 0.207/0.222 with FP64, 0.195/0.217
// with FP32 (88 instead of 108 VGPRs).  Non-periodic systems (closure rows,
// more live registers) gain most from FP32: the 512 x 512 x 256 wall-normal
// cd06stagg passes went from 0.43 / 0.37 ms to 0.23 / 0.13 ms.  Tuning
// choice.
template <int FAM, bool PER> constexpr bool rhs1_fp64() { return PER && Fam<FAM>::R > 2; }

// local banded solve, in place, factors fac[j*5 + {l1,l2,inv,u1,u2}]
// Scheduling fence every SB rows: the row coefficients are uniform scalar
// loads; without a fence the scheduler hoists all of them (hundreds of
// values) and spills scalar registers into vector lanes.
#define PDO_SCHED_FENCE(j, SB) \
    if (((j) % (SB)) == (SB) - 1) __builtin_amdgcn_sched_barrier(0)

template <int K, int MM, bool EXACT, typename T, typename PT>
__device__ __forceinline__ void local_solve(T (&d)[MM], PT fac, int m)
{
#pragma unroll
    for (int j = 1; j < MM; ++j) {
        if (EXACT || j < m) {
            T v = d[j] - fac[j * 5 + 0] * d[j - 1];
            if (K == 2 && j >= 2) v -= fac[j * 5 + 1] * d[j - 2];
            d[j] = v;
        }
        PDO_SCHED_FENCE(j, 8);
    }
#pragma unroll
    for (int j = MM - 1; j >= 0; --j) {
        if (EXACT || j < m) {
            T v = d[j] * fac[j * 5 + 2];
            if (j + 1 < MM && (EXACT || j + 1 < m)) v -= fac[j * 5 + 3] * d[j + 1];
            if (K == 2 && j + 2 < MM && (EXACT || j + 2 < m)) v -= fac[j * 5 + 4] * d[j + 2];
            d[j] = v;
        }
        PDO_SCHED_FENCE(j, 8);
    }
}

// the same for segments of MAXM rows, or MAXM + 1 when `ext` (MM = MAXM + 1)
template <int K, int MM, typename T, typename PT>
__device__ __forceinline__ void local_solve_x(T (&d)[MM], PT fac, bool ext)
{
#pragma unroll
    for (int j = 1; j < MM; ++j) {
        if (j < MAXM || ext) {
            T v = d[j] - fac[j * 5 + 0] * d[j - 1];
            if (K == 2 && j >= 2) v -= fac[j * 5 + 1] * d[j - 2];
            d[j] = v;
        }
        PDO_SCHED_FENCE(j, 8);
    }
#pragma unroll
    for (int j = MM - 1; j >= 0; --j) {
        if (j < MAXM || ext) {
            T v = d[j] * fac[j * 5 + 2];
            if (j + 1 < MM && (j + 1 < MAXM || ext)) v -= fac[j * 5 + 3] * d[j + 1];
            if (K == 2 && j + 2 < MM && (j + 2 < MAXM || ext)) v -= fac[j * 5 + 4] * d[j + 2];
            d[j] = v;
        }
        PDO_SCHED_FENCE(j, 8);
    }
}

template <int K, int MM, typename T, typename PT>
__device__ __forceinline__ void apply_spikes_x(T (&d)[MM], PT spk, const T (&yl)[2], const T (&yr)[2], bool ext)
{
#pragma unroll
    for (int j = 0; j < MM; ++j) {
        if (j < MAXM || ext) {
            T v = d[j];
#pragma unroll
            for (int q = 0; q < K; ++q) {
                v = fma(-spk[j * 4 + q], yl[q], v);
                v = fma(-spk[j * 4 + 2 + q], yr[q], v);
            }
            d[j] = v;
        }
        PDO_SCHED_FENCE(j, 8);
    }
}

// publish the first and last K rows of a segment (the last K are rows
// MAXM - K .. MAXM - 1, or one further when ext) at slots s*2K + r of an
// exchange buffer with NPW values per slot
template <int K, int NPW, int MM, typename T>
__device__ __forceinline__ void publish_ends(T *rb, const T (&d)[MM], int s, int pl, bool ext)
{
#pragma unroll
    for (int q = 0; q < K; ++q) {
        rb[(s * 2 * K + q) * NPW + pl] = d[q];
        T v = d[MAXM - K + q];
        if constexpr (MM > MAXM) v = ext ? d[MAXM + 1 - K + q] : v;
        rb[(s * 2 * K + K + q) * NPW + pl] = v;
    }
}

template <int K, int MM, bool EXACT, typename T, typename PT>
__device__ __forceinline__ void apply_spikes(T (&d)[MM], PT spk, const T (&yl)[2], const T (&yr)[2], int m)
{
#pragma unroll
    for (int j = 0; j < MM; ++j) {
        if (EXACT || j < m) {
            T v = d[j];
#pragma unroll
            for (int q = 0; q < K; ++q) {
                v = fma(-spk[j * 4 + q], yl[q], v);
                v = fma(-spk[j * 4 + 2 + q], yr[q], v);
            }
            d[j] = v;
        }
        PDO_SCHED_FENCE(j, 8);
    }
}

// ---------------------------------------------------------------------------
// raw buffer access: a uniform resource over a field plus a per-lane and a
// scalar byte offset, so that rows of a strided pencil (scalar offsets
// row * stride) cost no vector ALU work.  Byte offsets are 32-bit: the host
// sends fields of 4 GiB or more to the general kernel.
// ---------------------------------------------------------------------------
__device__ __forceinline__ __amdgpu_buffer_rsrc_t field_rsrc(const void *p)
{
    // stride 0, no range limit, 32-bit data format (raw dword access)
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), (short)0, -1, 0x00020000);
}
template <typename T> __device__ __forceinline__ T buf_ld(__amdgpu_buffer_rsrc_t r, unsigned vo, unsigned so);
template <> __device__ __forceinline__ float buf_ld<float>(__amdgpu_buffer_rsrc_t r, unsigned vo, unsigned so)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, (int)vo, (int)so, 0));
}
template <> __device__ __forceinline__ double buf_ld<double>(__amdgpu_buffer_rsrc_t r, unsigned vo, unsigned so)
{
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, (int)vo, (int)so, 0));
}
__device__ __forceinline__ void buf_st(float v, __amdgpu_buffer_rsrc_t r, unsigned vo, unsigned so)
{
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, (int)vo, (int)so, 0);
}
__device__ __forceinline__ void buf_st(double v, __amdgpu_buffer_rsrc_t r, unsigned vo, unsigned so)
{
    using U2 = unsigned __attribute__((ext_vector_type(2)));
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(U2, v), r, (int)vo, (int)so, 0);
}

// strided pencils: periodic windows and the stores through scalar row
// offsets (1) or per-row lane address arithmetic (0); tuning constants
#ifndef PDO_STRIDED_WINDOW_BUFFER
#define PDO_STRIDED_WINDOW_BUFFER 1
#endif
#ifndef PDO_STRIDED_STORE_BUFFER
#define PDO_STRIDED_STORE_BUFFER 1
#endif
// largest row stride (bytes) read with scalar row offsets.  Rows farther
// apart (the z pass: whole planes) are addressed per row, which spaces the
// loads of a window out: issued back to back, a window's 36-40 loads to
// rows megabytes apart ran the 512^3 cd06 z pass at 0.30 ms instead of 0.24.
#ifndef PDO_STRIDED_WINDOW_MAXSTRIDE
#define PDO_STRIDED_WINDOW_MAXSTRIDE 65536u
#endif
// epilogue operand rows: rows per load batch, and the largest stride read
// with scalar row offsets (512^3 z pass with an operand: 0.41 ms with
// scalar row offsets, 0.36 ms with per-row addresses)
#ifndef PDO_STRIDED_EPI_CHUNK
#define PDO_STRIDED_EPI_CHUNK 16
#endif
#ifndef PDO_STRIDED_EPI_MAXSTRIDE
#define PDO_STRIDED_EPI_MAXSTRIDE 65536u
#endif
// workgroup-to-pencil-group interleave of the strided kernel (1: none)
#ifndef PDO_STRIDED_SWIZZLE
#define PDO_STRIDED_SWIZZLE 1
#endif

// ---------------------------------------------------------------------------
// strided pencils (y, z): NPW pencils x P segments per workgroup
// ---------------------------------------------------------------------------
// NPW = 64: one wave per segment (segment index uniform across the wave);
// NPW = 32: two segments per wave (lanes 0-31 and 32-63), half the workgroup
// size, so two workgroups share a CU and overlap their load, solve and store
// phases.  PER (periodic) systems have identical segments: factors and spikes
// are uniform, and the reduced inverse is block circulant, so the interface
// values are sums over relative segment offsets d with uniform coefficients
// G_0[., d] and per-lane LDS addresses.
template <int FAM, int MODE, int NPW, bool PER, typename PT>
__device__ __forceinline__ void reduce_seg(const typename Mode<MODE>::TC *red, const ApplyArgs &A, int s, int pl,
                                           typename Mode<MODE>::TC (&yl)[2], typename Mode<MODE>::TC (&yr)[2],
                                           PT Gtab, int D)
{
    using TC = typename Mode<MODE>::TC;
    constexpr int K = Fam<FAM>::K;
    const int P = A.P, NR = 2 * K * P;
    yl[0] = yl[1] = yr[0] = yr[1] = TC(0);
    if constexpr (PER) {
        // only the segment offsets whose coupling is above the solve's
        // tolerance are summed: dd in [-D, D] (all P when 2D+1 >= P)
        const PT G = Gtab;
        const int nd = (2 * D + 1 < P) ? 2 * D + 1 : P;
        const int d0 = (2 * D + 1 < P) ? P - D : 0;
        // segment t = s + dd (mod P) advances with dd: a running index with
        // one wrap per pass, the 2K values of a segment at immediate offsets
        int t = s + d0;
        t -= (t >= P) ? P : 0;
        for (int id = 0; id < nd; ++id) {
            int dd = d0 + id;
            dd -= (dd >= P) ? P : 0;
            const TC *zp = red + t * 2 * K * NPW + pl;
            TC zv[2 * K];
#pragma unroll
            for (int r = 0; r < 2 * K; ++r) zv[r] = zp[r * NPW];
#pragma unroll
            for (int r = 0; r < 2 * K; ++r) {
#pragma unroll
                for (int q = 0; q < K; ++q) {
                    yl[q] = fma(G[q * NR + dd * 2 * K + r], zv[r], yl[q]);
                    yr[q] = fma(G[(K + q) * NR + dd * 2 * K + r], zv[r], yr[q]);
                }
            }
            t = (t == P - 1) ? 0 : t + 1;
        }
    } else {
        const TC *G = tab<TC>(A) + A.o_G + (size_t)s * 2 * K * NR;
        for (int t = 0; t < P; ++t) {
#pragma unroll
            for (int r = 0; r < 2 * K; ++r) {
                const TC zv = red[(t * 2 * K + r) * NPW + pl];
#pragma unroll
                for (int q = 0; q < K; ++q) {
                    yl[q] = fma(G[q * NR + t * 2 * K + r], zv, yl[q]);
                    yr[q] = fma(G[(K + q) * NR + t * 2 * K + r], zv, yr[q]);
                }
            }
        }
    }
}

// the solve of the strided kernel: RHS from this lane's window, local solve,
// interface exchange through `red` (LDS), spikes, and the FP64-residual
// refinement; leaves this lane's output rows in d.  XR = 1: the last segment
// of a non-periodic system of 32 P + 1 rows holds MAXM + 1 rows (`ext`);
// the others MAXM.
template <int FAM, int MODE, int NPW, bool PER, int XR>
__device__ __forceinline__ void strided_compute(const ApplyArgs &A,
                                                const typename Mode<MODE>::TS (&win)[MAXM + XR + 2 * Fam<FAM>::R],
                                                typename Mode<MODE>::TC *red, int s, int pl, int ty,
                                                const typename Mode<MODE>::TC *fac, const typename Mode<MODE>::TC *spk,
                                                typename Mode<MODE>::TC (&d)[MAXM + XR], bool ext)
{
    using TC = typename Mode<MODE>::TC;
    using F = Fam<FAM>;
    constexpr int K = F::K, MM = MAXM + XR;
    constexpr bool REFINE = Mode<MODE>::REFINE;
    const int m = MAXM, P = A.P;
    // rows of this segment
    const int ms = MAXM + (ext ? 1 : 0);
    // RHS of the first solve: with refinement its error is what the FP64
    // residual removes, so it may be formed in the solve's arithmetic (TR1 =
    // TC, fewer live registers) or in FP64, which the compiler shares with
    // the residual's identical stencil (no extra FP64 work, 32 more live
    // values); see rhs1_fp64()
    using TR1 = typename std::conditional<REFINE && rhs1_fp64<FAM, PER>(), double, TC>::type;
    TR1 cs[5];
#pragma unroll
    for (int q = 0; q < 5; ++q) cs[q] = (TR1)A.c[q];
#pragma unroll
    for (int j = 0; j < MM; ++j) {
        if (PER) d[j] = (TC)stencil_row<FAM, TR1>(win, j, cs);
        else if (j < MAXM || ext) d[j] = (TC)rhs_row<FAM, TR1, XR>(win, j, cs, A, ty, m, ext);
        else d[j] = TC(0);
    }

    const TC *Gs = tab<TC>(A) + A.o_G;
    TC yl[2], yr[2];

    local_solve_x<K, MM>(d, fac, ext);
    publish_ends<K, NPW>(red, d, s, pl, ext);
    __syncthreads();
    // coupling reach of the first solve (and of the only one, unrefined)
    const int D1 = (sizeof(TC) == 8) ? A.D64 : (REFINE ? A.D32R : A.D32);
    reduce_seg<FAM, MODE, NPW, PER>(red, A, s, pl, yl, yr, Gs, D1);
    apply_spikes_x<K, MM>(d, spk, yl, yr, ext);

    if constexpr (REFINE) {
        // residual r = B f - A y in FP64.  The neighbour rows it needs (rows
        // -K..-1 and ms..ms+K-1) are taken from the values the neighbouring
        // segments actually hold, published through LDS, so that the residual
        // belongs to the stored solution and the correction removes its error.
        TC *red2 = red + P * 2 * K * NPW;
        TC *red3 = red2 + P * 2 * K * NPW;
        publish_ends<K, NPW>(red3, d, s, pl, ext);
        __syncthreads();
        {
            const int pv = s > 0 ? s - 1 : (PER ? P - 1 : -1);
            const int nx = s < P - 1 ? s + 1 : (PER ? 0 : -1);
#pragma unroll
            for (int q = 0; q < K; ++q) {
                yl[q] = pv >= 0 ? red3[(pv * 2 * K + K + q) * NPW + pl] : TC(0);
                yr[q] = nx >= 0 ? red3[(nx * 2 * K + q) * NPW + pl] : TC(0);
            }
        }
        // periodic rows all share the interior band (kernel arguments); so do
        // the rows of a non-periodic system farther than NBMAX from its ends,
        // the others read theirs from the table
        const double *band = A.tab_d + A.o_lhs + ty * (m + A.xr) * 5;
        double cr[5];
#pragma unroll
        for (int q = 0; q < 5; ++q) cr[q] = A.c[q];
        TC e[MM];
#pragma unroll
        for (int j = 0; j < MM; ++j) {
            if (!PER && !(j < MAXM || ext)) { e[j] = TC(0); continue; }
            double acc = PER ? stencil_row<FAM, double>(win, j, cr) : rhs_row<FAM, double, XR>(win, j, cr, A, ty, m, ext);
            const bool tab_row = !PER && (((ty == 0 || ty == 3) && j < NBMAX) ||
                                          ((ty == 2 || ty == 3) && j >= MAXM - NBMAX && j >= ms - NBMAX));
#pragma unroll
            for (int o = -K; o <= K; ++o) {
                const int jj = j + o;
                TC y;
                if (jj < 0) y = yl[K + jj];
                else if (jj < MAXM) y = d[jj];
                else if (XR == 0) y = yr[jj - MAXM < 1 ? jj - MAXM : 1];
                else if (jj == MAXM) y = ext ? d[MM - 1] : yr[0];
                else y = ext ? yr[jj - MAXM - 1] : yr[jj - MAXM < 1 ? jj - MAXM : 1];
                const double yv = (double)y;
                acc = fma(-(tab_row ? band[j * 5 + o + 2] : A.band[o + 2]), yv, acc);
            }
            e[j] = (TC)acc;
        }
        local_solve_x<K, MM>(e, fac, ext);
        publish_ends<K, NPW>(red2, e, s, pl, ext);
        __syncthreads();
        TC el[2], er[2];
        reduce_seg<FAM, MODE, NPW, PER>(red2, A, s, pl, el, er, Gs, A.D32C);
        apply_spikes_x<K, MM>(e, spk, el, er, ext);
#pragma unroll
        for (int j = 0; j < MM; ++j) d[j] += e[j];
    }

}

// rows of this lane's segment to out (strided by se_out), optionally
// out = aux +- d; row MAXM too when ext
template <int MODE, bool EPI, int XR>
__device__ __forceinline__ void strided_store(const ApplyArgs &A, const typename Mode<MODE>::TC (&d)[MAXM + XR],
                                              long a, int s, bool ext)
{
    using TS = typename Mode<MODE>::TS;
    using TC = typename Mode<MODE>::TC;
    constexpr int MM = MAXM;
    const int m = MM;
    constexpr unsigned SZ = sizeof(TS);
    const __amdgpu_buffer_rsrc_t ro = field_rsrc(A.out);
    const __amdgpu_buffer_rsrc_t ra = field_rsrc(A.aux ? A.aux : A.out);
    const unsigned po = (unsigned)(a + (long)blockIdx.y * A.sb_out);
    const unsigned so = (unsigned)A.se_out;
    const unsigned vo = (po + (unsigned)(s * m) * so) * SZ;  // row 0 of this segment
    const unsigned ss = so * SZ;
    if (EPI) {
        // operand rows are read in chunks before the chunk's output rows are
        // written (the operand may be the output itself), so each chunk's
        // loads issue together without holding all 32 rows in registers
        constexpr int CH = sizeof(TS) == 8 ? 8 : PDO_STRIDED_EPI_CHUNK;
        const bool far = ss > PDO_STRIDED_EPI_MAXSTRIDE;
#pragma unroll
        for (int j0 = 0; j0 < MM; j0 += CH) {
            TS av[CH];
            if (far) {
#pragma unroll
                for (int j = 0; j < CH; ++j) av[j] = buf_ld<TS>(ra, vo + (unsigned)(j0 + j) * ss, 0u);
            } else {
#pragma unroll
                for (int j = 0; j < CH; ++j) av[j] = buf_ld<TS>(ra, vo, (unsigned)(j0 + j) * ss);
            }
#pragma unroll
            for (int j = 0; j < CH; ++j) {
                const TC v = A.epi == 1 ? (TC)av[j] + d[j0 + j] : (TC)av[j] - d[j0 + j];
                buf_st((TS)v, ro, vo, (unsigned)(j0 + j) * ss);
            }
        }
    } else if (PDO_STRIDED_STORE_BUFFER) {
#pragma unroll
        for (int j = 0; j < MM; ++j) buf_st((TS)d[j], ro, vo, (unsigned)j * ss);
    } else {
        TS *o = (TS *)A.out;
#pragma unroll
        for (int j = 0; j < MM; ++j) o[po + (unsigned)(s * m + j) * so] = (TS)d[j];
    }
    if constexpr (XR > 0) {
        // the extra row of the last segment (no epilogue on this path)
        if (ext) buf_st((TS)d[MAXM], ro, vo, (unsigned)MAXM * ss);
    }
    if (A.n_out > A.n_sys && s == 0) {
        // periodic edge layouts repeat row 0 at row n_sys
        const unsigned off = (unsigned)A.n_sys * ss;
        TC v = d[0];
        if (EPI) {
            const TC av = (TC)buf_ld<TS>(ra, vo, off);
            v = A.epi == 1 ? av + v : av - v;
        }
        buf_st((TS)v, ro, vo, off);
    }
}

// LB: workgroup size bound, 1024 (up to 16 segments) or 512 (up to 8, which
// leaves the compiler 256 vector registers per lane)
template <int FAM, int MODE, int NPW, bool PER, bool EPI, int XR, int LB>
__global__ void __launch_bounds__(LB) k_strided(ApplyArgs A)
{
    using TS = typename Mode<MODE>::TS;
    using TC = typename Mode<MODE>::TC;
    using F = Fam<FAM>;
    constexpr int K = F::K, R = F::R, MM = MAXM, W = MM + XR + 2 * R;
    static_assert(NPW == 64 || PER, "two segments per wave need uniform (periodic) tables");
    static_assert(XR == 0 || (!PER && !EPI), "extra rows: non-periodic systems without epilogue");
    const int m = MM, P = A.P;
    const int lane = threadIdx.x & 63;
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int pl = lane % NPW;
    const int s = (NPW == 64) ? wv : wv * 2 + (lane >> 5);
    // consecutive workgroups take pencil groups gridDim.x / SWZ apart, so the
    // workgroups in flight spread over the plane rather than one strip
    unsigned bx = blockIdx.x;
    if (PDO_STRIDED_SWIZZLE > 1 && gridDim.x % PDO_STRIDED_SWIZZLE == 0)
        bx = (bx % PDO_STRIDED_SWIZZLE) * (gridDim.x / PDO_STRIDED_SWIZZLE) + bx / PDO_STRIDED_SWIZZLE;
    const long a = (long)bx * NPW + pl;
    const bool act = a < A.na;
    const int ty = PER ? 1 : (P == 1 ? 3 : (s == 0 ? 0 : (s == P - 1 ? 2 : 1)));
    const bool ext = XR > 0 && s == P - 1;

    // byte offsets fit 32 bits (the host sends larger fields elsewhere)
    const __amdgpu_buffer_rsrc_t rs = field_rsrc(A.f);
    const unsigned pin = (unsigned)((act ? a : 0) + (long)blockIdx.y * A.sb_in);
    const unsigned se = (unsigned)A.se_in;
    constexpr unsigned SZ = sizeof(TS);
    TS win[W];
    if (PDO_STRIDED_WINDOW_BUFFER && PER && A.n_read == A.n_sys && se * SZ <= PDO_STRIDED_WINDOW_MAXSTRIDE) {
        // rows [s*m - R, s*m + m + R) of a periodic pencil: the halos wrap
        // as whole blocks, so each region has one per-lane base (already
        // wrapped) and row offsets that are scalar
        const unsigned ss = se * SZ;
        const unsigned vl = (pin + (unsigned)((s == 0 ? A.n_sys : s * m) - R) * se) * SZ;
        const unsigned vm = (pin + (unsigned)(s * m) * se) * SZ;
        const unsigned vr = (pin + (unsigned)(s == P - 1 ? 0 : (s + 1) * m) * se) * SZ;
#pragma unroll
        for (int w = 0; w < W; ++w) {
            if (w < R) win[w] = buf_ld<TS>(rs, vl, (unsigned)w * ss);
            else if (w < R + MM) win[w] = buf_ld<TS>(rs, vm, (unsigned)(w - R) * ss);
            else win[w] = buf_ld<TS>(rs, vr, (unsigned)(w - R - MM) * ss);
        }
    } else {
        const int r0 = s * m - R;
#pragma unroll
        for (int w = 0; w < W; ++w) {
            int r = r0 + w;
            bool ok = true;
            if (PER) r = r < 0 ? r + A.n_sys : (r >= A.n_read ? r - A.n_sys : r);
            else ok = (r >= 0) && (r < A.n_in);
            win[w] = ok ? buf_ld<TS>(rs, (pin + (unsigned)(ok ? r : 0) * se) * SZ, 0u) : TS(0);
        }
    }

    extern __shared__ unsigned char smem[];
    TC *red = (TC *)smem;
    const TC *fac = tab<TC>(A) + A.o_fac + ty * (m + A.xr) * 5;
    const TC *spk = tab<TC>(A) + A.o_spk + ty * (m + A.xr) * 4;
    TC d[MM + XR];
    strided_compute<FAM, MODE, NPW, PER, XR>(A, win, red, s, pl, ty, fac, spk, d, ext);
    if (!act) return;
    strided_store<MODE, EPI, XR>(A, d, a, s, ext);
}

// exchange buffers of the refined x kernel: 3 (one per exchange) or 1
// (shared, two more barriers; 8 KiB less LDS per workgroup for the
// pentadiagonal systems, which lets a fourth workgroup share a CU)
#ifndef PDO_XDIR_RED_SLOTS
#define PDO_XDIR_RED_SLOTS 1
#endif
static_assert(PDO_XDIR_RED_SLOTS == 1 || PDO_XDIR_RED_SLOTS == 3, "1 or 3 exchange buffers");

// ---------------------------------------------------------------------------
// contiguous pencils (x): TP pencils staged in LDS, lanes = (pencil, segment)
// ---------------------------------------------------------------------------
__device__ __forceinline__ int padrow(int r) { return r + (r >> 5); }  // one pad element per 32 rows

// the compute part of the x kernel: window from the LDS tile, RHS, local
// solve, interface exchange through `red`, spikes, and the FP64-residual
// refinement; leaves this lane's 32 output rows in d
template <int FAM, int MODE>
__device__ __forceinline__ void xdir_compute(const ApplyArgs &A, const typename Mode<MODE>::TS *tile,
                                             typename Mode<MODE>::TC *red, int TP, int LD, int tpa,
                                             const typename Mode<MODE>::TC *fac, const typename Mode<MODE>::TC *spk,
                                             const typename Mode<MODE>::TC *G, int tid,
                                             typename Mode<MODE>::TC (&d)[MAXM])
{
    constexpr bool PER = true;
    using TS = typename Mode<MODE>::TS;
    using TC = typename Mode<MODE>::TC;
    using F = Fam<FAM>;
    constexpr int K = F::K, R = F::R, MM = MAXM, W = MM + 2 * R;
    constexpr bool REFINE = Mode<MODE>::REFINE;
    const int m = MM, P = A.P, n = A.n_sys;
    const int s = tid % P, p = tid / P;
    const bool act = p < tpa;
    const int pc = act ? p : 0;
    const int ty = PER ? 1 : (P == 1 ? 3 : (s == 0 ? 0 : (s == P - 1 ? 2 : 1)));
    // window rows [s*m - R, s*m + m + R) of the periodic pencil; with the
    // padded pitch row r sits at r + r/32, so segment s starts at 33 s and
    // the halos wrap as whole blocks: three per-lane bases, constant offsets
    TS win[W];
    {
        const int P33 = n + (n >> 5);
        const TS *tm = tile + pc * LD + s * 33;
        const TS *tl = tile + pc * LD + (s == 0 ? P33 : s * 33) - R - 1;
        const TS *tr = tile + pc * LD + (s == P - 1 ? 0 : (s + 1) * 33);
#pragma unroll
        for (int w = 0; w < W; ++w) win[w] = w < R ? tl[w] : (w < R + MM ? tm[w - R] : tr[w - R - MM]);
    }
    // RHS of the first solve (see strided_compute)
    using TR1 = typename std::conditional<REFINE && rhs1_fp64<FAM, PER>(), double, TC>::type;
    TR1 cs[5];
#pragma unroll
    for (int q = 0; q < 5; ++q) cs[q] = (TR1)A.c[q];
#pragma unroll
    for (int j = 0; j < MM; ++j) d[j] = (TC)rhs_row<FAM, TR1>(win, j, cs, A, ty, m);

    const int NR = 2 * K * P;
    // exchange buffers hold value r of lane (p, s) at r * NT + p * P + s: the
    // lanes of a half-wave touch consecutive words (conflict-free)
    const int NT = TP * P;
    auto exchange = [&](TC *rb, TC (&v)[MM], TC (&yl)[2], TC (&yr)[2], int D) {
#pragma unroll
        for (int q = 0; q < K; ++q) {
            rb[q * NT + p * P + s] = v[q];
            rb[(K + q) * NT + p * P + s] = v[m - K + q];
        }
        __syncthreads();
        yl[0] = yl[1] = yr[0] = yr[1] = TC(0);
        const int nd = (2 * D + 1 < P) ? 2 * D + 1 : P;
        const int d0 = (2 * D + 1 < P) ? P - D : 0;
        // periodic systems are block circulant: G_s[.][t] = G_0[.][t - s].
        // The segment t = s + dd (mod P) advances with dd, so the lane keeps
        // a running LDS pointer (one wrap per pass) and the four values of a
        // segment are immediate offsets from it; dd is uniform.
        const TC *rp = rb + p * P;
        int t = s + d0;
        t -= (t >= P) ? P : 0;
        for (int id = 0; id < nd; ++id) {
            int dd = d0 + id;
            dd -= (dd >= P) ? P : 0;
            const TC *zp = rp + t;
            TC zv[2 * K];
#pragma unroll
            for (int r = 0; r < 2 * K; ++r) zv[r] = zp[r * NT];
#pragma unroll
            for (int r = 0; r < 2 * K; ++r) {
#pragma unroll
                for (int q = 0; q < K; ++q) {
                    yl[q] = fma(G[q * NR + dd * 2 * K + r], zv[r], yl[q]);
                    yr[q] = fma(G[(K + q) * NR + dd * 2 * K + r], zv[r], yr[q]);
                }
            }
            t = (t == P - 1) ? 0 : t + 1;
        }
    };

    TC yl[2], yr[2];
    local_solve<K, MM, true>(d, fac, m);
    exchange(red, d, yl, yr, (sizeof(TC) == 8) ? A.D64 : (REFINE ? A.D32R : A.D32));
    apply_spikes<K, MM, true>(d, spk, yl, yr, m);
    if constexpr (REFINE) {
        // neighbour rows of the residual from the values the neighbours hold
        TC *rb3 = red + (size_t)(PDO_XDIR_RED_SLOTS == 3 ? 2 : 0) * TP * P * 2 * K;
        if (PDO_XDIR_RED_SLOTS == 1) __syncthreads();  // the exchange's reads of the slot are done
        (void)ty;
#pragma unroll
        for (int q = 0; q < K; ++q) {
            rb3[q * NT + p * P + s] = d[q];
            rb3[(K + q) * NT + p * P + s] = d[m - K + q];
        }
        __syncthreads();
        {
            const int pv = s > 0 ? s - 1 : P - 1;
            const int nx = s < P - 1 ? s + 1 : 0;
#pragma unroll
            for (int q = 0; q < K; ++q) {
                yl[q] = rb3[(K + q) * NT + p * P + pv];
                yr[q] = rb3[q * NT + p * P + nx];
            }
        }
        double cr[5];
#pragma unroll
        for (int q = 0; q < 5; ++q) cr[q] = A.c[q];
        TC e[MM];
#pragma unroll
        for (int j = 0; j < MM; ++j) {
            double acc = stencil_row<FAM, double>(win, j, cr);
#pragma unroll
            for (int o = -K; o <= K; ++o) {
                const int jj = j + o;
                const double yv = jj < 0 ? (double)yl[K + jj] : (jj >= m ? (double)yr[jj - m] : (double)d[jj < 0 ? 0 : (jj >= MM ? MM - 1 : jj)]);
                acc = fma(-A.band[o + 2], yv, acc);
            }
            e[j] = (TC)acc;
        }
        local_solve<K, MM, true>(e, fac, m);
        TC el[2], er[2];
        if (PDO_XDIR_RED_SLOTS == 1) __syncthreads();  // the neighbour reads of the slot are done
        exchange(red + (size_t)(PDO_XDIR_RED_SLOTS == 3 ? 1 : 0) * TP * P * 2 * K, e, el, er, A.D32C);
        apply_spikes<K, MM, true>(e, spk, el, er, m);
#pragma unroll
        for (int j = 0; j < MM; ++j) d[j] += e[j];
    }
}

// write this lane's rows d through the LDS tile to coalesced 16-byte
// stores of A.out (optionally out = aux +- result); all threads take part
template <int MODE, bool EPI>
__device__ __forceinline__ void xdir_store(const ApplyArgs &A, typename Mode<MODE>::TS *tile,
                                           const typename Mode<MODE>::TC (&d)[MAXM], int TP, int LD, int tpa)
{
    using TS = typename Mode<MODE>::TS;
    using TC = typename Mode<MODE>::TC;
    using V = typename std::conditional<sizeof(TS) == 8, double2, float4>::type;
    constexpr int VW = 16 / sizeof(TS);
    constexpr int MM = MAXM;
    const int m = MM, P = A.P;
    const int tid = threadIdx.x;
    const int s = tid % P, p = tid / P;
    const bool act = p < tpa;
    const long pen0 = (long)blockIdx.x * TP;
    const int lgn = A.lgn;
    const int nvec = (tpa * A.n_in) / VW;
    __syncthreads();  // every window read of the tile is done
    if (act) {
#pragma unroll
        for (int j = 0; j < MM; ++j) tile[p * LD + padrow(s * m + j)] = (TS)d[j];
    }
    __syncthreads();
    TS *__restrict__ out = (TS *)A.out + pen0 * (long)A.n_out;
    const TS *__restrict__ aux = (const TS *)A.aux + (A.aux ? pen0 * (long)A.n_out : 0);
    // 8192 elements per workgroup tile: 32 / VW vectors per thread
    constexpr int NV = 32 / VW;
    V ops[NV];
    if (EPI) {
        // operand vectors first (the operand may alias the output)
#pragma unroll
        for (int i = 0; i < NV; ++i) {
            const int v = tid + i * 256;
            if (v < nvec) ops[i] = reinterpret_cast<const V *>(aux)[v];
        }
    }
#pragma unroll
    for (int i = 0; i < NV; ++i) {
        const int v = tid + i * 256;
        if (v >= nvec) break;
        const int e = v * VW, pp = e >> lgn, r = e & (A.n_in - 1);
        const TS *src = tile + pp * LD + padrow(r);
        V x;
        TS *xs = reinterpret_cast<TS *>(&x);
#pragma unroll
        for (int q = 0; q < VW; ++q) xs[q] = src[q];
        if (EPI) {
            const TS *ys = reinterpret_cast<const TS *>(&ops[i]);
#pragma unroll
            for (int q = 0; q < VW; ++q) {
                const TC t = A.epi == 1 ? (TC)ys[q] + (TC)xs[q] : (TC)ys[q] - (TC)xs[q];
                xs[q] = (TS)t;
            }
        }
        reinterpret_cast<V *>(out)[v] = x;
    }
}

template <int FAM, int MODE, bool EPI>
__global__ void __launch_bounds__(256) k_xdir(ApplyArgs A, int TP, int LD)
{
    // periodic pencils only: every segment has the same factors (type 1)
    extern __shared__ unsigned char smem[];
    constexpr bool PER = true;
    using TS = typename Mode<MODE>::TS;
    using TC = typename Mode<MODE>::TC;
    using F = Fam<FAM>;
    constexpr int K = F::K, R = F::R, MM = MAXM, W = MM + 2 * R;
    const int m = MM, P = A.P, n = A.n_sys;
    const int tid = threadIdx.x;
    const int s = tid % P, p = tid / P;
    const long pen0 = (long)blockIdx.x * TP;
    const long npen = A.na * A.nb;
    const int tpa = (int)((npen - pen0) < TP ? (npen - pen0) : TP);  // pencils of this workgroup
    TS *tile = (TS *)smem;
    TC *red = (TC *)(smem + (size_t)TP * LD * sizeof(TS));

    // coalesced staging of tpa contiguous pencils: 16-byte loads, pencil and
    // row from shifts (n_in is 32 * 2^k on this path)
    using V = typename std::conditional<sizeof(TS) == 8, double2, float4>::type;
    constexpr int VW = 16 / sizeof(TS);
    const int lgn = A.lgn;
    const TS *__restrict__ f = (const TS *)A.f + pen0 * (long)A.n_in;
    const int nvec = (tpa * A.n_in) / VW;
#pragma unroll 8
    for (int v = tid; v < nvec; v += 256) {
        const V x = reinterpret_cast<const V *>(f)[v];
        const int e = v * VW, pp = e >> lgn, r = e & (A.n_in - 1);
        TS *dst = tile + pp * LD + padrow(r);
        const TS *xs = reinterpret_cast<const TS *>(&x);
#pragma unroll
        for (int i = 0; i < VW; ++i) dst[i] = xs[i];
    }
    __syncthreads();

    const bool act = p < tpa;
    TC d[MM];
    const TC *tb = tab<TC>(A);
    xdir_compute<FAM, MODE>(A, tile, red, TP, LD, tpa, tb + A.o_fac + MM * 5, tb + A.o_spk + MM * 4, tb + A.o_G,
                            (int)threadIdx.x, d);
    xdir_store<MODE, EPI>(A, tile, d, TP, LD, tpa);

}

// ---------------------------------------------------------------------------
// general kernel: one lane per pencil, any n, any strides; FP64 arithmetic,
// intermediate solution in a global FP64 scratch laid out [row][pencil]
// ---------------------------------------------------------------------------
template <int FAM, typename TS>
__global__ void __launch_bounds__(256) k_general(ApplyArgs A, double *scratch)
{
    using F = Fam<FAM>;
    constexpr int K = F::K, R = F::R;
    const long npen = A.na * A.nb;
    const long pid = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (pid >= npen) return;
    const long ia = pid % A.na, ib = pid / A.na;
    const long pin = ia * A.sa_in + ib * A.sb_in;
    const long po = ia * A.sa_out + ib * A.sb_out;
    const int n = A.n_sys;
    const TS *__restrict__ f = (const TS *)A.f;
    const double *fac = A.tab_d + A.o_fac + (A.periodic ? 1 : 3) * n * 5;
    const double *spk = A.tab_d + A.o_spk + (A.periodic ? 1 : 3) * n * 4;
    auto in = [&](int r) -> double {
        if (A.periodic) r = r < 0 ? r + n : (r >= A.n_read ? r - n : r);
        else if (r < 0 || r >= A.n_in) return 0.0;
        return (double)f[pin + (long)r * A.se_in];
    };
    auto scr = [&](int r) -> double & { return scratch[(long)r * npen + pid]; };
    double w1 = 0, w2 = 0;
    for (int i = 0; i < n; ++i) {
        double d;
        if (!A.periodic && i < A.nbl) {
            d = 0;
            for (int q = 0; q < WBMAX; ++q) d += A.tab_d[A.o_sp + i * WBMAX + q] * in(q);
        } else if (!A.periodic && i >= n - A.nbr) {
            d = 0;
            const int jr = i - (n - A.nbr);
            for (int q = 0; q < WBMAX; ++q) d += A.tab_d[A.o_sp + (NBMAX + jr) * WBMAX + q] * in(A.n_in - WBMAX + q);
        } else {
            d = F::CENTER ? A.c[0] * in(i) : 0.0;
            for (int q = 1; q <= R; ++q) {
                const double hi = in(i + q + (F::SHIFT < 0 ? -1 : 0));
                const double lo = in(i - q + (F::SHIFT > 0 ? 1 : 0));
                d += A.c[q] * (F::SIGMA > 0 ? hi + lo : hi - lo);
            }
        }
        double v = d;
        if (i >= 1) v -= fac[i * 5 + 0] * w1;
        if (K == 2 && i >= 2) v -= fac[i * 5 + 1] * w2;
        scr(i) = v;
        w2 = w1;
        w1 = v;
    }
    double z1 = 0, z2 = 0;
    for (int i = n - 1; i >= 0; --i) {
        double v = scr(i) * fac[i * 5 + 2];
        if (i + 1 < n) v -= fac[i * 5 + 3] * z1;
        if (K == 2 && i + 2 < n) v -= fac[i * 5 + 4] * z2;
        scr(i) = v;
        z2 = z1;
        z1 = v;
    }
    double yl[2] = {0, 0}, yr[2] = {0, 0};
    if (A.periodic) {
        const int NR = 2 * K;
        double zh[4];
        for (int q = 0; q < K; ++q) { zh[q] = scr(q); zh[K + q] = scr(n - K + q); }
        for (int q = 0; q < K; ++q)
            for (int r = 0; r < NR; ++r) {
                yl[q] += A.tab_d[A.o_G + q * NR + r] * zh[r];
                yr[q] += A.tab_d[A.o_G + (K + q) * NR + r] * zh[r];
            }
    }
    TS *__restrict__ out = (TS *)A.out;
    const TS *__restrict__ aux = (const TS *)A.aux;
    double first = 0;
    for (int i = 0; i < n; ++i) {
        double v = scr(i);
        for (int q = 0; q < K; ++q) v -= spk[i * 4 + q] * yl[q] + spk[i * 4 + 2 + q] * yr[q];
        if (i == 0) first = v;
        const long idx = po + (long)i * A.se_out;
        if (A.epi == 1) v = (double)aux[idx] + v;
        else if (A.epi == 2) v = (double)aux[idx] - v;
        out[idx] = (TS)v;
    }
    if (A.n_out > n) {
        const long idx = po + (long)n * A.se_out;
        double v = first;
        if (A.epi == 1) v = (double)aux[idx] + v;
        else if (A.epi == 2) v = (double)aux[idx] - v;
        out[idx] = (TS)v;
    }
}

__global__ void k_zero_words(unsigned *p, long words)
{
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < words; i += (long)gridDim.x * blockDim.x) p[i] = 0u;
}

hipError_t launch_zero_words(void *p, long words, hipStream_t st)
{
    if (words <= 0) return hipSuccess;
    long blocks = (words + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(k_zero_words, dim3((unsigned)blocks), dim3(256), 0, st, (unsigned *)p, words);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
template <int FAM, int MODE, int NPW, bool PER, bool EPI, int XR = 0, int LB = 1024>
static hipError_t launch_strided_k(const ApplyArgs &A, hipStream_t st)
{
    using TC = typename Mode<MODE>::TC;
    constexpr int K = Fam<FAM>::K;
    dim3 grid((unsigned)((A.na + NPW - 1) / NPW), (unsigned)A.nb);
    dim3 block(NPW * A.P);
    if ((int)block.x > LB || A.xr != XR) return hipErrorInvalidValue;
    size_t lds = (size_t)A.P * 2 * K * NPW * sizeof(TC) * (Mode<MODE>::REFINE ? 3 : 1);
    hipLaunchKernelGGL((k_strided<FAM, MODE, NPW, PER, EPI, XR, LB>), grid, block, lds, st, A);
    return hipGetLastError();
}

// systems of 32 P + 1 rows on the register path: the edge systems of the
// staggered operators (tridiagonal families)
template <int FAM> constexpr bool has_extra_row() { return Fam<FAM>::K == 1; }

// epilogue operands (divergence accumulation, projection subtraction) are
// only used with the periodic cd06 first derivative
template <int FAM> constexpr bool has_epilogue() { return FAM == 0; }

// pencils per workgroup of the periodic strided kernel: 32 (two segments per
// wave) or 64 (one); tuning constant
#ifndef PDO_STRIDED_NPW_PER
#define PDO_STRIDED_NPW_PER 32
#endif

template <int FAM, int MODE>
static hipError_t launch_strided_t(const ApplyArgs &A, hipStream_t st)
{
    if (!A.periodic) {
        if (A.epi) return hipErrorInvalidValue;
        if (A.xr) {
            if constexpr (has_extra_row<FAM>()) {
                if (A.P <= 8) return launch_strided_k<FAM, MODE, 64, false, false, 1, 512>(A, st);
                return launch_strided_k<FAM, MODE, 64, false, false, 1, 1024>(A, st);
            }
            return hipErrorInvalidValue;
        }
        if (A.P <= 8) return launch_strided_k<FAM, MODE, 64, false, false, 0, 512>(A, st);
        return launch_strided_k<FAM, MODE, 64, false, false, 0, 1024>(A, st);
    }
    if (A.epi) {
        if constexpr (has_epilogue<FAM>()) {
            if (PDO_STRIDED_NPW_PER == 32 && A.P % 2 == 0) return launch_strided_k<FAM, MODE, 32, true, true>(A, st);
            return launch_strided_k<FAM, MODE, 64, true, true>(A, st);
        }
        return hipErrorInvalidValue;
    }
    if (PDO_STRIDED_NPW_PER == 32 && A.P % 2 == 0) return launch_strided_k<FAM, MODE, 32, true, false>(A, st);
    return launch_strided_k<FAM, MODE, 64, true, false>(A, st);
}

template <int FAM, int MODE>
static hipError_t launch_x_t(const ApplyArgs &A, hipStream_t st)
{
    using TS = typename Mode<MODE>::TS;
    using TC = typename Mode<MODE>::TC;
    constexpr int K = Fam<FAM>::K;
    const int P = A.P;
    int TP = 256 / P;
    // pencil pitch n + n/32 = 33 P: the P segments of consecutive pencils in
    // a half-wave start 33 s + 33 P p words in, all in distinct banks
    const int LD = A.n_in + (A.n_in >> 5);
    auto lds_of = [&](int tp) {
        return (size_t)tp * LD * sizeof(TS) +
               (size_t)tp * P * 2 * K * sizeof(TC) * (Mode<MODE>::REFINE ? PDO_XDIR_RED_SLOTS : 1);
    };
    while (TP > 1 && lds_of(TP) > 80 * 1024) TP >>= 1;
    if (TP * P != 256) return hipErrorInvalidValue;
    const long npen = A.na * A.nb;
    dim3 grid((unsigned)((npen + TP - 1) / TP));
    dim3 block(TP * P);
    if (A.epi) {
        if constexpr (has_epilogue<FAM>())
            hipLaunchKernelGGL((k_xdir<FAM, MODE, true>), grid, block, lds_of(TP), st, A, TP, LD);
        else
            return hipErrorInvalidValue;
    } else {
        hipLaunchKernelGGL((k_xdir<FAM, MODE, false>), grid, block, lds_of(TP), st, A, TP, LD);
    }
    return hipGetLastError();
}

template <int FAM, int MODE>
static hipError_t launch_general_t(const ApplyArgs &A, double *scratch, hipStream_t st)
{
    using TS = typename Mode<MODE>::TS;
    const long npen = A.na * A.nb;
    dim3 grid((unsigned)((npen + 255) / 256));
    hipLaunchKernelGGL((k_general<FAM, TS>), grid, dim3(256), 0, st, A, scratch);
    return hipGetLastError();
}

template <int FAM>
static hipError_t dispatch_mode(int shape, int mode, const ApplyArgs &A, double *scratch, hipStream_t st)
{
    switch (shape) {
    case SHAPE_STRIDED:
        switch (mode) {
        case MODE_F64: return launch_strided_t<FAM, MODE_F64>(A, st);
        case MODE_F32R: return launch_strided_t<FAM, MODE_F32R>(A, st);
        case MODE_F32S: return launch_strided_t<FAM, MODE_F32S>(A, st);
        case MODE_F32: return launch_strided_t<FAM, MODE_F32>(A, st);
        }
        break;
    case SHAPE_X:
        if constexpr (Fam<FAM>::SHIFT == 0) {
            switch (mode) {
            case MODE_F64: return launch_x_t<FAM, MODE_F64>(A, st);
            case MODE_F32R: return launch_x_t<FAM, MODE_F32R>(A, st);
            case MODE_F32S: return launch_x_t<FAM, MODE_F32S>(A, st);
            case MODE_F32: return launch_x_t<FAM, MODE_F32>(A, st);
            }
        }
        break;
    case SHAPE_GENERAL:
        if (mode == MODE_F64) return launch_general_t<FAM, MODE_F64>(A, scratch, st);
        return launch_general_t<FAM, MODE_F32S>(A, scratch, st);
    }
    return hipErrorInvalidValue;
}

hipError_t launch_compact(int family, int shape, int mode, const ApplyArgs &A, double *scratch, hipStream_t st)
{
    switch (family) {
    case 0: return dispatch_mode<0>(shape, mode, A, scratch, st);
    case 1: return dispatch_mode<1>(shape, mode, A, scratch, st);
    case 2: return dispatch_mode<2>(shape, mode, A, scratch, st);
    case 3: return dispatch_mode<3>(shape, mode, A, scratch, st);
    case 4: return dispatch_mode<4>(shape, mode, A, scratch, st);
    case 5: return dispatch_mode<5>(shape, mode, A, scratch, st);
    case 6: return dispatch_mode<6>(shape, mode, A, scratch, st);
    case 7: return dispatch_mode<7>(shape, mode, A, scratch, st);
    case 8: return dispatch_mode<8>(shape, mode, A, scratch, st);
    }
    return hipErrorInvalidValue;
}

}  // namespace pdo
