// Periodic FFT Poisson solver kernels for gfx950.
//
// A real field f(nx, ny, nz) (x fastest) is transformed in five passes, each
// one HBM read and one HBM write of the half spectrum (nx/2+1, ny, nz):
//   1. x   real-to-complex: nx/2-point complex FFT of the packed line of
//          (even, odd) samples, then the split step to the nx/2+1 modes
//   2. y   complex FFT of every y line
//   3. z   complex FFT, multiply by -1/|k'|^2 (modified wavenumbers), inverse
//          complex FFT -- one pass, the line never leaves LDS
//   4. y   inverse complex FFT
//   5. x   complex-to-real (inverse split + nx/2-point inverse FFT), scaled by
//          1/(nx ny nz)
// Each pass runs a mixed radix-8/4/2 Stockham FFT on a tile of lines: the
// first stage reads its operands straight from HBM, the intermediate stages
// go through LDS, and the last stage writes straight to HBM, so a 512-point
// line costs two LDS round trips.  Every thread holds exactly 8 complex
// values per stage (one radix-8, two radix-4 or four radix-2 butterflies).
// For y and z the tile is TX consecutive lines, so each global row access is
// a contiguous run of TX complex values; twiddles come from a table.
#include <hip/hip_runtime.h>

#include "pdo_fft.hpp"

namespace pdo {

// complex value; aligned to its size so that LDS and HBM accesses are single
// 8-byte (FP32) or 16-byte (FP64) operations rather than pairs of words
template <typename T> struct alignas(2 * sizeof(T)) C2 { T x, y; };

template <typename T> __device__ __forceinline__ C2<T> cmul(C2<T> a, C2<T> b)
{
    return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <typename T> __device__ __forceinline__ C2<T> cadd(C2<T> a, C2<T> b) { return {a.x + b.x, a.y + b.y}; }
template <typename T> __device__ __forceinline__ C2<T> csub(C2<T> a, C2<T> b) { return {a.x - b.x, a.y - b.y}; }
// multiply by SIGN*i
template <typename T, int SIGN> __device__ __forceinline__ C2<T> mul_i(C2<T> a)
{
    return SIGN > 0 ? C2<T>{-a.y, a.x} : C2<T>{a.y, -a.x};
}

template <typename T, int SIGN> __device__ __forceinline__ void dft2(C2<T> &a, C2<T> &b)
{
    const C2<T> t = csub(a, b);
    a = cadd(a, b);
    b = t;
}

// y_m = sum_r x_r (SIGN i)^(r m), in place, on x[o], x[o+s], x[o+2s], x[o+3s]
template <typename T, int SIGN> __device__ __forceinline__ void dft4(C2<T> &x0, C2<T> &x1, C2<T> &x2, C2<T> &x3)
{
    const C2<T> a0 = cadd(x0, x2), a1 = csub(x0, x2);
    const C2<T> b0 = cadd(x1, x3), b1 = mul_i<T, SIGN>(csub(x1, x3));
    x0 = cadd(a0, b0);
    x1 = cadd(a1, b1);
    x2 = csub(a0, b0);
    x3 = csub(a1, b1);
}

// 8-point DFT with exp(SIGN 2 pi i / 8): radix-2 split into two DFT4
template <typename T, int SIGN> __device__ __forceinline__ void dft8(C2<T> (&v)[8])
{
    const T h = (T)0.70710678118654752440084436210484903928;
    C2<T> a[4], b[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        a[k] = cadd(v[k], v[k + 4]);
        b[k] = csub(v[k], v[k + 4]);
    }
    // b_k *= W8^k
    b[1] = cmul(b[1], C2<T>{h, SIGN * h});
    b[2] = mul_i<T, SIGN>(b[2]);
    b[3] = cmul(b[3], C2<T>{-h, SIGN * h});
    dft4<T, SIGN>(a[0], a[1], a[2], a[3]);
    dft4<T, SIGN>(b[0], b[1], b[2], b[3]);
#pragma unroll
    for (int m = 0; m < 4; ++m) {
        v[2 * m] = a[m];
        v[2 * m + 1] = b[m];
    }
}

// One Stockham stage of radix R on one line of length L, current sub-size Ns.
// The thread owns base index jb in [0, L/8) and handles items
// j = jb + it*(L/8), it < 8/R.  LD(i) reads element i of the stage input,
// ST(i, v) writes element i of the stage output.  The caller synchronizes.
template <typename T, int R, int SIGN> struct Stage {
    template <class LD>
    __device__ __forceinline__ static void load(C2<T> (&v)[8], LD ld, const C2<T> *__restrict__ tw, int L, int Ns,
                                                int jb)
    {
        const int Q = L / R, L8 = L >> 3;
#pragma unroll
        for (int it = 0; it < 8 / R; ++it) {
            const int j = jb + it * L8;
            const int k = j & (Ns - 1);
            const int step = L / (Ns * R);
#pragma unroll
            for (int r = 0; r < R; ++r) {
                C2<T> x = ld(j + r * Q);
                if (r > 0 && Ns > 1) {  // the first stage's twiddles are all 1
                    C2<T> w = tw[(k * r * step) & (L - 1)];
                    if (SIGN > 0) w.y = -w.y;
                    x = cmul(x, w);
                }
                v[it * R + r] = x;
            }
        }
    }
    __device__ __forceinline__ static void butterfly(C2<T> (&v)[8])
    {
        if constexpr (R == 8) {
            dft8<T, SIGN>(v);
        } else if constexpr (R == 4) {
            dft4<T, SIGN>(v[0], v[1], v[2], v[3]);
            dft4<T, SIGN>(v[4], v[5], v[6], v[7]);
        } else {
#pragma unroll
            for (int it = 0; it < 4; ++it) dft2<T, SIGN>(v[2 * it], v[2 * it + 1]);
        }
    }
    template <class STO>
    __device__ __forceinline__ static void store(const C2<T> (&v)[8], STO st, int L, int Ns, int jb)
    {
        const int L8 = L >> 3;
#pragma unroll
        for (int it = 0; it < 8 / R; ++it) {
            const int j = jb + it * L8;
            const int k = j & (Ns - 1);
            const int base = (j - k) * R + k;
#pragma unroll
            for (int r = 0; r < R; ++r) st(base + r * Ns, v[it * R + r]);
        }
    }
};

// Full FFT of one line: the first stage reads through LDG, the last stage
// writes through STG, the intermediate stages live in the LDS line accessor
// `buf` (element i at buf(i)).  SRC_LDS: LDG reads the same LDS buffer, so the
// first stage also waits for every read before writing.  All threads of the
// block must call this (it synchronizes); threads with act == false only take
// part in the barriers.
template <typename T, int SIGN, bool SRC_LDS, int L, int NS, class LDG, class STG, class BUF>
__device__ __forceinline__ void line_fft_stage(LDG &ldg, STG &stg, BUF &buf, const C2<T> *__restrict__ tw, int jb,
                                               bool act)
{
    constexpr int R = (L / NS) >= 8 ? 8 : (L / NS);
    constexpr bool first = (NS == 1), last = (NS * R == L);
    C2<T> v[8];
    if (act) {
        if constexpr (first) {
            Stage<T, R, SIGN>::load(v, ldg, tw, L, NS, jb);
        } else {
            auto lds = [&](int i) { return buf(i); };
            Stage<T, R, SIGN>::load(v, lds, tw, L, NS, jb);
        }
        Stage<T, R, SIGN>::butterfly(v);
    }
    if (!first || SRC_LDS) __syncthreads();  // every LDS read of this stage is done
    if (act) {
        if constexpr (last) {
            Stage<T, R, SIGN>::store(v, stg, L, NS, jb);
        } else {
            auto sts = [&](int i, C2<T> x) { buf(i) = x; };
            Stage<T, R, SIGN>::store(v, sts, L, NS, jb);
        }
    }
    if constexpr (!last) {
        __syncthreads();  // stage output visible
        line_fft_stage<T, SIGN, SRC_LDS, L, NS * R>(ldg, stg, buf, tw, jb, act);
    }
}

// Full FFT of one line of L points: the first stage reads through LDG, the
// last stage writes through STG, the intermediate stages live in the LDS line
// accessor `buf` (element i at buf(i)).  SRC_LDS: LDG reads the same LDS
// buffer, so the first stage also waits for every read before writing.  All
// threads of the block must call this (it synchronizes); threads with
// act == false only take part in the barriers.
template <typename T, int SIGN, bool SRC_LDS, int L, class LDG, class STG, class BUF>
__device__ __forceinline__ void line_fft(LDG ldg, STG stg, BUF buf, const C2<T> *__restrict__ tw, int jb, bool act)
{
    line_fft_stage<T, SIGN, SRC_LDS, L, 1>(ldg, stg, buf, tw, jb, act);
}

template <typename TS, typename TC> struct CLoad;
template <> struct CLoad<double, double> {
    __device__ __forceinline__ static C2<double> ld(const double *p, long i)
    {
        const double2 x = reinterpret_cast<const double2 *>(p)[i];
        return {x.x, x.y};
    }
    __device__ __forceinline__ static void st(double *p, long i, C2<double> v)
    {
        reinterpret_cast<double2 *>(p)[i] = double2{v.x, v.y};
    }
};
template <typename TC> struct CLoadF {
    __device__ __forceinline__ static C2<TC> ld(const float *p, long i)
    {
        const float2 x = reinterpret_cast<const float2 *>(p)[i];
        return {(TC)x.x, (TC)x.y};
    }
    __device__ __forceinline__ static void st(float *p, long i, C2<TC> v)
    {
        reinterpret_cast<float2 *>(p)[i] = float2{(float)v.x, (float)v.y};
    }
};
template <> struct CLoad<float, float> : CLoadF<float> {};
template <> struct CLoad<float, double> : CLoadF<double> {};

// 1/x: correctly rounded in FP64; the hardware reciprocal (1 ulp) in FP32
__device__ __forceinline__ double recip(double x) { return 1.0 / x; }
__device__ __forceinline__ float recip(float x) { return __builtin_amdgcn_rcpf(x); }

// twiddles of the strided transforms staged in LDS (1) or read from the
// table in HBM / cache (0); tuning constant
#ifndef PDO_FFT_TW_LDS
#define PDO_FFT_TW_LDS 1
#endif

// the same for the x-line transforms; tuning constant
#ifndef PDO_FFT_X_TW_LDS
#define PDO_FFT_X_TW_LDS PDO_FFT_TW_LDS
#endif

// LDS rows of the strided transforms swizzled (1) or plain (0); tuning constant
#ifndef PDO_FFT_STRIDED_SWZ
#define PDO_FFT_STRIDED_SWZ 1
#endif

// ---------------------------------------------------------------------------
// passes 2-4: complex FFT of strided lines.  Line (a, b) starts at a + b*sb
// (complex units), element stride se; a in [0, na) is the contiguous index.
// Tile: TX = 1 << lgtx consecutive lines; thread = (line, jb), line fastest.
// MODE 0 forward, 1 inverse, 2 forward -> Poisson multiply -> inverse.
// ---------------------------------------------------------------------------
// XCD-aware tile order of a 2D grid of (line group, plane) tiles: workgroups
// are dispatched round-robin over the 8 XCDs in flattened order w, so XCD
// w % 8 gets the contiguous run of logical tiles below (a bijection for any
// grid).  Neighbouring line groups of a row then share one L2; with an odd
// row pitch (the (nx/2+1, ny, nz) spectra) a group's row segment straddles
// two 128-byte lines, each half read by a neighbour.
__device__ __forceinline__ void xcd_tile(unsigned &bx, unsigned &by)
{
    const unsigned gx = gridDim.x, nw = gx * gridDim.y, w = blockIdx.y * gx + blockIdx.x;
    const unsigned per = nw / 8, rem = nw % 8, xcd = w % 8, slot = w / 8;
    const unsigned lg = xcd < rem ? xcd * (per + 1) + slot : rem * (per + 1) + (xcd - rem) * per + slot;
    bx = lg % gx;
    by = lg / gx;
}

template <typename TS, typename TC, int MODE, int L, int TX>
__global__ void __launch_bounds__(1024) k_fft_strided(const TS *src, TS *dst, long na, long sb, int se,
                                                     const TC *__restrict__ twL, SpecMul pm)
{
    extern __shared__ unsigned char smem[];
    C2<TC> *tile = (C2<TC> *)smem;
    // TX is a compile-time constant so that the LDS addresses of a thread's
    // elements are one base register plus immediate offsets
    const int ln = threadIdx.x & (TX - 1), jb = threadIdx.x / TX;
    unsigned bx, by;
    xcd_tile(bx, by);
    const long a = (long)bx * TX + ln;
    const bool act = a < na;
    const long off = 2 * ((act ? a : 0) + (long)by * sb);
    const TS *const lin = src + off;
    TS *const lout = dst + off;
    const C2<TC> *tw = (const C2<TC> *)twL;
    const TC *azl = nullptr;
    if constexpr (PDO_FFT_TW_LDS) {
        // the twiddle table (L values) in LDS after the tile: every stage
        // reads it with broadcast LDS loads instead of vector memory loads;
        // the fused pass also stages its per-element factors az (L values)
        C2<TC> *twl = tile + TX * L;
        for (int i = threadIdx.x; i < L; i += TX * (L / 8)) twl[i] = tw[i];
        if constexpr (MODE == 2) {
            TC *a = (TC *)(twl + L);
            for (int i = threadIdx.x; i < L; i += TX * (L / 8)) a[i] = (TC)pm.az[i];
            azl = a;
        }
        __syncthreads();
        tw = twl;
    }
    using IO = CLoad<TS, TC>;
    auto ldg = [&](int i) { return IO::ld(lin, i * se); };
    auto stg = [&](int i, C2<TC> v) { IO::st(lout, i * se, v); };
    // element i of the line at row i ^ ((i >> 3) & SW): a half-wave spans
    // SW + 1 = 256 B / (TX complex) rows, and the first Stockham stage stores
    // rows 8 j + r, one per value of j in the half-wave; the swizzle sends
    // them to distinct parts of the 256-byte bank row
    constexpr int SLOTS = 256 / (int)sizeof(C2<TC>);
    constexpr int SW = (PDO_FFT_STRIDED_SWZ && SLOTS > TX) ? (SLOTS / TX - 1 < 7 ? SLOTS / TX - 1 : 7) : 0;
    auto buf = [&](int i) -> C2<TC> & { return tile[(i ^ ((i >> 3) & SW)) * TX + ln]; };
    if constexpr (MODE == 0) {
        line_fft<TC, -1, false, L>(ldg, stg, buf, tw, jb, act);
    } else if constexpr (MODE == 1) {
        line_fft<TC, +1, false, L>(ldg, stg, buf, tw, jb, act);
    } else {
        // forward into LDS, multiply by the spectral factor of (kx, ky, kz) =
        // (a, by, i), inverse from LDS to HBM
        // the factor is formed in the arithmetic of the transform (an FP64
        // division per point costs as much as the transforms in FP32)
        const int kx = (int)a, ky = (int)by;
        const bool line_only = pm.kind == 3 || pm.kind == 4;  // factors of the line index alone
        const TC fxy = (act && !line_only)
                           ? (TC)(pm.kind == 5 ? pm.ax[kx] : pm.kind == 2 ? pm.ax[kx] * pm.ay[ky] : pm.ax[kx] + pm.ay[ky])
                           : TC(0);
        const TC scale = (TC)pm.scale;
        auto sts = [&](int i, C2<TC> v) {
            TC fac;
            const TC az = PDO_FFT_TW_LDS ? azl[i] : (TC)pm.az[i];
            if (pm.kind == 3) {
                // i az: (x + i y) i a = -a y + i a x
                const TC s = az * scale;
                buf(i) = C2<TC>{-s * v.y, s * v.x};
                return;
            }
            if (pm.kind == 4) {
                fac = az;
            } else if (pm.kind == 2 || pm.kind == 5) {
                fac = fxy * az;
            } else {
                const TC ks = fxy + az;
                if (pm.kind == 0) fac = (kx == 0 && ky == 0 && i == 0) ? TC(0) : -recip(ks + TC(1.0e-20));
                else fac = (ks <= TC(1.0e-14)) ? TC(0) : -recip(ks);
            }
            fac *= scale;
            v.x *= fac;
            v.y *= fac;
            buf(i) = v;
        };
        line_fft<TC, -1, false, L>(ldg, sts, buf, tw, jb, act);
        __syncthreads();
        auto ldb = [&](int i) { return buf(i); };
        line_fft<TC, +1, true, L>(ldb, stg, buf, tw, jb, act);
    }
}

// LDS layout of an x line of L complex points.  An LDS access of 8-byte
// values is conflict-free when the 32 lanes of each half-wave fall in 32
// distinct slots (address mod 32).  The Stockham stages access a line
// contiguously (loads, the split/merge steps) and, in their stores, at
// 8 j + r and 64 (j / 8) + j % 8 + 8 r across the lanes j.
//   1: one pad element after every 8 (lines of 128 points or fewer, where a
//      half-wave spans several lines);
//   2: (L >= 256, a half-wave within one line) element i of 32-block b at
//      i ^ ((b & 7) | ((b >> 1) & 3) << 3): a permutation inside each block,
//      so contiguous runs stay conflict-free, and both store patterns spread
//      over all 32 slots.
// 0 is the plain layout.  Tuning constant.
#ifndef PDO_FFT_X_LAYOUT
#define PDO_FFT_X_LAYOUT 2
#endif
template <int L> __host__ __device__ constexpr int xslot(int i)
{
    if constexpr (PDO_FFT_X_LAYOUT == 2 && L >= 256) {
        const int b = i >> 5;
        return i ^ ((b & 7) | (((b >> 1) & 3) << 3));
    } else if constexpr (PDO_FFT_X_LAYOUT >= 1) {
        return i + (i >> 3);
    } else {
        return i;
    }
}
// line pitch (elements): past the last slot of element L (with L >= 256 a
// half-wave never spans two lines, so the lines need no relative offset)
template <int L> __host__ __device__ constexpr int xline_pitch()
{
    if constexpr (PDO_FFT_X_LAYOUT == 2 && L >= 256)
        return (L + 32) & ~31;
    else
        return xslot<L>(L) + 1;
}

// LDS bytes of an x-line workgroup: TXL lines and the staged twiddles (with
// split: also the L + 1 split/merge twiddles of the spectral-x kernels)
template <int L> static size_t xline_lds(int TXL, size_t ce, bool split = false)
{
    // with split, also the L + 1 wavenumbers (one value each, half a complex)
    return ((size_t)TXL * xline_pitch<L>() + (PDO_FFT_X_TW_LDS ? L + (split ? L + 1 + (L + 2) / 2 : 0) : 0)) * ce;
}

// the twiddle table of an x line (L values) staged in LDS after the tile of
// TXL lines; returns the LDS copy (or the table itself when staging is off).
// The first use is in the second Stockham stage, after a barrier.
template <typename TC, int L>
__device__ __forceinline__ const C2<TC> *x_twiddles(C2<TC> *tile, int TXL, const TC *twL)
{
    if constexpr (PDO_FFT_X_TW_LDS) {
        C2<TC> *twl = tile + (size_t)TXL * xline_pitch<L>();
        const C2<TC> *tw = (const C2<TC> *)twL;
        for (int i = threadIdx.x; i < L; i += blockDim.x) twl[i] = tw[i];
        return twl;
    } else {
        return (const C2<TC> *)twL;
    }
}

// the split/merge twiddles (L + 1 values) of the spectral-x kernels staged in
// LDS after the line twiddles; their first use follows a barrier
template <typename TC, int L>
__device__ __forceinline__ const C2<TC> *x_split_twiddles(C2<TC> *tile, int TXL, const TC *twN)
{
    if constexpr (PDO_FFT_X_TW_LDS) {
        C2<TC> *wl = tile + (size_t)TXL * xline_pitch<L>() + L;
        const C2<TC> *w = (const C2<TC> *)twN;
        for (int i = threadIdx.x; i <= L; i += blockDim.x) wl[i] = w[i];
        return wl;
    } else {
        return (const C2<TC> *)twN;
    }
}

// the modified wavenumbers k'(kx) (L + 1 values) of the spectral-x kernels
// staged in LDS, in the transform's arithmetic, after the split twiddles
template <typename TC, int L>
__device__ __forceinline__ const TC *x_wavenumbers(C2<TC> *tile, int TXL, const double *kp)
{
    if constexpr (PDO_FFT_X_TW_LDS) {
        TC *kl = (TC *)(tile + (size_t)TXL * xline_pitch<L>() + 2 * L + 1);
        for (int i = threadIdx.x; i <= L; i += blockDim.x) kl[i] = (TC)kp[i];
        return kl;
    } else {
        return nullptr;
    }
}

// Z_k = E_k + i O_k of the half-length inverse transform from the staged
// modes X_k (in place, one thread per pair (k, L-k)):
// E = (X_k + conj X_{L-k})/2, O = (X_k - conj X_{L-k}) conj(W^k)/2.  Only the
// real parts of the kx = 0 and kx = nx/2 modes enter (a Hermitian c2r);
// oddball != 0 drops kx = nx/2.
template <typename TC, int L, class BUF>
__device__ __forceinline__ void c2r_merge(BUF &buf, const C2<TC> *wN, int jb, int oddball)
{
    constexpr int L8 = L >> 3;
    for (int k = jb; k <= L / 2; k += L8) {
        const int kc = L - k;
        C2<TC> xk = buf(k), xc = buf(kc);
        if (k == 0) {
            xk.y = 0;
            xc.y = 0;
            if (oddball) xc.x = 0;
        }
        auto z = [&](C2<TC> a, C2<TC> b, int kk) {
            const C2<TC> ev = {(TC)0.5 * (a.x + b.x), (TC)0.5 * (a.y - b.y)};
            const C2<TC> df = {(TC)0.5 * (a.x - b.x), (TC)0.5 * (a.y + b.y)};
            const C2<TC> od = cmul(df, C2<TC>{wN[kk].x, -wN[kk].y});
            return C2<TC>{ev.x - od.y, ev.y + od.x};
        };
        const C2<TC> zk = z(xk, xc, k);
        const C2<TC> zc = z(xc, xk, kc);
        buf(k) = zk;
        if (k > 0 && kc != k) buf(kc) = zc;
    }
}

// X_k = E_k + W^k O_k of the real transform from the half-length one in buf,
// E = (Z_k + conj Z_{L-k})/2, O = (Z_k - conj Z_{L-k})/(2i)
template <typename TC, int L, class BUF>
__device__ __forceinline__ C2<TC> r2c_split(BUF &buf, const C2<TC> *wN, int k)
{
    const C2<TC> zk = buf(k == L ? 0 : k);
    const C2<TC> zc = buf(k == 0 ? 0 : L - k);
    const C2<TC> ev = {(TC)0.5 * (zk.x + zc.x), (TC)0.5 * (zk.y - zc.y)};
    const C2<TC> od = {(TC)0.5 * (zk.y + zc.y), (TC)-0.5 * (zk.x - zc.x)};
    return cadd(ev, cmul(wN[k], od));
}

// ---------------------------------------------------------------------------
// pass 1: x real-to-complex.  nlines lines of nx reals -> nx/2+1 complex.
// Tile: TXL lines; thread = (line, jb), jb fastest.
// ---------------------------------------------------------------------------
template <typename TS, typename TC, int L>
__global__ void __launch_bounds__(1024) k_fft_x_r2c(const TS *__restrict__ in, TS *__restrict__ out, long nlines,
                                                   const TC *__restrict__ twL, const TC *__restrict__ twN, int TXL,
                                                   int nxp)
{
    extern __shared__ unsigned char smem[];
    C2<TC> *tile = (C2<TC> *)smem;
    const C2<TC> *twl = x_twiddles<TC, L>(tile, TXL, twL);
    constexpr int nx = 2 * L, L8 = L >> 3, LDL = xline_pitch<L>(), nxh = L + 1;
    const int line = threadIdx.x / L8, jb = threadIdx.x - line * L8;
    const long l = (long)blockIdx.x * TXL + line;
    const bool act = (line < TXL) && (l < nlines);
    using IO = CLoad<TS, TC>;
    const TS *src = in + (act ? l : 0) * (long)nx;
    auto ldg = [&](int i) { return IO::ld(src, i); };  // packed z_j = (x_2j, x_2j+1)
    auto buf = [&](int i) -> C2<TC> & { return tile[line * LDL + xslot<L>(i)]; };
    auto sts = [&](int i, C2<TC> v) { buf(i) = v; };
    line_fft<TC, -1, false, L>(ldg, sts, buf, twl, jb, act);
    __syncthreads();
    const C2<TC> *wN = (const C2<TC> *)twN;
    TS *dst = out + (act ? l : 0) * (long)nxp * 2;
    if (act) {
        for (int k = jb; k < nxh; k += L8) IO::st(dst, k, r2c_split<TC, L>(buf, wN, k));
    }
}

// pass 5: x complex-to-real, scaled by `scale`.  Like a Hermitian c2r, only
// the real parts of the kx = 0 and kx = nx/2 modes enter; with oddball != 0
// the kx = nx/2 mode is dropped altogether.
template <typename TS, typename TC, int L>
__global__ void __launch_bounds__(1024) k_fft_x_c2r(const TS *__restrict__ in, TS *__restrict__ out, long nlines,
                                                   const TC *__restrict__ twL, const TC *__restrict__ twN, int TXL,
                                                   TC scale, int nxp, int oddball)
{
    extern __shared__ unsigned char smem[];
    C2<TC> *tile = (C2<TC> *)smem;
    const C2<TC> *twl = x_twiddles<TC, L>(tile, TXL, twL);
    constexpr int nx = 2 * L, L8 = L >> 3, LDL = xline_pitch<L>();
    const int line = threadIdx.x / L8, jb = threadIdx.x - line * L8;
    const long l = (long)blockIdx.x * TXL + line;
    const bool act = (line < TXL) && (l < nlines);
    using IO = CLoad<TS, TC>;
    const TS *src = in + (act ? l : 0) * (long)nxp * 2;
    auto buf = [&](int i) -> C2<TC> & { return tile[line * LDL + xslot<L>(i)]; };
    const C2<TC> *wN = (const C2<TC> *)twN;
    // stage the nx/2+1 modes of the line in LDS (coalesced, all loads of a
    // thread in flight together), then form Z_k = E_k + i O_k in place, one
    // thread per pair (k, L-k):
    // E = (X_k + conj X_{L-k})/2, O = (X_k - conj X_{L-k}) conj(W^k)/2
    if (act) {
        C2<TC> v[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) v[q] = IO::ld(src, jb + q * L8);
        C2<TC> vl = {};
        if (jb == 0) vl = IO::ld(src, L);
#pragma unroll
        for (int q = 0; q < 8; ++q) buf(jb + q * L8) = v[q];
        if (jb == 0) buf(L) = vl;
    }
    __syncthreads();
    if (act) c2r_merge<TC, L>(buf, wN, jb, oddball);
    __syncthreads();
    TS *dst = out + (act ? l : 0) * (long)nx;
    auto ldb = [&](int i) { return buf(i); };
    auto stg = [&](int i, C2<TC> v) {
        v.x *= scale;
        v.y *= scale;
        IO::st(dst, i, v);
    };
    line_fft<TC, +1, true, L>(ldb, stg, buf, twl, jb, act);
}

// The x transforms of the full-RHS Poisson solve with the x derivatives of
// its divergence and of its projection taken in spectral space.  On a
// periodic grid the cd06 first derivative is diagonal in Fourier space, with
// eigenvalue i k'(kx) on mode kx (k' the modified wavenumber, zero at kx = 0
// and kx = nx/2), so these passes form the same operators as a compact x
// solve, without the extra pass over HBM.
//
// r2c: X = FFT_x(div) + i k'(kx) FFT_x(ut), line by line.
template <typename TS, typename TC, int L>
__global__ void __launch_bounds__(1024) k_fft_x_r2c_dx(const TS *__restrict__ div, const TS *__restrict__ ut,
                                                      TS *__restrict__ out, long nlines, const TC *__restrict__ twL,
                                                      const TC *__restrict__ twN, int TXL, int nxp,
                                                      const double *__restrict__ kp)
{
    extern __shared__ unsigned char smem[];
    C2<TC> *tile = (C2<TC> *)smem;
    const C2<TC> *twl = x_twiddles<TC, L>(tile, TXL, twL);
    constexpr int nx = 2 * L, L8 = L >> 3, LDL = xline_pitch<L>(), nxh = L + 1;
    constexpr int NK = (nxh + L8 - 1) / L8;  // modes per thread
    const int line = threadIdx.x / L8, jb = threadIdx.x - line * L8;
    const long l = (long)blockIdx.x * TXL + line;
    const bool act = (line < TXL) && (l < nlines);
    using IO = CLoad<TS, TC>;
    auto buf = [&](int i) -> C2<TC> & { return tile[line * LDL + xslot<L>(i)]; };
    auto sts = [&](int i, C2<TC> v) { buf(i) = v; };
    const C2<TC> *wN = x_split_twiddles<TC, L>(tile, TXL, twN);
    const TC *kl = x_wavenumbers<TC, L>(tile, TXL, kp);
    auto kpr = [&](int k) { return PDO_FFT_X_TW_LDS ? kl[k] : (TC)kp[k]; };
    // i k' FFT(ut), held in registers
    C2<TC> du[NK];
    {
        const TS *src = ut + (act ? l : 0) * (long)nx;
        auto ldg = [&](int i) { return IO::ld(src, i); };
        line_fft<TC, -1, false, L>(ldg, sts, buf, twl, jb, act);
        __syncthreads();
#pragma unroll
        for (int q = 0; q < NK; ++q) {
            const int k = jb + q * L8;
            du[q] = C2<TC>{0, 0};
            if (act && k < nxh) {
                const C2<TC> x = r2c_split<TC, L>(buf, wN, k);
                const TC kk = kpr(k);
                du[q] = C2<TC>{-kk * x.y, kk * x.x};
            }
        }
        __syncthreads();  // every read of the buffer is done
    }
    const TS *src = div + (act ? l : 0) * (long)nx;
    auto ldg = [&](int i) { return IO::ld(src, i); };
    line_fft<TC, -1, false, L>(ldg, sts, buf, twl, jb, act);
    __syncthreads();
    TS *dst = out + (act ? l : 0) * (long)nxp * 2;
    if (act) {
#pragma unroll
        for (int q = 0; q < NK; ++q) {
            const int k = jb + q * L8;
            if (k < nxh) IO::st(dst, k, cadd(r2c_split<TC, L>(buf, wN, k), du[q]));
        }
    }
}

// c2r: p = scale IFFT_x(P) and u = ut - scale IFFT_x(i k'(kx) P), line by line.
template <typename TS, typename TC, int L>
__global__ void __launch_bounds__(1024) k_fft_x_c2r_dx(const TS *__restrict__ in, TS *__restrict__ p,
                                                      const TS *__restrict__ ut, TS *__restrict__ u, long nlines,
                                                      const TC *__restrict__ twL, const TC *__restrict__ twN, int TXL,
                                                      TC scale, int nxp, const double *__restrict__ kp)
{
    extern __shared__ unsigned char smem[];
    C2<TC> *tile = (C2<TC> *)smem;
    const C2<TC> *twl = x_twiddles<TC, L>(tile, TXL, twL);
    constexpr int nx = 2 * L, L8 = L >> 3, LDL = xline_pitch<L>();
    const int line = threadIdx.x / L8, jb = threadIdx.x - line * L8;
    const long l = (long)blockIdx.x * TXL + line;
    const bool act = (line < TXL) && (l < nlines);
    using IO = CLoad<TS, TC>;
    const TS *src = in + (act ? l : 0) * (long)nxp * 2;
    auto buf = [&](int i) -> C2<TC> & { return tile[line * LDL + xslot<L>(i)]; };
    auto ldb = [&](int i) { return buf(i); };
    const C2<TC> *wN = x_split_twiddles<TC, L>(tile, TXL, twN);
    const TC *kl = x_wavenumbers<TC, L>(tile, TXL, kp);
    auto kpr = [&](int k) { return PDO_FFT_X_TW_LDS ? kl[k] : (TC)kp[k]; };
    C2<TC> v[8];
    C2<TC> vl = {};
    if (act) {
#pragma unroll
        for (int q = 0; q < 8; ++q) v[q] = IO::ld(src, jb + q * L8);
        if (jb == 0) vl = IO::ld(src, L);
#pragma unroll
        for (int q = 0; q < 8; ++q) buf(jb + q * L8) = v[q];
        if (jb == 0) buf(L) = vl;
    }
    __syncthreads();
    if (act) c2r_merge<TC, L>(buf, wN, jb, 0);
    __syncthreads();
    TS *pl = p + (act ? l : 0) * (long)nx;
    auto stp = [&](int i, C2<TC> x) {
        x.x *= scale;
        x.y *= scale;
        IO::st(pl, i, x);
    };
    line_fft<TC, +1, true, L>(ldb, stp, buf, twl, jb, act);
    __syncthreads();  // every read of the buffer by the first transform is done
    // i k' P, from the modes still in registers
    if (act) {
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const TC kk = kpr(jb + q * L8);
            buf(jb + q * L8) = C2<TC>{-kk * v[q].y, kk * v[q].x};
        }
        if (jb == 0) {
            const TC kk = kpr(L);
            buf(L) = C2<TC>{-kk * vl.y, kk * vl.x};
        }
    }
    __syncthreads();
    if (act) c2r_merge<TC, L>(buf, wN, jb, 0);
    __syncthreads();
    const TS *ul = ut + (act ? l : 0) * (long)nx;
    TS *uo = u + (act ? l : 0) * (long)nx;
    auto stu = [&](int i, C2<TC> x) {
        const C2<TC> a = IO::ld(ul, i);
        IO::st(uo, i, C2<TC>{a.x - scale * x.x, a.y - scale * x.y});
    };
    line_fft<TC, +1, true, L>(ldb, stu, buf, twl, jb, act);
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
static int ilog2(int v)
{
    int l = 0;
    while ((1 << l) < v) l++;
    return l;
}

// largest workgroups of the line transforms (threads); tuning constants
#ifndef PDO_FFT_STRIDED_THREADS
#define PDO_FFT_STRIDED_THREADS 1024
#endif
#ifndef PDO_FFT_X_THREADS
#define PDO_FFT_X_THREADS 256
#endif

// the fused forward-multiply-inverse pass (MODE 2) may use smaller
// workgroups than the single transforms; tuning constant
#ifndef PDO_FFT_FUSED_THREADS
#define PDO_FFT_FUSED_THREADS 1024
#endif

// lines per workgroup of the strided transforms: PDO_FFT_STRIDED_THREADS
// (PDO_FFT_FUSED_THREADS for MODE 2) threads, 8 points each, at most 64 KiB
// of LDS
template <int L, size_t CE, int MODE> constexpr int strided_fft_lines()
{
    int tx = (MODE == 2 ? PDO_FFT_FUSED_THREADS : PDO_FFT_STRIDED_THREADS) / (L / 8);
    if (tx < 1) tx = 1;
    while (tx > 1 && (size_t)tx * L * CE > 64 * 1024) tx >>= 1;
    return tx;
}

template <typename TS, typename TC, int MODE, int L>
static void launch_strided_fft(const TS *src, TS *dst, long na, long sb, long se, const TC *tw, const SpecMul &pm,
                               int nb, hipStream_t st)
{
    constexpr size_t ce = 2 * sizeof(TC);
    constexpr int TX = strided_fft_lines<L, ce, MODE>();
    dim3 grid((unsigned)((na + TX - 1) / TX), (unsigned)nb);
    hipLaunchKernelGGL((k_fft_strided<TS, TC, MODE, L, TX>), grid, dim3(TX * (L / 8)),
                       (size_t)(TX + (PDO_FFT_TW_LDS ? 1 : 0)) * L * ce +
                           ((PDO_FFT_TW_LDS && MODE == 2) ? (size_t)L * sizeof(TC) : 0),
                       st, src,
                       dst, na, sb, (int)se, tw, pm);
}

template <typename TS, typename TC, int MODE>
static void strided_fft(int L, const TS *src, TS *dst, long na, long sb, long se, const TC *tw, const SpecMul &pm,
                        int nb, hipStream_t st)
{
    switch (L) {
    case 8: launch_strided_fft<TS, TC, MODE, 8>(src, dst, na, sb, se, tw, pm, nb, st); break;
    case 16: launch_strided_fft<TS, TC, MODE, 16>(src, dst, na, sb, se, tw, pm, nb, st); break;
    case 32: launch_strided_fft<TS, TC, MODE, 32>(src, dst, na, sb, se, tw, pm, nb, st); break;
    case 64: launch_strided_fft<TS, TC, MODE, 64>(src, dst, na, sb, se, tw, pm, nb, st); break;
    case 128: launch_strided_fft<TS, TC, MODE, 128>(src, dst, na, sb, se, tw, pm, nb, st); break;
    case 256: launch_strided_fft<TS, TC, MODE, 256>(src, dst, na, sb, se, tw, pm, nb, st); break;
    case 512: launch_strided_fft<TS, TC, MODE, 512>(src, dst, na, sb, se, tw, pm, nb, st); break;
    case 1024: launch_strided_fft<TS, TC, MODE, 1024>(src, dst, na, sb, se, tw, pm, nb, st); break;
    }
}

template <typename TS, typename TC, int L>
static void launch_x_fft(bool inverse, const TS *in, TS *out, long nlines, const TC *tw, const TC *twn, TC scale,
                         int nxp, int oddball, hipStream_t st)
{
    const size_t ce = 2 * sizeof(TC);
    int TXL = PDO_FFT_X_THREADS / (L / 8);
    if (TXL < 1) TXL = 1;
    while (TXL > 1 && (size_t)TXL * xline_pitch<L>() * ce > 64 * 1024) TXL >>= 1;
    dim3 grid((unsigned)((nlines + TXL - 1) / TXL));
    if (!inverse)
        hipLaunchKernelGGL((k_fft_x_r2c<TS, TC, L>), grid, dim3(TXL * (L / 8)), xline_lds<L>(TXL, ce), st, in,
                           out, nlines, tw, twn, TXL, nxp);
    else
        hipLaunchKernelGGL((k_fft_x_c2r<TS, TC, L>), grid, dim3(TXL * (L / 8)), xline_lds<L>(TXL, ce), st, in,
                           out, nlines, tw, twn, TXL, scale, nxp, oddball);
}

template <typename TS, typename TC>
static void x_fft(int L, bool inverse, const TS *in, TS *out, long nlines, const TC *tw, const TC *twn, TC scale,
                  int nxp, int oddball, hipStream_t st)
{
    switch (L) {
    case 8: launch_x_fft<TS, TC, 8>(inverse, in, out, nlines, tw, twn, scale, nxp, oddball, st); break;
    case 16: launch_x_fft<TS, TC, 16>(inverse, in, out, nlines, tw, twn, scale, nxp, oddball, st); break;
    case 32: launch_x_fft<TS, TC, 32>(inverse, in, out, nlines, tw, twn, scale, nxp, oddball, st); break;
    case 64: launch_x_fft<TS, TC, 64>(inverse, in, out, nlines, tw, twn, scale, nxp, oddball, st); break;
    case 128: launch_x_fft<TS, TC, 128>(inverse, in, out, nlines, tw, twn, scale, nxp, oddball, st); break;
    case 256: launch_x_fft<TS, TC, 256>(inverse, in, out, nlines, tw, twn, scale, nxp, oddball, st); break;
    case 512: launch_x_fft<TS, TC, 512>(inverse, in, out, nlines, tw, twn, scale, nxp, oddball, st); break;
    }
}

// the x passes with the spectral x derivative (k_fft_x_r2c_dx / _c2r_dx)
template <typename TS, typename TC, int L>
static void launch_x_fft_dx(bool inverse, const TS *a, const TS *b, TS *o1, TS *o2, long nlines, const TC *tw,
                            const TC *twn, TC scale, int nxp, const double *kp, hipStream_t st)
{
    const size_t ce = 2 * sizeof(TC);
    int TXL = PDO_FFT_X_THREADS / (L / 8);
    if (TXL < 1) TXL = 1;
    while (TXL > 1 && (size_t)TXL * xline_pitch<L>() * ce > 64 * 1024) TXL >>= 1;
    dim3 grid((unsigned)((nlines + TXL - 1) / TXL));
    if (!inverse)
        hipLaunchKernelGGL((k_fft_x_r2c_dx<TS, TC, L>), grid, dim3(TXL * (L / 8)), xline_lds<L>(TXL, ce, true), st, a,
                           b, o1, nlines, tw, twn, TXL, nxp, kp);
    else
        hipLaunchKernelGGL((k_fft_x_c2r_dx<TS, TC, L>), grid, dim3(TXL * (L / 8)), xline_lds<L>(TXL, ce, true), st, a,
                           o1, b, o2, nlines, tw, twn, TXL, scale, nxp, kp);
}

template <typename TS, typename TC>
static void x_fft_dx(int L, bool inverse, const TS *a, const TS *b, TS *o1, TS *o2, long nlines, const TC *tw,
                     const TC *twn, TC scale, int nxp, const double *kp, hipStream_t st)
{
    switch (L) {
    case 8: launch_x_fft_dx<TS, TC, 8>(inverse, a, b, o1, o2, nlines, tw, twn, scale, nxp, kp, st); break;
    case 16: launch_x_fft_dx<TS, TC, 16>(inverse, a, b, o1, o2, nlines, tw, twn, scale, nxp, kp, st); break;
    case 32: launch_x_fft_dx<TS, TC, 32>(inverse, a, b, o1, o2, nlines, tw, twn, scale, nxp, kp, st); break;
    case 64: launch_x_fft_dx<TS, TC, 64>(inverse, a, b, o1, o2, nlines, tw, twn, scale, nxp, kp, st); break;
    case 128: launch_x_fft_dx<TS, TC, 128>(inverse, a, b, o1, o2, nlines, tw, twn, scale, nxp, kp, st); break;
    case 256: launch_x_fft_dx<TS, TC, 256>(inverse, a, b, o1, o2, nlines, tw, twn, scale, nxp, kp, st); break;
    case 512: launch_x_fft_dx<TS, TC, 512>(inverse, a, b, o1, o2, nlines, tw, twn, scale, nxp, kp, st); break;
    }
}

template <typename TC> struct Tw {
    const TC *x, *xn, *y, *z;
    explicit Tw(const PoissonPlanDev &pl)
    {
        const bool dbl = sizeof(TC) == 8;
        x = (const TC *)(dbl ? (const void *)pl.tw_d[0] : (const void *)pl.tw_f[0]);
        xn = (const TC *)(dbl ? (const void *)pl.twn_d : (const void *)pl.twn_f);
        y = (const TC *)(dbl ? (const void *)pl.tw_d[1] : (const void *)pl.tw_f[1]);
        z = (const TC *)(dbl ? (const void *)pl.tw_d[2] : (const void *)pl.tw_f[2]);
    }
};

template <typename TS, typename TC>
static hipError_t run_poisson_t(const PoissonPlanDev &pl, const TS *rhs, TS *f, TS *spec, hipStream_t st)
{
    const int nx = pl.nx, ny = pl.ny, nz = pl.nz, nxh = nx / 2 + 1, L = nx / 2;
    const Tw<TC> tw(pl);
    const SpecMul pm{pl.kx2, pl.ky2, pl.kz2, pl.zero_policy, 1.0};
    const long nlines = (long)ny * nz;
    const TC scale = (TC)(1.0 / ((double)L * ny * nz));
    // half spectrum (nxh, ny, nz) stored with row pitch nxp (a multiple of 16
    // complex values) so every tile row starts on a 128-byte boundary
    const int nxp = pl.nxp;
    x_fft<TS, TC>(L, false, rhs, spec, nlines, tw.x, tw.xn, scale, nxp, 0, st);
    strided_fft<TS, TC, 0>(ny, spec, spec, (long)nxh, (long)nxp * ny, (long)nxp, tw.y, pm, nz, st);
    strided_fft<TS, TC, 2>(nz, spec, spec, (long)nxh, (long)nxp, (long)nxp * ny, tw.z, pm, ny, st);
    strided_fft<TS, TC, 1>(ny, spec, spec, (long)nxh, (long)nxp * ny, (long)nxp, tw.y, pm, nz, st);
    x_fft<TS, TC>(L, true, spec, f, nlines, tw.x, tw.xn, scale, nxp, 0, st);
    return hipGetLastError();
}

template <typename TS, typename TC>
static hipError_t run_poisson_dx_t(const PoissonPlanDev &pl, const TS *div, const TS *ut, TS *p, TS *ux, TS *spec,
                                   const double *kp, hipStream_t st)
{
    const int nx = pl.nx, ny = pl.ny, nz = pl.nz, nxh = nx / 2 + 1, L = nx / 2;
    const Tw<TC> tw(pl);
    const SpecMul pm{pl.kx2, pl.ky2, pl.kz2, pl.zero_policy, 1.0};
    const long nlines = (long)ny * nz;
    const TC scale = (TC)(1.0 / ((double)L * ny * nz));
    const int nxp = pl.nxp;
    x_fft_dx<TS, TC>(L, false, div, ut, spec, nullptr, nlines, tw.x, tw.xn, (TC)1, nxp, kp, st);
    strided_fft<TS, TC, 0>(ny, spec, spec, (long)nxh, (long)nxp * ny, (long)nxp, tw.y, pm, nz, st);
    strided_fft<TS, TC, 2>(nz, spec, spec, (long)nxh, (long)nxp, (long)nxp * ny, tw.z, pm, ny, st);
    strided_fft<TS, TC, 1>(ny, spec, spec, (long)nxh, (long)nxp * ny, (long)nxp, tw.y, pm, nz, st);
    x_fft_dx<TS, TC>(L, true, spec, ut, p, ux, nlines, tw.x, tw.xn, scale, nxp, kp, st);
    return hipGetLastError();
}

hipError_t run_poisson_dx(const PoissonPlanDev &pl, int mode, const void *div, const void *ut, void *p, void *ux,
                          void *spec, const double *kp, hipStream_t st)
{
    switch (mode) {
    case 0:
        return run_poisson_dx_t<double, double>(pl, (const double *)div, (const double *)ut, (double *)p,
                                                (double *)ux, (double *)spec, kp, st);
    case 1:
        return run_poisson_dx_t<float, float>(pl, (const float *)div, (const float *)ut, (float *)p, (float *)ux,
                                              (float *)spec, kp, st);
    case 2:
        return run_poisson_dx_t<float, double>(pl, (const float *)div, (const float *)ut, (float *)p, (float *)ux,
                                               (float *)spec, kp, st);
    }
    return hipErrorInvalidValue;
}

hipError_t run_poisson(const PoissonPlanDev &pl, int mode, const void *rhs, void *f, void *spec, hipStream_t st)
{
    switch (mode) {
    case 0: return run_poisson_t<double, double>(pl, (const double *)rhs, (double *)f, (double *)spec, st);
    case 1: return run_poisson_t<float, float>(pl, (const float *)rhs, (float *)f, (float *)spec, st);
    case 2: return run_poisson_t<float, double>(pl, (const float *)rhs, (float *)f, (float *)spec, st);
    }
    return hipErrorInvalidValue;
}

// 2D transforms of the (x, y) planes of (nx, ny, nplanes) data; the half
// spectrum is (nx/2+1, ny, nplanes) without padding.
template <typename TS, typename TC>
static hipError_t run_fft2_t(const PoissonPlanDev &pl, const TS *in, TS *out, int nplanes, hipStream_t st)
{
    const int nx = pl.nx, ny = pl.ny, nxh = nx / 2 + 1, L = nx / 2;
    const Tw<TC> tw(pl);
    const SpecMul pm{};
    x_fft<TS, TC>(L, false, in, out, (long)ny * nplanes, tw.x, tw.xn, (TC)1, nxh, 0, st);
    strided_fft<TS, TC, 0>(ny, out, out, (long)nxh, (long)nxh * ny, (long)nxh, tw.y, pm, nplanes, st);
    return hipGetLastError();
}

// 2D spectral filtering of real (nx, ny, nplanes) data by the separable
// factor t1(kx) t2(ky) (host-side 0/1 cutoffs, device arrays of nx/2+1 and
// ny): the x r2c pass into spec, then one fused pass per y line (forward
// transform, the factor, inverse transform, in place), then the x c2r pass
// into out.  Two passes fewer than fft2, a pointwise multiply and ifft2.
template <typename TS, typename TC>
static hipError_t run_filter2_t(const PoissonPlanDev &pl, const TS *in, TS *spec, TS *out, int nplanes,
                                const double *t1, const double *t2, hipStream_t st)
{
    const int nx = pl.nx, ny = pl.ny, nxh = nx / 2 + 1, L = nx / 2;
    const Tw<TC> tw(pl);
    x_fft<TS, TC>(L, false, in, spec, (long)ny * nplanes, tw.x, tw.xn, (TC)1, nxh, 0, st);
    // kind 5: factor t1[kx] t2[ky] of the line (kx) and element (ky) indices
    const SpecMul pm{t1, nullptr, t2, 5, 1.0};
    strided_fft<TS, TC, 2>(ny, spec, spec, (long)nxh, (long)nxh * ny, (long)nxh, tw.y, pm, nplanes, st);
    const TC scale = (TC)(1.0 / ((double)L * ny));
    x_fft<TS, TC>(L, true, spec, out, (long)ny * nplanes, tw.x, tw.xn, scale, nxh, 0, st);
    return hipGetLastError();
}

hipError_t run_filter2(const PoissonPlanDev &pl, int mode, const void *in, void *spec, void *out, int nplanes,
                       const double *t1, const double *t2, hipStream_t st)
{
    switch (mode) {
    case 0: return run_filter2_t<double, double>(pl, (const double *)in, (double *)spec, (double *)out, nplanes, t1, t2, st);
    case 1: return run_filter2_t<float, float>(pl, (const float *)in, (float *)spec, (float *)out, nplanes, t1, t2, st);
    case 2: return run_filter2_t<float, double>(pl, (const float *)in, (float *)spec, (float *)out, nplanes, t1, t2, st);
    }
    return hipErrorInvalidValue;
}

template <typename TS, typename TC>
static hipError_t run_ifft2_t(const PoissonPlanDev &pl, const TS *in, TS *work, TS *out, int nplanes, int oddball,
                              hipStream_t st)
{
    const int nx = pl.nx, ny = pl.ny, nxh = nx / 2 + 1, L = nx / 2;
    const Tw<TC> tw(pl);
    const SpecMul pm{};
    // the c2r pass returns L times the inverse transform: scale to 1/(nx ny)
    const TC scale = (TC)(1.0 / ((double)L * ny));
    strided_fft<TS, TC, 1>(ny, in, work, (long)nxh, (long)nxh * ny, (long)nxh, tw.y, pm, nplanes, st);
    x_fft<TS, TC>(L, true, work, out, (long)ny * nplanes, tw.x, tw.xn, scale, nxh, oddball, st);
    return hipGetLastError();
}

// forward z transform -> multiply -> inverse z transform of (nxh, ny, nz) complex data
template <typename TS, typename TC>
static hipError_t run_zmul_t(const PoissonPlanDev &pl, const TS *in, TS *out, const SpecMul &m, hipStream_t st)
{
    const int ny = pl.ny, nz = pl.nz, nxh = pl.nx / 2 + 1;
    const Tw<TC> tw(pl);
    strided_fft<TS, TC, 2>(nz, in, out, (long)nxh, (long)nxh, (long)nxh * ny, tw.z, m, ny, st);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// PadePoisson with walls in z (PadePoisson.F90:472-603), one workgroup per TX
// (kx) lines of a ky: the 2 nz-point even extension of f (cells, nz points:
// nz-1 .. 0, 0 .. nz-1) and odd extension of w (edges, nz+1 points:
// -w[nz-1] .. -w[1], w[0] .. w[nz]) are transformed in LDS, then
//   P = -(F + i kcm W) / |k|^2   (|k|^2 = ax[kx] + ay[ky] + az[kz], <= 1e-14: 0)
//   W = W - i kcp P
// with kcm, kcp = k3' exp(-+ i k3 dz / 2), both are transformed back (scale
// 1 / L), f takes the second half of P and w points nz-1 .. L-1 of W, zero on
// the two walls.  f and w are updated in place.  UPD false (getPressure,
// :834-871): only P goes back to f, w is read only.
// ---------------------------------------------------------------------------
template <typename TS, typename TC, int L, int TX, bool UPD>
__global__ void __launch_bounds__(1024) k_pp_walls(TS *f, TS *w, long na, long sb, int se, const TC *__restrict__ twL,
                                                  const double *__restrict__ ax, const double *__restrict__ ay,
                                                  const double *__restrict__ az, const double *__restrict__ kcm,
                                                  const double *__restrict__ kcp)
{
    extern __shared__ unsigned char smem[];
    constexpr int H = L / 2;
    C2<TC> *tF = (C2<TC> *)smem;
    C2<TC> *tW = tF + TX * L;
    C2<TC> *twl = tW + TX * L;
    const int ln = threadIdx.x & (TX - 1), jb = threadIdx.x / TX;
    unsigned bx, by;
    xcd_tile(bx, by);
    const long a = (long)bx * TX + ln;
    const bool act = a < na;
    const long off = 2 * ((act ? a : 0) + (long)by * sb);
    TS *const lf = f + off;
    TS *const lw = w + off;
    for (int i = threadIdx.x; i < L; i += TX * (L / 8)) twl[i] = ((const C2<TC> *)twL)[i];
    __syncthreads();
    using IO = CLoad<TS, TC>;
    constexpr int SLOTS = 256 / (int)sizeof(C2<TC>);
    constexpr int SW = (PDO_FFT_STRIDED_SWZ && SLOTS > TX) ? (SLOTS / TX - 1 < 7 ? SLOTS / TX - 1 : 7) : 0;
    auto bufF = [&](int i) -> C2<TC> & { return tF[(i ^ ((i >> 3) & SW)) * TX + ln]; };
    auto bufW = [&](int i) -> C2<TC> & { return tW[(i ^ ((i >> 3) & SW)) * TX + ln]; };
    // forward transforms of the extensions into LDS
    auto ldF = [&](int i) { return IO::ld(lf, (i < H ? H - 1 - i : i - H) * se); };
    auto ldW = [&](int i) {
        if (i < H - 1) {
            const C2<TC> v = IO::ld(lw, (H - 1 - i) * se);
            return C2<TC>{-v.x, -v.y};
        }
        return IO::ld(lw, (i - (H - 1)) * se);
    };
    auto stF = [&](int i, C2<TC> v) { bufF(i) = v; };
    auto stW = [&](int i, C2<TC> v) { bufW(i) = v; };
    line_fft<TC, -1, false, L>(ldF, stF, bufF, (const C2<TC> *)twl, jb, act);
    __syncthreads();
    line_fft<TC, -1, false, L>(ldW, stW, bufW, (const C2<TC> *)twl, jb, act);
    __syncthreads();
    // the pressure and the w update, 8 modes per thread
    if (act) {
        const TC fxy = (TC)(ax[a] + ay[by]);
#pragma unroll
        for (int r = 0; r < 8; ++r) {
            const int i = jb * 8 + r;
            const C2<TC> F = bufF(i), W = bufW(i);
            const TC ks = fxy + (TC)az[i];
            const TC kinv = (ks <= TC(1.0e-14)) ? TC(0) : recip(ks);
            // P = -(F + i cm W) kinv
            const C2<TC> cw = cmul(C2<TC>{(TC)kcm[2 * i], (TC)kcm[2 * i + 1]}, W);
            const C2<TC> Pv{-(F.x - cw.y) * kinv, -(F.y + cw.x) * kinv};
            // W -= i cp P
            bufF(i) = Pv;
            if constexpr (UPD) {
                const C2<TC> cpp = cmul(C2<TC>{(TC)kcp[2 * i], (TC)kcp[2 * i + 1]}, Pv);
                bufW(i) = C2<TC>{W.x + cpp.y, W.y - cpp.x};
            }
        }
    }
    __syncthreads();
    const TC scale = TC(1) / TC(L);
    auto lbF = [&](int i) { return bufF(i); };
    auto lbW = [&](int i) { return bufW(i); };
    auto sgF = [&](int i, C2<TC> v) {
        if (i >= H) IO::st(lf, (i - H) * se, C2<TC>{v.x * scale, v.y * scale});
    };
    auto sgW = [&](int i, C2<TC> v) {
        if (i >= H - 1) {
            const int k = i - (H - 1);
            const bool wall = (k == 0 || k == H);
            IO::st(lw, k * se, wall ? C2<TC>{TC(0), TC(0)} : C2<TC>{v.x * scale, v.y * scale});
        }
    };
    line_fft<TC, +1, true, L>(lbF, sgF, bufF, (const C2<TC> *)twl, jb, act);
    if constexpr (UPD) {
        __syncthreads();
        line_fft<TC, +1, true, L>(lbW, sgW, bufW, (const C2<TC> *)twl, jb, act);
    }
}

template <int L, size_t CE> constexpr int pp_walls_lines()
{
    int tx = 512 / (L / 8);
    if (tx < 1) tx = 1;
    while (tx > 1 && (2 * (size_t)tx + 1) * L * CE > 64 * 1024) tx >>= 1;
    return tx;
}

template <typename TS, typename TC, int L>
static void launch_pp_walls(TS *f, TS *w, long na, long sb, long se, const TC *tw, const PadeWallTables &t, int nb,
                            bool upd, hipStream_t st)
{
    constexpr size_t ce = 2 * sizeof(TC);
    // two line tiles and the twiddles: at most 64 KiB
    constexpr int TX = pp_walls_lines<L, ce>();
    const size_t lds = (2 * (size_t)TX + 1) * L * ce;
    dim3 grid((unsigned)((na + TX - 1) / TX), (unsigned)nb);
    if (upd)
        hipLaunchKernelGGL((k_pp_walls<TS, TC, L, TX, true>), grid, dim3(TX * (L / 8)), lds, st, f, w, na, sb, (int)se,
                           tw, t.ax, t.ay, t.az, t.kcm, t.kcp);
    else
        hipLaunchKernelGGL((k_pp_walls<TS, TC, L, TX, false>), grid, dim3(TX * (L / 8)), lds, st, f, w, na, sb,
                           (int)se, tw, t.ax, t.ay, t.az, t.kcm, t.kcp);
}

template <typename TS, typename TC>
static hipError_t run_pp_walls_t(const PoissonPlanDev &pl, TS *f, TS *w, const PadeWallTables &t, bool upd,
                                 hipStream_t st)
{
    const int ny = pl.ny, L = pl.nz, nxh = pl.nx / 2 + 1;
    const Tw<TC> tw(pl);
    const long na = nxh, sb = nxh, se = (long)nxh * ny;
    switch (L) {
    case 8: launch_pp_walls<TS, TC, 8>(f, w, na, sb, se, tw.z, t, ny, upd, st); break;
    case 16: launch_pp_walls<TS, TC, 16>(f, w, na, sb, se, tw.z, t, ny, upd, st); break;
    case 32: launch_pp_walls<TS, TC, 32>(f, w, na, sb, se, tw.z, t, ny, upd, st); break;
    case 64: launch_pp_walls<TS, TC, 64>(f, w, na, sb, se, tw.z, t, ny, upd, st); break;
    case 128: launch_pp_walls<TS, TC, 128>(f, w, na, sb, se, tw.z, t, ny, upd, st); break;
    case 256: launch_pp_walls<TS, TC, 256>(f, w, na, sb, se, tw.z, t, ny, upd, st); break;
    case 512: launch_pp_walls<TS, TC, 512>(f, w, na, sb, se, tw.z, t, ny, upd, st); break;
    case 1024: launch_pp_walls<TS, TC, 1024>(f, w, na, sb, se, tw.z, t, ny, upd, st); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t run_pp_walls(const PoissonPlanDev &pl, int mode, void *f, void *w, const PadeWallTables &t, bool upd,
                        hipStream_t st)
{
    switch (mode) {
    case 0: return run_pp_walls_t<double, double>(pl, (double *)f, (double *)w, t, upd, st);
    case 1: return run_pp_walls_t<float, float>(pl, (float *)f, (float *)w, t, upd, st);
    case 2: return run_pp_walls_t<float, double>(pl, (float *)f, (float *)w, t, upd, st);
    }
    return hipErrorInvalidValue;
}

hipError_t run_fft2(const PoissonPlanDev &pl, int mode, const void *in, void *out, int nplanes, hipStream_t st)
{
    switch (mode) {
    case 0: return run_fft2_t<double, double>(pl, (const double *)in, (double *)out, nplanes, st);
    case 1: return run_fft2_t<float, float>(pl, (const float *)in, (float *)out, nplanes, st);
    case 2: return run_fft2_t<float, double>(pl, (const float *)in, (float *)out, nplanes, st);
    }
    return hipErrorInvalidValue;
}

hipError_t run_ifft2(const PoissonPlanDev &pl, int mode, const void *in, void *work, void *out, int nplanes,
                     int oddball, hipStream_t st)
{
    switch (mode) {
    case 0:
        return run_ifft2_t<double, double>(pl, (const double *)in, (double *)work, (double *)out, nplanes, oddball,
                                           st);
    case 1:
        return run_ifft2_t<float, float>(pl, (const float *)in, (float *)work, (float *)out, nplanes, oddball, st);
    case 2:
        return run_ifft2_t<float, double>(pl, (const float *)in, (float *)work, (float *)out, nplanes, oddball, st);
    }
    return hipErrorInvalidValue;
}

hipError_t run_zmul(const PoissonPlanDev &pl, int mode, const void *in, void *out, const SpecMul &m, hipStream_t st)
{
    switch (mode) {
    case 0: return run_zmul_t<double, double>(pl, (const double *)in, (double *)out, m, st);
    case 1: return run_zmul_t<float, float>(pl, (const float *)in, (float *)out, m, st);
    case 2: return run_zmul_t<float, double>(pl, (const float *)in, (float *)out, m, st);
    }
    return hipErrorInvalidValue;
}

// x lines of the half spectrum (pitch nxp complex): v *= i kv[kx] (imag) or kv[kx]
template <typename TS>
__global__ void __launch_bounds__(256) k_xline_mul(TS *spec, int nxp, int nxh, long nlines, const double *kv, int imag)
{
    const long total = (long)nxh * nlines;
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (long)gridDim.x * blockDim.x) {
        const long line = e / nxh;
        const int kx = (int)(e - line * nxh);
        TS *p = spec + 2 * (line * nxp + kx);
        const double k = kv[kx];
        const double re = (double)p[0], im = (double)p[1];
        if (imag) {
            p[0] = (TS)(-k * im);
            p[1] = (TS)(k * re);
        } else {
            p[0] = (TS)(k * re);
            p[1] = (TS)(k * im);
        }
    }
}

template <typename TS, typename TC>
static hipError_t run_fourier_deriv_t(const PoissonPlanDev &pl, int dir, int imag, const TS *in, TS *out, TS *work,
                                      long na, long nb, const double *kv, hipStream_t st)
{
    const Tw<TC> tw(pl);
    if (dir == 1) {
        // r2c lines, multiply, c2r: the c2r returns L times the inverse transform
        const int nx = pl.nx, L = nx / 2, nxh = L + 1, nxp = pl.nxp;
        const long nlines = na * nb;
        x_fft<TS, TC>(L, false, in, work, nlines, tw.x, tw.xn, (TC)1, nxp, 0, st);
        long blocks = ((long)nxh * nlines + 255) / 256;
        if (blocks > 65536) blocks = 65536;
        hipLaunchKernelGGL((k_xline_mul<TS>), dim3((unsigned)blocks), dim3(256), 0, st, work, nxp, nxh, nlines, kv,
                           imag);
        x_fft<TS, TC>(L, true, work, out, nlines, tw.x, tw.xn, (TC)(1.0 / L), nxp, 0, st);
        return hipGetLastError();
    }
    // y / z: consecutive pairs along the contiguous index read as one complex
    // value (the operator is real and maps real lines to real lines, so the
    // real and imaginary parts come back as the two derivatives); one fused
    // forward -> multiply -> inverse pass along the direction
    SpecMul pm{};
    pm.az = kv;
    pm.kind = imag ? 3 : 4;
    if (dir == 2) {
        const int n = pl.ny;
        pm.scale = 1.0 / n;
        strided_fft<TS, TC, 2>(n, in, out, na / 2, (long)(na / 2) * n, na / 2, tw.y, pm, (int)nb, st);
    } else {
        const int n = pl.nz;
        pm.scale = 1.0 / n;
        strided_fft<TS, TC, 2>(n, in, out, na * nb / 2, 0, na * nb / 2, tw.z, pm, 1, st);
    }
    return hipGetLastError();
}

hipError_t run_fourier_deriv(const PoissonPlanDev &pl, int mode, int dir, int imag, const void *in, void *out,
                             void *work, long na, long nb, const double *kv, hipStream_t st)
{
    switch (mode) {
    case 0:
        return run_fourier_deriv_t<double, double>(pl, dir, imag, (const double *)in, (double *)out, (double *)work,
                                                   na, nb, kv, st);
    case 1:
        return run_fourier_deriv_t<float, float>(pl, dir, imag, (const float *)in, (float *)out, (float *)work, na,
                                                 nb, kv, st);
    case 2:
        return run_fourier_deriv_t<float, double>(pl, dir, imag, (const float *)in, (float *)out, (float *)work, na,
                                                  nb, kv, st);
    }
    return hipErrorInvalidValue;
}

}  // namespace pdo
