// Learned SGS closure: a pointwise MLP 10 -> 64 -> 64 -> 64 -> 6 (GELU on the
// hidden layers), trained and evaluated with hand-written gfx950 kernels:
// BF16 weights and activations, FP32 accumulation on the matrix cores
// (v_mfma_f32_16x16x16_bf16), FP32 master parameters updated by Adam.
//
// Training: a workgroup (4 waves) owns a tile of 64 samples and runs the whole
// network on it with every weight matrix staged in LDS; wave w owns rows 16w..16w+15
// of the activation tile.  The backward kernel forms per-workgroup weight
// gradients (the batch is the contraction dimension of dW = dZ^T H) that a
// second kernel sums in a fixed order, so training is deterministic.
#include <hip/hip_runtime.h>

#include <atomic>

#include "pdo_mlp.hpp"

namespace pdo {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int TILE = 64;     // samples per workgroup
constexpr int HS = 64 + 8;   // LDS row pitch of activation tiles (bf16)

__device__ __forceinline__ unsigned short to_bf16(float x)
{
    unsigned u = __float_as_uint(x);
    u += 0x7FFFu + ((u >> 16) & 1u);  // round to nearest even
    return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float from_bf16(unsigned short b) { return __uint_as_float((unsigned)b << 16); }

__device__ __forceinline__ float gelu_f(float z)
{
    const float t = tanhf(0.7978845608028654f * (z + 0.044715f * z * z * z));
    return 0.5f * z * (1.0f + t);
}
__device__ __forceinline__ float gelu_grad_f(float z)
{
    const float t = tanhf(0.7978845608028654f * (z + 0.044715f * z * z * z));
    const float du = 0.7978845608028654f * (1.0f + 3.0f * 0.044715f * z * z);
    return 0.5f * (1.0f + t) + 0.5f * z * (1.0f - t * t) * du;
}

// One 16x16 output tile of this wave: D(m, n) += sum_k A(m, k) B(k, n),
// K a multiple of 16.  Operand layout of the 16x16x16 BF16 MFMA: lane l
// holds A(l % 16, 4 (l / 16) + i) and B(4 (l / 16) + i, l % 16), i < 4; the
// result D(4 (l / 16) + i, l % 16) lands in element i of lane l.
template <class AF, class BF>
__device__ __forceinline__ f32x4 mma16(AF a, BF b, int K, f32x4 acc)
{
    const int lane = threadIdx.x & 63;
    const int r = lane & 15, g = lane >> 4;
    for (int k0 = 0; k0 < K; k0 += 16) {
        bf16x4 av, bv;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            av[i] = (short)a(r, k0 + 4 * g + i);
            bv[i] = (short)b(k0 + 4 * g + i, r);
        }
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, bv, acc, 0, 0, 0);
    }
    return acc;
}

struct Layers {
    int in[4] = {MLP_IN_PAD, MLP_HID, MLP_HID, MLP_HID};
    int out[4] = {MLP_HID, MLP_HID, MLP_HID, MLP_OUT_PAD};
    int w[4] = {MLP_OFF_W1, MLP_OFF_W2, MLP_OFF_W3, MLP_OFF_W4};
    int b[4] = {MLP_OFF_B1, MLP_OFF_B2, MLP_OFF_B3, MLP_OFF_B4};
};

// stage every weight matrix (BF16) of the network in LDS
__device__ void stage_weights(unsigned short *sW, const unsigned short *Wb)
{
    for (int i = threadIdx.x; i < MLP_NPARAM; i += blockDim.x) sW[i] = Wb[i];
}

// ---------------------------------------------------------------------------
// forward: X (B, 10) FP32 -> Y (B, 6) FP32; when cache != null the BF16 layer
// inputs H0 (B, 16), H1..H3 (B, 64) and pre-activations Z1..Z3 (B, 64) are
// kept for the backward pass
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_mlp_forward(const float *__restrict__ X, float *__restrict__ Y,
                                                    const unsigned short *__restrict__ Wb,
                                                    const float *__restrict__ P, long B, MlpCache cache)
{
    __shared__ unsigned short sW[MLP_NPARAM];
    __shared__ unsigned short sH[2][TILE * HS];
    const Layers L;
    const long row0 = (long)blockIdx.x * TILE;
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63, r = lane & 15, g = lane >> 4;
    stage_weights(sW, Wb);
    for (int i = threadIdx.x; i < TILE * MLP_IN_PAD; i += blockDim.x) {
        const int m = i / MLP_IN_PAD, k = i % MLP_IN_PAD;
        const long s = row0 + m;
        const float x = (s < B && k < MLP_IN) ? X[s * MLP_IN + k] : 0.0f;
        const unsigned short hb = to_bf16(x);
        sH[0][m * HS + k] = hb;
        if (cache.H[0] && s < B) cache.H[0][s * MLP_IN_PAD + k] = hb;
    }
    __syncthreads();
    int cur = 0;
    for (int l = 0; l < 4; ++l) {
        const unsigned short *Wl = sW + L.w[l];
        const int K = L.in[l], N = L.out[l];
        const unsigned short *Hc = sH[cur];
        for (int nt = 0; nt < N / 16; ++nt) {
            f32x4 acc = {0, 0, 0, 0};
            acc = mma16([&](int m, int k) { return Hc[(16 * w + m) * HS + k]; },
                        [&](int k, int n) { return Wl[(16 * nt + n) * K + k]; }, K, acc);
            const int col = 16 * nt + r;
            const float bias = P[L.b[l] + col];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int m = 16 * w + 4 * g + i;
                const long s = row0 + m;
                const float z = acc[i] + bias;
                if (l < 3) {
                    const unsigned short zb = to_bf16(z);
                    const unsigned short hb = to_bf16(gelu_f(from_bf16(zb)));
                    sH[cur ^ 1][m * HS + col] = hb;
                    if (cache.Z[l] && s < B) {
                        cache.Z[l][s * MLP_HID + col] = zb;
                        cache.H[l + 1][s * MLP_HID + col] = hb;
                    }
                } else if (col < MLP_OUT && s < B) {
                    Y[s * MLP_OUT + col] = z;
                }
            }
        }
        __syncthreads();
        cur ^= 1;
    }
}

// ---------------------------------------------------------------------------
// inference forward (no cache), the closure's evaluation on whole fields.
// Each wave runs 16 samples at a time through the network in the transposed
// form Z^T = W H^T: the weights are the MFMA A operand and the activations
// the B operand, so lane l ends a 16-row output tile holding
// Z(16 nt + 4 (l / 16) + i, sample l % 16), i < 4 -- exactly the B operand
// of the next layer's k step nt.  Activations therefore never leave the
// registers.  The BF16 weights of all four layers are staged in LDS once per
// (persistent) workgroup in fragment order, lane l's 8 bytes of a fragment
// holding W(16 nt + l % 16, 16 ks + 4 (l / 16) + i), i < 4, so each A
// operand is one conflict-free ds_read_b64 and the kernel needs ~66 VGPRs
// (held in registers instead, the weights take 80 VGPRs and cut the waves per
// SIMD from 7 to 2, which measured 17% slower).  The biases, staged in LDS
// too, seed the accumulators.  BF16 rounding is the hardware's
// round-to-nearest-even conversion (the values of to_bf16 for finite
// inputs).  GELU is evaluated as z sigma(2 y) = z / (1 + 2^(-2 y log2 e)),
// y = sqrt(2/pi) (z + 0.044715 z^3), identical to 0.5 z (1 + tanh y); with
// the hardware exp2 and reciprocal it is within a few FP32 ulp of the tanhf
// form, far below the BF16 rounding of the activation.
// ---------------------------------------------------------------------------
typedef __bf16 bf16v4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ bf16x4 pack_bf16(f32x4 v)
{
    return __builtin_bit_cast(bf16x4, __builtin_convertvector(v, bf16v4));
}
__device__ __forceinline__ f32x4 unpack_bf16(bf16x4 v)
{
    f32x4 f;
#pragma unroll
    for (int i = 0; i < 4; ++i) f[i] = from_bf16((unsigned short)v[i]);
    return f;
}

__device__ __forceinline__ float gelu_fast(float z)
{
    constexpr float c1 = (float)(-2.0 * 0.7978845608028654 * 1.4426950408889634);
    constexpr float c3 = (float)(-2.0 * 0.7978845608028654 * 1.4426950408889634 * 0.044715);
    const float a = z * __builtin_fmaf(z * z, c3, c1);
    return z * __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(a));
}

// bias + BF16 pre-activation + GELU + BF16 activation of one output tile
__device__ __forceinline__ bf16x4 hidden_act(f32x4 z)
{
    const f32x4 zr = unpack_bf16(pack_bf16(z));
    f32x4 h;
#pragma unroll
    for (int i = 0; i < 4; ++i) h[i] = gelu_fast(zr[i]);
    return pack_bf16(h);
}

constexpr int INFER_WAVES = 4;  // waves per workgroup
constexpr int INFER_NFRAG = 4 + 16 + 16 + 4;  // A-operand fragments of the four layers

// fragment f of the A operands: layer 1 nt (f < 4), layer 2 (nt, ks) (4 + 4 nt + ks),
// layer 3 (20 + 4 nt + ks), layer 4 ks (36 + ks); lane l holds
// W(16 nt + l % 16, 16 ks + 4 (l / 16) + i), i < 4
__device__ __forceinline__ bf16x4 weight_frag(const unsigned short *Wb, int f, int lane)
{
    int off, K, nt, ks;
    if (f < 4) off = MLP_OFF_W1, K = MLP_IN_PAD, nt = f, ks = 0;
    else if (f < 20) off = MLP_OFF_W2, K = MLP_HID, nt = (f - 4) >> 2, ks = (f - 4) & 3;
    else if (f < 36) off = MLP_OFF_W3, K = MLP_HID, nt = (f - 20) >> 2, ks = (f - 20) & 3;
    else off = MLP_OFF_W4, K = MLP_HID, nt = 0, ks = f - 36;
    const unsigned short *q = Wb + off + (16 * nt + (lane & 15)) * K + 16 * ks + 4 * (lane >> 4);
    bf16x4 v;
#pragma unroll
    for (int i = 0; i < 4; ++i) v[i] = (short)q[i];
    return v;
}

__global__ void __launch_bounds__(64 * INFER_WAVES)
k_mlp_infer(const float *__restrict__ X, float *__restrict__ Y, const unsigned short *__restrict__ Wb,
            const float *__restrict__ P, long B, long ntiles)
{
    constexpr int NB = 3 * MLP_HID + MLP_OUT_PAD;
    __shared__ __align__(16) float sB[NB];
    __shared__ __align__(16) bf16x4 sW[INFER_NFRAG * 64];
    for (int i = threadIdx.x; i < NB; i += blockDim.x) {
        const int l = i / MLP_HID;
        const int off = l == 0 ? MLP_OFF_B1 : l == 1 ? MLP_OFF_B2 : l == 2 ? MLP_OFF_B3 : MLP_OFF_B4;
        sB[i] = P[off + i - l * MLP_HID];
    }
    const int lane = threadIdx.x & 63, g = lane >> 4;
    for (int i = threadIdx.x; i < INFER_NFRAG * 64; i += blockDim.x) sW[i] = weight_frag(Wb, i >> 6, i & 63);
    __syncthreads();
    auto W = [&](int f) { return sW[f * 64 + lane]; };
    auto bias = [&](int l, int nt) { return *(const f32x4 *)(sB + l * MLP_HID + 16 * nt + 4 * g); };
    const long stride = (long)gridDim.x * INFER_WAVES;
    for (long t = (long)blockIdx.x * INFER_WAVES + (threadIdx.x >> 6); t < ntiles; t += stride) {
        // keeps the LDS weight reads inside the loop (hoisted, they would
        // occupy the registers the LDS placement is meant to free)
        asm volatile("" ::: "memory");
        const long s = 16 * t + (lane & 15);  // this lane's sample
        // B operand of layer 1: X(s, 4 g + i), zero beyond the 10 inputs and the batch
        f32x4 xv = {0, 0, 0, 0};
        if (s < B) {
#pragma unroll
            for (int i = 0; i < 4; ++i)
                if (4 * g + i < MLP_IN) xv[i] = X[s * MLP_IN + 4 * g + i];
        }
        const bf16x4 x = pack_bf16(xv);
        bf16x4 h[4], h2[4];
#pragma unroll
        for (int nt = 0; nt < 4; ++nt)
            h[nt] = hidden_act(__builtin_amdgcn_mfma_f32_16x16x16bf16_1k(W(nt), x, bias(0, nt), 0, 0, 0));
#pragma unroll
        for (int nt = 0; nt < 4; ++nt) {
            f32x4 acc = bias(1, nt);
#pragma unroll
            for (int ks = 0; ks < 4; ++ks)
                acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(W(4 + 4 * nt + ks), h[ks], acc, 0, 0, 0);
            h2[nt] = hidden_act(acc);
        }
#pragma unroll
        for (int nt = 0; nt < 4; ++nt) {
            f32x4 acc = bias(2, nt);
#pragma unroll
            for (int ks = 0; ks < 4; ++ks)
                acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(W(20 + 4 * nt + ks), h2[ks], acc, 0, 0, 0);
            h[nt] = hidden_act(acc);
        }
        f32x4 y = bias(3, 0);
#pragma unroll
        for (int ks = 0; ks < 4; ++ks) y = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(W(36 + ks), h[ks], y, 0, 0, 0);
        // lane (r, g) holds outputs 4 g .. 4 g + 3 of sample s; 6 are used
        if (s < B && g < 2) {
#pragma unroll
            for (int i = 0; i < 4; ++i)
                if (4 * g + i < MLP_OUT) Y[s * MLP_OUT + 4 * g + i] = y[i];
        }
    }
}

// ---------------------------------------------------------------------------
// backward: per-workgroup gradient partials of the mean squared error
// loss = mean over B x 6 of (Y - T)^2, or of any loss whose gradient with
// respect to the outputs is dY; partial[blockIdx.x][MLP_NPARAM]
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_mlp_backward(const float *__restrict__ Y, const float *__restrict__ T,
                                                     const float *__restrict__ dY,
                                                     const unsigned short *__restrict__ Wb, long B,
                                                     MlpCache cache, float *__restrict__ partial)
{
    __shared__ unsigned short sW[MLP_NPARAM];
    __shared__ unsigned short sD[TILE * HS];   // dZ (BF16), [sample][feature]
    __shared__ float sDf[TILE * (64 + 1)];     // dZ (FP32) for the bias sums
    __shared__ unsigned short sH[TILE * HS];   // layer input H, [sample][feature]
    __shared__ unsigned short sZ[TILE * HS];   // pre-activation of the layer below
    const Layers L;
    const long row0 = (long)blockIdx.x * TILE;
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63, r = lane & 15, g = lane >> 4;
    float *out = partial + (size_t)blockIdx.x * MLP_NPARAM;
    stage_weights(sW, Wb);
    const float scale = 2.0f / (float)(B * MLP_OUT);
    for (int i = threadIdx.x; i < TILE * MLP_OUT_PAD; i += blockDim.x) {
        const int m = i / MLP_OUT_PAD, n = i % MLP_OUT_PAD;
        const long s = row0 + m;
        // dL/dY: of the mean squared error against T, or given (dY)
        float d = 0.0f;
        if (s < B && n < MLP_OUT)
            d = dY ? dY[s * MLP_OUT + n] : scale * (Y[s * MLP_OUT + n] - T[s * MLP_OUT + n]);
        sDf[m * 65 + n] = d;
        sD[m * HS + n] = to_bf16(d);
    }
    for (int l = 3; l >= 0; --l) {
        const int K = L.in[l], N = L.out[l];
        const unsigned short *Wl = sW + L.w[l];
        // layer input H_l (and Z_l of the layer below, for the GELU derivative)
        const unsigned short *Hg = cache.H[l];
        const unsigned short *Zg = l > 0 ? cache.Z[l - 1] : nullptr;
        for (int i = threadIdx.x; i < TILE * K; i += blockDim.x) {
            const int m = i / K, k = i % K;
            const long s = row0 + m;
            sH[m * HS + k] = s < B ? Hg[s * K + k] : (unsigned short)0;
            if (Zg) sZ[m * HS + k] = s < B ? Zg[s * K + k] : (unsigned short)0;
        }
        __syncthreads();
        // bias gradient: column sums of dZ over the tile
        for (int n = threadIdx.x; n < N; n += blockDim.x) {
            float acc = 0.0f;
            for (int m = 0; m < TILE; ++m) acc += sDf[m * 65 + n];
            out[L.b[l] + n] = acc;
        }
        // weight gradient dW (N x K) = dZ^T (N x TILE) * H (TILE x K)
        for (int t = w; t < (N / 16) * (K / 16); t += 4) {
            const int mt = t / (K / 16), nt = t % (K / 16);
            f32x4 acc = {0, 0, 0, 0};
            acc = mma16([&](int m, int k) { return sD[k * HS + 16 * mt + m]; },
                        [&](int k, int n) { return sH[k * HS + 16 * nt + n]; }, TILE, acc);
#pragma unroll
            for (int i = 0; i < 4; ++i) out[L.w[l] + (16 * mt + 4 * g + i) * K + 16 * nt + r] = acc[i];
        }
        if (l == 0) break;
        // dH (TILE x K) = dZ (TILE x N) * W (N x K); dZ_below = dH * gelu'(Z)
        f32x4 dh[4];
        for (int nt = 0; nt < K / 16; ++nt) {
            f32x4 acc = {0, 0, 0, 0};
            dh[nt] = mma16([&](int m, int k) { return sD[(16 * w + m) * HS + k]; },
                           [&](int k, int n) { return Wl[k * K + 16 * nt + n]; }, N, acc);
        }
        __syncthreads();  // every read of sD / sDf for this layer is done
        for (int nt = 0; nt < K / 16; ++nt) {
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int m = 16 * w + 4 * g + i, col = 16 * nt + r;
                const float d = dh[nt][i] * gelu_grad_f(from_bf16(sZ[m * HS + col]));
                sDf[m * 65 + col] = d;
                sD[m * HS + col] = to_bf16(d);
            }
        }
        __syncthreads();
    }
}

// grads[p] = sum over the nblk partials, in block order
__global__ void __launch_bounds__(256) k_mlp_reduce(const float *__restrict__ partial, int nblk, float *grads)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= MLP_NPARAM) return;
    float acc = 0.0f;
    for (int b = 0; b < nblk; ++b) acc += partial[(size_t)b * MLP_NPARAM + p];
    grads[p] = acc;
}

// Adam on the FP32 master parameters, then the BF16 copy used by the kernels
__global__ void __launch_bounds__(256) k_mlp_adam(float *P, float *M, float *V, const float *grads,
                                                 unsigned short *Wb, float lr, float b1, float b2, float eps,
                                                 float c1, float c2)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= MLP_NPARAM) return;
    const float gr = grads[p];
    const float m = b1 * M[p] + (1.0f - b1) * gr;
    const float v = b2 * V[p] + (1.0f - b2) * gr * gr;
    M[p] = m;
    V[p] = v;
    const float np = P[p] - lr * (m / c1) / (sqrtf(v / c2) + eps);
    P[p] = np;
    Wb[p] = to_bf16(np);
}

__global__ void k_mlp_to_bf16(const float *P, unsigned short *Wb)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < MLP_NPARAM) Wb[p] = to_bf16(P[p]);
}

static unsigned nblocks(long B) { return (unsigned)((B + TILE - 1) / TILE); }

hipError_t launch_mlp_forward(const float *X, float *Y, const unsigned short *Wb, const float *P, long B,
                              const MlpCache &cache, hipStream_t st)
{
    if (B <= 0) return hipSuccess;
    if (!cache.H[0]) {
        // inference: as many persistent workgroups as are resident at once
        // (cached per device)
        constexpr int MAXDEV = 64;
        static std::atomic<int> resident[MAXDEV];
        int dev = 0;
        hipError_t e = hipGetDevice(&dev);
        if (e != hipSuccess) return e;
        int res = (dev >= 0 && dev < MAXDEV) ? resident[dev].load(std::memory_order_relaxed) : 0;
        if (!res) {
            int ncu = 0, per_cu = 0;
            e = hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev);
            if (e == hipSuccess)
                e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_mlp_infer, 64 * INFER_WAVES, 0);
            if (e != hipSuccess) return e;
            res = ncu * (per_cu > 0 ? per_cu : 1);
            if (dev >= 0 && dev < MAXDEV) resident[dev].store(res, std::memory_order_relaxed);
        }
        const long ntiles = (B + 15) / 16;  // 16 samples per wave
        const long need = (ntiles + INFER_WAVES - 1) / INFER_WAVES;
        const long nwg = need < res ? need : res;
        hipLaunchKernelGGL(k_mlp_infer, dim3((unsigned)nwg), dim3(64 * INFER_WAVES), 0, st, X, Y, Wb, P, B, ntiles);
        return hipGetLastError();
    }
    hipLaunchKernelGGL(k_mlp_forward, dim3(nblocks(B)), dim3(256), 0, st, X, Y, Wb, P, B, cache);
    return hipGetLastError();
}

hipError_t launch_mlp_backward(const float *Y, const float *T, const float *dY, const unsigned short *Wb, long B,
                               const MlpCache &cache, float *partial, float *grads, hipStream_t st)
{
    if (B <= 0) return hipSuccess;
    const unsigned nb = nblocks(B);
    hipLaunchKernelGGL(k_mlp_backward, dim3(nb), dim3(256), 0, st, Y, T, dY, Wb, B, cache, partial);
    hipLaunchKernelGGL(k_mlp_reduce, dim3((MLP_NPARAM + 255) / 256), dim3(256), 0, st, partial, (int)nb, grads);
    return hipGetLastError();
}

hipError_t launch_mlp_adam(float *P, float *M, float *V, const float *grads, unsigned short *Wb, float lr, float b1,
                           float b2, float eps, float c1, float c2, hipStream_t st)
{
    hipLaunchKernelGGL(k_mlp_adam, dim3((MLP_NPARAM + 255) / 256), dim3(256), 0, st, P, M, V, grads, Wb, lr, b1, b2,
                       eps, c1, c2);
    return hipGetLastError();
}

hipError_t launch_mlp_to_bf16(const float *P, unsigned short *Wb, hipStream_t st)
{
    hipLaunchKernelGGL(k_mlp_to_bf16, dim3((MLP_NPARAM + 255) / 256), dim3(256), 0, st, P, Wb);
    return hipGetLastError();
}

}  // namespace pdo
