// Learned SGS closure network (pdo_mlp.hip): sizes, parameter layout, kernels.
#pragma once

#include <hip/hip_runtime.h>

namespace pdo {

// 10 inputs (padded to 16) -> 64 -> 64 -> 64 -> 6 outputs (padded to 16).
// FP32 parameter vector, row-major W[out][in] per layer, padding entries 0:
constexpr int MLP_IN = 10, MLP_IN_PAD = 16, MLP_HID = 64, MLP_OUT = 6, MLP_OUT_PAD = 16;
constexpr int MLP_OFF_W1 = 0;                                    // 64 x 16
constexpr int MLP_OFF_B1 = MLP_OFF_W1 + MLP_HID * MLP_IN_PAD;    // 64
constexpr int MLP_OFF_W2 = MLP_OFF_B1 + MLP_HID;                 // 64 x 64
constexpr int MLP_OFF_B2 = MLP_OFF_W2 + MLP_HID * MLP_HID;       // 64
constexpr int MLP_OFF_W3 = MLP_OFF_B2 + MLP_HID;                 // 64 x 64
constexpr int MLP_OFF_B3 = MLP_OFF_W3 + MLP_HID * MLP_HID;       // 64
constexpr int MLP_OFF_W4 = MLP_OFF_B3 + MLP_HID;                 // 16 x 64
constexpr int MLP_OFF_B4 = MLP_OFF_W4 + MLP_OUT_PAD * MLP_HID;   // 16
constexpr int MLP_NPARAM = MLP_OFF_B4 + MLP_OUT_PAD;             // 10448

// BF16 activations kept by a training forward pass for the backward pass:
// H[0] (B, 16) inputs, H[1..3] (B, 64) hidden outputs, Z[0..2] (B, 64)
// hidden pre-activations; all null for inference
struct MlpCache {
    unsigned short *H[4];
    unsigned short *Z[3];
};

hipError_t launch_mlp_forward(const float *X, float *Y, const unsigned short *Wb, const float *P, long B,
                              const MlpCache &cache, hipStream_t st);
// dY == nullptr: gradient of the mean squared error of Y against T;
// otherwise dY is dL/dY (B, MLP_OUT) and Y, T are unused
hipError_t launch_mlp_backward(const float *Y, const float *T, const float *dY, const unsigned short *Wb, long B,
                               const MlpCache &cache, float *partial, float *grads, hipStream_t st);
hipError_t launch_mlp_adam(float *P, float *M, float *V, const float *grads, unsigned short *Wb, float lr, float b1,
                           float b2, float eps, float c1, float c2, hipStream_t st);
hipError_t launch_mlp_to_bf16(const float *P, unsigned short *Wb, hipStream_t st);

}  // namespace pdo
