// szx_mlp.hpp -- host interface of the rate-model MLP kernels (szx_mlp.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace szxamd {

constexpr int MLP_IN = 24;
constexpr int MLP_HIDDEN = 64;
// offsets into the FP32 parameter array
constexpr int MLP_W1 = 0;                                   // [64][24]
constexpr int MLP_B1 = MLP_W1 + MLP_HIDDEN * MLP_IN;        // [64]
constexpr int MLP_W2 = MLP_B1 + MLP_HIDDEN;                 // [64][64] (out, in)
constexpr int MLP_B2 = MLP_W2 + MLP_HIDDEN * MLP_HIDDEN;    // [64]
constexpr int MLP_W3 = MLP_B2 + MLP_HIDDEN;                 // [64]
constexpr int MLP_B3 = MLP_W3 + MLP_HIDDEN;                 // scalar
constexpr int MLP_PARAMS = MLP_B3 + 1;                      // 5825
constexpr int MLP_MAX_WG = 256;

// y = MLP(X) for n samples (X: [n][24] FP32).
void mlp_forward_launch(const float *P, const float *X, uint64_t n, float *y, hipStream_t st);

// Gradient of mean((y - T)^2) over n samples into grad[MLP_PARAMS] and the
// loss into *loss; ws holds mlp_grad_workspace_bytes().
uint64_t mlp_grad_workspace_bytes();
void mlp_grad_launch(const float *P, const float *X, const float *T, uint64_t n, float *grad,
                     float *loss, uint8_t *ws, hipStream_t st);

// One bias-corrected Adam step (step counts from 1) on the FP32 parameters.
void mlp_adam_launch(float *P, const float *grad, float *m, float *v, float lr, float b1, float b2,
                     float eps, int step, hipStream_t st);

}  // namespace szxamd
