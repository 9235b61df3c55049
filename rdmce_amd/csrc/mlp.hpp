// Learned root-cost model: sizes and kernel launchers (mlp.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace rdmce {

constexpr int kFeat = 8;
constexpr int kH1 = 32;
constexpr int kH2 = 32;
// Trainable floats: W1, b1, W2, b2, W3, b3.
constexpr int kTrainFloats = kFeat * kH1 + kH1 + kH1 * kH2 + kH2 + kH2 + 1;
// Plus the input normalisation (mean, 1/std per feature).
constexpr int kParamFloats = kTrainFloats + 2 * kFeat;

hipError_t launch_features(const uint64_t *rowoff, const uint32_t *col, const uint32_t *core, float inv_degen,
                           const uint64_t *roots, uint64_t n_roots, float *out, hipStream_t s);
hipError_t launch_forward(const float *params, const float *X, uint64_t n, float *out, hipStream_t s);
uint32_t backward_blocks(uint64_t count);
hipError_t launch_backward(const float *params, const float *X, const float *Y, const uint32_t *idx, uint64_t count,
                           float *partial, float *loss_partial, float *grad, float *loss, hipStream_t s);
hipError_t launch_adam(float *params, const float *grad, float *m, float *v, float lr, int step, hipStream_t s);

}  // namespace rdmce
