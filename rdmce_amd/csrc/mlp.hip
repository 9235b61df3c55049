// Learned root-cost model: per-root features and a small FP32 MLP
// (8 -> 32 -> 32 -> 1, ReLU) trained with Adam, all in HIP.
//
// The model predicts log(1 + cost) of an edge root from cheap structural
// features; the scheduler orders roots by predicted cost and splits the
// heaviest before launch.  It never changes results, only the order and
// the eager-split flags (see mce_kernel.hip).
//
// Backward is deterministic: each block writes its partial gradient for a
// fixed slice of the minibatch, and a second kernel sums the partials in
// block order.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "mlp.hpp"

namespace rdmce {
namespace {

__device__ __forceinline__ uint64_t lower_bound_u32(const uint32_t *a, uint64_t lo, uint64_t hi, uint32_t key) {
  while (lo < hi) {
    uint64_t mid = (lo + hi) >> 1;
    if (a[mid] < key) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

// Features of edge root (v0, v1), v0 < v1:
//   log1p(deg v0), log1p(deg v1), core v0 / degeneracy, core v1 / degeneracy,
//   log1p(outdeg v0), log1p(outdeg v1), log1p(t), log1p(t)^2
// where t = |N+(v1) ^ N(v0)| counts the triangles (v0, v1, w) with w above
// v1 - the size of the root's candidate set P.  One thread per root.
__global__ __launch_bounds__(256) void features_kernel(const uint64_t *rowoff, const uint32_t *col,
                                                       const uint32_t *core, float inv_degen, const uint64_t *roots,
                                                       uint64_t n_roots, float *out) {
  uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n_roots) return;
  uint32_t v0 = (uint32_t)(roots[r] >> 32), v1 = (uint32_t)roots[r];
  uint64_t b0 = rowoff[v0], e0 = rowoff[v0 + 1], b1 = rowoff[v1], e1 = rowoff[v1 + 1];
  uint64_t f0 = lower_bound_u32(col, b0, e0, v0 + 1), f1 = lower_bound_u32(col, b1, e1, v1 + 1);
  // t: entries of N+(v1) found in N(v0) above v1
  uint64_t lo = lower_bound_u32(col, b0, e0, v1 + 1);
  uint32_t t = 0;
  for (uint64_t i = f1; i < e1 && lo < e0; ++i) {
    lo = lower_bound_u32(col, lo, e0, col[i]);
    if (lo < e0 && col[lo] == col[i]) ++t;
  }
  float lt = log1pf((float)t);
  float *o = out + r * kFeat;
  o[0] = log1pf((float)(e0 - b0));
  o[1] = log1pf((float)(e1 - b1));
  o[2] = (float)core[v0] * inv_degen;
  o[3] = (float)core[v1] * inv_degen;
  o[4] = log1pf((float)(e0 - f0));
  o[5] = log1pf((float)(e1 - f1));
  o[6] = lt;
  o[7] = lt * lt;
}

// Parameter layout (floats): W1[kFeat][kH1] b1[kH1] W2[kH1][kH2] b2[kH2]
// W3[kH2] b3, plus per-feature normalisation mean[kFeat] inv_std[kFeat]
// (not trained).
struct Params {
  const float *W1, *b1, *W2, *b2, *W3, *b3, *mean, *istd;
};
__device__ __forceinline__ Params params_at(const float *p) {
  Params q;
  q.W1 = p;
  q.b1 = q.W1 + kFeat * kH1;
  q.W2 = q.b1 + kH1;
  q.b2 = q.W2 + kH1 * kH2;
  q.W3 = q.b2 + kH2;
  q.b3 = q.W3 + kH2;
  q.mean = q.b3 + 1;
  q.istd = q.mean + kFeat;
  return q;
}

// Forward for one sample; keeps activations for backward.
__device__ __forceinline__ float forward_one(const Params &q, const float *x, float *xn, float *h1, float *h2) {
  // Fully unrolled: the activation arrays stay in registers, indexed by
  // constants only.
#pragma unroll
  for (int i = 0; i < kFeat; ++i) xn[i] = (x[i] - q.mean[i]) * q.istd[i];
#pragma unroll
  for (int j = 0; j < kH1; ++j) {
    float s = q.b1[j];
#pragma unroll
    for (int i = 0; i < kFeat; ++i) s = fmaf(xn[i], q.W1[i * kH1 + j], s);
    h1[j] = s > 0.f ? s : 0.f;
  }
#pragma unroll
  for (int j = 0; j < kH2; ++j) {
    float s = q.b2[j];
#pragma unroll
    for (int i = 0; i < kH1; ++i) s = fmaf(h1[i], q.W2[i * kH2 + j], s);
    h2[j] = s > 0.f ? s : 0.f;
  }
  float y = q.b3[0];
  for (int i = 0; i < kH2; ++i) y = fmaf(h2[i], q.W3[i], y);
  return y;
}

__global__ __launch_bounds__(256) void forward_kernel(const float *params, const float *X, uint64_t n, float *out) {
  __shared__ float sp[kParamFloats];
  for (int i = threadIdx.x; i < kParamFloats; i += blockDim.x) sp[i] = params[i];
  __syncthreads();
  uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  Params q = params_at(sp);
  float xn[kFeat], h1[kH1], h2[kH2];
  out[r] = forward_one(q, X + r * kFeat, xn, h1, h2);
}

// Partial gradient of the mean squared error over samples idx[0 .. count)
// (scaled by 1/count): block b takes samples [b*kRows, (b+1)*kRows), one per
// thread; the activations and back-propagated errors of its samples are
// staged in LDS and every gradient entry is then a dot product over the
// block's samples, summed in sample order (deterministic, no atomics).
constexpr int kRows = 64;

__global__ __launch_bounds__(kRows) void backward_partial_kernel(const float *params, const float *X, const float *Y,
                                                                 const uint32_t *idx, uint64_t count, float scale,
                                                                 float *partial, float *loss_partial) {
  __shared__ float sp[kParamFloats];
  __shared__ float s_xn[kRows * kFeat], s_h1[kRows * kH1], s_h2[kRows * kH2];
  __shared__ float s_d1[kRows * kH1], s_d2[kRows * kH2], s_dy[kRows], s_loss[kRows];
  for (int i = threadIdx.x; i < kParamFloats; i += blockDim.x) sp[i] = params[i];
  __syncthreads();
  const Params q = params_at(sp);
  const int t = threadIdx.x;
  const uint64_t si = (uint64_t)blockIdx.x * kRows + t;
  float xn[kFeat], h1[kH1], h2[kH2], d1[kH1], d2[kH2];
  float dy = 0.f, lossv = 0.f;
  for (int i = 0; i < kFeat; ++i) xn[i] = 0.f;
  for (int i = 0; i < kH1; ++i) h1[i] = d1[i] = 0.f;
  for (int i = 0; i < kH2; ++i) h2[i] = d2[i] = 0.f;
  if (si < count) {
    const uint64_t r = idx ? idx[si] : si;
    float y = forward_one(q, X + r * kFeat, xn, h1, h2);
    float d = y - Y[r];
    lossv = d * d;
    dy = 2.f * d * scale;
#pragma unroll
    for (int j = 0; j < kH2; ++j) d2[j] = h2[j] > 0.f ? dy * q.W3[j] : 0.f;
#pragma unroll
    for (int i = 0; i < kH1; ++i) {
      float acc = 0.f;
#pragma unroll
      for (int j = 0; j < kH2; ++j) acc = fmaf(q.W2[i * kH2 + j], d2[j], acc);
      d1[i] = h1[i] > 0.f ? acc : 0.f;
    }
  }
  for (int i = 0; i < kFeat; ++i) s_xn[t * kFeat + i] = xn[i];
  for (int i = 0; i < kH1; ++i) {
    s_h1[t * kH1 + i] = h1[i];
    s_d1[t * kH1 + i] = d1[i];
  }
  for (int i = 0; i < kH2; ++i) {
    s_h2[t * kH2 + i] = h2[i];
    s_d2[t * kH2 + i] = d2[i];
  }
  s_dy[t] = dy;
  s_loss[t] = lossv;
  __syncthreads();
  float *out = partial + (uint64_t)blockIdx.x * kTrainFloats;
  constexpr int oW1 = 0, ob1 = oW1 + kFeat * kH1, oW2 = ob1 + kH1, ob2 = oW2 + kH1 * kH2, oW3 = ob2 + kH2,
                ob3 = oW3 + kH2;
  for (int e = t; e < kTrainFloats; e += kRows) {
    float acc = 0.f;
    if (e < ob1) {
      const int i = (e - oW1) / kH1, j = (e - oW1) % kH1;
      for (int s = 0; s < kRows; ++s) acc = fmaf(s_xn[s * kFeat + i], s_d1[s * kH1 + j], acc);
    } else if (e < oW2) {
      const int j = e - ob1;
      for (int s = 0; s < kRows; ++s) acc += s_d1[s * kH1 + j];
    } else if (e < ob2) {
      const int i = (e - oW2) / kH2, j = (e - oW2) % kH2;
      for (int s = 0; s < kRows; ++s) acc = fmaf(s_h1[s * kH1 + i], s_d2[s * kH2 + j], acc);
    } else if (e < oW3) {
      const int j = e - ob2;
      for (int s = 0; s < kRows; ++s) acc += s_d2[s * kH2 + j];
    } else if (e < ob3) {
      const int i = e - oW3;
      for (int s = 0; s < kRows; ++s) acc = fmaf(s_dy[s], s_h2[s * kH2 + i], acc);
    } else {
      for (int s = 0; s < kRows; ++s) acc += s_dy[s];
    }
    out[e] = acc;
  }
  if (t == 0) {
    float sum = 0.f;
    for (int s = 0; s < kRows; ++s) sum += s_loss[s];
    loss_partial[blockIdx.x] = sum;
  }
}

// grad[i] = sum over blocks (in order) of partial[b][i].
__global__ __launch_bounds__(256) void reduce_partials_kernel(const float *partial, uint32_t blocks, float *grad,
                                                              const float *loss_partial, float *loss) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < kTrainFloats) {
    float s = 0.f;
    for (uint32_t b = 0; b < blocks; ++b) s += partial[(uint64_t)b * kTrainFloats + i];
    grad[i] = s;
  }
  if (i == 0) {
    float s = 0.f;
    for (uint32_t b = 0; b < blocks; ++b) s += loss_partial[b];
    *loss = s;
  }
}

// Adam (Kingma & Ba 2015) on the trainable prefix of params.
__global__ __launch_bounds__(256) void adam_kernel(float *params, const float *grad, float *m, float *v, float lr,
                                                   float b1, float b2, float eps, float bc1, float bc2) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= kTrainFloats) return;
  float g = grad[i];
  float mi = b1 * m[i] + (1.f - b1) * g;
  float vi = b2 * v[i] + (1.f - b2) * g * g;
  m[i] = mi;
  v[i] = vi;
  params[i] -= lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
}

}  // namespace

hipError_t launch_features(const uint64_t *rowoff, const uint32_t *col, const uint32_t *core, float inv_degen,
                           const uint64_t *roots, uint64_t n_roots, float *out, hipStream_t s) {
  if (!n_roots) return hipSuccess;
  hipLaunchKernelGGL(features_kernel, dim3((uint32_t)((n_roots + 255) / 256)), dim3(256), 0, s, rowoff, col, core,
                     inv_degen, roots, n_roots, out);
  return hipGetLastError();
}

hipError_t launch_forward(const float *params, const float *X, uint64_t n, float *out, hipStream_t s) {
  if (!n) return hipSuccess;
  hipLaunchKernelGGL(forward_kernel, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, s, params, X, n, out);
  return hipGetLastError();
}

uint32_t backward_blocks(uint64_t count) { return (uint32_t)((count + kRows - 1) / kRows); }

hipError_t launch_backward(const float *params, const float *X, const float *Y, const uint32_t *idx, uint64_t count,
                           float *partial, float *loss_partial, float *grad, float *loss, hipStream_t s) {
  uint32_t blocks = backward_blocks(count);
  if (!blocks) return hipSuccess;
  hipLaunchKernelGGL(backward_partial_kernel, dim3(blocks), dim3(kRows), 0, s, params, X, Y, idx, count,
                     1.0f / (float)count, partial, loss_partial);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(reduce_partials_kernel, dim3((kTrainFloats + 255) / 256), dim3(256), 0, s, partial, blocks,
                     grad, loss_partial, loss);
  return hipGetLastError();
}

hipError_t launch_adam(float *params, const float *grad, float *m, float *v, float lr, int step, hipStream_t s) {
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  float bc1 = 1.f - powf(b1, (float)step), bc2 = 1.f - powf(b2, (float)step);
  hipLaunchKernelGGL(adam_kernel, dim3((kTrainFloats + 255) / 256), dim3(256), 0, s, params, grad, m, v, lr, b1, b2,
                     eps, bc1, bc2);
  return hipGetLastError();
}

}  // namespace rdmce
