// C ABI of the learned root-cost model (include/rdmce.h, "Learned cost
// model"): root features on the GPU, FP32 MLP training (Adam) and inference
// with the HIP kernels of mlp.hip.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "capi_internal.hpp"
#include "mlp.hpp"
#include "rng.hpp"

using namespace rdmce;

struct mce_cost_model {
  std::vector<float> params;  // kParamFloats
};

namespace {
mce_status mfail(mce_status s, const std::string &msg) {
  set_last_error(msg);
  return s;
}

#define MTRY(expr)                                                                                 \
  do {                                                                                             \
    hipError_t e_ = (expr);                                                                        \
    if (e_ != hipSuccess) return mfail(MCE_ERR_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

// Device buffers released on scope exit.
struct DevBufs {
  std::vector<void *> ptrs;
  ~DevBufs() {
    for (void *p : ptrs) (void)hipFree(p);
  }
  template <class T>
  hipError_t alloc(T **p, uint64_t count) {
    hipError_t e = hipMalloc((void **)p, std::max<uint64_t>(count, 1) * sizeof(T));
    if (e == hipSuccess) ptrs.push_back(*p);
    return e;
  }
};

void check_device(int32_t device, mce_status &st) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n == 0) st = mfail(MCE_ERR_DEVICE, "no HIP device");
  else if (device < 0 || device >= n) st = mfail(MCE_ERR_ARG, "device ordinal out of range");
}
}  // namespace

extern "C" {

mce_status mce_root_features(mce_graph *g, int32_t device, float *out, uint64_t cap, uint64_t *n_roots) {
  if (!g || !n_roots) return mfail(MCE_ERR_ARG, "null argument");
  const HostGraph &G = g->g;
  // Canonical edge roots (v0 << 32 | v1), v0 < v1, ascending.
  std::vector<uint64_t> roots;
  roots.reserve(G.m);
  for (uint64_t v = 0; v < G.n; ++v)
    for (uint64_t i = G.rowoff[v]; i < G.rowoff[v + 1]; ++i)
      if (G.col[i] > v) roots.push_back((v << 32) | G.col[i]);
  *n_roots = roots.size();
  if (!out) return MCE_OK;
  if (cap < roots.size() * kFeat) return mfail(MCE_ERR_ARG, "feature buffer too small");
  if (roots.empty()) return MCE_OK;
  if (G.core.size() != G.n) {
    std::vector<uint32_t> order, core;
    degeneracy_order(G, order, core);
    g->g.core.swap(core);
  }
  mce_status st = MCE_OK;
  check_device(device, st);
  if (st) return st;
  MTRY(hipSetDevice(device));
  DevBufs d;
  uint64_t *d_row = nullptr, *d_roots = nullptr;
  uint32_t *d_col = nullptr, *d_core = nullptr;
  float *d_out = nullptr;
  MTRY(d.alloc(&d_row, G.n + 1));
  MTRY(d.alloc(&d_col, G.col.size()));
  MTRY(d.alloc(&d_core, G.n));
  MTRY(d.alloc(&d_roots, roots.size()));
  MTRY(d.alloc(&d_out, roots.size() * kFeat));
  MTRY(hipMemcpy(d_row, G.rowoff.data(), (G.n + 1) * 8, hipMemcpyHostToDevice));
  MTRY(hipMemcpy(d_col, G.col.data(), G.col.size() * 4, hipMemcpyHostToDevice));
  MTRY(hipMemcpy(d_core, g->g.core.data(), G.n * 4, hipMemcpyHostToDevice));
  MTRY(hipMemcpy(d_roots, roots.data(), roots.size() * 8, hipMemcpyHostToDevice));
  float inv_degen = 1.0f / (float)std::max<uint64_t>(1, G.degeneracy);
  MTRY(launch_features(d_row, d_col, d_core, inv_degen, d_roots, roots.size(), d_out, nullptr));
  MTRY(hipMemcpy(out, d_out, roots.size() * kFeat * 4, hipMemcpyDeviceToHost));
  return MCE_OK;
}

uint32_t mce_cost_model_num_params(void) { return (uint32_t)kParamFloats; }
uint32_t mce_cost_model_num_trainable(void) { return (uint32_t)kTrainFloats; }

mce_status mce_cost_model_create(uint64_t seed, mce_cost_model **out) {
  if (!out) return mfail(MCE_ERR_ARG, "null argument");
  auto m = new mce_cost_model();
  m->params.assign(kParamFloats, 0.f);
  // He-uniform initialisation of the weights, zero biases, identity
  // normalisation.
  Xoshiro256ss rng(seed);
  auto init = [&](float *w, int fan_in, int count) {
    float lim = std::sqrt(6.0f / (float)fan_in);
    for (int i = 0; i < count; ++i) w[i] = (float)((2.0 * rng.next_double() - 1.0) * lim);
  };
  float *p = m->params.data();
  init(p, kFeat, kFeat * kH1);
  init(p + kFeat * kH1 + kH1, kH1, kH1 * kH2);
  init(p + kFeat * kH1 + kH1 + kH1 * kH2 + kH2, kH2, kH2);
  float *istd = p + kTrainFloats + kFeat;
  for (int i = 0; i < kFeat; ++i) istd[i] = 1.f;
  *out = m;
  return MCE_OK;
}

mce_status mce_cost_model_get_params(const mce_cost_model *m, float *out, uint64_t cap) {
  if (!m || !out || cap < (uint64_t)kParamFloats) return mfail(MCE_ERR_ARG, "bad argument");
  std::memcpy(out, m->params.data(), kParamFloats * 4);
  return MCE_OK;
}

mce_status mce_cost_model_set_params(mce_cost_model *m, const float *in, uint64_t n) {
  if (!m || !in || n != (uint64_t)kParamFloats) return mfail(MCE_ERR_ARG, "bad argument");
  std::memcpy(m->params.data(), in, kParamFloats * 4);
  return MCE_OK;
}

mce_status mce_cost_model_fit_normalizer(mce_cost_model *m, const float *X, uint64_t n) {
  if (!m || (!X && n)) return mfail(MCE_ERR_ARG, "bad argument");
  float *mean = m->params.data() + kTrainFloats, *istd = mean + kFeat;
  for (int f = 0; f < kFeat; ++f) {
    double s = 0, s2 = 0;
    for (uint64_t i = 0; i < n; ++i) {
      double x = X[i * kFeat + f];
      s += x;
      s2 += x * x;
    }
    double mu = n ? s / n : 0, var = n ? s2 / n - mu * mu : 1;
    mean[f] = (float)mu;
    istd[f] = (float)(1.0 / std::sqrt(std::max(var, 1e-12)));
  }
  return MCE_OK;
}

// Mean squared error and its gradient over all n samples (trainable
// parameters only).
mce_status mce_cost_model_gradient(const mce_cost_model *m, int32_t device, const float *X, const float *Y,
                                   uint64_t n, float *grad, float *loss) {
  if (!m || !X || !Y || !grad || !loss || n == 0) return mfail(MCE_ERR_ARG, "bad argument");
  mce_status st = MCE_OK;
  check_device(device, st);
  if (st) return st;
  MTRY(hipSetDevice(device));
  DevBufs d;
  float *dp, *dx, *dy, *dpart, *dlp, *dg, *dl;
  uint32_t blocks = backward_blocks(n);
  MTRY(d.alloc(&dp, kParamFloats));
  MTRY(d.alloc(&dx, n * kFeat));
  MTRY(d.alloc(&dy, n));
  MTRY(d.alloc(&dpart, (uint64_t)blocks * kTrainFloats));
  MTRY(d.alloc(&dlp, blocks));
  MTRY(d.alloc(&dg, kTrainFloats));
  MTRY(d.alloc(&dl, 1));
  MTRY(hipMemcpy(dp, m->params.data(), kParamFloats * 4, hipMemcpyHostToDevice));
  MTRY(hipMemcpy(dx, X, n * kFeat * 4, hipMemcpyHostToDevice));
  MTRY(hipMemcpy(dy, Y, n * 4, hipMemcpyHostToDevice));
  MTRY(launch_backward(dp, dx, dy, nullptr, n, dpart, dlp, dg, dl, nullptr));
  MTRY(hipMemcpy(grad, dg, kTrainFloats * 4, hipMemcpyDeviceToHost));
  MTRY(hipMemcpy(loss, dl, 4, hipMemcpyDeviceToHost));
  *loss /= (float)n;
  return MCE_OK;
}

mce_status mce_cost_model_predict(const mce_cost_model *m, int32_t device, const float *X, uint64_t n, float *out) {
  if (!m || (!X && n) || (!out && n)) return mfail(MCE_ERR_ARG, "bad argument");
  if (n == 0) return MCE_OK;
  mce_status st = MCE_OK;
  check_device(device, st);
  if (st) return st;
  MTRY(hipSetDevice(device));
  DevBufs d;
  float *dp, *dx, *dout;
  MTRY(d.alloc(&dp, kParamFloats));
  MTRY(d.alloc(&dx, n * kFeat));
  MTRY(d.alloc(&dout, n));
  MTRY(hipMemcpy(dp, m->params.data(), kParamFloats * 4, hipMemcpyHostToDevice));
  MTRY(hipMemcpy(dx, X, n * kFeat * 4, hipMemcpyHostToDevice));
  MTRY(launch_forward(dp, dx, n, dout, nullptr));
  MTRY(hipMemcpy(out, dout, n * 4, hipMemcpyDeviceToHost));
  return MCE_OK;
}

// Minibatch Adam on the mean squared error; the samples are reshuffled
// every epoch (xoshiro256** seeded by `seed` + epoch).
mce_status mce_cost_model_train(mce_cost_model *m, int32_t device, const float *X, const float *Y, uint64_t n,
                                uint32_t epochs, uint32_t batch, float lr, uint64_t seed, float *final_loss) {
  if (!m || !X || !Y || n == 0 || batch == 0 || epochs == 0) return mfail(MCE_ERR_ARG, "bad argument");
  if (n >= (1ull << 32)) return mfail(MCE_ERR_UNSUPPORTED, "more than 2^32-1 samples");
  mce_status st = MCE_OK;
  check_device(device, st);
  if (st) return st;
  MTRY(hipSetDevice(device));
  DevBufs d;
  float *dp, *dx, *dy, *dpart, *dlp, *dg, *dl, *dm, *dv;
  uint32_t *didx;
  uint32_t blocks = backward_blocks(batch);
  MTRY(d.alloc(&dp, kParamFloats));
  MTRY(d.alloc(&dx, n * kFeat));
  MTRY(d.alloc(&dy, n));
  MTRY(d.alloc(&didx, n));
  MTRY(d.alloc(&dpart, (uint64_t)blocks * kTrainFloats));
  MTRY(d.alloc(&dlp, blocks));
  MTRY(d.alloc(&dg, kTrainFloats));
  const uint64_t steps_per_epoch = (n + batch - 1) / batch;
  MTRY(d.alloc(&dl, steps_per_epoch));
  MTRY(d.alloc(&dm, kTrainFloats));
  MTRY(d.alloc(&dv, kTrainFloats));
  MTRY(hipMemcpy(dp, m->params.data(), kParamFloats * 4, hipMemcpyHostToDevice));
  MTRY(hipMemcpy(dx, X, n * kFeat * 4, hipMemcpyHostToDevice));
  MTRY(hipMemcpy(dy, Y, n * 4, hipMemcpyHostToDevice));
  MTRY(hipMemset(dm, 0, kTrainFloats * 4));
  MTRY(hipMemset(dv, 0, kTrainFloats * 4));
  std::vector<uint32_t> idx(n);
  int step = 0;
  double epoch_loss = 0;
  for (uint32_t ep = 0; ep < epochs; ++ep) {
    for (uint64_t i = 0; i < n; ++i) idx[i] = (uint32_t)i;
    Xoshiro256ss rng(stream_seed(seed, ep));
    for (uint64_t i = n - 1; i > 0; --i) std::swap(idx[i], idx[rng.next_below(i + 1)]);
    MTRY(hipMemcpy(didx, idx.data(), n * 4, hipMemcpyHostToDevice));
    epoch_loss = 0;
    for (uint64_t b = 0, k = 0; b < n; b += batch, ++k) {
      uint64_t cnt = std::min<uint64_t>(batch, n - b);
      MTRY(launch_backward(dp, dx, dy, didx + b, cnt, dpart, dlp, dg, dl + k, nullptr));
      MTRY(launch_adam(dp, dg, dm, dv, lr, ++step, nullptr));
    }
    // per-step loss sums of this epoch (unscaled squared errors)
    std::vector<float> ls(steps_per_epoch);
    MTRY(hipMemcpy(ls.data(), dl, steps_per_epoch * 4, hipMemcpyDeviceToHost));
    for (float l : ls) epoch_loss += l;
  }
  MTRY(hipMemcpy(m->params.data(), dp, kTrainFloats * 4, hipMemcpyDeviceToHost));
  if (final_loss) *final_loss = (float)(epoch_loss / (double)n);
  for (float v : m->params)
    if (!std::isfinite(v)) return mfail(MCE_ERR_INTERNAL, "training diverged (non-finite weights)");
  return MCE_OK;
}

void mce_cost_model_free(mce_cost_model *m) { delete m; }

}  // extern "C"
