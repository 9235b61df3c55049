// Enumeration that allocates nothing: mce_solve_device_csr_noalloc
// (include/rdmce.h).  For callers that cannot allocate device memory - code
// under HIP graph capture, tight serving loops - every scratch buffer of the
// one-root-per-wavefront kernel (the derived root list, counters, the clique
// size histogram, sharing records, per-wave data regions and stacks) lives in
// one file-scope device array.  Each device holds its own copy of the array,
// at its own address, from module load until the process ends; the host
// reaches it only through the address hipGetSymbolAddress returns for that
// device.  Every call on a device shares that device's array, so calls that
// overlap on one device race.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <string>

#include "capi_internal.hpp"
#include "mce_device.hpp"

using namespace rdmce;

namespace {

constexpr uint64_t kScratchBytes = MCE_NOALLOC_SCRATCH_BYTES;
static __device__ __align__(256) uint8_t scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
constexpr uint32_t kWavesPerCu = 4;     // few waves, large per-wave regions
constexpr uint32_t kDeferCap = 1024;
constexpr uint32_t kScanThreads = 1024;

// Counters in the array (u64 each).
enum Ctr { cCount, cHash, cNodes, cNext, cList, cNDeferred, cDeferNeed, cDeferStackNeed, cShared, cActive, cErr,
           cWords, cBuild, cPublished, cOwner, cNum };

std::mutex g_mu;
uint8_t *g_scratch[kMaxDevices] = {};
uint32_t g_cus[kMaxDevices] = {};
uint32_t g_lds_cu[kMaxDevices] = {};
uint32_t g_lds_block[kMaxDevices] = {};

// First higher neighbour of each vertex: outoff[v] = index in col of the
// first entry of row v above v.
__global__ void outoff_kernel(uint64_t n, const uint64_t *rowoff, const uint32_t *col, uint64_t *outoff) {
  for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += (uint64_t)gridDim.x * blockDim.x) {
    uint64_t lo = rowoff[v], hi = rowoff[v + 1];
    while (lo < hi) {
      const uint64_t mid = (lo + hi) >> 1;
      if (col[mid] <= (uint32_t)v) lo = mid + 1;
      else hi = mid;
    }
    outoff[v] = lo;
  }
}

// rfirst[v] = sum of the out-degrees of the vertices below v (one
// workgroup: each thread sums a contiguous chunk, then the chunk sums are
// scanned in LDS).
__global__ __launch_bounds__(kScanThreads) void scan_kernel(uint64_t n, const uint64_t *rowoff, const uint64_t *outoff,
                                                            uint64_t *rfirst) {
  __shared__ uint64_t part[kScanThreads];
  const uint32_t t = threadIdx.x;
  const uint64_t chunk = (n + kScanThreads - 1) / kScanThreads;
  const uint64_t b = std::min<uint64_t>(n, t * chunk), e = std::min<uint64_t>(n, b + chunk);
  uint64_t s = 0;
  for (uint64_t v = b; v < e; ++v) s += rowoff[v + 1] - outoff[v];
  part[t] = s;
  __syncthreads();
  if (t == 0) {
    uint64_t acc = 0;
    for (uint32_t i = 0; i < kScanThreads; ++i) {
      const uint64_t x = part[i];
      part[i] = acc;
      acc += x;
    }
  }
  __syncthreads();
  uint64_t acc = part[t];
  for (uint64_t v = b; v < e; ++v) {
    rfirst[v] = acc;
    acc += rowoff[v + 1] - outoff[v];
  }
}

// Canonical edge roots v0 << 32 | v1 (v0 < v1), ascending.
// (An asymmetric CSR could hold more than m entries above the diagonal; those
// are not written.)
__global__ void roots_kernel(uint64_t n, uint64_t m, const uint64_t *rowoff, const uint64_t *outoff,
                             const uint32_t *col, const uint64_t *rfirst, uint64_t *roots) {
  for (uint64_t v = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; v < n; v += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t o = outoff[v], e = rowoff[v + 1], r = rfirst[v];
    for (uint64_t k = 0; o + k < e && r + k < m; ++k) roots[r + k] = (v << 32) | col[o + k];
  }
}

// out[0..3] = count, hash, tree nodes, error flags; out[4 ..] = histogram.
__global__ void copy_out_kernel(const unsigned long long *ctr, const unsigned long long *hist, uint32_t hist_len,
                                uint64_t *out) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < 4 + hist_len; i += gridDim.x * blockDim.x) {
    uint64_t v;
    if (i == 0) v = ctr[cCount];
    else if (i == 1) v = ctr[cHash];
    else if (i == 2) v = ctr[cNodes];
    else if (i == 3) v = ctr[cErr] | (ctr[cNDeferred] ? MCE_NOALLOC_ERR_SCRATCH : 0);
    else v = hist[i - 4];
    out[i] = v;
  }
}

mce_status fail(mce_status s, const std::string &msg) {
  set_last_error(msg);
  return s;
}

uint64_t up(uint64_t x, uint64_t a) { return (x + a - 1) / a * a; }

}  // namespace

extern "C" mce_status mce_solve_device_csr_noalloc(int32_t device, uint64_t n, uint64_t nnz, const uint64_t *rowoff,
                                                    const uint32_t *col, const uint64_t *label_hash,
                                                    uint32_t hist_len, void *stream, uint64_t *out) {
  if (device < 0 || device >= kMaxDevices) return fail(MCE_ERR_ARG, "device ordinal out of range");
  if (!rowoff || !col || !label_hash || !out || (n == 0 && nnz != 0))
    return fail(MCE_ERR_ARG, "null device array");
  if (nnz % 2 || n >= (1ull << 32)) return fail(MCE_ERR_ARG, "not a symmetric CSR (odd entry count) or too many vertices");
  if (hist_len < 3 || hist_len > 4096) return fail(MCE_ERR_ARG, "hist_len must be in [3, 4096]");
  if (hipSetDevice(device) != hipSuccess) return fail(MCE_ERR_DEVICE, "hipSetDevice failed");
  uint8_t *base = nullptr;
  uint32_t cus = 0, lds_cu = 0, lds_block = 0;
  {
    // Once per device, with that device current: the array's address on it.
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_scratch[device]) {
      void *p = nullptr;
      hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
      if (e != hipSuccess || !p)
        return fail(MCE_ERR_DEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
      hipDeviceProp_t prop;
      e = hipGetDeviceProperties(&prop, device);
      if (e != hipSuccess) return fail(MCE_ERR_DEVICE, std::string("hipGetDeviceProperties: ") + hipGetErrorString(e));
      g_scratch[device] = (uint8_t *)p;
      g_cus[device] = (uint32_t)prop.multiProcessorCount;
      g_lds_cu[device] = (uint32_t)prop.sharedMemPerMultiprocessor;
      g_lds_block[device] = (uint32_t)prop.sharedMemPerBlock;
    }
    base = g_scratch[device];
    cus = g_cus[device];
    lds_cu = g_lds_cu[device];
    lds_block = g_lds_block[device];
  }
  const uint64_t m = nnz / 2;
  const uint32_t waves = std::max<uint32_t>(1, cus * kWavesPerCu);
  // Layout of the array.
  uint64_t off = 0;
  auto take = [&](uint64_t bytes) {
    const uint64_t o = off;
    off = up(off + bytes, 256);
    return o;
  };
  const uint64_t o_ctr = take(8ull * cNum), o_hist = take(8ull * hist_len), o_outoff = take(8 * (n + 1)),
                 o_rfirst = take(8 * (n + 1)), o_roots = take(8 * std::max<uint64_t>(m, 1)),
                 o_defer = take(8ull * kDeferCap), o_recs = take(sizeof(ShareRec) * (uint64_t)waves),
                 o_users = take(8ull * ((waves + 2) / 2)), o_work = take(8ull * ((waves + 63) / 64)),
                 o_idle = take(4ull * kIdleCopies * kIdleStride);
  if (off >= kScratchBytes) return fail(MCE_ERR_UNSUPPORTED, "graph too large for the no-alloc scratch array");
  const uint64_t per_wave = (kScratchBytes - off) / waves / 512 * 256;  // data region and stack, each
  if (per_wave < 4096)
    return fail(MCE_ERR_UNSUPPORTED, "graph too large for the no-alloc scratch array: " +
                                         std::to_string(kScratchBytes - off) + " bytes left for " +
                                         std::to_string(waves) + " waves");
  const uint64_t o_slots = take(per_wave * waves), o_stacks = take(per_wave * waves);
  if (off > kScratchBytes) return fail(MCE_ERR_INTERNAL, "no-alloc scratch layout");

  hipStream_t s = (hipStream_t)stream;
  unsigned long long *ctr = (unsigned long long *)(base + o_ctr);
  unsigned long long *hist = (unsigned long long *)(base + o_hist);
  uint64_t *outoff = (uint64_t *)(base + o_outoff), *rfirst = (uint64_t *)(base + o_rfirst);
  uint64_t *roots = (uint64_t *)(base + o_roots);
  // Resets with fill kernels (no memset nodes if the caller captures this).
  hipError_t e = launch_fill_zero(ctr, cNum, s);
  if (e == hipSuccess) e = launch_fill_zero(hist, hist_len, s);
  if (e == hipSuccess) e = launch_fill_zero(base + o_recs, sizeof(ShareRec) / 8 * (uint64_t)waves, s);
  if (e == hipSuccess) e = launch_fill_zero(base + o_users, (waves + 2) / 2, s);
  if (e == hipSuccess) e = launch_fill_zero(base + o_work, (waves + 63) / 64, s);
  if (e == hipSuccess) e = launch_fill_zero(base + o_idle, (uint64_t)kIdleCopies * kIdleStride / 2, s);
  if (e != hipSuccess) return fail(MCE_ERR_DEVICE, std::string("reset: ") + hipGetErrorString(e));
  if (m) {
    const uint32_t blocks = (uint32_t)std::min<uint64_t>(4096, (n + 255) / 256);
    hipLaunchKernelGGL(outoff_kernel, dim3(blocks), dim3(256), 0, s, n, rowoff, col, outoff);
    hipLaunchKernelGGL(scan_kernel, dim3(1), dim3(kScanThreads), 0, s, n, rowoff, (const uint64_t *)outoff, rfirst);
    hipLaunchKernelGGL(roots_kernel, dim3(blocks), dim3(256), 0, s, n, m, rowoff, (const uint64_t *)outoff, col,
                       (const uint64_t *)rfirst, roots);
    e = hipGetLastError();
    if (e != hipSuccess) return fail(MCE_ERR_DEVICE, std::string("root list: ") + hipGetErrorString(e));
    KArgs a{};
    a.rowoff = rowoff;
    a.outoff = outoff;
    a.col = col;
    a.label_hash = label_hash;
    a.n = n;
    a.roots = roots;
    a.n_roots_canon = m;
    a.n_dispatch = m;
    a.reverse = 1;  // high core first, BkPivotBitBalance.cuh:429-431
    a.root_mode = MCE_ROOT_EDGE;
    a.count = ctr + cCount;
    a.hash = ctr + cHash;
    a.total_nodes = ctr + cNodes;
    a.search_words = ctr + cWords;
    a.build_bytes = ctr + cBuild;
    a.next_root = ctr + cNext;
    a.list_cursor = ctr + cList;
    a.n_deferred = ctr + cNDeferred;
    a.deferred_need = ctr + cDeferNeed;
    a.deferred_stack_need = ctr + cDeferStackNeed;
    a.shared_tasks = ctr + cShared;
    a.published_levels = ctr + cPublished;
    a.owner_tasks = ctr + cOwner;
    a.active = (unsigned int *)(ctr + cActive);
    a.err = (unsigned int *)(ctr + cErr);
    a.hist = hist;
    a.hist_len = hist_len;
    a.slots = base + o_slots;
    a.slot_bytes = per_wave;
    a.stacks = base + o_stacks;
    a.stack_bytes = per_wave;
    a.recs = (ShareRec *)(base + o_recs);
    a.users = (unsigned int *)(base + o_users);
    a.has_work = (unsigned long long *)(base + o_work);
    a.idle_hint = (unsigned int *)(base + o_idle);
    a.share_bound = 24;
    a.n_waves = waves;
    a.lds_tier = 1;
    uint64_t lds = (uint64_t)(lds_cu ? lds_cu : (160u << 10)) / kWavesPerCu;
    lds = lds > 1024 ? lds - 1024 : 0;
    lds = std::min<uint64_t>(lds, (lds_block ? lds_block : (64u << 10)) - 2048);
    a.lds_bytes = (uint32_t)(lds / 256 * 256);
    a.word_path = 1;
    a.deferred = (uint64_t *)(base + o_defer);
    a.deferred_cap = kDeferCap;
    e = launch_mce(a, waves, kWavesPerCu, s);
    if (e != hipSuccess) return fail(MCE_ERR_DEVICE, std::string("enumeration: ") + hipGetErrorString(e));
  }
  hipLaunchKernelGGL(copy_out_kernel, dim3(1), dim3(256), 0, s, (const unsigned long long *)ctr,
                     (const unsigned long long *)hist, hist_len, out);
  e = hipGetLastError();
  if (e != hipSuccess) return fail(MCE_ERR_DEVICE, std::string("copy out: ") + hipGetErrorString(e));
  return MCE_OK;
}
