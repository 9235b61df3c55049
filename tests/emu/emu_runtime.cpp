// Host emulation of the HIP runtime for rdmce_amd/csrc (test infrastructure;
// see hip/hip_runtime.h).  Launches run synchronously: every thread of every
// block becomes an OS thread when the launch is small enough (the persistent
// enumeration kernel needs all its waves at once), else the blocks and
// threads of a kernel without collectives run one after another on the
// calling thread.
#include "hip/hip_runtime.h"

#include <linux/futex.h>
#include <pthread.h>
#include <sched.h>
#include <sys/mman.h>
#include <sys/syscall.h>
#include <time.h>
#include <unistd.h>

#include <execinfo.h>
#include <signal.h>

#include <chrono>
#include <map>
#include <memory>
#include <mutex>

thread_local dim3 blockIdx, threadIdx, blockDim, gridDim;

namespace {

constexpr size_t kPage = 4096;

// ---- allocations that end at a PROT_NONE page --------------------------
struct Mapping {
  void *base;
  size_t total;
};
std::mutex g_alloc_mu;
std::map<void *, Mapping> g_allocs;

void *guarded_alloc(size_t bytes, int fill) {
  const size_t body = (bytes + kPage - 1) / kPage * kPage;
  const size_t total = body + kPage;
  void *base = mmap(nullptr, total, PROT_READ | PROT_WRITE, MAP_PRIVATE | MAP_ANONYMOUS | MAP_NORESERVE, -1, 0);
  if (base == MAP_FAILED) return nullptr;
  mprotect((char *)base + body, kPage, PROT_NONE);
  // 256-byte aligned start as close to the guard page as alignment allows
  const size_t span = (bytes + 255) / 256 * 256;
  char *p = (char *)base + body - span;
  memset(p, fill, span);
  std::lock_guard<std::mutex> g(g_alloc_mu);
  g_allocs[p] = Mapping{base, total};
  return p;
}

void guarded_free(void *p) {
  if (!p) return;
  Mapping m;
  {
    std::lock_guard<std::mutex> g(g_alloc_mu);
    auto it = g_allocs.find(p);
    if (it == g_allocs.end()) emu_fail("free of a pointer that was not allocated");
    m = it->second;
    g_allocs.erase(it);
  }
  munmap(m.base, m.total);
}

// ---- waves -------------------------------------------------------------
struct Wave {
  std::atomic<uint32_t> count{0};
  std::atomic<uint32_t> gen{0};
  uint32_t lanes = 64;
  uint64_t buf[2][64];
  uint64_t seq[2][64];  // each lane's count of collectives, to catch divergent ones
  uint64_t site[2][64];  // with EMU_CHECK_SITES=1: a hash of each lane's call stack
  void *frames[2][64][12];
  int nframes[2][64];
};

thread_local Wave *t_wave = nullptr;
thread_local int t_lane = 0;
thread_local int t_phase = 0;
thread_local uint64_t t_seq = 0;
std::atomic<int> g_failing{0};  // set by the first lane that reports a failure
thread_local uint64_t *t_lds = nullptr;

double barrier_timeout_s() {
  static double t = [] {
    const char *e = getenv("EMU_BARRIER_TIMEOUT");
    return e ? atof(e) : 600.0;
  }();
  return t;
}

void wave_barrier(Wave *w) {
  const uint32_t g = w->gen.load(std::memory_order_acquire);
  if (w->count.fetch_add(1, std::memory_order_acq_rel) + 1 == w->lanes) {
    w->count.store(0, std::memory_order_relaxed);
    w->gen.fetch_add(1, std::memory_order_acq_rel);
    syscall(SYS_futex, (uint32_t *)&w->gen, FUTEX_WAKE_PRIVATE, INT32_MAX, nullptr, nullptr, 0);
    return;
  }
  for (int i = 0; i < 64; ++i) {
    if (w->gen.load(std::memory_order_acquire) != g) return;
    sched_yield();
  }
  const auto t0 = std::chrono::steady_clock::now();
  while (w->gen.load(std::memory_order_acquire) == g) {
    timespec ts{0, 50 * 1000 * 1000};
    syscall(SYS_futex, (uint32_t *)&w->gen, FUTEX_WAIT_PRIVATE, g, &ts, nullptr, 0);
    const double dt = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (dt > barrier_timeout_s()) emu_fail("wave barrier timeout: a collective that not every lane reaches");
  }
}

// Every collective (exchange or barrier) goes through here: all lanes of the
// wave must arrive having executed the same number of collectives, as lanes
// of one wave do on the GPU when collectives sit in uniform control flow.
bool check_sites() {
  static bool on = getenv("EMU_CHECK_SITES") && atoi(getenv("EMU_CHECK_SITES")) != 0;
  return on;
}

uint64_t *exchange(uint64_t v) {
  if (!t_wave) emu_fail("cross-lane operation in a launch run without waves");
  const int ph = t_phase;
  t_phase ^= 1;
  ++t_seq;
  t_wave->buf[ph][t_lane] = v;
  t_wave->seq[ph][t_lane] = t_seq;
  if (check_sites()) {
    void **fr = t_wave->frames[ph][t_lane];
    const int n = backtrace(fr, 12);
    t_wave->nframes[ph][t_lane] = n;
    uint64_t h = 1469598103934665603ull;
    for (int i = 0; i < n; ++i) h = (h ^ (uint64_t)(uintptr_t)fr[i]) * 1099511628211ull;
    t_wave->site[ph][t_lane] = h;
  }
  wave_barrier(t_wave);
  for (uint32_t i = 0; i < t_wave->lanes; ++i)
    if (t_wave->seq[ph][i] != t_seq) emu_fail("divergent collective: lanes of the wave disagree on which one this is");
  if (check_sites())
    for (uint32_t i = 0; i < t_wave->lanes; ++i)
      if (t_wave->site[ph][i] != t_wave->site[ph][t_lane]) {
        if (g_failing.load()) {
          for (;;) pause();
        }
        char msg[128];
        snprintf(msg, sizeof msg, "divergent collective: lane %u is at another call site; its frames:\n", i);
        if (write(2, msg, strlen(msg)) < 0) {
        }
        backtrace_symbols_fd(t_wave->frames[ph][i], t_wave->nframes[ph][i], 2);
        emu_fail("divergent collective (this lane's frames follow)");
      }
  return t_wave->buf[ph];
}

// ---- streams, events, graphs -------------------------------------------
struct StreamState {
  bool capturing = false;
  std::vector<EmuLaunch> rec;
};

}  // namespace

struct EmuStream {
  StreamState st;
};
struct EmuEvent {
  std::chrono::steady_clock::time_point t;
};
struct EmuGraph {
  std::vector<EmuLaunch> launches;
};

[[noreturn]] void emu_fail(const char *what) {
  if (g_failing.exchange(1)) {  // another lane is already reporting
    for (;;) pause();
  }
  fprintf(stderr, "emu: %s (block %u thread %u lane %d)\n", what, blockIdx.x, threadIdx.x, t_lane);
  void *frames[32];
  const int n = backtrace(frames, 32);
  backtrace_symbols_fd(frames, n, 2);
  fflush(stderr);
  abort();
}

namespace {
// A fault names the faulting address and the allocation that ends nearest
// below it (a guard-page hit means an access past that allocation's end).
void on_fault(int sig, siginfo_t *si, void *) {
  char msg[256];
  const uintptr_t a = (uintptr_t)si->si_addr;
  uintptr_t best = 0;
  size_t best_span = 0;
  for (const auto &kv : g_allocs) {  // no lock: the process is going down
    const uintptr_t p = (uintptr_t)kv.first;
    const size_t span = (uintptr_t)kv.second.base + kv.second.total - 4096 - p;
    if (p <= a && p > best) {
      best = p;
      best_span = span;
    }
  }
  const int n = snprintf(msg, sizeof msg,
                         "emu: signal %d at address %p (block %u thread %u lane %d); nearest allocation below: %p of "
                         "%zu bytes (offset %zu)\n",
                         sig, si->si_addr, blockIdx.x, threadIdx.x, t_lane, (void *)best, best_span,
                         best ? (size_t)(a - best) : 0);
  if (write(2, msg, n) < 0) {
  }
  void *frames[32];
  const int k = backtrace(frames, 32);
  backtrace_symbols_fd(frames, k, 2);
  _exit(139);
}

struct FaultHandlers {
  FaultHandlers() {
    struct sigaction sa;
    memset(&sa, 0, sizeof sa);
    sa.sa_sigaction = on_fault;
    sa.sa_flags = SA_SIGINFO | SA_ONSTACK;
    sigaction(SIGSEGV, &sa, nullptr);
    sigaction(SIGBUS, &sa, nullptr);
  }
} g_fault_handlers;
}  // namespace

uint32_t emu_lane() { return (uint32_t)t_lane; }

uint64_t emu_ballot(bool pred) {
  const uint64_t *b = exchange(pred ? 1 : 0);
  uint64_t m = 0;
  for (uint32_t i = 0; i < t_wave->lanes; ++i)
    if (b[i]) m |= 1ull << i;
  return m;
}

// v_readlane_b32 takes its lane select from an SGPR (or a constant): one
// index for the whole wave.  Each lane passes its value and its index; lanes
// that disagree on the index are a program the hardware would run
// differently (with the first active lane's index), so they fail here.
uint32_t emu_readlane(uint32_t v, uint32_t lane) {
  if (lane >= 64) emu_fail("readlane index >= 64");
  const uint64_t *b = exchange(((uint64_t)lane << 32) | v);
  for (uint32_t i = 0; i < t_wave->lanes; ++i)
    if ((uint32_t)(b[i] >> 32) != lane) emu_fail("readlane with a lane index that differs between lanes");
  if (lane >= t_wave->lanes) emu_fail("readlane of a lane the wave does not have");
  return (uint32_t)b[lane];
}

uint32_t emu_readfirstlane(uint32_t v) { return (uint32_t)exchange(v)[0]; }

uint32_t emu_bpermute(int addr, uint32_t v) {  // ds_bpermute: the value of lane (addr / 4) mod 64
  const uint64_t *b = exchange(v);
  const uint32_t src = ((uint32_t)addr >> 2) & 63;
  if (src >= t_wave->lanes) emu_fail("bpermute of a lane the wave does not have");
  return (uint32_t)b[src];
}

uint32_t emu_update_dpp(uint32_t old, uint32_t src, int ctrl, int row_mask, int bank_mask, bool bound_ctrl) {
  const uint64_t *b = exchange(src);
  const int l = t_lane, row = l >> 4, rl = l & 15;
  if (!((row_mask >> row) & 1) || !((bank_mask >> (rl >> 2)) & 1)) return old;
  int s = -1;
  if (ctrl >= 0x111 && ctrl <= 0x11F) {  // row_shr:n
    const int n = ctrl & 0xF;
    if (rl >= n) s = l - n;
  } else if (ctrl >= 0x101 && ctrl <= 0x10F) {  // row_shl:n
    const int n = ctrl & 0xF;
    if (rl + n < 16) s = l + n;
  } else if (ctrl == 0x142) {  // row_bcast:15: lane 15 of the previous row
    if (row >= 1) s = row * 16 - 1;
  } else if (ctrl == 0x143) {  // row_bcast:31: lane 31 to rows 2 and 3
    if (row >= 2) s = 31;
  } else {
    emu_fail("DPP control not emulated");
  }
  if (s < 0 || s >= (int)t_wave->lanes) return bound_ctrl ? 0u : old;
  return (uint32_t)b[s];
}

uint32_t emu_mbcnt_lo(uint32_t mask, uint32_t acc) {
  const int l = t_lane;
  const uint32_t m = l >= 32 ? mask : (l == 0 ? 0u : (mask & ((1u << l) - 1)));
  return acc + (uint32_t)__builtin_popcount(m);
}

uint32_t emu_mbcnt_hi(uint32_t mask, uint32_t acc) {
  const int l = t_lane;
  if (l < 32) return acc;
  const int k = l - 32;
  const uint32_t m = k == 0 ? 0u : (mask & ((1u << k) - 1));
  return acc + (uint32_t)__builtin_popcount(m);
}

void emu_wave_sync() {
  if (!t_wave) emu_fail("barrier in a launch run without waves");
  std::atomic_thread_fence(std::memory_order_seq_cst);
  (void)exchange(0);
  std::atomic_thread_fence(std::memory_order_seq_cst);
}

void emu_sleep() { usleep(20); }

uint64_t emu_clock64() {
  return (uint64_t)std::chrono::duration_cast<std::chrono::nanoseconds>(
             std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

uint64_t (*emu_lds_base())[] {
  if (!t_lds) emu_fail("LDS access in a launch without dynamic LDS");
  return (uint64_t(*)[])t_lds;
}

// ---- launches ------------------------------------------------------------
namespace {

size_t max_concurrent() {
  const char *e = getenv("EMU_MAX_THREADS");
  return e ? (size_t)atoll(e) : 4096;
}

struct ThreadArg {
  const EmuLaunch *L;
  uint32_t b, t;
  Wave *w;
  uint64_t *lds;
};

void *thread_main(void *p) {
  ThreadArg *a = (ThreadArg *)p;
  blockIdx = dim3(a->b);
  threadIdx = dim3(a->t);
  blockDim = a->L->block;
  gridDim = a->L->grid;
  t_wave = a->w;
  t_lane = (int)(a->t % 64);
  t_phase = 0;
  t_seq = 0;
  t_lds = a->lds;
  a->L->body();
  return nullptr;
}

}  // namespace

void emu_run(const EmuLaunch &L) {
  const uint32_t nb = L.grid.x, nt = L.block.x;
  if (L.grid.y != 1 || L.grid.z != 1 || L.block.y != 1 || L.block.z != 1) emu_fail("only 1-D launches are emulated");
  const size_t total = (size_t)nb * nt;
  std::vector<uint64_t *> lds(nb, nullptr);
  if (L.shmem)
    for (uint32_t b = 0; b < nb; ++b) lds[b] = (uint64_t *)guarded_alloc(L.shmem, 0xA5);
  if (total <= max_concurrent()) {
    const uint32_t wpb = (nt + 63) / 64;
    std::unique_ptr<Wave[]> waves(new Wave[(size_t)nb * wpb]);
    for (uint32_t b = 0; b < nb; ++b)
      for (uint32_t k = 0; k < wpb; ++k) waves[(size_t)b * wpb + k].lanes = std::min<uint32_t>(64, nt - 64 * k);
    std::vector<ThreadArg> args(total);
    std::vector<pthread_t> th(total);
    pthread_attr_t attr;
    pthread_attr_init(&attr);
    pthread_attr_setstacksize(&attr, 1 << 20);
    for (uint32_t b = 0; b < nb; ++b)
      for (uint32_t t = 0; t < nt; ++t) {
        ThreadArg &a = args[(size_t)b * nt + t];
        a = ThreadArg{&L, b, t, &waves[(size_t)b * wpb + t / 64], lds[b]};
        if (pthread_create(&th[(size_t)b * nt + t], &attr, thread_main, &a) != 0) emu_fail("pthread_create failed");
      }
    for (size_t i = 0; i < total; ++i) pthread_join(th[i], nullptr);
    pthread_attr_destroy(&attr);
  } else {
    Wave *saved_w = t_wave;
    uint64_t *saved_lds = t_lds;
    for (uint32_t b = 0; b < nb; ++b)
      for (uint32_t t = 0; t < nt; ++t) {
        blockIdx = dim3(b);
        threadIdx = dim3(t);
        blockDim = L.block;
        gridDim = L.grid;
        t_wave = nullptr;
        t_lane = (int)(t % 64);
        t_lds = lds[b];
        L.body();
      }
    t_wave = saved_w;
    t_lds = saved_lds;
  }
  for (uint64_t *p : lds) guarded_free(p);
}

bool emu_capture(hipStream_t s, const EmuLaunch &L) {
  if (!s || !s->st.capturing) return false;
  s->st.rec.push_back(L);
  return true;
}

// ---- runtime API ---------------------------------------------------------
hipError_t hipMalloc(void **p, size_t bytes) {
  *p = guarded_alloc(bytes ? bytes : 1, 0xA5);
  return *p ? hipSuccess : hipErrorOutOfMemory;
}
hipError_t hipFree(void *p) {
  guarded_free(p);
  return hipSuccess;
}
hipError_t hipMemcpy(void *dst, const void *src, size_t bytes, hipMemcpyKind) {
  if (bytes) memcpy(dst, src, bytes);
  return hipSuccess;
}
hipError_t hipMemcpyAsync(void *dst, const void *src, size_t bytes, hipMemcpyKind k, hipStream_t s) {
  if (s && s->st.capturing) emu_fail("memcpy inside a captured stream is not emulated");
  return hipMemcpy(dst, src, bytes, k);
}
hipError_t hipMemset(void *p, int v, size_t bytes) {
  if (bytes) memset(p, v, bytes);
  return hipSuccess;
}
hipError_t hipMemGetInfo(size_t *free_b, size_t *total_b) {
  *free_b = 2ull << 30;
  *total_b = 4ull << 30;
  return hipSuccess;
}
hipError_t hipSetDevice(int) { return hipSuccess; }
hipError_t hipGetDeviceCount(int *n) {
  *n = 1;
  return hipSuccess;
}
hipError_t hipGetDeviceProperties(hipDeviceProp_t *prop, int) {
  memset(prop, 0, sizeof(*prop));
  strcpy(prop->name, "host emulation");
  strcpy(prop->gcnArchName, "emu");
  const char *e = getenv("EMU_CUS");
  prop->multiProcessorCount = e ? atoi(e) : 1;
  prop->sharedMemPerMultiprocessor = 160 << 10;
  prop->sharedMemPerBlock = 160 << 10;
  prop->totalGlobalMem = 4ull << 30;
  prop->clockRate = 1000000;
  prop->warpSize = 64;
  return hipSuccess;
}
hipError_t hipGetLastError() { return hipSuccess; }
const char *hipGetErrorString(hipError_t) { return "emulated HIP error"; }
hipError_t hipDeviceSynchronize() { return hipSuccess; }
hipError_t hipStreamCreateWithFlags(hipStream_t *s, unsigned) {
  *s = new EmuStream();
  return hipSuccess;
}
hipError_t hipStreamDestroy(hipStream_t s) {
  delete s;
  return hipSuccess;
}
hipError_t hipStreamSynchronize(hipStream_t) { return hipSuccess; }
hipError_t hipEventCreate(hipEvent_t *e) {
  *e = new EmuEvent();
  (*e)->t = std::chrono::steady_clock::now();
  return hipSuccess;
}
hipError_t hipEventDestroy(hipEvent_t e) {
  delete e;
  return hipSuccess;
}
hipError_t hipEventRecord(hipEvent_t e, hipStream_t) {
  e->t = std::chrono::steady_clock::now();
  return hipSuccess;
}
hipError_t hipEventSynchronize(hipEvent_t) { return hipSuccess; }
hipError_t hipEventElapsedTime(float *ms, hipEvent_t a, hipEvent_t b) {
  *ms = std::chrono::duration<float, std::milli>(b->t - a->t).count();
  return hipSuccess;
}
hipError_t hipStreamBeginCapture(hipStream_t s, hipStreamCaptureMode) {
  if (!s) return hipErrorInvalidValue;
  s->st.capturing = true;
  s->st.rec.clear();
  return hipSuccess;
}
hipError_t hipStreamEndCapture(hipStream_t s, hipGraph_t *g) {
  if (!s || !s->st.capturing) return hipErrorInvalidValue;
  *g = new EmuGraph{s->st.rec};
  s->st.capturing = false;
  s->st.rec.clear();
  return hipSuccess;
}
hipError_t hipGraphInstantiate(hipGraphExec_t *e, hipGraph_t g, hipGraphNode_t *, char *, size_t) {
  *e = new EmuGraph{g->launches};
  return hipSuccess;
}
hipError_t hipGraphLaunch(hipGraphExec_t e, hipStream_t) {
  for (const EmuLaunch &L : e->launches) emu_run(L);
  return hipSuccess;
}
hipError_t hipGraphDestroy(hipGraph_t g) {
  delete g;
  return hipSuccess;
}
hipError_t hipGraphExecDestroy(hipGraphExec_t e) {
  delete e;
  return hipSuccess;
}
