// Host emulation of the parts of the HIP runtime and of the wave64 device
// builtins that rdmce_amd/csrc uses.  Test infrastructure only: it lets the
// library's kernels run on the CPU (one OS thread per lane, the 64 lanes of
// a block form one wave) so that kernel changes can be checked for
// out-of-bounds accesses and wrong results without a GPU.  Device memory and
// each block's LDS end at a guard page, so an access past the end of an
// allocation stops the process.
//
// Cross-lane builtins (ballot, readlane, DPP, readfirstlane) and fences are
// collectives here: every lane of the wave must reach them, as they would in
// uniform control flow on the GPU; one that only some lanes reach stops the
// process after a timeout with a message naming the wave.
#pragma once

#include <algorithm>
#include <atomic>
#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <string>
#include <vector>

#define __device__
#define __host__
#define __global__
#define __forceinline__ inline
#define __launch_bounds__(...)
#define __shared__
// kernel-only attributes have no host meaning
#define amdgpu_waves_per_eu(...) unused
#define amdgpu_flat_work_group_size(...) unused

typedef int hipError_t;
enum : int { hipSuccess = 0, hipErrorInvalidValue = 1, hipErrorOutOfMemory = 2 };
typedef struct EmuStream *hipStream_t;
typedef struct EmuEvent *hipEvent_t;
typedef struct EmuGraph *hipGraph_t;
typedef struct EmuGraph *hipGraphExec_t;
typedef struct EmuGraphNode *hipGraphNode_t;
enum hipMemcpyKind {
  hipMemcpyHostToHost = 0,
  hipMemcpyHostToDevice = 1,
  hipMemcpyDeviceToHost = 2,
  hipMemcpyDeviceToDevice = 3,
  hipMemcpyDefault = 4
};
enum hipStreamCaptureMode {
  hipStreamCaptureModeGlobal = 0,
  hipStreamCaptureModeThreadLocal = 1,
  hipStreamCaptureModeRelaxed = 2
};
#define hipStreamNonBlocking 1u

struct uint4 {
  uint32_t x, y, z, w;
};

struct dim3 {
  uint32_t x, y, z;
  dim3(uint32_t a = 1, uint32_t b = 1, uint32_t c = 1) : x(a), y(b), z(c) {}
};

struct hipDeviceProp_t {
  char name[256];
  char gcnArchName[256];
  int multiProcessorCount;
  size_t sharedMemPerMultiprocessor;
  size_t sharedMemPerBlock;
  size_t totalGlobalMem;
  int clockRate;
  int warpSize;
};

// ---- runtime API (emu_runtime.cpp) ----
hipError_t hipMalloc(void **p, size_t bytes);
template <class T>
inline hipError_t hipMalloc(T **p, size_t bytes) {
  return hipMalloc((void **)p, bytes);
}
hipError_t hipFree(void *p);
hipError_t hipMemcpy(void *dst, const void *src, size_t bytes, hipMemcpyKind kind);
hipError_t hipMemcpyAsync(void *dst, const void *src, size_t bytes, hipMemcpyKind kind, hipStream_t s = nullptr);
hipError_t hipMemset(void *p, int v, size_t bytes);
hipError_t hipMemGetInfo(size_t *free_b, size_t *total_b);
hipError_t hipSetDevice(int d);
hipError_t hipGetDeviceCount(int *n);
hipError_t hipGetDeviceProperties(hipDeviceProp_t *prop, int d);
hipError_t hipGetLastError();
const char *hipGetErrorString(hipError_t e);
hipError_t hipDeviceSynchronize();
hipError_t hipStreamCreateWithFlags(hipStream_t *s, unsigned flags);
hipError_t hipStreamDestroy(hipStream_t s);
hipError_t hipStreamSynchronize(hipStream_t s);
hipError_t hipEventCreate(hipEvent_t *e);
hipError_t hipEventDestroy(hipEvent_t e);
hipError_t hipEventRecord(hipEvent_t e, hipStream_t s = nullptr);
hipError_t hipEventSynchronize(hipEvent_t e);
hipError_t hipEventElapsedTime(float *ms, hipEvent_t a, hipEvent_t b);
hipError_t hipStreamBeginCapture(hipStream_t s, hipStreamCaptureMode mode);
hipError_t hipStreamEndCapture(hipStream_t s, hipGraph_t *g);
hipError_t hipGraphInstantiate(hipGraphExec_t *e, hipGraph_t g, hipGraphNode_t *err_node, char *log, size_t n);
hipError_t hipGraphLaunch(hipGraphExec_t e, hipStream_t s);
hipError_t hipGraphDestroy(hipGraph_t g);
hipError_t hipGraphExecDestroy(hipGraphExec_t e);

// ---- kernel launches ----
struct EmuLaunch {
  dim3 grid, block;
  size_t shmem;
  std::function<void()> body;
};
// Runs a launch now (all threads of all blocks at once when they are few
// enough, else block after block on one thread: kernels without collectives).
void emu_run(const EmuLaunch &L);
// Records the launch when `s` is being captured; returns whether it did.
bool emu_capture(hipStream_t s, const EmuLaunch &L);

template <class... P, class... A>
inline hipError_t emu_launch(void (*k)(P...), dim3 grid, dim3 block, size_t shmem, hipStream_t s, A... args) {
  EmuLaunch L{grid, block, shmem, [=]() { k(static_cast<P>(args)...); }};
  if (!emu_capture(s, L)) emu_run(L);
  return hipSuccess;
}
#define hipLaunchKernelGGL(k, g, b, sh, st, ...) emu_launch(k, g, b, sh, st, ##__VA_ARGS__)

// ---- per-thread launch state and device builtins ----
extern thread_local dim3 blockIdx, threadIdx, blockDim, gridDim;
uint32_t emu_lane();
uint64_t emu_ballot(bool pred);
uint32_t emu_readlane(uint32_t v, uint32_t lane);
uint32_t emu_readfirstlane(uint32_t v);
uint32_t emu_bpermute(int addr, uint32_t v);
uint32_t emu_update_dpp(uint32_t old, uint32_t src, int ctrl, int row_mask, int bank_mask, bool bound_ctrl);
uint32_t emu_mbcnt_lo(uint32_t mask, uint32_t acc);
uint32_t emu_mbcnt_hi(uint32_t mask, uint32_t acc);
void emu_wave_sync();  // barrier over the wave's lanes + memory fence
void emu_sleep();
uint64_t emu_clock64();
uint64_t (*emu_lds_base())[];  // this block's dynamic LDS
[[noreturn]] void emu_fail(const char *what);

inline int emu_ctz(unsigned long long x) {
  if (x == 0) emu_fail("count trailing zeros of 0");
  return __builtin_ctzll(x);
}

#define __lane_id() emu_lane()
#define __ballot(x) emu_ballot((x))
#define __popcll(x) __builtin_popcountll((unsigned long long)(x))
#define __builtin_ctzll(x) emu_ctz((x))
#define __builtin_amdgcn_readlane(v, l) emu_readlane((uint32_t)(v), (uint32_t)(l))
#define __builtin_amdgcn_readfirstlane(v) emu_readfirstlane((uint32_t)(v))
#define __builtin_amdgcn_ds_bpermute(a, v) ((int)emu_bpermute((a), (uint32_t)(v)))
#define __builtin_amdgcn_update_dpp(o, s, c, rm, bm, bc) emu_update_dpp((o), (s), (c), (rm), (bm), (bc))
#define __builtin_amdgcn_mbcnt_lo(m, a) emu_mbcnt_lo((m), (a))
#define __builtin_amdgcn_mbcnt_hi(m, a) emu_mbcnt_hi((m), (a))
#define __builtin_amdgcn_wave_barrier() emu_wave_sync()
#define __builtin_amdgcn_fence(order, scope) emu_wave_sync()
#define __builtin_amdgcn_s_sleep(n) emu_sleep()
#define __syncthreads() emu_wave_sync()
#define clock64() emu_clock64()
#define lds_dyn (*emu_lds_base())

// ---- atomics (sequentially consistent on the host) ----
#define __HIP_MEMORY_SCOPE_WORKGROUP 2
#define __HIP_MEMORY_SCOPE_AGENT 3
using emu_as3_u64 = __attribute__((address_space(3))) uint64_t;
using emu_as3_ull = __attribute__((address_space(3))) unsigned long long;
template <class T>
inline T *emu_gp(T *p) {
  return p;
}
inline uint64_t *emu_gp(emu_as3_u64 *p) { return (uint64_t *)p; }
inline unsigned long long *emu_gp(emu_as3_ull *p) { return (unsigned long long *)p; }
using emu_as3_u32 = __attribute__((address_space(3))) uint32_t;
inline uint32_t *emu_gp(emu_as3_u32 *p) { return (uint32_t *)p; }

#define __hip_atomic_fetch_add(p, v, o, s) __atomic_fetch_add(emu_gp(p), (v), __ATOMIC_SEQ_CST)
#define __hip_atomic_fetch_sub(p, v, o, s) __atomic_fetch_sub(emu_gp(p), (v), __ATOMIC_SEQ_CST)
#define __hip_atomic_fetch_or(p, v, o, s) __atomic_fetch_or(emu_gp(p), (v), __ATOMIC_SEQ_CST)
#define __hip_atomic_load(p, o, s) __atomic_load_n(emu_gp(p), __ATOMIC_SEQ_CST)
#define __hip_atomic_store(p, v, o, s) __atomic_store_n(emu_gp(p), (v), __ATOMIC_SEQ_CST)

template <class T, class U>
inline T atomicAdd(T *p, U v) {
  return __atomic_fetch_add(p, (T)v, __ATOMIC_SEQ_CST);
}
template <class T, class U>
inline T atomicSub(T *p, U v) {
  return __atomic_fetch_sub(p, (T)v, __ATOMIC_SEQ_CST);
}
template <class T, class U>
inline T atomicOr(T *p, U v) {
  return __atomic_fetch_or(p, (T)v, __ATOMIC_SEQ_CST);
}
template <class T, class U>
inline T atomicAnd(T *p, U v) {
  return __atomic_fetch_and(p, (T)v, __ATOMIC_SEQ_CST);
}
template <class T, class U>
inline T atomicMax(T *p, U v) {
  T cur = __atomic_load_n(p, __ATOMIC_SEQ_CST);
  while (cur < (T)v && !__atomic_compare_exchange_n(p, &cur, (T)v, false, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST)) {
  }
  return cur;
}
