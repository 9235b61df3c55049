// Host-side parallel helpers: a fork-join parallel_for over index ranges and
// a parallel LSD radix sort for 64-bit keys.  Thread count: RDMCE_THREADS,
// else OMP_NUM_THREADS, else hardware_concurrency, capped at 64.
#pragma once
#include <algorithm>
#include <cstdint>
#include <cstdlib>
#include <functional>
#include <thread>
#include <vector>

namespace rdmce {

inline int host_threads() {
  static int cached = 0;
  if (cached) return cached;
  int t = 0;
  if (const char *e = std::getenv("RDMCE_THREADS")) t = std::atoi(e);
  if (t <= 0) {
    if (const char *e = std::getenv("OMP_NUM_THREADS")) t = std::atoi(e);
  }
  if (t <= 0) t = (int)std::thread::hardware_concurrency();
  if (t <= 0) t = 1;
  cached = std::min(t, 64);
  return cached;
}

// Calls fn(tid, begin, end) on nthreads contiguous chunks of [0, n).
inline void parallel_chunks(uint64_t n, const std::function<void(int, uint64_t, uint64_t)> &fn,
                            int nthreads = 0) {
  if (nthreads <= 0) nthreads = host_threads();
  if (n < 4096 || nthreads == 1) {
    fn(0, 0, n);
    return;
  }
  nthreads = (int)std::min<uint64_t>(nthreads, n);
  std::vector<std::thread> th;
  th.reserve(nthreads);
  for (int t = 0; t < nthreads; ++t) {
    uint64_t b = n * t / nthreads, e = n * (t + 1) / nthreads;
    th.emplace_back([&fn, t, b, e] { fn(t, b, e); });
  }
  for (auto &x : th) x.join();
}

// Dynamic scheduling: fn(i) for i in [0, n), grabbed in blocks of `grain`.
void parallel_for_dynamic(uint64_t n, uint64_t grain, const std::function<void(uint64_t)> &fn,
                          int nthreads = 0);

// Sorts keys ascending.  Only the low `key_bits` bits are examined (the
// caller guarantees higher bits are zero).
void radix_sort_u64(std::vector<uint64_t> &keys, int key_bits = 64);

}  // namespace rdmce
