#include "parallel.hpp"

#include <atomic>
#include <cstring>

namespace rdmce {

void parallel_for_dynamic(uint64_t n, uint64_t grain, const std::function<void(uint64_t)> &fn,
                          int nthreads) {
  if (nthreads <= 0) nthreads = host_threads();
  if (grain == 0) grain = 1;
  if (n <= grain || nthreads == 1) {
    for (uint64_t i = 0; i < n; ++i) fn(i);
    return;
  }
  std::atomic<uint64_t> next{0};
  auto work = [&] {
    for (;;) {
      uint64_t b = next.fetch_add(grain, std::memory_order_relaxed);
      if (b >= n) return;
      uint64_t e = std::min(n, b + grain);
      for (uint64_t i = b; i < e; ++i) fn(i);
    }
  };
  std::vector<std::thread> th;
  for (int t = 0; t < nthreads; ++t) th.emplace_back(work);
  for (auto &x : th) x.join();
}

void radix_sort_u64(std::vector<uint64_t> &keys, int key_bits) {
  const uint64_t n = keys.size();
  if (n < 2) return;
  if (n < (1u << 16)) {
    std::sort(keys.begin(), keys.end());
    return;
  }
  constexpr int kBits = 11;
  constexpr int kBuckets = 1 << kBits;
  int nthreads = std::min<int>(host_threads(), (int)(n / 65536) + 1);
  std::vector<uint64_t> tmp(n);
  uint64_t *src = keys.data(), *dst = tmp.data();
  std::vector<uint64_t> hist((size_t)nthreads * kBuckets);
  for (int shift = 0; shift < key_bits; shift += kBits) {
    std::fill(hist.begin(), hist.end(), 0);
    parallel_chunks(
        n,
        [&](int t, uint64_t b, uint64_t e) {
          uint64_t *h = &hist[(size_t)t * kBuckets];
          for (uint64_t i = b; i < e; ++i) ++h[(src[i] >> shift) & (kBuckets - 1)];
        },
        nthreads);
    // Skip a pass whose digit is constant over all keys.
    bool trivial = false;
    for (int d = 0; d < kBuckets && !trivial; ++d) {
      uint64_t tot = 0;
      for (int t = 0; t < nthreads; ++t) tot += hist[(size_t)t * kBuckets + d];
      if (tot == n) trivial = true;
      else if (tot != 0) break;
    }
    if (trivial) continue;
    // Exclusive offsets in (digit, thread) order keep the sort stable.
    uint64_t run = 0;
    for (int d = 0; d < kBuckets; ++d)
      for (int t = 0; t < nthreads; ++t) {
        uint64_t c = hist[(size_t)t * kBuckets + d];
        hist[(size_t)t * kBuckets + d] = run;
        run += c;
      }
    parallel_chunks(
        n,
        [&](int t, uint64_t b, uint64_t e) {
          uint64_t *h = &hist[(size_t)t * kBuckets];
          for (uint64_t i = b; i < e; ++i) dst[h[(src[i] >> shift) & (kBuckets - 1)]++] = src[i];
        },
        nthreads);
    std::swap(src, dst);
  }
  if (src != keys.data()) std::memcpy(keys.data(), src, n * sizeof(uint64_t));
}

}  // namespace rdmce
