// Device-side data shared by the enumeration kernels and their launcher.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace rdmce {

// Maximum P size handled (64 words of 64 bits, one word per lane).
constexpr uint32_t kMaxP = 4096;
constexpr uint32_t kMaxW = kMaxP / 64;
// Per-wave LDS histogram of clique sizes below this bound.
constexpr uint32_t kLdsHist = 64;

// Error flags raised by the kernels (bit mask in KArgs::err).
constexpr uint32_t kErrPTooLarge = 1u;       // a root has |P| > kMaxP
constexpr uint32_t kErrHistOverflow = 2u;    // clique larger than hist_len - 1
constexpr uint32_t kErrDeferFull = 4u;       // deferred-root list full
constexpr uint32_t kErrListFull = 8u;        // clique list capacity exceeded

struct KArgs {
  // Graph (CSR in HBM).
  const uint64_t *rowoff;
  const uint32_t *col;
  const uint64_t *label_hash;   // splitmix64(label(v)) per vertex
  const uint32_t *labels;       // original ids (clique list)
  uint64_t n;
  // Roots: canonical descriptors (edge: v0<<32|v1, vertex: v).
  const uint64_t *roots;
  uint64_t n_roots_canon;
  // Dispatch: entry i of this pass handles canonical root
  //   perm ? perm[i] : (reverse ? n_dispatch-1-i : i).
  const uint64_t *perm;
  uint64_t n_dispatch;
  int32_t reverse;
  int32_t root_mode;            // 1 vertex, 2 edge
  // Results.
  unsigned long long *count;
  unsigned long long *hash;
  unsigned long long *hist;     // hist_len entries
  uint32_t hist_len;
  unsigned long long *total_nodes;
  unsigned long long *root_nodes;   // per canonical root, nullable
  unsigned long long *root_cycles;  // per canonical root, nullable
  uint32_t *list;                   // nullable
  unsigned long long *list_cursor;
  uint64_t list_cap;
  // Work distribution.
  unsigned long long *next_root;
  // Per-wave HBM scratch.
  uint8_t *slots;
  uint64_t slot_bytes;
  // Roots whose scratch need exceeds slot_bytes are deferred to a pass with
  // larger slots: canonical ids, count, and the largest need seen.
  uint64_t *deferred;
  unsigned long long *n_deferred;
  uint64_t deferred_cap;
  unsigned long long *deferred_need;
  unsigned int *err;
};

hipError_t launch_mce(const KArgs &a, uint32_t waves, hipStream_t stream);

}  // namespace rdmce
