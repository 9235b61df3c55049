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
constexpr uint32_t kIdleCopies = 64;  // copies of the idle count, one per 128-byte line
constexpr uint32_t kIdleStride = 32;  // u32 between copies
constexpr uint32_t kPollSteps = 128;  // descents (tree nodes) between polls for idle waves (a power of two)

// Error flags raised by the kernels (bit mask in KArgs::err).
constexpr uint32_t kErrPTooLarge = 1u;       // a root has |P| > kMaxP
constexpr uint32_t kErrHistOverflow = 2u;    // clique larger than hist_len - 1
constexpr uint32_t kErrDeferFull = 4u;       // deferred-root list full
constexpr uint32_t kErrListFull = 8u;        // clique list capacity exceeded
constexpr uint32_t kErrPackStack = 16u;      // packed search: record stack full

// One root's subproblem, as every wavefront working on it sees it.  The
// pointers are HBM copies (the owner may work on an LDS copy of bm and gP).
struct Sub {
  uint32_t p, xq, W, Wx, rows_pad, LS;
  uint32_t base_size, v0, v1, bm_owner;
  uint32_t levels, graph;  // graph: index within a batch (0 outside batches)
  uint64_t base_key, canon;
  const uint32_t *Plist;  // p entries, P0 ascending
  const uint64_t *bm;     // word-major bitmap, W * rows_pad words
  const uint64_t *gP;     // splitmix64(label) of each P0 vertex, p words
};

// Exposure record of one wavefront: at most one published level at a time.
// ctl = epoch << 32 | ntot << 16 | next; a claim is fetch_add(ctl, 1) and
// succeeds when next < ntot (task index = next).
struct alignas(128) ShareRec {
  unsigned long long ctl;
  unsigned int pending;   // published tasks not yet materialized
  unsigned int epoch;     // owner's copy of the epoch
  const uint64_t *snap;   // frozen level record in the owner's stack
  Sub sub;
};

struct KArgs {
  // Graph (CSR in HBM).
  const uint64_t *rowoff;
  const uint64_t *outoff;       // per vertex: index in col of its first higher neighbour
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
  const uint8_t *eager;         // per dispatch index: expose the top level at once
  uint64_t n_dispatch;
  int32_t reverse;
  int32_t root_mode;            // 1 vertex, 2 edge
  // Results.
  unsigned long long *count;
  unsigned long long *hash;
  unsigned long long *hist;     // hist_len entries
  uint32_t hist_len;
  unsigned long long *total_nodes;
  unsigned long long *search_words;  // bitmap words read by the search
  unsigned long long *build_bytes;   // bytes moved building subproblems
  unsigned long long *root_nodes;   // per canonical root, nullable
  unsigned long long *root_cycles;  // per canonical root, nullable
  uint32_t *list;                   // nullable
  unsigned long long *list_cursor;
  uint64_t list_cap;
  // Work distribution.
  unsigned long long *next_root;
  // Per-wave HBM scratch: data (candidate lists + bitmap) and DFS stack.
  uint8_t *slots;
  uint64_t slot_bytes;
  uint8_t *stacks;
  uint64_t stack_bytes;
  // Sharing (root-down exposure + stealing).
  ShareRec *recs;                   // one per wave
  unsigned int *users;              // per wave: other waves reading its data region
  unsigned long long *has_work;     // bit per wave: record has unclaimed tasks
  unsigned int *active;             // waves running a root or a task
  unsigned int *idle_hint;          // kIdleCopies copies (kIdleStride apart) of the number of waves looking for work
  unsigned long long *shared_tasks; // tasks taken by a wave other than the publisher
  unsigned long long *published_levels;  // levels exposed
  unsigned long long *owner_tasks;       // tasks of exposed levels run by their publisher
  int32_t share_bound;              // |P| threshold for exposing a level; <= 0: off
  uint32_t n_waves;
  // LDS tier: subproblems whose bitmap, label hashes and stack fit in
  // lds_bytes are searched entirely in the wave's LDS.
  int32_t lds_tier;
  uint32_t lds_bytes;
  // Register paths for roots with |P0|, |X0'| <= 64 or <= 128.
  int32_t word_path;
  // Roots whose scratch need exceeds the slots are deferred to a pass with
  // larger slots: canonical ids, count, and the largest needs seen.
  uint64_t *deferred;
  unsigned long long *n_deferred;
  uint64_t deferred_cap;
  unsigned long long *deferred_need;        // data region bytes
  unsigned long long *deferred_stack_need;  // stack bytes
  unsigned int *err;
  // Batches: graph index of each canonical root (nullable) and per-graph
  // results (count, hash, hist with `g_hist_stride` entries per graph).
  const uint32_t *root_graph;
  unsigned long long *g_count;
  unsigned long long *g_hash;
  unsigned long long *g_hist;
  uint32_t g_hist_stride;
  // Dispatch count read by the kernel at launch instead of n_dispatch
  // (nullable): the length of a root list that an earlier launch of the same
  // run appended to.
  const unsigned long long *n_dispatch_dev;
  // Packed search (pack_kernel.hip): per-wave record stack of pk_cap records,
  // pk_slots subproblems resident per wave in LDS, and the roots too wide for
  // the launched variant, appended in the order met.
  uint64_t *pk_stack;
  uint64_t pk_cap;
  uint32_t pk_slots;
  uint32_t pk_thr;                  // stack depth above which only the top record is expanded
  uint32_t *pk_plist;               // per wave pk_slots * kPackMaxP P0 vertex ids (clique lists)
  uint64_t *wide;
  unsigned long long *n_wide;
  // Vertex bitmaps kept for the next launch (nullable): a run of edge roots
  // that met wide roots copies its vertex bitmap to pk_pool (pk_pool_words
  // words; pk_pool_used taken, may overshoot) and names the copy's word
  // offset (~0: none) in wide_h beside each of those roots; the next launch
  // reads the same list as pk_whandle, one entry per dispatch entry.
  uint64_t *wide_h;
  const uint64_t *pk_whandle;
  uint64_t *pk_pool;
  uint64_t pk_pool_words;
  unsigned long long *pk_pool_used;
  // Edge roots in runs sharing their first vertex: run r covers dispatch
  // entries [pk_gstart[r], pk_gstart[r + 1]) (nullable: single entries);
  // rfirst[v] = canonical id of the first edge root of vertex v.
  const uint64_t *pk_gstart;
  uint64_t pk_ngroups;
  const uint64_t *rfirst;
  // Packed search: dispatch entries flagged in `eager` (roots predicted
  // heavy) are appended here for the one-root-per-wave kernel, which splits
  // them across waves.
  uint64_t *heavy;
  unsigned long long *n_heavy;
};

// Packed search variants: |P0| <= 64 NW and |X0'| <= 64 NX.
constexpr uint32_t kPackMaxP = 128;
constexpr uint32_t kPackMaxSlots = 8;
// Record words of variant (nw, nx): P XP V R (nw each), XX (nx), key, meta.
constexpr uint32_t pack_rec_words(uint32_t nw, uint32_t nx) { return 4 * nw + nx + 2; }
// LDS words of one resident subproblem: rows of P0 and X0' over P0 (nw words
// each), columns of the X0' block per P0 vertex (nx words), label hashes.
constexpr uint32_t pack_slot_words(uint32_t nw, uint32_t nx) { return 64 * (nw * nw + 2 * nw * nx + nw); }
// Clique sizes below this bound are counted per resident subproblem (batches).
constexpr uint32_t pack_hist_words(uint32_t nw) { return 64 * nw + 3; }
// Per-slot bookkeeping words (canon, key, p|xq, base|graph, v0|v1, nodes,
// refs, count, hash).
constexpr uint32_t kPackMetaWords = 9;
// LDS bytes of a wave of the packed search.
constexpr uint64_t pack_lds_bytes(uint32_t nw, uint32_t nx, uint32_t slots, bool batch) {
  return 8ull * (kLdsHist + 32 + slots * (kPackMetaWords + pack_slot_words(nw, nx) + (batch ? pack_hist_words(nw) : 0)));
}

// waves_per_cu selects the kernel variant compiled for that occupancy
// (<= 16: 4 waves per SIMD, <= 20: 5, <= 24: 6, else 8).
hipError_t launch_mce(const KArgs &a, uint32_t waves, uint32_t waves_per_cu, hipStream_t stream);
// Packed search: variant 0 = (NW, NX) = (1, 1), 1 = (2, 2); waves_per_cu
// selects the kernel compiled for that occupancy (<= 16: 4 waves per SIMD,
// <= 20: 5, <= 24: 6, else 8).
hipError_t launch_pack(const KArgs &a, int variant, uint32_t waves, uint32_t waves_per_cu, hipStream_t stream);
hipError_t launch_copy_probe(const void *src, void *dst, uint64_t bytes, hipStream_t stream);
// Zero-fills n 64-bit words with a kernel (safe inside captured hipGraphs,
// unlike memset nodes).
hipError_t launch_fill_zero(void *p, uint64_t n_words, hipStream_t stream);

}  // namespace rdmce
