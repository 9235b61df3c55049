// Device helpers shared by the enumeration kernels (mce_kernel.hip, the
// one-wave-per-root search, and pack_kernel.hip, the packed search): wave64
// cross-lane operations, sorted-list searches, the candidate lists of a root
// and the matching of bitmap rows against adjacency lists.
#pragma once
#include "mce_device.hpp"
#include "rng.hpp"

namespace rdmce {
namespace {

__device__ __forceinline__ uint32_t lane() { return __lane_id(); }

__device__ __forceinline__ uint64_t readlane64(uint64_t v, uint32_t l) {
  uint32_t lo = __builtin_amdgcn_readlane((uint32_t)v, l);
  uint32_t hi = __builtin_amdgcn_readlane((uint32_t)(v >> 32), l);
  return ((uint64_t)hi << 32) | lo;
}

// Wave-wide reductions with DPP (row_shr 1/2/4/8 inside each 16-lane row,
// then row_bcast 15/31 across rows); the result is read from lane 63.
__device__ __forceinline__ uint32_t wave_min_u32(uint32_t v) {
  const uint32_t id = 0xFFFFFFFFu;
  uint32_t t;
  t = __builtin_amdgcn_update_dpp(id, v, 0x111, 0xF, 0xF, false); v = v < t ? v : t;
  t = __builtin_amdgcn_update_dpp(id, v, 0x112, 0xF, 0xF, false); v = v < t ? v : t;
  t = __builtin_amdgcn_update_dpp(id, v, 0x114, 0xF, 0xF, false); v = v < t ? v : t;
  t = __builtin_amdgcn_update_dpp(id, v, 0x118, 0xF, 0xF, false); v = v < t ? v : t;
  t = __builtin_amdgcn_update_dpp(id, v, 0x142, 0xA, 0xF, false); v = v < t ? v : t;
  t = __builtin_amdgcn_update_dpp(id, v, 0x143, 0xC, 0xF, false); v = v < t ? v : t;
  return __builtin_amdgcn_readlane(v, 63);
}
__device__ __forceinline__ uint32_t wave_max_u32(uint32_t v) {
  uint32_t t;
  t = __builtin_amdgcn_update_dpp(0u, v, 0x111, 0xF, 0xF, false); v = v > t ? v : t;
  t = __builtin_amdgcn_update_dpp(0u, v, 0x112, 0xF, 0xF, false); v = v > t ? v : t;
  t = __builtin_amdgcn_update_dpp(0u, v, 0x114, 0xF, 0xF, false); v = v > t ? v : t;
  t = __builtin_amdgcn_update_dpp(0u, v, 0x118, 0xF, 0xF, false); v = v > t ? v : t;
  t = __builtin_amdgcn_update_dpp(0u, v, 0x142, 0xA, 0xF, false); v = v > t ? v : t;
  t = __builtin_amdgcn_update_dpp(0u, v, 0x143, 0xC, 0xF, false); v = v > t ? v : t;
  return __builtin_amdgcn_readlane(v, 63);
}
// Population count of a set whose words are held one per lane (lanes < W).
__device__ __forceinline__ uint32_t set_popc(uint64_t wk, uint32_t W) {
  uint32_t s = 0;
  for (uint32_t k = 0; k < W; ++k) s += (uint32_t)__popcll(readlane64(wk, k));
  return s;
}
// Number of set bits of `mask` below this lane.
__device__ __forceinline__ uint32_t rank_below(uint64_t mask) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}
// Orders this wave's earlier global stores before its later loads.
__device__ __forceinline__ void wave_fence() { __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup"); }

__device__ __forceinline__ uint64_t lower_bound(const uint32_t *a, uint64_t lo, uint64_t hi, uint32_t key) {
  while (lo < hi) {
    uint64_t mid = (lo + hi) >> 1;
    if (a[mid] < key) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

__device__ __forceinline__ uint64_t low_mask(uint64_t bits) {  // bits in [0, 64]
  return bits >= 64 ? ~0ull : ((1ull << bits) - 1);
}
__device__ __forceinline__ uint64_t align_up(uint64_t x, uint64_t a) { return (x + a - 1) / a * a; }

// Calls on_hit(j) for every j in [j0, p), ascending, with P[j] in the sorted
// run col[b, e); stops early when on_hit returns true.  Binary search of the
// shorter side into the longer when their lengths differ by 8x or more, a
// linear merge otherwise.
template <class F>
__device__ __forceinline__ void match_run(const uint32_t *col, uint64_t b, uint64_t e, const uint32_t *P, uint32_t j0,
                                          uint32_t p, F on_hit) {
  if (b >= e || j0 >= p) return;
  const uint64_t n = e - b;
  const uint32_t m = p - j0;
  if ((uint64_t)m * 8 < n) {
    uint64_t pos = b;
    for (uint32_t j = j0; j < p && pos < e; ++j) {
      const uint32_t key = P[j];
      pos = lower_bound(col, pos, e, key);
      if (pos < e && col[pos] == key && on_hit(j)) return;
    }
  } else if (n * 8 < (uint64_t)m) {
    uint32_t lo = j0;
    for (uint64_t i = b; i < e && lo < p; ++i) {
      const uint32_t w = col[i];
      uint32_t x = lo, z = p;
      while (x < z) {
        uint32_t mid = (x + z) >> 1;
        if (P[mid] < w) x = mid + 1;
        else z = mid;
      }
      lo = x;
      if (x < p && P[x] == w) {
        if (on_hit(x)) return;
        lo = x + 1;
      }
    }
  } else {
    uint64_t i = b;
    uint32_t j = j0;
    while (i < e && j < p) {
      const uint32_t x = col[i], y = P[j];
      if (x < y) ++i;
      else if (x > y) ++j;
      else {
        if (on_hit(j)) return;
        ++i;
        ++j;
      }
    }
  }
}

// Relaxed 64-bit OR into a bitmap word (HBM or LDS through a flat pointer).
__device__ __forceinline__ void or64(uint64_t *p, uint64_t v) {
  __hip_atomic_fetch_or(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Per-wave accumulators (lane 0 flushes them at exit).  `words` counts the
// 64-bit bitmap words the search reads; `build` the bytes the subproblem
// build moves (both adjacency lists, the candidate list, the bitmap).
struct Acc {
  unsigned long long count;
  unsigned long long hash;
  unsigned long long words;
  unsigned long long build;
  unsigned long long scan;  // per lane: adjacency entries of the bitmap rows' runs
};

// Sum over the wave of a per-lane 64-bit value (uniform result).
__device__ __forceinline__ uint64_t wave_sum64(uint64_t v) {
  uint64_t s = 0;
  for (uint32_t k = 0; k < 64; ++k) s += readlane64(v, k);
  return s;
}

__device__ __forceinline__ unsigned int ld_relaxed(unsigned int *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ unsigned long long ld_relaxed(unsigned long long *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ unsigned int ld_acquire(unsigned int *p) {
  return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
}
// Sleeps about 256 * n cycles (s_sleep takes an immediate operand).
__device__ __forceinline__ void nap(uint32_t n) {
  for (uint32_t i = 0; i < n; ++i) __builtin_amdgcn_s_sleep(4);
}
__device__ __forceinline__ uint32_t bcast32(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ uint64_t bcast64(uint64_t v) {
  return ((uint64_t)__builtin_amdgcn_readfirstlane((uint32_t)(v >> 32)) << 32) |
         __builtin_amdgcn_readfirstlane((uint32_t)v);
}

// LDS pointers (address space 3): the LDS tiers and the per-wave clique-size
// histogram.
using lds64 = __attribute__((address_space(3))) uint64_t;
using ldsu64 = __attribute__((address_space(3))) unsigned long long;
using ldsu32 = __attribute__((address_space(3))) uint32_t;

// Inclusive prefix sum / maximum over the 64 lanes (DPP row_shr 1/2/4/8
// inside each 16-lane row, then row_bcast 15/31 across rows).
__device__ __forceinline__ uint32_t wave_scan_add_u32(uint32_t v) {
  uint32_t t;
  t = __builtin_amdgcn_update_dpp(0u, v, 0x111, 0xF, 0xF, false); v += t;
  t = __builtin_amdgcn_update_dpp(0u, v, 0x112, 0xF, 0xF, false); v += t;
  t = __builtin_amdgcn_update_dpp(0u, v, 0x114, 0xF, 0xF, false); v += t;
  t = __builtin_amdgcn_update_dpp(0u, v, 0x118, 0xF, 0xF, false); v += t;
  t = __builtin_amdgcn_update_dpp(0u, v, 0x142, 0xA, 0xF, false); v += t;
  t = __builtin_amdgcn_update_dpp(0u, v, 0x143, 0xC, 0xF, false); v += t;
  return v;
}
__device__ __forceinline__ uint32_t wave_scan_max_u32(uint32_t v) {
  uint32_t t;
  t = __builtin_amdgcn_update_dpp(0u, v, 0x111, 0xF, 0xF, false); v = v > t ? v : t;
  t = __builtin_amdgcn_update_dpp(0u, v, 0x112, 0xF, 0xF, false); v = v > t ? v : t;
  t = __builtin_amdgcn_update_dpp(0u, v, 0x114, 0xF, 0xF, false); v = v > t ? v : t;
  t = __builtin_amdgcn_update_dpp(0u, v, 0x118, 0xF, 0xF, false); v = v > t ? v : t;
  t = __builtin_amdgcn_update_dpp(0u, v, 0x142, 0xA, 0xF, false); v = v > t ? v : t;
  t = __builtin_amdgcn_update_dpp(0u, v, 0x143, 0xC, 0xF, false); v = v > t ? v : t;
  return v;
}
// Relaxed 64-bit OR into an LDS word.
__device__ __forceinline__ void or64_lds(lds64 *p, uint64_t v) {
  __hip_atomic_fetch_or(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// Candidate lists of root `canon`: P ascending, then X ascending, written to
// Pl (capacity ccap entries; when p + x > ccap only the counts are exact).
// Sets s.v0, s.v1, s.base_size and s.base_key; returns `above` - every P
// vertex lies above it, every X vertex below it.  In degeneracy order every
// vertex has at most `degeneracy` higher-ordered neighbours (their run in
// col starts at outoff[v]), so edge roots intersect short out-lists:
//   P = N+(v1) & N+(v0) (all of P lies above v1 > v0),
//   X = N(v0) & N-(v1) below v1 (only its emptiness when P is empty).
__device__ __forceinline__ uint32_t root_lists(const KArgs &a, uint64_t canon, uint32_t *Pl, uint64_t ccap, Sub &s,
                                               uint64_t &p, uint64_t &x, Acc &acc) {
  const uint32_t l = lane();
  const uint64_t desc = a.roots[canon];
  const uint32_t *col = a.col;
  p = 0;
  x = 0;
  uint32_t above;
  if (a.root_mode == 2) {
    s.v0 = (uint32_t)(desc >> 32);
    s.v1 = (uint32_t)desc;
    s.base_size = 2;
    s.base_key = a.label_hash[s.v0] + a.label_hash[s.v1];
    above = s.v1;
    const uint64_t b0 = a.rowoff[s.v0], e0 = a.rowoff[s.v0 + 1], o0 = a.outoff[s.v0];
    const uint64_t b1 = a.rowoff[s.v1], e1 = a.rowoff[s.v1 + 1], o1 = a.outoff[s.v1];
    // P = N+(v1) & (N+(v0) above v1): iterate the shorter run, binary-search
    // the longer (BkPivotBitBalance.cuh:449-469 builds it from N(v0) & N(v1)).
    {
      const uint64_t s0 = lower_bound(col, o0, e0, s.v1 + 1);
      uint64_t ab = o1, ae = e1, bb = s0, be = e0;
      if (ae - ab > be - bb) {
        uint64_t t0 = ab, t1 = ae;
        ab = bb;
        ae = be;
        bb = t0;
        be = t1;
      }
      for (uint64_t base = ab; base < ae; base += 64) {
        const uint64_t i = base + l;
        uint32_t val = 0;
        bool found = false;
        if (i < ae) {
          val = col[i];
          const uint64_t pos = lower_bound(col, bb, be, val);
          found = pos < be && col[pos] == val;
        }
        const uint64_t m = __ballot(found);
        if (found && p + rank_below(m) < ccap) Pl[p + rank_below(m)] = val;
        p += __popcll(m);
      }
      acc.build += 4 * ((e1 - o1) + (e0 - s0));
    }
    // X = (N(v0) below v1) & N-(v1); with P empty only its emptiness matters.
    {
      uint64_t ab = b0, ae = lower_bound(col, b0, e0, s.v1), bb = b1, be = o1;
      if (ae - ab > be - bb) {
        uint64_t t0 = ab, t1 = ae;
        ab = bb;
        ae = be;
        bb = t0;
        be = t1;
      }
      acc.build += 4 * ((ae - ab) + (be - bb));
      for (uint64_t base = ab; base < ae; base += 64) {
        const uint64_t i = base + l;
        uint32_t val = 0;
        bool found = false;
        if (i < ae) {
          val = col[i];
          const uint64_t pos = lower_bound(col, bb, be, val);
          found = pos < be && col[pos] == val;
        }
        const uint64_t m = __ballot(found);
        if (found && p + x + rank_below(m) < ccap) Pl[p + x + rank_below(m)] = val;
        x += __popcll(m);
        if (p == 0 && x) break;
      }
    }
  } else {
    s.v0 = (uint32_t)desc;
    s.v1 = 0;
    s.base_size = 1;
    s.base_key = a.label_hash[s.v0];
    above = s.v0;
    const uint64_t b = a.rowoff[s.v0], e = a.rowoff[s.v0 + 1], o = a.outoff[s.v0];
    p = e - o;
    x = o - b;
    for (uint64_t i = l; i < p && i < ccap; i += 64) Pl[i] = col[o + i];
    for (uint64_t i = l; i < x && p + i < ccap; i += 64) Pl[p + i] = col[b + i];
    acc.build += 4 * (e - b);
  }
  return above;
}

// X' = the vertices of X with a neighbour in P (N+(x) above `above` meets
// P), compacted in place at the front of X; returns |X'|
// (graph_gpu.cuh:90-106 drops those rows after building them).
__device__ __forceinline__ uint64_t compact_x(const KArgs &a, const uint32_t *P, uint64_t p, uint32_t *X, uint64_t x,
                                              uint32_t above) {
  const uint32_t l = lane();
  const uint32_t *col = a.col;
  uint64_t xq = 0;
  for (uint64_t base = 0; base < x; base += 64) {
    const uint64_t i = base + l;
    bool keep = false;
    uint32_t xv = 0;
    if (i < x) {
      xv = X[i];
      const uint64_t e = a.rowoff[xv + 1];
      const uint64_t rb = lower_bound(col, a.outoff[xv], e, above + 1);
      match_run(col, rb, e, P, 0, (uint32_t)p, [&](uint32_t) {
        keep = true;
        return true;
      });
    }
    const uint64_t m = __ballot(keep);
    if (keep) X[xq + rank_below(m)] = xv;
    xq += __popcll(m);
  }
  return xq;
}

// Bitmap rows of a subproblem with rows P (ascending) then X' (rows
// p .. rows-1): for each P row i, every j > i with P[j] in N+(P[i]) calls
// pp(i, j) (the P x P block is symmetric: the callee sets both bits); for
// each X' row i, the words of N+(X'[i - p]) & P (above `above`) call
// xw(i, w, bits) once per non-empty word w.  One lane matches one row's
// adjacency run against P; with fewer than 64 rows, g = 2 or 4 lanes share a
// row, each matching one share of its P range (a row's cost is a chain of
// dependent loads into adjacency lists; splitting the range shortens the
// chain), so callees must set bits with atomic ORs.  With `claim` (an LDS
// word), lanes take the items after their first one from a shared counter,
// so a lane with short rows goes on to further rows.  Returns this lane's
// share of the adjacency entries in the runs matched (each row's run counted
// once, by its first lane).
template <class PP, class XW>
__device__ __forceinline__ uint64_t match_rows(const KArgs &a, const uint32_t *P, uint64_t p, const uint32_t *X,
                                               uint64_t rows, uint32_t above, PP pp, XW xw, ldsu32 *claim = nullptr) {
  const uint32_t l = lane();
  const uint32_t *col = a.col;
  uint64_t scanned = 0;
  const uint32_t g = rows <= 16 ? 4u : (rows <= 32 ? 2u : 1u);
  const uint32_t gs = g == 4 ? 2u : (g == 2 ? 1u : 0u);
  const uint64_t items = rows << gs;
  if (claim) {  // items past the first 64 go to whichever lane is free first
    if (l == 0) *claim = 64;
    wave_fence();
  }
  for (uint64_t t = l; t < items;
       t = claim ? (uint64_t)__hip_atomic_fetch_add(claim, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) : t + 64) {
    const uint64_t i = t >> gs;
    const uint32_t sh = (uint32_t)t & (g - 1);
    const uint32_t j0 = i < p ? (uint32_t)i + 1 : 0u;  // the row's P range [j0, p)
    const uint32_t q = ((uint32_t)p - j0 + g - 1) >> gs;
    const uint32_t jlo = j0 + sh * q;
    const uint32_t jhi = jlo + q < (uint32_t)p ? jlo + q : (uint32_t)p;
    if (jlo >= jhi) continue;
    if (i < p) {
      const uint32_t u = P[i];
      const uint64_t e = a.rowoff[u + 1];
      const uint64_t rb = sh ? lower_bound(col, a.outoff[u], e, P[jlo]) : a.outoff[u];
      if (!sh) scanned += e - rb;
      match_run(col, rb, e, P, jlo, jhi, [&](uint32_t j) {
        pp(i, j);
        return false;
      });
    } else {
      const uint32_t u = X[i - p];
      const uint64_t e = a.rowoff[u + 1];
      const uint64_t rb = lower_bound(col, a.outoff[u], e, sh ? P[jlo] : above + 1);
      if (!sh) scanned += e - rb;
      uint32_t curw = ~0u;
      uint64_t accw = 0;
      match_run(col, rb, e, P, jlo, jhi, [&](uint32_t b) {
        const uint32_t w = b >> 6;
        if (w != curw) {
          if (curw != ~0u) xw(i, curw, accw);
          curw = w;
          accw = 0;
        }
        accw |= 1ull << (b & 63);
        return false;
      });
      if (curw != ~0u) xw(i, curw, accw);
    }
  }
  return scanned;
}

}  // namespace
}  // namespace rdmce
