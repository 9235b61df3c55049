// Maximal clique enumeration kernel for gfx950 (wave64).
//
// One wavefront owns one root at a time (persistent grid, roots pulled from
// a global atomic counter - RDMCE/src/kernels/BkPivotBitBalance.cuh:405-412,
// 510-542).  For a root it builds the subproblem and runs the bitmap
// Bron-Kerbosch search with pivoting and aggressive node generation
// (BkPivotBitBalance.cuh:72-403), with these MI355X-specific choices:
//
//  * Bitmap layout is word-major: word k of row u lives at
//    bm[k * rows_pad + u], so the lane-per-row scans of a node read 64
//    consecutive 8-byte words per instruction (one 512 B segment).
//  * Candidate sets are explicit.  Besides R and P (bit j = j-th vertex of
//    P0, the root's initial candidate list) each level keeps XP (excluded
//    vertices of P0) and XX (excluded rows of X0', one bit per row).  The
//    union P | XP | XX is exactly the set of rows u with R a subset of N(u)
//    that the reference rescans over all rows at every node, so the pivot
//    scan of a node touches only its candidates, and the leaf test is
//    "XP and XX empty".
//  * Words of R, P, XP and V sit one per lane in VGPRs (|P| <= 4096, at most
//    64 words); a scan broadcasts them with v_readlane (SGPR operands).
//  * Pivot ties are broken by the lowest row index (the reference breaks
//    them by lane, a function of warp width), so the search tree and the
//    per-root node counts are properties of the graph alone.
//  * Scratch is sized per root from its actual p and x'; a root that does
//    not fit the wave's slot is deferred to a second pass with larger slots
//    instead of being dropped (the reference drops it,
//    BkPivotBitBalance.cuh:480-487).
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

__device__ __forceinline__ uint64_t wave_min_u64(uint64_t v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    uint64_t o = __shfl_xor(v, off, 64);
    v = o < v ? o : v;
  }
  return v;
}
__device__ __forceinline__ uint32_t wave_sum_u32(uint32_t v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}
__device__ __forceinline__ uint32_t wave_max_u32(uint32_t v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    uint32_t o = __shfl_xor(v, off, 64);
    v = o > v ? o : v;
  }
  return v;
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

// Calls on_hit(b) for every index b (ascending) of the sorted list P[0..p)
// with P[b] in N(u); stops early when on_hit returns true.  Binary search of
// P into N(u) when N(u) is much longer, else a linear merge.
template <class F>
__device__ __forceinline__ void for_each_p_in_nu(const KArgs &a, uint32_t u, const uint32_t *P, uint32_t p, F on_hit) {
  const uint32_t *nu = a.col + a.rowoff[u];
  uint64_t du = a.rowoff[u + 1] - a.rowoff[u];
  uint64_t pos = lower_bound(nu, 0, du, P[0]);
  if ((uint64_t)p * 16 < du - pos) {
    for (uint32_t b = 0; b < p && pos < du; ++b) {
      uint32_t key = P[b];
      pos = lower_bound(nu, pos, du, key);
      if (pos < du && nu[pos] == key && on_hit(b)) return;
    }
  } else {
    uint32_t b = 0;
    while (pos < du && b < p) {
      uint32_t x = nu[pos], y = P[b];
      if (x < y) ++pos;
      else if (x > y) ++b;
      else {
        if (on_hit(b)) return;
        ++pos;
        ++b;
      }
    }
  }
}

struct WaveAcc {
  unsigned long long count;
  unsigned long long hash;
};

__device__ __forceinline__ void emit(const KArgs &a, WaveAcc &acc, unsigned long long *s_hist, uint32_t size,
                                     uint64_t key, uint64_t rk, uint32_t W, const uint32_t *Plist, uint32_t v0,
                                     uint32_t v1, uint32_t base_size) {
  if (lane() == 0) {
    acc.count++;
    acc.hash += splitmix64_mix(key);
    if (size < kLdsHist) s_hist[size]++;
    else if (size < a.hist_len) atomicAdd(&a.hist[size], 1ull);
    else {
      atomicAdd(&a.hist[a.hist_len - 1], 1ull);
      atomicOr(a.err, kErrHistOverflow);
    }
  }
  if (a.list) {
    if (lane() == 0) {
      unsigned long long pos = atomicAdd(a.list_cursor, (unsigned long long)(size + 1));
      if (pos + size + 1 > a.list_cap) {
        atomicOr(a.err, kErrListFull);
      } else {
        uint32_t *dst = a.list + pos;
        uint32_t k = 0;
        dst[k++] = size;
        dst[k++] = a.labels[v0];
        if (base_size == 2) dst[k++] = a.labels[v1];
        for (uint32_t w = 0; w < W; ++w) {
          uint64_t bits = readlane64(rk, w);
          while (bits) {
            uint32_t b = __builtin_ctzll(bits);
            bits &= bits - 1;
            dst[k++] = a.labels[Plist[w * 64 + b]];
          }
        }
      }
    }
  }
}

// Level record layout (u64 words): R[W] P[W] XP[W] V[W] XX[Wx] key meta1 meta2 pad
//   meta1 = rsize | xxcnt << 32, meta2 = xlo | xhi << 32.
struct Sub {
  uint32_t p, xq, W, Wx, rows_pad, LS;
  const uint32_t *Plist;  // p entries (P0 ascending)
  uint64_t *bm;
  uint64_t *stk;
};

// Depth-first search of one root.  Returns the number of tree nodes.
__device__ uint64_t search(const KArgs &a, const Sub &s, WaveAcc &acc, unsigned long long *s_hist, uint64_t base_key,
                           uint32_t base_size, uint32_t v0, uint32_t v1) {
  const uint32_t W = s.W, p = s.p, rp = s.rows_pad, LS = s.LS;
  const uint32_t l = lane();
  const bool wl = l < W;
  const uint64_t *bm = s.bm;
  uint64_t nodes = 0;

  // Level 0: R = {}, P = P0, XP = {}, XX = X0'.
  uint64_t rk = 0, pk = wl ? low_mask((uint64_t)p - (uint64_t)l * 64) : 0, xpk = 0, vk = 0;
  if (wl && (uint64_t)l * 64 >= p) pk = 0;
  uint64_t key = base_key;
  uint32_t rsize = 0, xxcnt = s.xq, xlo = 0, xhi = s.Wx;
  {
    uint64_t *rec = s.stk;
    for (uint32_t j = l; j < s.Wx; j += 64) rec[4 * W + j] = low_mask((uint64_t)s.xq - (uint64_t)j * 64);
    if (wl) {
      rec[l] = rk;
      rec[W + l] = pk;
      rec[2 * W + l] = xpk;
    }
    wave_fence();
  }
  int level = 0;
  // state machine: 0 = evaluate node, 1 = next child, 2 = backtrack
  int st = 0;
  for (;;) {
    uint64_t *rec = s.stk + (uint64_t)level * LS;
    uint64_t *XX = rec + 4 * W;
    if (st == 0) {
      ++nodes;
      uint32_t psz = wave_sum_u32((uint32_t)__popcll(pk));
      if (psz == 0) {
        if (__ballot(xpk != 0) == 0 && xxcnt == 0)
          emit(a, acc, s_hist, base_size + rsize, key, rk, W, s.Plist, v0, v1, base_size);
        st = 2;
        continue;
      }
      // Pivot: argmin over candidate rows u of (|P \ N(u)|, u).
      uint64_t best = ~0ull;
      for (uint32_t j = 0; j < W; ++j) {
        uint64_t cw = readlane64(pk | xpk, j);
        if (cw == 0) continue;
        if ((cw >> l) & 1) {
          uint32_t u = j * 64 + l;
          uint32_t cnt = 0;
          for (uint32_t k = 0; k < W; ++k) cnt += __popcll(readlane64(pk, k) & ~bm[(uint64_t)k * rp + u]);
          uint64_t kk = ((uint64_t)cnt << 32) | u;
          best = kk < best ? kk : best;
        }
      }
      for (uint32_t j = xlo; j < xhi; ++j) {
        uint64_t cw = XX[j];
        if (cw == 0) continue;
        if ((cw >> l) & 1) {
          uint32_t u = p + j * 64 + l;
          uint32_t cnt = 0;
          for (uint32_t k = 0; k < W; ++k) cnt += __popcll(readlane64(pk, k) & ~bm[(uint64_t)k * rp + u]);
          uint64_t kk = ((uint64_t)cnt << 32) | u;
          best = kk < best ? kk : best;
        }
      }
      best = wave_min_u64(best);
      if ((best >> 32) == 0) {  // an excluded row covers P: no maximal clique below
        st = 2;
        continue;
      }
      uint32_t piv = (uint32_t)best;
      vk = wl ? (pk & ~bm[(uint64_t)l * rp + piv]) : 0;
      if (wl) rec[3 * W + l] = vk;
      st = 1;
      continue;
    }
    if (st == 1) {
      uint64_t vm = __ballot(vk != 0);
      if (vm == 0) {
        st = 2;
        continue;
      }
      uint32_t kc = (uint32_t)__builtin_ctzll(vm);
      uint32_t b = (uint32_t)__builtin_ctzll(readlane64(vk, kc));
      uint32_t c = kc * 64 + b;
      uint64_t cbit = 1ull << b;
      if (l == kc) {
        vk &= ~cbit;
        pk &= ~cbit;
        xpk |= cbit;
        rec[W + l] = pk;
        rec[2 * W + l] = xpk;
        rec[3 * W + l] = vk;
      }
      // Child: R' = R + c, P' = P & N(c), XP' = XP & N(c).
      uint64_t rowc = wl ? bm[(uint64_t)l * rp + c] : 0;
      uint64_t r2 = rk | (l == kc ? cbit : 0);
      uint64_t p2 = pk & rowc;
      uint64_t x2 = xpk & rowc;
      uint64_t key2 = key + a.label_hash[s.Plist[c]];
      uint32_t rs2 = rsize + 1;
      // Aggressive node generation: every i in P' adjacent to all of P' \ {i}
      // moves to R' (BkPivotBitBalance.cuh:351-382).
      uint64_t mk = 0;
      uint32_t psz2 = wave_sum_u32((uint32_t)__popcll(p2));
      if (psz2 > 1) {
        for (uint32_t j = 0; j < W; ++j) {
          uint64_t w = readlane64(p2, j);
          if (w == 0) continue;
          bool is = (w >> l) & 1;
          uint32_t cnt = 0;
          if (is) {
            uint32_t i = j * 64 + l;
            for (uint32_t k = 0; k < W; ++k) cnt += __popcll(readlane64(p2, k) & ~bm[(uint64_t)k * rp + i]);
          }
          uint64_t m = __ballot(is && cnt == 1);
          if (l == j) mk = m;
        }
        if (__ballot(mk != 0)) {
          r2 |= mk;
          p2 &= ~mk;
          for (uint32_t j = 0; j < W; ++j) {
            uint64_t w = readlane64(mk, j);
            while (w) {
              uint32_t bb = (uint32_t)__builtin_ctzll(w);
              w &= w - 1;
              uint32_t mm = j * 64 + bb;
              if (wl) x2 &= bm[(uint64_t)l * rp + mm];
              key2 += a.label_hash[s.Plist[mm]];
              rs2++;
            }
          }
        }
      }
      bool anym = __ballot(mk != 0) != 0;
      // XX' = rows x of XX adjacent to c and to every moved vertex.
      uint64_t *rec2 = rec + LS;
      uint64_t *XX2 = rec2 + 4 * W;
      uint32_t cw = c >> 6, cb = c & 63;
      uint32_t cnt2 = 0, nlo = xhi, nhi = xlo;
      for (uint32_t j = xlo; j < xhi; ++j) {
        uint64_t w = XX[j];
        uint64_t nw = 0;
        if (w != 0) {
          bool hit = false;
          if ((w >> l) & 1) {
            uint32_t u = p + j * 64 + l;
            hit = (bm[(uint64_t)cw * rp + u] >> cb) & 1;
            if (hit && anym) {
              for (uint32_t k = 0; k < W && hit; ++k) hit = (readlane64(mk, k) & ~bm[(uint64_t)k * rp + u]) == 0;
            }
          }
          nw = __ballot(hit);
        }
        if (l == 0) XX2[j] = nw;
        if (nw) {
          cnt2 += __popcll(nw);
          nlo = j < nlo ? j : nlo;
          nhi = j + 1;
        }
      }
      if (cnt2 == 0) nlo = nhi = 0;
      if (wl) {
        rec2[l] = r2;
        rec2[W + l] = p2;
        rec2[2 * W + l] = x2;
      }
      if (l == 0) {
        uint64_t *meta = rec2 + 4 * W + s.Wx;
        meta[0] = key2;
        meta[1] = (uint64_t)rs2 | ((uint64_t)cnt2 << 32);
        meta[2] = (uint64_t)nlo | ((uint64_t)nhi << 32);
      }
      wave_fence();
      ++level;
      rk = r2;
      pk = p2;
      xpk = x2;
      vk = 0;
      key = key2;
      rsize = rs2;
      xxcnt = cnt2;
      xlo = nlo;
      xhi = nhi;
      st = 0;
      continue;
    }
    // st == 2: backtrack to the parent level.
    if (level == 0) break;
    --level;
    rec = s.stk + (uint64_t)level * LS;
    if (wl) {
      rk = rec[l];
      pk = rec[W + l];
      xpk = rec[2 * W + l];
      vk = rec[3 * W + l];
    } else {
      rk = pk = xpk = vk = 0;
    }
    if (level == 0) {
      key = base_key;
      rsize = 0;
      xxcnt = s.xq;
      xlo = 0;
      xhi = s.Wx;
    } else {
      const uint64_t *meta = rec + 4 * W + s.Wx;
      key = meta[0];
      uint64_t m1 = meta[1], m2 = meta[2];
      rsize = (uint32_t)m1;
      xxcnt = (uint32_t)(m1 >> 32);
      xlo = (uint32_t)m2;
      xhi = (uint32_t)(m2 >> 32);
    }
    st = 1;
  }
  return nodes;
}

__device__ void defer_root(const KArgs &a, uint64_t canon, uint64_t need) {
  if (lane() == 0) {
    unsigned long long pos = atomicAdd(a.n_deferred, 1ull);
    if (pos < a.deferred_cap) a.deferred[pos] = canon;
    else atomicOr(a.err, kErrDeferFull);
    atomicMax(a.deferred_need, (unsigned long long)need);
  }
}

// Scratch need of a root before its bitmap is built (x instead of x', p+1
// levels): an upper bound on what it needs in any pass.
__device__ __forceinline__ uint64_t need_bound(uint64_t c, uint64_t p, uint64_t x) {
  uint64_t W = (p + 63) / 64, Wx = (x + 63) / 64;
  uint64_t bm_off = align_up(4 * c, 256);
  uint64_t st_off = bm_off + align_up(8 * W * align_up(p + x, 64), 256);
  return st_off + (p + 1) * (4 * W + Wx + 4) * 8;
}

// Builds and searches one root.  Returns tree nodes.
__device__ uint64_t solve_root(const KArgs &a, uint8_t *slot, uint64_t canon, WaveAcc &acc,
                               unsigned long long *s_hist) {
  const uint32_t l = lane();
  const uint64_t desc = a.roots[canon];
  uint32_t *C = (uint32_t *)slot;
  const uint64_t ccap = a.slot_bytes / 4;
  uint32_t v0, v1 = 0, base_size;
  uint64_t base_key;
  uint64_t c = 0, x = 0;
  const uint32_t *Plist;
  if (a.root_mode == 2) {
    v0 = (uint32_t)(desc >> 32);
    v1 = (uint32_t)desc;
    base_size = 2;
    base_key = a.label_hash[v0] + a.label_hash[v1];
    // C = N(v0) ^ N(v1): lanes take 64 entries of the shorter list at a time
    // and binary-search them in the longer one (BkPivotBitBalance.cuh:449-460).
    const uint32_t *A = a.col + a.rowoff[v0], *B = a.col + a.rowoff[v1];
    uint64_t na = a.rowoff[v0 + 1] - a.rowoff[v0], nb = a.rowoff[v1 + 1] - a.rowoff[v1];
    if (na > nb) {
      const uint32_t *t = A;
      A = B;
      B = t;
      uint64_t tn = na;
      na = nb;
      nb = tn;
    }
    for (uint64_t base = 0; base < na; base += 64) {
      uint64_t i = base + l;
      uint32_t val = 0;
      bool found = false;
      if (i < na) {
        val = A[i];
        uint64_t pos = lower_bound(B, 0, nb, val);
        found = pos < nb && B[pos] == val;
      }
      uint64_t m = __ballot(found);
      if (found) {
        uint64_t at = c + rank_below(m);
        if (at < ccap) C[at] = val;
      }
      c += __popcll(m);
      x += __popcll(__ballot(found && val < v1));
    }
  } else {
    v0 = (uint32_t)desc;
    base_size = 1;
    base_key = a.label_hash[v0];
    const uint32_t *N = a.col + a.rowoff[v0];
    c = a.rowoff[v0 + 1] - a.rowoff[v0];
    x = lower_bound(N, 0, c, v0);
    for (uint64_t i = l; i < c && i < ccap; i += 64) C[i] = N[i];
  }
  const uint64_t p = c - x;
  if (p == 0) {
    if (x == 0) emit(a, acc, s_hist, base_size, base_key, 0, 0, nullptr, v0, v1, base_size);
    return 0;
  }
  if (p > kMaxP) {
    if (l == 0) atomicOr(a.err, kErrPTooLarge);
    return 0;
  }
  if (c > ccap) {  // candidate list alone exceeds the slot
    defer_root(a, canon, need_bound(c, p, x));
    return 0;
  }
  wave_fence();
  Plist = C + x;
  // X0' = X0 rows with at least one neighbour in P0, compacted in place
  // (graph_gpu.cuh:90-106).
  uint64_t xq = 0;
  for (uint64_t base = 0; base < x; base += 64) {
    uint64_t i = base + l;
    bool keep = false;
    uint32_t xv = 0;
    if (i < x) {
      xv = C[i];
      for_each_p_in_nu(a, xv, Plist, (uint32_t)p, [&](uint32_t) {
        keep = true;
        return true;
      });
    }
    uint64_t m = __ballot(keep);
    if (keep) C[xq + rank_below(m)] = xv;
    xq += __popcll(m);
  }
  wave_fence();
  const uint32_t W = (uint32_t)((p + 63) / 64);
  const uint32_t Wx = (uint32_t)((xq + 63) / 64);
  const uint64_t rows = p + xq;
  const uint32_t rows_pad = (uint32_t)align_up(rows, 64);
  const uint64_t bm_off = align_up(4 * c, 256);
  const uint64_t st_off = bm_off + align_up(8ull * W * rows_pad, 256);
  const uint32_t LS = 4 * W + Wx + 4;
  if (st_off + (p + 1) * LS * 8 > a.slot_bytes) {
    // Bitmap too large for the slot, or the stack bound p + 1 levels is:
    // build it only when the bitmap fits (the exact depth bound follows).
    if (st_off > a.slot_bytes) {
      defer_root(a, canon, st_off + (p + 1) * LS * 8);
      return 0;
    }
  }
  uint64_t *bm = (uint64_t *)(slot + bm_off);
  // Bitmap (word-major): row i is P0[i] for i < p, else X0'[i - p].
  for (uint64_t i = l; i < (uint64_t)W * rows_pad; i += 64) bm[i] = 0;
  wave_fence();
  uint32_t maxdeg = 0;
  for (uint64_t i = l; i < rows; i += 64) {
    uint32_t u = i < p ? Plist[i] : C[i - p];
    uint32_t curw = ~0u, deg = 0;
    uint64_t accw = 0;
    for_each_p_in_nu(a, u, Plist, (uint32_t)p, [&](uint32_t b) {
      uint32_t w = b >> 6;
      if (w != curw) {
        if (curw != ~0u) bm[(uint64_t)curw * rows_pad + i] = accw;
        curw = w;
        accw = 0;
      }
      accw |= 1ull << (b & 63);
      ++deg;
      return false;
    });
    if (curw != ~0u) bm[(uint64_t)curw * rows_pad + i] = accw;
    if (i < p) maxdeg = deg > maxdeg ? deg : maxdeg;
  }
  maxdeg = wave_max_u32(maxdeg);
  wave_fence();
  const uint64_t levels = (p < (uint64_t)maxdeg + 1 ? p : (uint64_t)maxdeg + 1) + 1;
  const uint64_t need = st_off + levels * LS * 8;
  if (need > a.slot_bytes) {
    defer_root(a, canon, need);
    return 0;
  }
  Sub s;
  s.p = (uint32_t)p;
  s.xq = (uint32_t)xq;
  s.W = W;
  s.Wx = Wx;
  s.rows_pad = rows_pad;
  s.LS = LS;
  s.Plist = Plist;
  s.bm = bm;
  s.stk = (uint64_t *)(slot + st_off);
  return search(a, s, acc, s_hist, base_key, base_size, v0, v1);
}

__global__ __launch_bounds__(64) void mce_kernel(KArgs a) {
  __shared__ unsigned long long s_hist[kLdsHist];
  const uint32_t l = lane();
  for (uint32_t i = l; i < kLdsHist; i += 64) s_hist[i] = 0;
  __syncthreads();
  const uint64_t wave = blockIdx.x;
  uint8_t *slot = a.slots + wave * a.slot_bytes;
  WaveAcc acc{0, 0};
  unsigned long long nodes_total = 0;
  for (;;) {
    unsigned long long i = 0;
    if (l == 0) i = atomicAdd(a.next_root, 1ull);
    i = ((unsigned long long)__builtin_amdgcn_readfirstlane((uint32_t)(i >> 32)) << 32) |
        __builtin_amdgcn_readfirstlane((uint32_t)i);
    if (i >= a.n_dispatch) break;
    uint64_t canon = a.perm ? a.perm[i] : (a.reverse ? a.n_dispatch - 1 - i : i);
    uint64_t t0 = clock64();
    uint64_t nodes = solve_root(a, slot, canon, acc, s_hist);
    nodes_total += nodes;
    if (l == 0) {
      if (a.root_nodes && nodes) atomicAdd(&a.root_nodes[canon], (unsigned long long)nodes);
      if (a.root_cycles) atomicAdd(&a.root_cycles[canon], (unsigned long long)(clock64() - t0));
    }
  }
  __syncthreads();
  if (l == 0) {
    if (acc.count) atomicAdd(a.count, acc.count);
    if (acc.hash) atomicAdd(a.hash, acc.hash);
    if (nodes_total) atomicAdd(a.total_nodes, nodes_total);
  }
  for (uint32_t i = l; i < kLdsHist && i < a.hist_len; i += 64)
    if (s_hist[i]) atomicAdd(&a.hist[i], s_hist[i]);
}

}  // namespace

hipError_t launch_mce(const KArgs &a, uint32_t waves, hipStream_t stream) {
  hipLaunchKernelGGL(mce_kernel, dim3(waves), dim3(64), 0, stream, a);
  return hipGetLastError();
}

}  // namespace rdmce
