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
#include <type_traits>

#include "mce_common.hpp"

// Dynamic LDS of a wave (the launch gives its size).
extern __shared__ uint64_t lds_dyn[];

namespace rdmce {
namespace {

__device__ __forceinline__ void emit(const KArgs &a, const Sub &s, Acc &acc, ldsu64 *s_hist,
                                     uint32_t size, uint64_t key, uint64_t rk) {
  if (lane() == 0) {
    acc.count++;
    acc.hash += splitmix64_mix(key);
    if (size < kLdsHist) s_hist[size]++;
    else if (a.root_graph) {
      if (size < a.g_hist_stride) atomicAdd(&a.g_hist[(uint64_t)s.graph * a.g_hist_stride + size], 1ull);
      else atomicOr(a.err, kErrHistOverflow);
    } else if (size < a.hist_len) atomicAdd(&a.hist[size], 1ull);
    else {
      atomicAdd(&a.hist[a.hist_len - 1], 1ull);
      atomicOr(a.err, kErrHistOverflow);
    }
  }
  if (a.list) {
    // entry: size, the root's vertices, then R's members in P0 order; lane l
    // writes the member of bit l of each word of R
    const uint32_t l = lane();
    unsigned long long pos = 0;
    if (l == 0) pos = atomicAdd(a.list_cursor, (unsigned long long)(size + 1));
    pos = bcast64(pos);
    if (pos + size + 1 > a.list_cap) {
      if (l == 0) atomicOr(a.err, kErrListFull);
    } else {
      uint32_t *dst = a.list + pos;
      if (l == 0) {
        dst[0] = size;
        dst[1] = a.labels[s.v0];
        if (s.base_size == 2) dst[2] = a.labels[s.v1];
      }
      uint32_t k = 1 + s.base_size;
      for (uint32_t w = 0; w < s.W; ++w) {
        const uint64_t bits = readlane64(rk, w);
        if ((bits >> l) & 1) dst[k + rank_below(bits)] = a.labels[s.Plist[w * 64 + l]];
        k += (uint32_t)__popcll(bits);
      }
    }
  }
}



// Level record layout (u64 words): R[W] P[W] XP[W] V[W] XX[Wx] meta[4]
//   meta[0] = key (sum of label hashes of the clique so far)
//   meta[1] = rsize | xxcnt << 32
//   meta[2] = xlo | xhi << 32   (range of possibly non-zero XX words)
// Words of R, P, XP and V are held one per lane (lane k = word k).
struct Lvl {
  uint64_t rk, pk, xpk, vk;
  uint64_t key;
  uint32_t rsize, xxcnt, xlo, xhi;
};

template <class Ptr>
__device__ __forceinline__ void store_rpx(Ptr rec, const Sub &s, const Lvl &L) {
  const uint32_t l = lane(), W = s.W;
  if (l < W) {
    rec[l] = L.rk;
    rec[W + l] = L.pk;
    rec[2 * W + l] = L.xpk;
  }
  if (l == 0) {
    rec[4 * W + s.Wx] = L.key;
    rec[4 * W + s.Wx + 1] = (uint64_t)L.rsize | ((uint64_t)L.xxcnt << 32);
    rec[4 * W + s.Wx + 2] = (uint64_t)L.xlo | ((uint64_t)L.xhi << 32);
  }
}

template <class Ptr>
__device__ __forceinline__ Lvl load_level(Ptr rec, const Sub &s) {
  const uint32_t l = lane(), W = s.W;
  Lvl L;
  if (l < W) {
    L.rk = rec[l];
    L.pk = rec[W + l];
    L.xpk = rec[2 * W + l];
    L.vk = rec[3 * W + l];
  } else {
    L.rk = L.pk = L.xpk = L.vk = 0;
  }
  L.key = rec[4 * W + s.Wx];
  uint64_t m1 = rec[4 * W + s.Wx + 1], m2 = rec[4 * W + s.Wx + 2];
  L.rsize = (uint32_t)m1;
  L.xxcnt = (uint32_t)(m1 >> 32);
  L.xlo = (uint32_t)m2;
  L.xhi = (uint32_t)(m2 >> 32);
  return L;
}

// |S \ N(u)| for this lane's bitmap row u when `act` (0 otherwise), with
// the words of S held one per lane (lanes < W).  The broadcasts of S's words
// are issued by the whole wave; only the row loads depend on `act`.
template <class BmP>
__device__ __forceinline__ uint32_t miss_count(BmP bm, uint32_t rp, uint32_t u, uint64_t sk, uint32_t W, bool act) {
  uint32_t cnt = 0;
  for (uint32_t k = 0; k < W; ++k) {
    const uint64_t sw = readlane64(sk, k);
    if (act) cnt += (uint32_t)__popcll(sw & ~(uint64_t)bm[(uint64_t)k * rp + u]);
  }
  return cnt;
}

// Child of candidate c (a vertex of P): R' = R + c, P' = P & N(c),
// XP' = XP & N(c), XX' = rows of XX adjacent to c, followed by aggressive
// node generation: every i in P' adjacent to all of P' \ {i} moves to R'
// (BkPivotBitBalance.cuh:351-382); XP' and XX' are then restricted to the
// common neighbours of the moved vertices.  Writes the child's record to
// `out` and returns it.
template <class BmP, class GpP, class XxP, class OutP>
__device__ __forceinline__ Lvl make_child(const Sub &s, const Lvl &par, XxP par_xx, uint32_t c, OutP out, BmP bm, GpP gP,
                          Acc &acc) {
  const uint32_t l = lane(), W = s.W, rp = s.rows_pad, p = s.p;
  const bool wl = l < W;
  const uint32_t kc = c >> 6;
  const uint64_t cbit = 1ull << (c & 63);
  uint64_t rowc = wl ? bm[(uint64_t)l * rp + c] : 0;
  acc.words += W;
  Lvl ch;
  ch.rk = par.rk | (l == kc ? cbit : 0);
  ch.pk = par.pk & rowc;
  ch.xpk = par.xpk & rowc;
  ch.vk = 0;
  ch.key = par.key + gP[c];
  ch.rsize = par.rsize + 1;
  uint64_t mk = 0;
  uint32_t psz = set_popc(ch.pk, W);
  if (psz > 1) {
    for (uint32_t j = 0; j < W; ++j) {
      uint64_t w = readlane64(ch.pk, j);
      if (w == 0) continue;
      acc.words += (uint64_t)__popcll(w) * W;
      const bool is = (w >> l) & 1;
      const uint32_t cnt = miss_count(bm, rp, j * 64 + l, ch.pk, W, is);
      uint64_t m = __ballot(is && cnt == 1);
      if (l == j) mk = m;
    }
    if (__ballot(mk != 0)) {
      ch.rk |= mk;
      ch.pk &= ~mk;
      for (uint32_t j = 0; j < W; ++j) {
        uint64_t w = readlane64(mk, j);
        while (w) {
          uint32_t bb = (uint32_t)__builtin_ctzll(w);
          w &= w - 1;
          uint32_t mm = j * 64 + bb;
          acc.words += W;
          if (wl) ch.xpk &= bm[(uint64_t)l * rp + mm];
          ch.key += gP[mm];
          ch.rsize++;
        }
      }
    }
  }
  const bool anym = __ballot(mk != 0) != 0;
  const uint32_t cb = c & 63;
  uint32_t cnt2 = 0, nlo = par.xhi, nhi = par.xlo;
  for (uint32_t j = par.xlo; j < par.xhi; ++j) {
    uint64_t w = par_xx[j];
    uint64_t nw = 0;
    if (w != 0) {
      acc.words += (uint64_t)__popcll(w) * (anym ? W + 1 : 1);
      const uint32_t u = p + j * 64 + l;
      bool hit = ((w >> l) & 1) && ((bm[(uint64_t)kc * rp + u] >> cb) & 1);
      if (anym) hit = miss_count(bm, rp, u, mk, W, hit) == 0 && hit;
      nw = __ballot(hit);
    }
    if (l == 0) out[4 * W + j] = nw;
    if (nw) {
      cnt2 += __popcll(nw);
      nlo = j < nlo ? j : nlo;
      nhi = j + 1;
    }
  }
  if (cnt2 == 0) nlo = nhi = 0;
  ch.xxcnt = cnt2;
  ch.xlo = nlo;
  ch.xhi = nhi;
  store_rpx(out, s, ch);
  wave_fence();
  return ch;
}

// Tries to take one task from record r (lane 0 only).  Returns the task
// index or -1; *last is set when it was the final task of the publication.
__device__ __forceinline__ int claim_task(ShareRec *r, bool *last) {
  unsigned long long c = ld_relaxed(&r->ctl);
  if ((c & 0xFFFF) >= ((c >> 16) & 0xFFFF)) return -1;
  unsigned long long old = __hip_atomic_fetch_add(&r->ctl, 1ull, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
  uint32_t next = (uint32_t)(old & 0xFFFF), ntot = (uint32_t)((old >> 16) & 0xFFFF);
  if (next >= ntot) return -1;
  *last = next + 1 == ntot;
  return (int)next;
}

// Child for task `idx` of a frozen level record: the idx-th lowest remaining
// candidate t of V, with P and XP as sequential DFS would hold them when it
// reaches t (candidates of V below t moved from P to XP) - the same child
// the publisher would have generated (BkPivotBitBalance.cuh:40-70).
template <class BmP, class GpP, class SnapP, class OutP>
__device__ __forceinline__ Lvl materialize(const Sub &s, SnapP snap, uint32_t idx, OutP out, BmP bm, GpP gP, Acc &acc) {
  const uint32_t l = lane(), W = s.W;
  const bool wl = l < W;
  Lvl par = load_level(snap, s);
  uint32_t cnt = wl ? (uint32_t)__popcll(par.vk) : 0;
  uint32_t before = 0, jt = 0;
  for (uint32_t j = 0; j < W; ++j) {
    uint32_t cj = __builtin_amdgcn_readlane(cnt, j);
    if (before + cj > idx) {
      jt = j;
      break;
    }
    before += cj;
  }
  uint64_t bits = readlane64(par.vk, jt);
  for (uint32_t k = before; k < idx; ++k) bits &= bits - 1;
  uint32_t bt = (uint32_t)__builtin_ctzll(bits);
  uint64_t below = l < jt ? par.vk : (l == jt ? (par.vk & low_mask(bt)) : 0);
  par.pk &= ~below;
  par.xpk |= below;
  return make_child(s, par, snap + 4 * W, jt * 64 + bt, out, bm, gP, acc);
}

__device__ __forceinline__ void publish_bit(const KArgs &a, uint32_t w, bool set) {
  if (set) atomicOr(&a.has_work[w >> 6], 1ull << (w & 63));
  else atomicAnd(&a.has_work[w >> 6], ~(1ull << (w & 63)));
}

// The number of waves looking for work is kept in kIdleCopies copies, one
// per 128-byte line: a wave changing state updates every copy (one vector
// atomic, lane i on copy i), and a searching wave polls only copy me % 64,
// so the polls of thousands of waves do not queue on one address.
__device__ __forceinline__ void idle_add(const KArgs &a, int d) {
  atomicAdd(&a.idle_hint[lane() * kIdleStride], (unsigned int)d);
}
__device__ __forceinline__ bool any_idle(const KArgs &a, uint32_t me) {
  return ld_relaxed(&a.idle_hint[(me % kIdleCopies) * kIdleStride]) != 0;
}

// Evaluates node L (the record at `rec`): a leaf is emitted; otherwise the
// pivot is chosen and the candidates P \ N(pivot) set in L.vk and the
// record.  Returns |P| for a branching node, 0 for a leaf or a node whose P
// an excluded row covers (L.vk empty).
template <class BmP, class GpP, class StP>
__device__ __forceinline__ uint32_t eval_node(const KArgs &a, const Sub &s, Lvl &L, BmP bm, StP rec, Acc &acc,
                                              ldsu64 *s_hist) {
  const uint32_t W = s.W, p = s.p, rp = s.rows_pad;
  const uint32_t l = lane();
  const bool wl = l < W;
  const uint32_t psz = set_popc(L.pk, W);
  L.vk = 0;
  if (psz == 0) {
    if (__ballot(L.xpk != 0) == 0 && L.xxcnt == 0) emit(a, s, acc, s_hist, s.base_size + L.rsize, L.key, L.rk);
    return 0;
  }
  // Pivot: argmin over candidate rows u of (|P \ N(u)|, u).  Each lane
  // visits its rows in increasing u, so a strict < keeps its lowest u.
  uint32_t best_c = ~0u, best_u = ~0u;
  for (uint32_t j = 0; j < W; ++j) {
    uint64_t cw = readlane64(L.pk | L.xpk, j);
    if (cw == 0) continue;
    acc.words += (uint64_t)__popcll(cw) * W;
    const bool act = (cw >> l) & 1;
    const uint32_t u = j * 64 + l;
    const uint32_t cnt = miss_count(bm, rp, u, L.pk, W, act);
    if (act && cnt < best_c) {
      best_c = cnt;
      best_u = u;
    }
  }
  StP XX = rec + 4 * W;
  for (uint32_t j = L.xlo; j < L.xhi; ++j) {
    uint64_t cw = XX[j];
    if (cw == 0) continue;
    acc.words += (uint64_t)__popcll(cw) * W;
    const bool act = (cw >> l) & 1;
    const uint32_t u = p + j * 64 + l;
    const uint32_t cnt = miss_count(bm, rp, u, L.pk, W, act);
    if (act && cnt < best_c) {
      best_c = cnt;
      best_u = u;
    }
  }
  const uint32_t min_c = wave_min_u32(best_c);
  if (min_c == 0) return 0;  // an excluded row covers P: no maximal clique below
  const uint32_t piv = wave_min_u32(best_c == min_c ? best_u : ~0u);
  L.vk = wl ? (L.pk & ~bm[(uint64_t)l * rp + piv]) : 0;
  acc.words += W;
  if (wl) rec[3 * W + l] = L.vk;
  return psz;
}

// Depth-first search from the record at level 0 of `stk` (a root or a
// materialized task).  bm, gP and stk are either this wave's LDS copies or
// HBM; stk_g is this wave's HBM stack arena (frozen levels are copied there
// when the stack lives in LDS, so other waves can read them).  One pass of
// the loop per tree node: evaluate the node, backtrack to the deepest level
// with a candidate left (or take a task of a frozen level), descend.
// Exposes levels root-down while other waves are idle
// (BkPivotBitBalance.cuh:87-138): when a branching node has |P| >=
// share_bound, the shallowest not yet exposed ancestor level is frozen and
// its remaining candidates become tasks; the owner then never backtracks
// into a frozen level but claims its remaining tasks like any other wave.
// Returns the number of nodes this wave evaluated.
template <class BmP, class GpP, class StP>
__device__ __forceinline__ uint64_t run_dfs(const KArgs &a, const Sub &s, BmP bm, GpP gP, StP stk, uint64_t *stk_g, Acc &acc,
                            ldsu64 *s_hist, uint32_t me, bool &published, bool eager = false,
                            bool hbm_pending = false) {
  constexpr bool kStkLds = !std::is_same<StP, uint64_t *>::value;
  const uint32_t W = s.W, p = s.p, rp = s.rows_pad, LS = s.LS;
  const uint32_t l = lane();
  const bool wl = l < W;
  // Lstar (deepest level worth exposing) stays -1 without sharing, which
  // keeps every exposure test false.
  const int share_at = a.share_bound > 0 ? a.share_bound : 0x7fffffff;
  const bool eager_top = eager && a.share_bound > 0;
  ShareRec *my = a.recs + me;
  uint64_t nodes = 0;
  Lvl L = load_level(stk, s);
  int level = 0;
  int S = -1, Lstar = -1;  // levels <= S are frozen; Lstar: deepest level worth exposing
  uint32_t tick = 0;
  for (;;) {
    ++nodes;
    const uint32_t psz = eval_node<BmP, GpP, StP>(a, s, L, bm, stk + (uint64_t)level * LS, acc, s_hist);
    if ((int)psz >= share_at || (eager_top && level == 0 && psz > 0)) Lstar = level;
    bool task = false;
    while (__ballot(L.vk != 0) == 0) {  // backtrack
      if (S >= 0 && level - 1 <= S) {
        // The parent level is frozen: serve its remaining tasks.
        int idx = -1;
        bool last = false;
        if (l == 0) idx = claim_task(my, &last);
        idx = (int)bcast32((uint32_t)idx);
        if (idx < 0) {
          // All tasks taken; wait until the takers have copied their part of
          // the frozen levels out.
          if (l == 0) {
            uint32_t n = 1;
            while (ld_acquire(&my->pending) != 0) {
              nap(n);
              n = n < 8 ? n * 2 : 8;
            }
          }
          wave_fence();
          return nodes;
        }
        if (bcast32(last) && l == 0) publish_bit(a, me, false);
        if (l == 0) atomicAdd(a.owner_tasks, 1ull);
        level = S + 1;
        L = materialize(s, stk + (uint64_t)S * LS, (uint32_t)idx, stk + (uint64_t)level * LS, bm, gP, acc);
        if (l == 0) __hip_atomic_fetch_sub(&my->pending, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        Lstar = Lstar < S ? Lstar : S;
        task = true;
        break;
      }
      if (level == 0) return nodes;
      --level;
      L = load_level(stk + (uint64_t)level * LS, s);
      Lstar = Lstar < level ? Lstar : level;
    }
    if (task) continue;
    // Exposure (an eager root, split before launch, exposes its top level at
    // once; otherwise exposure waits for idle waves, polled every kPollSteps
    // descents - a poll is a round trip past this XCD's L2).
    if (S < Lstar && S + 1 < level && ((eager_top && S < 0) || ((++tick & (kPollSteps - 1)) == 0))) {
      bool go = false;
      if (l == 0) go = ((eager_top && S < 0) || any_idle(a, me)) && ld_relaxed(&my->pending) == 0;
      if (bcast32(go)) {
        const uint32_t lp = (uint32_t)(S + 1);
        StP src = stk + (uint64_t)lp * LS;
        uint32_t n = set_popc(wl ? (uint64_t)src[3 * W + l] : 0ull, W);
        if (n > 0) {
          if (hbm_pending) {  // bitmap and label hashes exist only in LDS so far
            uint64_t *dbm = (uint64_t *)s.bm, *dgp = (uint64_t *)s.gP;
            const uint64_t nb = (uint64_t)W * rp;
            for (uint64_t i = l; i < nb; i += 64) dbm[i] = bm[i];
            for (uint64_t i = l; i < p; i += 64) dgp[i] = gP[i];
            hbm_pending = false;
          }
          const uint64_t *snap_g;
          if (kStkLds) {
            uint64_t *dst = stk_g + (uint64_t)lp * LS;
            for (uint32_t i = l; i < LS; i += 64) dst[i] = src[i];
            snap_g = dst;
          } else {
            snap_g = (const uint64_t *)src;
          }
          if (l == 0) {
            __hip_atomic_store(&my->pending, n, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            my->snap = snap_g;
            my->sub = s;
            uint32_t ep = my->epoch + 1;
            my->epoch = ep;
            __hip_atomic_store(&my->ctl, ((unsigned long long)ep << 32) | ((unsigned long long)n << 16),
                               __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
            publish_bit(a, me, true);
            atomicAdd(a.published_levels, 1ull);
          }
          published = true;
        }
        S = (int)lp;
      }
    }
    // Descend into the lowest candidate.
    StP rec = stk + (uint64_t)level * LS;
    const uint64_t vm = __ballot(L.vk != 0);
    const uint32_t kc = (uint32_t)__builtin_ctzll(vm);
    const uint32_t b = (uint32_t)__builtin_ctzll(readlane64(L.vk, kc));
    const uint64_t cbit = 1ull << b;
    if (l == kc) {
      L.vk &= ~cbit;
      L.pk &= ~cbit;
      L.xpk |= cbit;
      rec[W + l] = L.pk;
      rec[2 * W + l] = L.xpk;
      rec[3 * W + l] = L.vk;
    }
    L = make_child(s, L, rec + 4 * W, kc * 64 + b, rec + LS, bm, gP, acc);
    ++level;
  }
}

// ---------------------------------------------------------------------------
// Register paths: |P0| <= 64 NW and |X0'| <= 64 NX (W == NW, Wx <= NX; NW,
// NX = 1, 1 or 2, 1).  Each lane keeps, for the whole root, the bitmap rows
// it is asked about: P[j][k] = word k of the row of P0 vertex 64 j + l,
// X[j][k] = word k of the row of X0' vertex 64 j + l (bits over P0), g[j] =
// label hash of P0 vertex 64 j + l.  R, P, XP and V are NW uniform words, XX
// NX, and a node reads no memory:
//   pivot   one masked popcount per lane and row slot and one wave min over
//           |P \ N(u)| << 16 | u, the (count, lowest row) argmin of the scans
//           of run_dfs;
//   child   P' = P & N(c), XP' = XP & N(c); vertex i of P' moves to R' when
//           P' & ~N(i) == {i} (one ballot per slot); XP' and XX' keep the
//           rows that contain every moved vertex (and c, for XX') - the
//           bitmap is symmetric on P0, so x in N(m) iff m in N(x).
// Arrays are indexed by compile-time constants or through a uniform select
// (no dynamic register indexing).  Level records keep the layout of the
// other tiers (R P XP V XX key meta, words 0 .. 4 NW + Wx + 2 of LS), so
// publication, task claims and thieves are unchanged.  A record is written
// only when the search descends from its level (by lane 0; other lanes read
// it after a wave_fence), so leaves write nothing.  Same
// search tree, node counts and accounting as run_dfs.
// ---------------------------------------------------------------------------
template <int NW, int NX>
struct LvlR {
  uint64_t r[NW], p[NW], x[NW], v[NW], xx[NX];
  uint64_t key;
  uint32_t rsize;
};

template <int NW, int NX>
struct RowsR {
  uint64_t P[NW][NW], X[NX][NW], g[NW];
};

// Element j (uniform) of a per-slot array.
template <int NW>
__device__ __forceinline__ uint64_t pick(const uint64_t (&a)[NW], uint32_t j) {
  if constexpr (NW == 1) {
    return a[0];
  } else {
    return j ? a[1] : a[0];
  }
}

// `v` (uniform) on lane ln, `old` elsewhere.
__device__ __forceinline__ uint64_t writelane64(uint64_t old, uint64_t v, uint32_t ln) {
  return lane() == ln ? v : old;
}

// Whole record of a level, written by lane 0 (its words are read back by
// lanes 0 .. 4 NW + Wx + 2 after a wave_fence).  XX words: [xlo, xhi) spans
// the non-zero ones (both 0 when XX is empty), as make_child writes them.
template <int NW, int NX, class Ptr>
__device__ __forceinline__ void store_lvlr(Ptr rec, uint32_t Wx, const LvlR<NW, NX> &L) {
  uint32_t xc = 0, xlo = 0, xhi = 0;
#pragma unroll
  for (int j = NX - 1; j >= 0; --j)
    if (L.xx[j]) {
      xc += (uint32_t)__popcll(L.xx[j]);
      xlo = (uint32_t)j;
      if (!xhi) xhi = (uint32_t)j + 1;
    }
  if (lane() == 0) {
#pragma unroll
    for (int k = 0; k < NW; ++k) {
      rec[k] = L.r[k];
      rec[NW + k] = L.p[k];
      rec[2 * NW + k] = L.x[k];
      rec[3 * NW + k] = L.v[k];
    }
#pragma unroll
    for (int j = 0; j < NX; ++j)
      if ((uint32_t)j < Wx) rec[4 * NW + j] = L.xx[j];
    rec[4 * NW + Wx] = L.key;
    rec[4 * NW + Wx + 1] = (uint64_t)L.rsize | ((uint64_t)xc << 32);
    rec[4 * NW + Wx + 2] = (uint64_t)xlo | ((uint64_t)xhi << 32);
  }
}

// P, XP and V of a level whose record exists (lane 0).
template <int NW, int NX, class Ptr>
__device__ __forceinline__ void store_pxvr(Ptr rec, const LvlR<NW, NX> &L) {
  if (lane() == 0) {
#pragma unroll
    for (int k = 0; k < NW; ++k) {
      rec[NW + k] = L.p[k];
      rec[2 * NW + k] = L.x[k];
      rec[3 * NW + k] = L.v[k];
    }
  }
}

template <int NW, int NX, class Ptr>
__device__ __forceinline__ LvlR<NW, NX> load_lvlr(Ptr rec, uint32_t Wx) {
  const uint64_t w = lane() < 4 * NW + Wx + 3 ? (uint64_t)rec[lane()] : 0ull;
  LvlR<NW, NX> L;
#pragma unroll
  for (int k = 0; k < NW; ++k) {
    L.r[k] = readlane64(w, k);
    L.p[k] = readlane64(w, NW + k);
    L.x[k] = readlane64(w, 2 * NW + k);
    L.v[k] = readlane64(w, 3 * NW + k);
  }
  L.key = readlane64(w, 4 * NW + Wx);
  L.rsize = __builtin_amdgcn_readlane((uint32_t)w, 4 * NW + Wx + 1);
  const uint32_t xlo = __builtin_amdgcn_readlane((uint32_t)w, 4 * NW + Wx + 2);
  const uint32_t xhi = __builtin_amdgcn_readlane((uint32_t)(w >> 32), 4 * NW + Wx + 2);
#pragma unroll
  for (int j = 0; j < NX; ++j) {  // a word outside [xlo, xhi) may be stale
    const uint64_t xw = readlane64(w, 4 * NW + j);
    L.xx[j] = ((uint32_t)j < Wx && (uint32_t)j >= xlo && (uint32_t)j < xhi) ? xw : 0ull;
  }
  return L;
}

// Child of candidate c, as make_child computes it.
template <int NW, int NX>
__device__ __forceinline__ LvlR<NW, NX> child_r(const LvlR<NW, NX> &par, uint32_t c, const RowsR<NW, NX> &rw,
                                                Acc &acc) {
  const uint32_t l = lane(), jc = c >> 6, bc = c & 63;
  LvlR<NW, NX> ch;
  uint32_t psz = 0;
#pragma unroll
  for (int k = 0; k < NW; ++k) {
    const uint64_t rc = readlane64(jc ? rw.P[NW - 1][k] : rw.P[0][k], bc);
    ch.r[k] = par.r[k] | ((uint32_t)k == jc ? 1ull << bc : 0ull);
    ch.p[k] = par.p[k] & rc;
    ch.x[k] = par.x[k] & rc;
    ch.v[k] = 0;
    psz += (uint32_t)__popcll(ch.p[k]);
  }
  ch.key = par.key + readlane64(pick<NW>(rw.g, jc), bc);
  ch.rsize = par.rsize + 1;
  acc.words += NW;
  uint64_t mk[NW];
  bool anym = false;
#pragma unroll
  for (int k = 0; k < NW; ++k) mk[k] = 0;
  if (psz > 1) {
    acc.words += (uint64_t)psz * NW;
#pragma unroll
    for (int j = 0; j < NW; ++j) {
      const bool is = (ch.p[j] >> l) & 1;
      uint32_t cnt = 0;
#pragma unroll
      for (int k = 0; k < NW; ++k) cnt += (uint32_t)__popcll(ch.p[k] & ~rw.P[j][k]);
      mk[j] = __ballot(is && cnt == 1);
      anym = anym || mk[j] != 0;
    }
    if (anym) {
      uint32_t nm = 0;
#pragma unroll
      for (int k = 0; k < NW; ++k) {
        ch.r[k] |= mk[k];
        ch.p[k] &= ~mk[k];
        nm += (uint32_t)__popcll(mk[k]);
      }
#pragma unroll
      for (int j = 0; j < NW; ++j) {
        bool all = true;
#pragma unroll
        for (int k = 0; k < NW; ++k) all = all && (rw.P[j][k] & mk[k]) == mk[k];
        ch.x[j] &= __ballot(all);
        for (uint64_t m = mk[j]; m; m &= m - 1) ch.key += readlane64(rw.g[j], (uint32_t)__builtin_ctzll(m));
      }
      ch.rsize += nm;
      acc.words += (uint64_t)nm * NW;
    }
  }
  uint32_t xc = 0;
#pragma unroll
  for (int j = 0; j < NX; ++j) xc += (uint32_t)__popcll(par.xx[j]);
  acc.words += (uint64_t)xc * (anym ? NW + 1 : 1);
#pragma unroll
  for (int j = 0; j < NX; ++j) {
    bool keep = (pick<NW>(rw.X[j], jc) >> bc) & 1;
#pragma unroll
    for (int k = 0; k < NW; ++k) keep = keep && (rw.X[j][k] & mk[k]) == mk[k];
    ch.xx[j] = par.xx[j] & __ballot(keep);
  }
  return ch;
}

// Evaluates node L: a leaf is emitted; otherwise the pivot is chosen and L.v
// set to the candidates P \ N(pivot).  Returns |P| for a branching node, 0
// for a leaf or a node whose P an excluded row covers (no candidates).
template <int NW, int NX>
__device__ __forceinline__ uint32_t eval_r(const KArgs &a, const Sub &s, LvlR<NW, NX> &L, const RowsR<NW, NX> &rw,
                                           Acc &acc, ldsu64 *s_hist) {
  const uint32_t l = lane(), p = s.p;
  uint32_t psz = 0;
  uint64_t xany = 0, cw = 0, xw = 0;
#pragma unroll
  for (int k = 0; k < NW; ++k) {
    psz += (uint32_t)__popcll(L.p[k]);
    xany |= L.x[k];
    cw += (uint64_t)__popcll(L.p[k] | L.x[k]);
    L.v[k] = 0;
  }
#pragma unroll
  for (int j = 0; j < NX; ++j) {
    xany |= L.xx[j];
    xw += (uint64_t)__popcll(L.xx[j]);
  }
  if (psz == 0) {
    if (xany == 0) {
      uint64_t rk = 0;
#pragma unroll
      for (int k = 0; k < NW; ++k) rk = writelane64(rk, L.r[k], k);
      emit(a, s, acc, s_hist, s.base_size + L.rsize, L.key, rk);
    }
    return 0;
  }
  acc.words += (cw + xw) * NW;
  uint32_t best = ~0u;
#pragma unroll
  for (int j = 0; j < NW; ++j) {
    uint32_t cp = 0;
#pragma unroll
    for (int k = 0; k < NW; ++k) cp += (uint32_t)__popcll(L.p[k] & ~rw.P[j][k]);
    const uint32_t kp = (((L.p[j] | L.x[j]) >> l) & 1) ? (cp << 16) | (64u * j + l) : ~0u;
    best = kp < best ? kp : best;
  }
#pragma unroll
  for (int j = 0; j < NX; ++j) {
    uint32_t cx = 0;
#pragma unroll
    for (int k = 0; k < NW; ++k) cx += (uint32_t)__popcll(L.p[k] & ~rw.X[j][k]);
    const uint32_t kx = ((L.xx[j] >> l) & 1) ? (cx << 16) | (p + 64u * j + l) : ~0u;
    best = kx < best ? kx : best;
  }
  const uint32_t k = wave_min_u32(best);
  if ((k >> 16) == 0) return 0;  // an excluded row covers P: no maximal clique below
  acc.words += NW;               // the pivot's row
  const uint32_t piv = k & 0xFFFFu;
  const bool in_p = piv < p;
  const uint32_t row = in_p ? piv : piv - p, js = row >> 6, ls = row & 63;
#pragma unroll
  for (int kk = 0; kk < NW; ++kk) {
    const uint64_t pr = in_p ? (js ? rw.P[NW - 1][kk] : rw.P[0][kk]) : (js ? rw.X[NX - 1][kk] : rw.X[0][kk]);
    L.v[kk] = L.p[kk] & ~readlane64(pr, ls);
  }
  return psz;
}

// One pass of the loop per tree node: evaluate the node, backtrack to the
// deepest level with a candidate left (or take a task of a frozen level),
// descend.  Exposure of levels to idle waves (polled every kPollSteps
// descents) and the serving of a frozen level's tasks follow run_dfs.  The
// stack is addressed through a flat pointer (LDS or HBM, stk_lds telling
// which): a node touches it at most once, and one instantiation per width
// keeps the kernel's code and register allocation within budget.  Lane 0
// writes the records (no per-lane select of uniform words), and every read
// of one by other lanes follows a wave_fence; the backtrack that follows a
// store comes a whole subtree later, so the fence rarely waits.
template <int NW, int NX>
__device__ __forceinline__ uint64_t run_dfsr(const KArgs &a, const Sub &s, const uint64_t *bm, const uint64_t *gP,
                                             uint64_t *stk, bool stk_lds, uint64_t *stk_g, Acc &acc, ldsu64 *s_hist,
                                             uint32_t me, bool &published, bool eager, bool hbm_pending) {
  const uint32_t p = s.p, rp = s.rows_pad, LS = s.LS, Wx = s.Wx;
  const uint32_t l = lane();
  // Lstar (deepest level worth exposing) stays -1 without sharing, which
  // keeps every exposure test false.
  const int share_at = a.share_bound > 0 ? a.share_bound : 0x7fffffff;
  const bool eager_top = eager && a.share_bound > 0;
  ShareRec *my = a.recs + me;
  RowsR<NW, NX> rw;
#pragma unroll
  for (int j = 0; j < NW; ++j) {
    const uint32_t u = 64u * j + l;
#pragma unroll
    for (int k = 0; k < NW; ++k) rw.P[j][k] = u < p ? bm[(uint64_t)k * rp + u] : 0ull;
    rw.g[j] = u < p ? gP[u] : 0ull;
  }
#pragma unroll
  for (int j = 0; j < NX; ++j) {
    const uint32_t u = 64u * j + l;
#pragma unroll
    for (int k = 0; k < NW; ++k) rw.X[j][k] = u < s.xq ? bm[(uint64_t)k * rp + p + u] : 0ull;
  }
  LvlR<NW, NX> L = load_lvlr<NW, NX>(stk, Wx);
  int level = 0, S = -1, Lstar = -1;
  uint32_t tick = 0;
  bool fresh = true;  // the current level's record is not written yet
  uint64_t nodes = 0;
  for (;;) {
    ++nodes;
    const uint32_t psz = eval_r<NW, NX>(a, s, L, rw, acc, s_hist);
    if ((int)psz >= share_at || (eager_top && level == 0 && psz > 0)) Lstar = level;
    uint32_t c;
    bool task = false;
    for (;;) {  // backtrack while no candidate is left
      uint64_t vany = 0;
#pragma unroll
      for (int k = 0; k < NW; ++k) vany |= L.v[k];
      if (vany) break;
      if (S >= 0 && level - 1 <= S) {
        // The parent level is frozen: serve its remaining tasks.
        int idx = -1;
        bool last = false;
        if (l == 0) idx = claim_task(my, &last);
        idx = (int)bcast32((uint32_t)idx);
        if (idx < 0) {
          // All tasks taken; wait until the takers have copied their part of
          // the frozen levels out.
          if (l == 0) {
            uint32_t n = 1;
            while (ld_acquire(&my->pending) != 0) {
              nap(n);
              n = n < 8 ? n * 2 : 8;
            }
          }
          wave_fence();
          return nodes;
        }
        if (bcast32(last) && l == 0) publish_bit(a, me, false);
        if (l == 0) atomicAdd(a.owner_tasks, 1ull);
        // candidate idx of the frozen V, with P and XP as sequential DFS
        // holds them when it reaches it (as materialize does)
        wave_fence();
        L = load_lvlr<NW, NX>(stk + (uint64_t)S * LS, Wx);
        uint32_t before = 0, jt = 0;
        if constexpr (NW > 1) {
          const uint32_t n0 = (uint32_t)__popcll(L.v[0]);
          if ((uint32_t)idx >= n0) {
            jt = 1;
            before = n0;
          }
        }
        uint64_t bits = pick<NW>(L.v, jt);
        for (uint32_t k = before; k < (uint32_t)idx; ++k) bits &= bits - 1;
        const uint32_t bt = (uint32_t)__builtin_ctzll(bits);
        c = 64u * jt + bt;
#pragma unroll
        for (int k = 0; k < NW; ++k) {
          const uint64_t below = (uint32_t)k < jt ? L.v[k] : ((uint32_t)k == jt ? L.v[k] & ((1ull << bt) - 1) : 0ull);
          L.p[k] &= ~below;
          L.x[k] |= below;
        }
        if (l == 0) __hip_atomic_fetch_sub(&my->pending, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        Lstar = Lstar < S ? Lstar : S;
        level = S;
        task = true;
        break;
      }
      if (level == 0) return nodes;
      --level;
      wave_fence();  // lane 0 wrote the record
      L = load_lvlr<NW, NX>(stk + (uint64_t)level * LS, Wx);
      fresh = false;
      Lstar = Lstar < level ? Lstar : level;
    }
    if (!task) {
      // Exposure: the shallowest level above this one that is not frozen.
      if (S < Lstar && S + 1 < level && ((eager_top && S < 0) || ((++tick & (kPollSteps - 1)) == 0))) {
        bool go = false;
        if (l == 0) go = ((eager_top && S < 0) || any_idle(a, me)) && ld_relaxed(&my->pending) == 0;
        if (bcast32(go)) {
          const uint32_t lp = (uint32_t)(S + 1);
          const uint64_t *src = stk + (uint64_t)lp * LS;
          wave_fence();  // lane 0 wrote the record
          const uint64_t sw = l < LS ? src[l] : 0ull;
          uint32_t n = 0;
#pragma unroll
          for (int k = 0; k < NW; ++k) n += (uint32_t)__popcll(readlane64(sw, 3 * NW + k));
          if (n > 0) {
            if (hbm_pending) {  // bitmap and label hashes exist only in LDS so far
              uint64_t *dbm = (uint64_t *)s.bm, *dgp = (uint64_t *)s.gP;
              const uint64_t nb = (uint64_t)NW * rp;
              for (uint64_t i = l; i < nb; i += 64) dbm[i] = bm[i];
              for (uint64_t i = l; i < p; i += 64) dgp[i] = gP[i];
              hbm_pending = false;
            }
            const uint64_t *snap_g = src;
            if (stk_lds) {
              uint64_t *dst = stk_g + (uint64_t)lp * LS;
              if (l < LS) dst[l] = sw;
              snap_g = dst;
            }
            if (l == 0) {
              __hip_atomic_store(&my->pending, n, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
              my->snap = snap_g;
              my->sub = s;
              uint32_t ep = my->epoch + 1;
              my->epoch = ep;
              __hip_atomic_store(&my->ctl, ((unsigned long long)ep << 32) | ((unsigned long long)n << 16),
                                 __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
              publish_bit(a, me, true);
              atomicAdd(a.published_levels, 1ull);
            }
            published = true;
          }
          S = (int)lp;
        }
      }
      // Descend into the lowest candidate.
      uint32_t jt = 0;
      if constexpr (NW > 1) jt = L.v[0] ? 0u : 1u;
      const uint32_t bt = (uint32_t)__builtin_ctzll(pick<NW>(L.v, jt));
      c = 64u * jt + bt;
#pragma unroll
      for (int k = 0; k < NW; ++k) {
        const uint64_t cbit = (uint32_t)k == jt ? 1ull << bt : 0ull;
        L.v[k] &= ~cbit;
        L.p[k] &= ~cbit;
        L.x[k] |= cbit;
      }
      uint64_t *rec = stk + (uint64_t)level * LS;
      if (fresh) store_lvlr<NW, NX>(rec, Wx, L);
      else store_pxvr<NW, NX>(rec, L);
    }
    L = child_r<NW, NX>(L, c, rw, acc);
    fresh = true;
    ++level;
  }
}

__device__ void defer_root(const KArgs &a, uint64_t canon, uint64_t data_need, uint64_t stack_need) {
  if (lane() == 0) {
    unsigned long long pos = atomicAdd(a.n_deferred, 1ull);
    if (pos < a.deferred_cap) a.deferred[pos] = canon;
    else atomicOr(a.err, kErrDeferFull);
    atomicMax(a.deferred_need, (unsigned long long)data_need);
    atomicMax(a.deferred_stack_need, (unsigned long long)stack_need);
  }
}

// LDS words a subproblem needs for the LDS tier: bitmap, label hashes,
// `levels` stack records.
__device__ __forceinline__ uint64_t lds_words(const Sub &s) {
  return (uint64_t)s.W * s.rows_pad + s.p + (uint64_t)s.levels * s.LS;
}

// Copies bitmap and label hashes of s from HBM into the LDS tier layout
// [bm | gP | stack]; returns the stack base.
__device__ __forceinline__ lds64 *stage_lds(const Sub &s, lds64 *lds) {
  const uint32_t l = lane();
  const uint64_t nb = (uint64_t)s.W * s.rows_pad;
  for (uint64_t i = l; i < nb; i += 64) lds[i] = s.bm[i];
  for (uint64_t i = l; i < s.p; i += 64) lds[nb + i] = s.gP[i];
  return lds + nb + s.p;
}

// Batches: moves what this wave accumulated since the last flush (count,
// hash, LDS histogram) to graph g's results.
__device__ __forceinline__ void flush_graph(const KArgs &a, uint32_t g, Acc &acc, ldsu64 *s_hist) {
  const uint32_t l = lane();
  wave_fence();  // lane 0's histogram updates before every lane reads its entries
  if (l == 0) {
    if (acc.count) atomicAdd(&a.g_count[g], acc.count);
    if (acc.hash) atomicAdd(&a.g_hash[g], acc.hash);
    acc.count = 0;
    acc.hash = 0;
  }
  for (uint32_t i = l; i < kLdsHist; i += 64) {
    unsigned long long v = s_hist[i];
    if (v) {
      if (i < a.g_hist_stride) atomicAdd(&a.g_hist[(uint64_t)g * a.g_hist_stride + i], v);
      else atomicOr(a.err, kErrHistOverflow);
      s_hist[i] = 0;
    }
  }
  __builtin_amdgcn_wave_barrier();
}

// Degeneracy of the graph on P0 (rows 0..p-1 of an LDS bitmap), p <= 256:
// repeatedly removes a vertex of minimum remaining degree.  Every clique in
// P0 has at most degeneracy + 1 vertices, which bounds the depth of the
// search (the stack) more tightly than the maximum degree does.
__device__ __forceinline__ uint32_t p0_degeneracy(const lds64 *T, uint32_t rp, uint32_t p, uint32_t W) {
  const uint32_t l = lane();
  uint32_t d[4];
  bool al[4];
#pragma unroll
  for (uint32_t j = 0; j < 4; ++j) {
    const uint32_t u = j * 64 + l;
    al[j] = u < p;
    d[j] = 0;
    if (al[j])
      for (uint32_t k = 0; k < W; ++k) d[j] += (uint32_t)__popcll(T[(uint64_t)k * rp + u]);
  }
  uint32_t degen = 0;
  for (uint32_t left = p; left > 0; --left) {
    uint32_t key = ~0u;
#pragma unroll
    for (uint32_t j = 0; j < 4; ++j) {
      const uint32_t kj = (d[j] << 16) | (j * 64 + l);
      if (al[j] && kj < key) key = kj;
    }
    const uint32_t m = wave_min_u32(key);
    const uint32_t v = m & 0xFFFF, dv = m >> 16;
    degen = dv > degen ? dv : degen;
    if (dv + 1 >= left) break;  // the remaining vertices form a clique
    const uint64_t kv = v >> 6, vb = 1ull << (v & 63);
#pragma unroll
    for (uint32_t j = 0; j < 4; ++j) {
      if (j * 64 + l == v) al[j] = false;
      if (al[j] && (T[kv * rp + j * 64 + l] & vb)) d[j]--;
    }
  }
  return degen;
}

// Builds root `canon` and searches it.  Returns tree nodes evaluated by this
// wave.
//
// In degeneracy order every vertex has at most `degeneracy` higher-ordered
// neighbours, so the build touches only short out-lists N+(.) (their start
// is outoff[v]) wherever it can:
//   P  = N+(v1) & N+(v0) (edge roots; all of P lies above v1 > v0),
//   X  = N(v0) & N-(v1) below v1 (only its emptiness when P is empty),
//   X' = rows x of X with N+(x) & P non-empty (x < v1 < every P vertex),
//   bitmap rows: P x P from N+(u) for u in P (set symmetrically), X' rows
//   from N+(x) above v1.
// The bitmap is built in this wave's LDS when it fits, else in its HBM data
// region; an LDS-built bitmap is copied to HBM only when stealing needs it.
__device__ __forceinline__ uint64_t solve_root(const KArgs &a, uint8_t *slot, uint64_t *stk_g, lds64 *lds, uint32_t me,
                               uint64_t canon, Acc &acc, ldsu64 *s_hist, bool &published,
                               bool eager) {
  const uint32_t l = lane();
  uint32_t *Pl = (uint32_t *)slot;  // P list, then X (later X') list
  const uint64_t ccap = a.slot_bytes / 4;
  Sub s;
  s.canon = canon;
  s.bm_owner = me;
  s.graph = a.root_graph ? a.root_graph[canon] : 0;
  uint64_t p = 0, x = 0;
  // every P vertex lies above `above`, every X vertex below it
  const uint32_t above = root_lists(a, canon, Pl, ccap, s, p, x, acc);
  if (p == 0) {
    if (x == 0) {
      s.W = 0;
      s.Plist = nullptr;
      emit(a, s, acc, s_hist, s.base_size, s.base_key, 0);
    }
    return 0;
  }
  if (p > kMaxP) {
    if (l == 0) atomicOr(a.err, kErrPTooLarge);
    return 0;
  }
  const uint64_t W = (p + 63) / 64;
  if (p + x > ccap) {  // candidate lists alone exceed the data region
    const uint64_t c = p + x;
    defer_root(a, canon, align_up(4 * c, 256) + align_up(8 * W * align_up(c, 64), 256) + align_up(8 * p, 256),
               (p + 1) * (4 * W + (x + 63) / 64 + 4) * 8);
    return 0;
  }
  wave_fence();
  const uint32_t *P = Pl;
  uint32_t *X = Pl + p;
  // X' = X rows with a neighbour in P, compacted in place (graph_gpu.cuh:90-106).
  const uint64_t xq = compact_x(a, P, p, X, x, above);
  wave_fence();
  const uint32_t Wx = (uint32_t)((xq + 63) / 64);
  const uint64_t rows = p + xq;
  const uint32_t rows_pad = (uint32_t)align_up(rows, 64);
  const uint64_t bm_off = align_up(4 * (p + x), 256);
  const uint64_t gp_off = bm_off + align_up(8ull * W * rows_pad, 256);
  const uint64_t data_need = gp_off + align_up(8 * p, 256);
  const uint32_t LS = (uint32_t)(4 * W + Wx + 4);
  if (data_need > a.slot_bytes) {
    defer_root(a, canon, data_need, (p + 1) * LS * 8);
    return 0;
  }
  uint64_t *bm_g = (uint64_t *)(slot + bm_off);
  uint64_t *gp_g = (uint64_t *)(slot + gp_off);
  const uint64_t nbw = W * rows_pad;
  const bool in_lds = a.lds_tier && (nbw + p) * 8 <= a.lds_bytes;
  acc.build += 4 * (p + x) + 8 * nbw;
  // Bitmap (word-major): row i is P[i] for i < p, else X'[i - p].
  uint32_t maxdeg = 0;
  // One build path through flat pointers (LDS or HBM resolved at run time):
  // a second, LDS-typed instantiation would double this code in the kernel
  // and its register allocation.
  auto build = [&](uint64_t *T, uint64_t *G) {
    for (uint64_t i = l; i < nbw; i += 64) T[i] = 0;
    for (uint64_t i = l; i < p; i += 64) G[i] = a.label_hash[P[i]];
    wave_fence();
    acc.scan += match_rows(a, P, p, X, rows, above,
               [&](uint64_t i, uint32_t j) {
                 or64(&T[(uint64_t)(j >> 6) * rows_pad + i], 1ull << (j & 63));
                 or64(&T[(uint64_t)(i >> 6) * rows_pad + j], 1ull << (i & 63));
               },
               [&](uint64_t i, uint32_t w, uint64_t bits) { or64(&T[(uint64_t)w * rows_pad + i], bits); });
    wave_fence();
    for (uint64_t i = l; i < p; i += 64) {
      uint32_t d = 0;
      for (uint64_t w = 0; w < W; ++w) d += (uint32_t)__popcll(T[w * rows_pad + i]);
      maxdeg = d > maxdeg ? d : maxdeg;
    }
    maxdeg = wave_max_u32(maxdeg);
  };
  build(in_lds ? (uint64_t *)lds : bm_g, in_lds ? (uint64_t *)(lds + nbw) : gp_g);
  // Depth bound: a clique inside P0 has at most maxdeg + 1 vertices, and at
  // most degeneracy + 1 (computed when that can move the stack into LDS).
  uint64_t levels = (p < (uint64_t)maxdeg + 1 ? p : (uint64_t)maxdeg + 1) + 1;
  if (in_lds && p <= 256 && (nbw + p + levels * LS) * 8 > a.lds_bytes) {
    const uint64_t dg = p0_degeneracy((const lds64 *)lds, rows_pad, (uint32_t)p, (uint32_t)W);
    levels = (p < dg + 1 ? p : dg + 1) + 1;
  }
  if (levels * LS * 8 > a.stack_bytes) {
    defer_root(a, canon, data_need, levels * LS * 8);
    return 0;
  }
  s.p = (uint32_t)p;
  s.xq = (uint32_t)xq;
  s.W = (uint32_t)W;
  s.Wx = Wx;
  s.rows_pad = rows_pad;
  s.LS = LS;
  s.levels = (uint32_t)levels;
  s.Plist = P;
  s.bm = bm_g;
  s.gP = gp_g;
  // |P0| <= 64 or <= 128 and |X0'| <= 64: the register paths (rows in
  // registers).
  const bool reg1 = a.word_path && W == 1 && Wx <= 1;
  const bool reg2 = a.word_path && W == 2 && Wx <= 1;
  // Level 0: R = {}, P = P0, XP = {}, XX = X0'.
  Lvl L0;
  L0.rk = 0;
  L0.pk = l < W ? low_mask(p - (uint64_t)l * 64) : 0;
  L0.xpk = 0;
  L0.vk = 0;
  L0.key = s.base_key;
  L0.rsize = 0;
  L0.xxcnt = (uint32_t)xq;
  L0.xlo = 0;
  L0.xhi = Wx;
  if (reg1 || reg2) {
    // The stack in LDS when it fits beside the bitmap, as for run_dfs.
    const bool sl = in_lds && lds_words(s) * 8 <= a.lds_bytes;
    uint64_t *stk = sl ? (uint64_t *)(lds + nbw + p) : stk_g;
    for (uint32_t j = l; j < Wx; j += 64) stk[4 * W + j] = low_mask(xq - (uint64_t)j * 64);
    store_rpx(stk, s, L0);
    wave_fence();
    const uint64_t *bmp = in_lds ? (const uint64_t *)lds : bm_g, *gpp = in_lds ? (const uint64_t *)(lds + nbw) : gp_g;
    if (reg1)
      return run_dfsr<1, 1>(a, s, bmp, gpp, stk, sl, stk_g, acc, s_hist, me, published, eager, in_lds);
    return run_dfsr<2, 1>(a, s, bmp, gpp, stk, sl, stk_g, acc, s_hist, me, published, eager, in_lds);
  }
  if (in_lds && lds_words(s) * 8 <= a.lds_bytes) {
    lds64 *st = lds + nbw + p;
    for (uint32_t j = l; j < Wx; j += 64) st[4 * W + j] = low_mask(xq - (uint64_t)j * 64);
    store_rpx(st, s, L0);
    wave_fence();
    return run_dfs(a, s, (const lds64 *)lds, (const lds64 *)(lds + nbw), st, stk_g, acc, s_hist, me, published,
                   eager, true);
  }
  for (uint32_t j = l; j < Wx; j += 64) stk_g[4 * W + j] = low_mask(xq - (uint64_t)j * 64);
  store_rpx(stk_g, s, L0);
  wave_fence();
  if (in_lds)  // bitmap and label hashes in LDS, the stack (too deep for LDS) in HBM
    return run_dfs(a, s, (const lds64 *)lds, (const lds64 *)(lds + nbw), stk_g, stk_g, acc, s_hist, me, published,
                   eager, true);
  return run_dfs(a, s, (const uint64_t *)bm_g, (const uint64_t *)gp_g, stk_g, stk_g, acc, s_hist, me, published,
                 eager);
}

// Looks for a published task (preferring waves on this XCD, which share
// this wave's L2) and runs it.  Returns true when a task was run.
__device__ __forceinline__ bool steal(const KArgs &a, uint32_t me, uint64_t *stk_g, lds64 *lds, Acc &acc,
                      ldsu64 *s_hist, bool &is_idle) {
  const uint32_t l = lane();
  const uint32_t nwords = (a.n_waves + 63) / 64;
  const uint64_t xcd = 0x0101010101010101ull << (me & 7);
  int victim = -1;
  for (uint32_t base = 0; base < nwords && victim < 0; base += 64) {
    uint32_t j = base + l;
    uint64_t w = j < nwords ? ld_relaxed(&a.has_work[j]) : 0;
    if (j == (me >> 6)) w &= ~(1ull << (me & 63));
    uint64_t near = __ballot((w & xcd) != 0);
    uint64_t any = __ballot(w != 0);
    uint64_t pick = near ? near : any;
    if (pick) {
      // rotate so that thieves starting from different waves spread out
      uint32_t start = (me >> 6) & 63;
      uint64_t rot = (pick >> start) | (start ? pick << (64 - start) : 0);
      uint32_t jl = ((uint32_t)__builtin_ctzll(rot) + start) & 63;
      uint64_t word = readlane64(w, jl);
      if (near) word &= xcd;
      uint32_t sb = (me + 1) & 63;
      uint64_t r2 = (word >> sb) | (sb ? word << (64 - sb) : 0);
      uint32_t bit = ((uint32_t)__builtin_ctzll(r2) + sb) & 63;
      victim = (int)((base + jl) * 64 + bit);
    }
  }
  if (victim < 0 || (uint32_t)victim >= a.n_waves) return false;
  ShareRec *r = a.recs + victim;
  int idx = -1;
  bool last = false;
  if (l == 0) idx = claim_task(r, &last);
  idx = (int)bcast32((uint32_t)idx);
  if (idx < 0) return false;
  if (l == 0 && last) publish_bit(a, (uint32_t)victim, false);
  if (is_idle) {  // looking for work -> working
    idle_add(a, -1);
    if (l == 0) atomicAdd(a.active, 1u);
    is_idle = false;
  }
  Sub s = r->sub;
  const uint64_t *snap = r->snap;
  if (l == 0) atomicAdd(&a.users[s.bm_owner], 1u);
  // Tiers as for a root: bitmap, label hashes and stack in LDS; or bitmap
  // and label hashes in LDS with the stack in HBM; or all in HBM.
  const uint64_t nb = (uint64_t)s.W * s.rows_pad;
  const bool in_lds = a.lds_tier && lds_words(s) * 8 <= a.lds_bytes;
  const bool bm_lds = a.lds_tier && (nb + s.p) * 8 <= a.lds_bytes;
  lds64 *st = nullptr;
  if (bm_lds) {
    st = stage_lds(s, lds);
    wave_fence();
    if (in_lds)
      materialize(s, snap, (uint32_t)idx, st, (const lds64 *)lds, (const lds64 *)(lds + nb), acc);
    else
      materialize(s, snap, (uint32_t)idx, stk_g, (const lds64 *)lds, (const lds64 *)(lds + nb), acc);
  } else {
    materialize(s, snap, (uint32_t)idx, stk_g, s.bm, s.gP, acc);
  }
  if (l == 0) __hip_atomic_fetch_sub(&r->pending, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
  uint64_t t0 = clock64();
  bool pub_unused = false;
  uint64_t nodes;
  if (in_lds)
    nodes = run_dfs(a, s, (const lds64 *)lds, (const lds64 *)(lds + nb), st, stk_g, acc, s_hist, me, pub_unused);
  else if (bm_lds)
    nodes = run_dfs(a, s, (const lds64 *)lds, (const lds64 *)(lds + nb), stk_g, stk_g, acc, s_hist, me,
                    pub_unused);
  else
    nodes = run_dfs(a, s, s.bm, s.gP, stk_g, stk_g, acc, s_hist, me, pub_unused);
  if (a.root_graph) flush_graph(a, s.graph, acc, s_hist);
  if (l == 0) {
    if (a.root_nodes && nodes) atomicAdd(&a.root_nodes[s.canon], (unsigned long long)nodes);
    if (a.root_cycles) atomicAdd(&a.root_cycles[s.canon], (unsigned long long)(clock64() - t0));
    atomicAdd(a.total_nodes, (unsigned long long)nodes);
    atomicAdd(a.shared_tasks, 1ull);
    __hip_atomic_fetch_sub(&a.users[s.bm_owner], 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
  }
  return true;
}

__device__ __forceinline__ void mce_main(const KArgs &a) {
  // The wave's LDS: the clique-size histogram of sizes < kLdsHist, then the
  // area of the LDS tiers (lds_bytes).
  ldsu64 *s_hist = (ldsu64 *)lds_dyn;
  lds64 *lds = (lds64 *)lds_dyn + kLdsHist;
  const uint32_t l = lane();
  for (uint32_t i = l; i < kLdsHist; i += 64) s_hist[i] = 0;
  __syncthreads();
  const uint32_t me = blockIdx.x;
  uint8_t *slot = a.slots + (uint64_t)me * a.slot_bytes;
  uint64_t *stk = (uint64_t *)(a.stacks + (uint64_t)me * a.stack_bytes);
  Acc acc{0, 0, 0, 0, 0};
  unsigned long long nodes_total = 0;
  // A wave is working (counted in `active`) until it finds neither a root
  // nor a task; it is then looking for work (counted in the idle copies)
  // until it gets either.  Only these transitions touch the shared
  // counters, not every root.  The run ends for a wave that is looking for
  // work when no roots are left and no wave is working: only working waves
  // publish tasks, and a thief becomes working before the publisher it
  // took from can finish.
  const uint64_t n_dispatch = a.n_dispatch_dev ? ld_relaxed((unsigned long long *)a.n_dispatch_dev) : a.n_dispatch;
  bool roots_left = true, is_idle = false;
  bool published = false;  // others may be reading this wave's data region
  uint32_t backoff = 1;
  if (l == 0) atomicAdd(a.active, 1u);
  for (;;) {
    bool did = false;
    if (roots_left) {
      bool free_slot = true;
      if (published) {
        if (l == 0) free_slot = ld_relaxed(&a.users[me]) == 0;
        free_slot = bcast32(free_slot);
        if (free_slot) {
          if (l == 0) (void)ld_acquire(&a.users[me]);  // order our rebuild after their reads
          published = false;
        }
      }
      if (free_slot) {
        unsigned long long i = 0;
        if (l == 0) i = atomicAdd(a.next_root, 1ull);
        i = bcast64(i);
        if (i < n_dispatch) {
          if (is_idle) {
            idle_add(a, -1);
            if (l == 0) atomicAdd(a.active, 1u);
            is_idle = false;
          }
          uint64_t canon = a.perm ? a.perm[i] : (a.reverse ? n_dispatch - 1 - i : i);
          uint64_t t0 = clock64();
          const bool eager = a.eager && a.eager[i];
          uint64_t nodes = solve_root(a, slot, stk, lds, me, canon, acc, s_hist, published, eager);
          nodes_total += nodes;
          if (a.root_graph) flush_graph(a, a.root_graph[canon], acc, s_hist);
          if (l == 0) {
            if (a.root_nodes && nodes) atomicAdd(&a.root_nodes[canon], (unsigned long long)nodes);
            if (a.root_cycles) atomicAdd(&a.root_cycles[canon], (unsigned long long)(clock64() - t0));
          }
          did = true;
          backoff = 1;
        } else {
          roots_left = false;
        }
      }
    }
    if (did) continue;
    if (!is_idle) {  // working -> looking for work
      idle_add(a, 1);
      if (l == 0) atomicSub(a.active, 1u);
      is_idle = true;
    }
    if (a.share_bound > 0 && steal(a, me, stk, lds, acc, s_hist, is_idle)) {
      backoff = 1;
      continue;
    }
    bool done = false;
    if (l == 0) done = !roots_left && ld_relaxed(a.active) == 0;
    if (bcast32(done)) break;
    if (l == 0) nap(backoff);
    backoff = backoff < 64 ? backoff * 2 : 64;
  }
  __syncthreads();
  acc.build += 4 * wave_sum64(acc.scan);
  if (l == 0) {
    if (acc.count) atomicAdd(a.count, acc.count);
    if (acc.hash) atomicAdd(a.hash, acc.hash);
    if (nodes_total) atomicAdd(a.total_nodes, nodes_total);
    if (acc.words) atomicAdd(a.search_words, acc.words);
    if (acc.build) atomicAdd(a.build_bytes, acc.build);
  }
  for (uint32_t i = l; i < kLdsHist && i < a.hist_len; i += 64)
    if (s_hist[i]) atomicAdd(&a.hist[i], s_hist[i]);
}

// One persistent kernel per occupancy target: the register budget that
// lets WPE waves share a SIMD (512 VGPRs / WPE).  The search is latency-bound
// (dependent LDS and L2 round trips per tree node), so more resident waves
// hide more latency, at the price of LDS per wave and a few spilled values.
template <int WPE>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(WPE, 8))) void mce_kernel(KArgs a) {
  mce_main(a);
}

}  // namespace

hipError_t launch_mce(const KArgs &a, uint32_t waves, uint32_t waves_per_cu, hipStream_t stream) {
  const size_t lds = kLdsHist * 8 + (a.lds_tier ? a.lds_bytes : 0);
  if (waves_per_cu <= 16)
    hipLaunchKernelGGL(mce_kernel<4>, dim3(waves), dim3(64), lds, stream, a);
  else if (waves_per_cu <= 20)
    hipLaunchKernelGGL(mce_kernel<5>, dim3(waves), dim3(64), lds, stream, a);
  else if (waves_per_cu <= 24)
    hipLaunchKernelGGL(mce_kernel<6>, dim3(waves), dim3(64), lds, stream, a);
  else
    hipLaunchKernelGGL(mce_kernel<8>, dim3(waves), dim3(64), lds, stream, a);
  return hipGetLastError();
}

}  // namespace rdmce

namespace rdmce {
namespace {
// Streaming copy used to calibrate hardware traffic counters (FETCH_SIZE,
// WRITE_SIZE) against a kernel of known byte count.
__global__ __launch_bounds__(256) void copy_probe_kernel(const uint4 *__restrict__ src, uint4 *__restrict__ dst,
                                                         uint64_t n) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}
}  // namespace

namespace {
__global__ __launch_bounds__(256) void fill_zero_kernel(uint64_t *p, uint64_t n) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = 0;
}
}  // namespace

hipError_t launch_fill_zero(void *p, uint64_t n_words, hipStream_t stream) {
  if (n_words == 0) return hipSuccess;
  uint32_t blocks = (uint32_t)std::min<uint64_t>(2048, (n_words + 255) / 256);
  hipLaunchKernelGGL(fill_zero_kernel, dim3(blocks), dim3(256), 0, stream, (uint64_t *)p, n_words);
  return hipGetLastError();
}

hipError_t launch_copy_probe(const void *src, void *dst, uint64_t bytes, hipStream_t stream) {
  uint64_t n = bytes / 16;
  uint32_t blocks = (uint32_t)std::min<uint64_t>(8192, (n + 255) / 256 + 1);
  hipLaunchKernelGGL(copy_probe_kernel, dim3(blocks), dim3(256), 0, stream, (const uint4 *)src, (uint4 *)dst, n);
  return hipGetLastError();
}
}  // namespace rdmce
