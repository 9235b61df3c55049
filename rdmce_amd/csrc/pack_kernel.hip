// Packed search for gfx950: many small roots per wavefront, one tree node
// per lane.
//
// The one-wave-per-root kernel (mce_kernel.hip, the reference's layout:
// BkPivotBitBalance.cuh:510-608) spends a whole wave on one tree node at a
// time, so a node of a root with |P0| <= 64 costs hundreds of wave
// instructions for a few words of set work.  Here a wave keeps up to
// pk_slots subproblems resident in its LDS and a stack of tree-node records
// in HBM; every pass of its loop hands up to 64 pending tasks to the 64
// lanes, and each lane computes one tree node on its own:
//
//   task    (record, candidate t of the record's V): the child of t as the
//           sequential search makes it, P' = (P \ {v in V : v < t}) & N(t),
//           XP' = (XP | {v in V : v < t}) & N(t), XX' = XX & cols(t)
//           (BkPivotBitBalance.cuh:40-70: the k-th candidate of a published
//           level, which is what the reference's thieves materialise),
//           followed by aggressive node generation (:351-382);
//   node    a leaf is counted when XP' and XX' are empty; otherwise the
//           pivot is the (|P' \ N(u)|, lowest row) argmin over the rows u
//           of P' | XP' | XX' (:229-257), and a node whose P' an excluded
//           row covers is pruned; a branching node pushes its record, whose
//           candidates V = P' \ N(pivot) become |V| tasks.
//
// Every child is a function of its parent record and its candidate alone,
// so the order in which lanes take tasks never changes the search tree:
// node counts, cliques, hash and histogram equal the one-wave-per-root
// kernel's and the oracle's.  Dispatch takes tasks from the top of the stack
// (the newest records), which keeps the stack near depth x 64 records;
// beyond pk_thr records only the top record is expanded (plain depth-first
// order), which bounds it by pk_thr + 64 (64 NW + 3) records.
//
// A subproblem occupies an LDS slot from its build until its last record is
// consumed:  rows of P0 and of X0' over P0 (NW words), the columns of the X0'
// block for each P0 vertex (NX words: XX' = XX & cols(t) is one load), and
// the label hashes of P0.  Roots with |P0| > 64 NW or |X0'| > 64 NX are
// appended to a.wide for the next launch of the run.  When a.wide_h is set,
// the vertex bitmap of a run that met such roots is copied to a.pk_pool
// (while it has room) and the next launch derives each of them from that
// copy (a.pk_whandle) instead of building it again.  A derived root with
// too many X0' rows whose top node one of them prunes (it covers P0) is
// settled there, as one tree node, and goes nowhere.
#include "mce_common.hpp"

extern __shared__ uint64_t lds_dyn[];

namespace rdmce {
namespace {

constexpr uint32_t kRecRoot = 1u << 24;  // meta flag: the record's one task is its root's top node

template <int NW, int NX>
struct PackCfg {
  static constexpr uint32_t RS = pack_rec_words(NW, NX);
  static constexpr uint32_t SW = pack_slot_words(NW, NX);
  static constexpr uint32_t HS = pack_hist_words(NW);
  // record fields (each a column of pk_cap words in the wave's stack):
  // P XP V (NW each), XX (NX), key, meta, R (NW, read and written only for
  // clique lists).  meta = first << 32 | root flag | slot << 16 | |R|, where
  // the record's remaining candidates are the members of V from `first` on.
  static constexpr uint32_t oP = 0, oX = NW, oV = 2 * NW, oXX = 3 * NW, oKey = 3 * NW + NX, oMeta = 3 * NW + NX + 1,
                            oR = 3 * NW + NX + 2;
  // slot word offsets
  static constexpr uint32_t sProw = 0, sXrow = 64 * NW * NW, sXcol = sXrow + 64 * NX * NW, sG = sXcol + 64 * NW * NX;
};

// Position of the k-th (0-based) set bit of w; k < popcount(w).
__device__ __forceinline__ uint32_t select64(uint64_t w, uint32_t k) {
  uint32_t pos = 0, c;
  c = (uint32_t)__popcll(w & 0xFFFFFFFFull);
  if (k >= c) { k -= c; w >>= 32; pos = 32; }
  c = (uint32_t)__popcll(w & 0xFFFFull);
  if (k >= c) { k -= c; w >>= 16; pos += 16; }
  c = (uint32_t)__popcll(w & 0xFFull);
  if (k >= c) { k -= c; w >>= 8; pos += 8; }
  c = (uint32_t)__popcll(w & 0xFull);
  if (k >= c) { k -= c; w >>= 4; pos += 4; }
  c = (uint32_t)__popcll(w & 0x3ull);
  if (k >= c) { k -= c; w >>= 2; pos += 2; }
  if (k >= (uint32_t)(w & 1)) pos += 1;
  return pos;
}

// The per-wave LDS areas (word offsets from lds_dyn).
struct PackLds {
  ldsu64 *hist;     // kLdsHist clique sizes (outside batches)
  lds64 *map;       // 64 u32: dispatch map
  lds64 *meta;      // kPackMetaWords x S
  lds64 *bhist;     // batches: HS x S per-slot clique sizes
  lds64 *slots;     // SW x S
};

// Slot bookkeeping, kPackMetaWords arrays of S words.
enum PackMeta { mCanon, mKey, mPX, mBase, mV, mNodes, mRefs, mCount, mHash };

template <int NW, int NX>
struct PackWave {
  using C = PackCfg<NW, NX>;
  const KArgs &a;
  PackLds L;
  uint32_t S;
  uint32_t me;

  __device__ __forceinline__ lds64 *m(PackMeta f, uint32_t s) const { return L.meta + (uint32_t)f * S + s; }
  __device__ __forceinline__ lds64 *slot(uint32_t s) const { return L.slots + s * C::SW; }
};

// Counts one maximal clique of slot s (size, key = sum of label hashes),
// with R's members given by rb (bits over P0).  Per lane.
template <int NW, int NX, bool LISTS>
__device__ __forceinline__ void pk_emit(const PackWave<NW, NX> &w, uint32_t s, uint32_t size, uint64_t key,
                                        const uint64_t (&rb)[NW], uint64_t &cnt, uint64_t &hsh) {
  using C = PackCfg<NW, NX>;
  const KArgs &a = w.a;
  const uint64_t h = splitmix64_mix(key);
  if (a.root_graph) {
    __hip_atomic_fetch_add(w.m(mCount, s), (uint64_t)1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    __hip_atomic_fetch_add(w.m(mHash, s), h, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    __hip_atomic_fetch_add(w.L.bhist + s * C::HS + size, (uint64_t)1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
  } else {
    cnt++;
    hsh += h;
    if (size < kLdsHist) {
      __hip_atomic_fetch_add(w.L.hist + size, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    } else if (size < a.hist_len) {
      atomicAdd(&a.hist[size], 1ull);
    } else {
      atomicAdd(&a.hist[a.hist_len - 1], 1ull);
      atomicOr(a.err, kErrHistOverflow);
    }
  }
  if (LISTS && a.list) {
    // entry: size, the root's vertices, then R's members in P0 order
    const unsigned long long pos = atomicAdd(a.list_cursor, (unsigned long long)(size + 1));
    if (pos + size + 1 > a.list_cap) {
      atomicOr(a.err, kErrListFull);
    } else {
      uint32_t *dst = a.list + pos;
      const uint64_t bg = *w.m(mBase, s), vv = *w.m(mV, s);
      const uint32_t *pl = a.pk_plist + ((uint64_t)w.me * w.S + s) * kPackMaxP;
      dst[0] = size;
      dst[1] = a.labels[(uint32_t)vv];
      uint32_t k = 1;
      if ((uint32_t)bg == 2) dst[++k] = a.labels[(uint32_t)(vv >> 32)];
      for (int q = 0; q < NW; ++q)
        for (uint64_t r = rb[q]; r; r &= r - 1) dst[++k] = a.labels[pl[q * 64 + __builtin_ctzll(r)]];
    }
  }
}

// Completes slot s once its P0 rows, X0' rows, label hashes (and clique-list
// ids) are in place: the columns of the X0' block and the bookkeeping.
template <int NW, int NX>
__device__ __forceinline__ void pk_seal(const PackWave<NW, NX> &w, uint32_t s, uint64_t p, uint64_t xq, uint64_t canon,
                                        uint64_t base_key, uint32_t base_size, uint32_t graph, uint32_t v0,
                                        uint32_t v1) {
  using C = PackCfg<NW, NX>;
  const KArgs &a = w.a;
  const uint32_t l = lane();
  lds64 *sl = w.slot(s);
  wave_fence();
  // Columns of the X0' block: bit x of column j is bit j of row x.  Lane l
  // holds row 64 b + l of block b; one ballot per P0 vertex and block.
#pragma unroll
  for (int b = 0; b < NX; ++b) {
    uint64_t row[NW];
#pragma unroll
    for (int q = 0; q < NW; ++q) row[q] = (64u * b + l < xq) ? (uint64_t)sl[C::sXrow + (64u * b + l) * NW + q] : 0ull;
#pragma unroll
    for (int q = 0; q < NW; ++q) {
      uint64_t mine = 0;
      for (uint32_t j = 0; j < 64; ++j) {
        const uint64_t col = __ballot((row[q] >> j) & 1);
        if (l == j) mine = col;
      }
      if (64u * q + l < p) sl[C::sXcol + (64u * q + l) * NX + b] = mine;
    }
  }
  if (l == 0) {
    *w.m(mCanon, s) = canon;
    *w.m(mKey, s) = base_key;
    *w.m(mPX, s) = p | (xq << 32);
    *w.m(mBase, s) = (uint64_t)base_size | ((uint64_t)graph << 32);
    *w.m(mV, s) = (uint64_t)v0 | ((uint64_t)v1 << 32);
    *w.m(mNodes, s) = 0;
    *w.m(mRefs, s) = 1;  // the root record pushed next
    *w.m(mCount, s) = 0;
    *w.m(mHash, s) = 0;
  }
  if (a.root_graph)
    for (uint32_t i = l; i < C::HS; i += 64) w.L.bhist[s * C::HS + i] = 0;
  wave_fence();
}

// Counts the root's own vertex set {v0, v1} (or {v0}) as a maximal clique:
// a root whose candidate and excluded sets are both empty.  Lane 0.
template <int NW, int NX>
__device__ __forceinline__ void pk_emit_base(const PackWave<NW, NX> &w, uint64_t base_key, uint32_t base_size,
                                             uint32_t graph, uint32_t v0, uint32_t v1, uint64_t &cnt, uint64_t &hsh) {
  const KArgs &a = w.a;
  if (lane() != 0) return;
  const uint64_t h = splitmix64_mix(base_key);
  if (a.root_graph) {
    atomicAdd(&a.g_count[graph], 1ull);
    atomicAdd(&a.g_hash[graph], h);
    atomicAdd(&a.g_hist[(uint64_t)graph * a.g_hist_stride + base_size], 1ull);
  } else {
    cnt++;
    hsh += h;
    __hip_atomic_fetch_add(w.L.hist + base_size, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
  }
  if (a.list) {
    const unsigned long long pos = atomicAdd(a.list_cursor, (unsigned long long)(base_size + 1));
    if (pos + base_size + 1 > a.list_cap) {
      atomicOr(a.err, kErrListFull);
    } else {
      a.list[pos] = base_size;
      a.list[pos + 1] = a.labels[v0];
      if (base_size == 2) a.list[pos + 2] = a.labels[v1];
    }
  }
}

// Builds root `canon` into slot s (the whole wave).  Returns 1 when the root
// occupies the slot, 0 when it needs no search (its 2-clique, if maximal,
// is counted here), 2 when it is too wide for this variant (the caller hands
// it to the next launch).
template <int NW, int NX>
__device__ __forceinline__ uint32_t pk_build(const PackWave<NW, NX> &w, uint8_t *scratch, uint32_t s, uint64_t canon,
                                             Acc &acc, uint64_t &cnt, uint64_t &hsh) {
  using C = PackCfg<NW, NX>;
  const KArgs &a = w.a;
  const uint32_t l = lane();
  Sub sub;
  sub.canon = canon;
  sub.graph = a.root_graph ? a.root_graph[canon] : 0;
  uint32_t *Pl = (uint32_t *)scratch;
  const uint64_t ccap = a.slot_bytes / 4;
  uint64_t p = 0, x = 0;
  const uint32_t above = root_lists(a, canon, Pl, ccap, sub, p, x, acc);
  if (p == 0) {
    if (x == 0) pk_emit_base(w, sub.base_key, sub.base_size, sub.graph, sub.v0, sub.v1, cnt, hsh);
    return 0;
  }
  if (p > 64u * NW || p + x > ccap) {
    return 2;
  }
  wave_fence();
  const uint32_t *P = Pl;
  uint32_t *X = Pl + p;
  const uint64_t xq = compact_x(a, P, p, X, x, above);
  wave_fence();
  if (xq > 64u * NX) {
    return 2;
  }
  const uint64_t rows = p + xq;
  // the bitmap tiers' accounting: both lists and a W x rows_pad bitmap
  acc.build += 4 * (p + x) + 8ull * ((p + 63) / 64) * align_up(rows, 64);
  lds64 *sl = w.slot(s);
  for (uint32_t i = l; i < C::sXcol; i += 64) sl[i] = 0;  // P0 rows and X0' rows
  for (uint32_t i = l; i < p; i += 64) sl[C::sG + i] = a.label_hash[P[i]];
  if (a.list) {
    uint32_t *pl = a.pk_plist + ((uint64_t)w.me * w.S + s) * kPackMaxP;
    for (uint32_t i = l; i < p; i += 64) pl[i] = P[i];
  }
  wave_fence();
  acc.scan += match_rows(
      a, P, p, X, rows, above,
      [&](uint64_t i, uint32_t j) {
        or64_lds(sl + C::sProw + i * NW + (j >> 6), 1ull << (j & 63));
        or64_lds(sl + C::sProw + (uint64_t)j * NW + (i >> 6), 1ull << (i & 63));
      },
      [&](uint64_t i, uint32_t wd, uint64_t bits) { or64_lds(sl + C::sXrow + (i - p) * NW + wd, bits); });
  wave_fence();
  pk_seal<NW, NX>(w, s, p, xq, canon, sub.base_key, sub.base_size, sub.graph, sub.v0, sub.v1);
  return 1;
}

// ---------------------------------------------------------------------------
// Edge roots derived from their first vertex's bitmap.
//
// Every edge root (v0, u_k) with u_k the k-th vertex of U = N+(v0) is a
// restriction of the vertex root v0 (BkPivotBitBalance.cuh:414-427 builds it
// as the LEVEL_1 root): P0 = {u_j : j > k, u_j ~ u_k}, X0 = {x ~ v0 : x < v0,
// x ~ u_k} then {u_j : j < k, u_j ~ u_k} (ascending vertex order), and X0'
// keeps the rows of X0 with a neighbour in P0.  A wave that takes a run of
// edge roots of one v0 therefore builds the vertex root's bitmap once, in its
// scratch region (rows U then X0v' = the x ~ v0 below v0 with a neighbour in
// U, word-major), and derives each edge root from it by bit extraction:
// rows of the edge root are the vertex rows restricted to P0's columns and
// compressed to P0's indices.  The subproblem (rows, order, label hashes) is
// exactly the one pk_build makes, so the search tree is unchanged.
// ---------------------------------------------------------------------------
constexpr uint32_t kGroupMaxW = 16;  // |U| <= 1024
constexpr uint32_t kPoolHdr = 8;     // kept vertex bitmap: v0 d0 W0 xq0 rp o0 rfirst lh0, then T, cm and lh

struct PkGroup {
  uint32_t v0, d0, W0, xq0, rp;
  uint64_t o0;          // U = col[o0, o0 + d0)
  uint64_t rfirst;      // canonical id of the edge root (v0, U[0])
  const uint64_t *T;    // W0 x rp words: rows 0..d0-1 (U), d0..d0+xq0-1 (X0v')
  const uint64_t *cm;   // W0 words: OR of the X0v' rows (U vertices with a neighbour below v0)
  uint64_t *wide;       // the run's roots too wide for this variant, kept until the run ends
  const uint64_t *lh;   // d0 words: splitmix64(label) of each member of U
  uint64_t lh0;         // splitmix64(label(v0))
};

// Builds the bitmap of vertex root v0 into `scratch` (the whole wave), with
// room for `run` wide root ids after it; `claim` is an LDS word for handing
// out rows.  Returns false when it does not fit (the caller then builds each
// edge root on its own).
__device__ __forceinline__ bool pk_group_build(const KArgs &a, uint8_t *scratch, uint32_t v0, uint64_t run, PkGroup &G,
                                               Acc &acc, ldsu32 *claim) {
  const uint32_t l = lane();
  const uint64_t b0 = a.rowoff[v0], e0 = a.rowoff[(uint64_t)v0 + 1], o0 = a.outoff[v0];
  const uint64_t d0 = e0 - o0, x0 = o0 - b0;
  const uint64_t W0 = (d0 + 63) / 64;
  if (W0 == 0 || W0 > kGroupMaxW) return false;
  const uint64_t t_off = align_up(4 * x0, 256);
  const uint64_t rp_max = align_up(d0 + x0, 64);
  const uint64_t cm_off = t_off + align_up(8 * W0 * rp_max, 256);
  const uint64_t wide_off = cm_off + align_up(8 * W0, 256);
  const uint64_t lh_off = wide_off + align_up(8 * run, 256);
  if (lh_off + 8 * d0 > a.slot_bytes) return false;
  uint32_t *Xs = (uint32_t *)scratch;
  for (uint64_t i = l; i < x0; i += 64) Xs[i] = a.col[b0 + i];
  wave_fence();
  const uint32_t *U = a.col + o0;
  const uint64_t xq0 = compact_x(a, U, d0, Xs, x0, v0);
  wave_fence();
  const uint64_t rows = d0 + xq0, rp = align_up(rows, 64);
  uint64_t *T = (uint64_t *)(scratch + t_off);
  uint64_t *cm = (uint64_t *)(scratch + cm_off);
  for (uint64_t i = l; i < W0 * rp; i += 64) T[i] = 0;
  wave_fence();
  acc.scan += match_rows(
      a, U, d0, Xs, rows, v0,
      [&](uint64_t i, uint32_t j) {
        or64(&T[(uint64_t)(j >> 6) * rp + i], 1ull << (j & 63));
        or64(&T[(uint64_t)(i >> 6) * rp + j], 1ull << (i & 63));
      },
      [&](uint64_t i, uint32_t wd, uint64_t bits) { or64(&T[(uint64_t)wd * rp + i], bits); }, claim);
  wave_fence();
  if (l < W0) {
    uint64_t o = 0;
    for (uint64_t x = 0; x < xq0; ++x) o |= T[(uint64_t)l * rp + d0 + x];
    cm[l] = o;
  }
  wave_fence();
  acc.build += 4 * (e0 - b0) + 4 * (d0 + x0) + 8 * W0 * rp;
  G.v0 = v0;
  G.d0 = (uint32_t)d0;
  G.W0 = (uint32_t)W0;
  G.xq0 = (uint32_t)xq0;
  G.rp = (uint32_t)rp;
  G.o0 = o0;
  G.rfirst = a.rfirst[v0];
  G.T = T;
  G.cm = cm;
  G.wide = (uint64_t *)(scratch + wide_off);
  uint64_t *lh = (uint64_t *)(scratch + lh_off);
  for (uint64_t j = l; j < d0; j += 64) lh[j] = a.label_hash[a.col[o0 + j]];
  wave_fence();
  G.lh = lh;
  G.lh0 = a.label_hash[v0];
  return true;
}

// Bits of `row` at the set positions of the mask (held one word per lane,
// lanes < W0), compressed in order into NW words.  The mask is uniform, so
// every lane runs the same loop.
template <int NW>
__device__ __forceinline__ void pext_row(const PkGroup &G, uint64_t mk, uint32_t row, bool act, uint64_t (&out)[NW]) {
#pragma unroll
  for (int q = 0; q < NW; ++q) out[q] = 0;
  uint32_t c = 0;
  for (uint32_t q = 0; q < G.W0; ++q) {
    uint64_t m = readlane64(mk, q);
    if (m == 0) continue;
    const uint64_t rw = act ? G.T[(uint64_t)q * G.rp + row] : 0ull;
    for (; m; m &= m - 1, ++c) {
      const uint64_t bit = (rw >> __builtin_ctzll(m)) & 1;
      if constexpr (NW == 1) {
        out[0] |= bit << c;
      } else {
        if (c < 64) out[0] |= bit << c;
        else out[1] |= bit << (c - 64);
      }
    }
  }
}

// Derives edge root `canon` = (v0, u_k) of group G into slot s (the whole
// wave).  Returns as pk_build does, or 3 when the root's top node is pruned
// by an X0' row covering P0 (its search words in `pruned`).
template <int NW, int NX>
__device__ __forceinline__ uint32_t pk_derive(const PackWave<NW, NX> &w, const PkGroup &G, uint32_t s, uint64_t canon,
                                              Acc &acc, uint64_t &cnt, uint64_t &hsh, uint32_t &pruned) {
  using C = PackCfg<NW, NX>;
  const KArgs &a = w.a;
  const uint32_t l = lane();
  const uint32_t k = (uint32_t)(canon - G.rfirst);
  const uint32_t uk = a.list ? a.col[G.o0 + k] : 0u;  // the root's v1 (clique lists only)
  const uint32_t graph = a.root_graph ? a.root_graph[canon] : 0;
  const uint64_t base_key = G.lh0 + G.lh[k];
  // Row k of U, split at k: P0 above it, the U part of X0 below it (one
  // word per lane).
  const uint32_t q = l;
  const uint64_t ak = q < G.W0 ? G.T[(uint64_t)q * G.rp + k] : 0ull;
  const uint32_t kq = k >> 6, kb = k & 63;
  const uint64_t above = q > kq ? ~0ull : (q == kq ? (kb == 63 ? 0ull : ~0ull << (kb + 1)) : 0ull);
  const uint64_t below = q < kq ? ~0ull : (q == kq ? low_mask(kb) : 0ull);
  const uint64_t pk = ak & above, xu = ak & below;
  const uint32_t p = __builtin_amdgcn_readlane(wave_scan_add_u32((uint32_t)__popcll(pk)), 63);
  uint64_t words_read = G.W0;
  if (p == 0) {
    const bool xlow = (readlane64(q < G.W0 ? G.cm[q] : 0ull, kq) >> kb) & 1;
    if (!xlow && __ballot(xu != 0) == 0) pk_emit_base(w, base_key, 2, graph, G.v0, uk, cnt, hsh);
    acc.build += 8 * (words_read + 1);
    return 0;
  }
  if (p > 64u * NW) {
    return 2;
  }
  // X0' rows, in vertex order: rows x of X0v' adjacent to u_k, then U rows
  // j < k adjacent to u_k, each kept when it has a neighbour in P0.  Their
  // row numbers go to a list in the slot's column area (rewritten below).
  lds64 *sl = w.slot(s);
  ldsu32 *xl = (ldsu32 *)(sl + C::sXcol);
  uint32_t xq = 0;
  for (uint32_t base = 0; base < G.xq0; base += 64) {
    const uint32_t x = base + l;
    const bool cand = x < G.xq0 && ((G.T[(uint64_t)kq * G.rp + G.d0 + x] >> kb) & 1);
    bool keep = false;
    for (uint32_t z = 0; z < G.W0; ++z) {  // uniform loop: the mask word is a readlane
      const uint64_t mz = readlane64(pk, z);
      if (cand && !keep) keep = (G.T[(uint64_t)z * G.rp + G.d0 + x] & mz) != 0;
    }
    const uint64_t m = __ballot(keep);
    if (keep && xq + rank_below(m) < 64u * NX) xl[xq + rank_below(m)] = G.d0 + x;
    xq += (uint32_t)__popcll(m);
  }
  words_read += (G.xq0 + 63) / 64 * 64;
  const uint32_t nxu = __builtin_amdgcn_readlane(wave_scan_add_u32((uint32_t)__popcll(xu)), 63);
  for (uint32_t base = 0; base < nxu; base += 64) {
    // lane l takes the (base + l)-th U row below k adjacent to u_k
    const uint32_t t = base + l;
    uint32_t j = 0, before = 0;
    for (uint32_t z = 0; z <= kq; ++z) {
      const uint64_t wz = readlane64(xu, z);
      const uint32_t cz = (uint32_t)__popcll(wz);
      if (t >= before && t < before + cz) j = 64 * z + select64(wz, t - before);
      before += cz;
    }
    bool keep = false;
    for (uint32_t z = 0; z < G.W0; ++z) {
      const uint64_t mz = readlane64(pk, z);
      if (t < nxu && !keep) keep = (G.T[(uint64_t)z * G.rp + j] & mz) != 0;
    }
    const uint64_t m = __ballot(keep);
    if (keep && xq + rank_below(m) < 64u * NX) xl[xq + rank_below(m)] = j;
    xq += (uint32_t)__popcll(m);
  }
  words_read += (uint64_t)nxu * G.W0;
  if (xq > 64u * NX) {
    // Too many X0' rows for this variant.  When one of them covers P0, the
    // root's top node is pruned (no maximal clique below it): one node, no
    // slot.  Otherwise the root goes to the next launch.
    bool cov = false;
    for (uint32_t base = 0; base < G.xq0; base += 64) {
      const uint32_t x = base + l;
      const bool cand = x < G.xq0 && ((G.T[(uint64_t)kq * G.rp + G.d0 + x] >> kb) & 1);
      uint32_t c = 0;
      for (uint32_t z = 0; z < G.W0; ++z) {
        const uint64_t mz = readlane64(pk, z);
        if (cand) c += (uint32_t)__popcll(G.T[(uint64_t)z * G.rp + G.d0 + x] & mz);
      }
      cov |= cand && c == p;
    }
    for (uint32_t base = 0; base < nxu; base += 64) {
      const uint32_t t = base + l;
      uint32_t j = 0, before = 0;
      for (uint32_t z = 0; z <= kq; ++z) {
        const uint64_t wz = readlane64(xu, z);
        const uint32_t cz = (uint32_t)__popcll(wz);
        if (t >= before && t < before + cz) j = 64 * z + select64(wz, t - before);
        before += cz;
      }
      uint32_t c = 0;
      for (uint32_t z = 0; z < G.W0; ++z) {
        const uint64_t mz = readlane64(pk, z);
        if (t < nxu) c += (uint32_t)__popcll(G.T[(uint64_t)z * G.rp + j] & mz);
      }
      cov |= t < nxu && c == p;
    }
    if (__ballot(cov) == 0) return 2;
    words_read += (uint64_t)(G.xq0 + nxu) * G.W0;
    acc.build += 8 * words_read;
    pruned = (p + xq) * ((p + 63) / 64);  // the search words of evaluating the top node
    return 3;
  }
  wave_fence();
  // Row numbers (in G) of the edge root's rows: P0 (the set bits of pk in
  // order) and X0' (the list).
  uint32_t prow_id[NW], xrow_id[NX];
#pragma unroll
  for (int r = 0; r < NW; ++r) {
    const uint32_t t = 64u * r + l;
    uint32_t j = 0, before = 0;
    for (uint32_t z = 0; z < G.W0; ++z) {
      const uint64_t wz = readlane64(pk, z);
      const uint32_t cz = (uint32_t)__popcll(wz);
      if (t >= before && t < before + cz) j = 64 * z + select64(wz, t - before);
      before += cz;
    }
    prow_id[r] = j;
  }
#pragma unroll
  for (int r = 0; r < NX; ++r) {
    const uint32_t t = 64u * r + l;
    xrow_id[r] = t < xq ? (uint32_t)xl[t] : 0u;
  }
  wave_fence();  // the list is read before the column area is rewritten
  for (uint32_t i = l; i < C::sXcol; i += 64) sl[i] = 0;
#pragma unroll
  for (int r = 0; r < NW; ++r) {
    const uint32_t t = 64u * r + l;
    uint64_t o[NW];
    pext_row<NW>(G, pk, prow_id[r], t < p, o);
    if (t < p) {
#pragma unroll
      for (int z = 0; z < NW; ++z) sl[C::sProw + t * NW + z] = o[z];
      sl[C::sG + t] = G.lh[prow_id[r]];
      if (a.list) a.pk_plist[((uint64_t)w.me * w.S + s) * kPackMaxP + t] = a.col[G.o0 + prow_id[r]];
    }
  }
#pragma unroll
  for (int r = 0; r < NX; ++r) {
    const uint32_t t = 64u * r + l;
    uint64_t o[NW];
    pext_row<NW>(G, pk, xrow_id[r], t < xq, o);
    if (t < xq) {
#pragma unroll
      for (int z = 0; z < NW; ++z) sl[C::sXrow + t * NW + z] = o[z];
    }
  }
  words_read += (uint64_t)(p + xq) * G.W0;
  acc.build += 8 * words_read + 8ull * NW * (p + xq);
  wave_fence();
  pk_seal<NW, NX>(w, s, p, xq, canon, base_key, 2, graph, G.v0, uk);
  return 1;
}

// Moves what slot s accumulated (batches: count, hash, clique sizes; root
// statistics: its node count) to the results.  Per lane (lane s).
template <int NW, int NX>
__device__ __forceinline__ void pk_finish(const PackWave<NW, NX> &w, uint32_t s) {
  using C = PackCfg<NW, NX>;
  const KArgs &a = w.a;
  const uint64_t canon = *w.m(mCanon, s);
  if (a.root_nodes) atomicAdd(&a.root_nodes[canon], (unsigned long long)*w.m(mNodes, s));
  if (a.root_graph) {
    const uint32_t g = (uint32_t)(*w.m(mBase, s) >> 32);
    const uint64_t c = *w.m(mCount, s), h = *w.m(mHash, s);
    if (c) atomicAdd(&a.g_count[g], (unsigned long long)c);
    if (h) atomicAdd(&a.g_hash[g], (unsigned long long)h);
    for (uint32_t i = 0; i < C::HS; ++i) {
      const uint64_t v = w.L.bhist[s * C::HS + i];
      if (v) {
        if (i < a.g_hist_stride) atomicAdd(&a.g_hist[(uint64_t)g * a.g_hist_stride + i], (unsigned long long)v);
        else atomicOr(a.err, kErrHistOverflow);
      }
    }
  }
}

template <int NW, int NX, bool LISTS>
__device__ __forceinline__ void pack_main(const KArgs &a) {
  using C = PackCfg<NW, NX>;
  const uint32_t l = lane();
  const uint32_t S = a.pk_slots;
  lds64 *base = (lds64 *)lds_dyn;
  PackWave<NW, NX> w{a, {}, S, (uint32_t)blockIdx.x};
  w.L.hist = (ldsu64 *)base;
  w.L.map = base + kLdsHist;
  w.L.meta = w.L.map + 32;
  w.L.bhist = w.L.meta + kPackMetaWords * S;
  w.L.slots = w.L.bhist + (a.root_graph ? C::HS * S : 0);
  for (uint32_t i = l; i < kLdsHist; i += 64) w.L.hist[i] = 0;
  wave_fence();

  const uint64_t n_dispatch = a.n_dispatch_dev ? ld_relaxed((unsigned long long *)a.n_dispatch_dev) : a.n_dispatch;
  uint8_t *scratch = a.slots + (uint64_t)w.me * a.slot_bytes;
  // The wave's record stack, field-major: field f of record r is
  // stk[f * cap + r], so lanes touching neighbouring records touch
  // neighbouring words.
  uint64_t *stk = a.pk_stack + (uint64_t)w.me * a.pk_cap * C::RS;
  const uint64_t cap = a.pk_cap;
  constexpr bool lists = LISTS;  // clique lists: R is kept in the records
  auto fld = [stk, cap](uint32_t f, uint32_t r) -> uint64_t & { return stk[(uint64_t)f * cap + r]; };
  ldsu32 *map32 = (ldsu32 *)w.L.map;
  Acc acc{0, 0, 0, 0, 0};
  uint64_t cnt = 0, hsh = 0, words = 0, nodes = 0;  // per lane
  uint32_t top = 0;                                  // records in the stack
  uint64_t pending = 0;                              // candidates (tasks) in the stack
  uint32_t busy = 0;                                 // occupied slots
  const uint32_t full = S >= 32 ? ~0u : ((1u << S) - 1);
  bool roots_left = true;
  // Dispatch: runs of consecutive dispatch entries (edge roots of one v0,
  // a.pk_gstart) or single entries; a run of two or more edge roots shares
  // the vertex bitmap of its v0 (pk_group_build).
  uint64_t gcur = 0, gend = 0;
  bool grouped = false;
  uint32_t nwide = 0;  // wide roots of the current grouped run, in G.wide
  PkGroup G{};
  for (;;) {
    // Refill: build roots into free slots while fewer than 64 tasks wait.
    while (roots_left && pending < 64 && busy != full) {
      if (gcur == gend) {
        if (nwide) {  // the finished run's wide roots go out with a copy of its vertex bitmap
          wave_fence();
          const uint64_t tw = (uint64_t)G.W0 * G.rp, need = kPoolHdr + tw + G.W0 + G.d0;
          unsigned long long at = 0, hp = 0;
          if (l == 0) {
            at = atomicAdd(a.n_wide, (unsigned long long)nwide);
            hp = atomicAdd(a.pk_pool_used, (unsigned long long)need);
          }
          at = bcast64(at);
          hp = bcast64(hp);
          const bool kept = hp + need <= a.pk_pool_words;
          if (kept) {
            uint64_t *dst = a.pk_pool + hp;
            if (l < kPoolHdr) {
              const uint64_t hv[kPoolHdr] = {G.v0, G.d0, G.W0, G.xq0, G.rp, G.o0, G.rfirst, G.lh0};
              uint64_t v = 0;
#pragma unroll
              for (uint32_t f = 0; f < kPoolHdr; ++f) v = f == l ? hv[f] : v;
              dst[l] = v;
            }
            for (uint64_t i = l; i < tw; i += 64) dst[kPoolHdr + i] = G.T[i];
            for (uint64_t i = l; i < G.W0; i += 64) dst[kPoolHdr + tw + i] = G.cm[i];
            for (uint64_t i = l; i < G.d0; i += 64) dst[kPoolHdr + tw + G.W0 + i] = G.lh[i];
            acc.build += 8 * need;
          }
          for (uint32_t i = l; i < nwide; i += 64) {
            a.wide[at + i] = G.wide[i];
            a.wide_h[at + i] = kept ? hp : ~0ull;
          }
          nwide = 0;
        }
        unsigned long long gi = 0;
        if (l == 0) gi = atomicAdd(a.next_root, 1ull);
        gi = bcast64(gi);
        if (a.pk_gstart ? gi >= a.pk_ngroups : gi >= n_dispatch) {
          roots_left = false;
          break;
        }
        gcur = a.pk_gstart ? a.pk_gstart[gi] : gi;
        gend = a.pk_gstart ? a.pk_gstart[gi + 1] : gi + 1;
        grouped = false;
        if (gend - gcur >= 2 && a.root_mode == 2) {
          const uint64_t c0 = a.perm ? a.perm[gcur] : (a.reverse ? n_dispatch - 1 - gcur : gcur);
          grouped = pk_group_build(a, scratch, (uint32_t)(a.roots[c0] >> 32), a.wide_h ? gend - gcur : 0, G, acc,
                                   map32);
        }
      }
      const uint64_t i = gcur++;
      const uint64_t canon = a.perm ? a.perm[i] : (a.reverse ? n_dispatch - 1 - i : i);
      if (a.eager && a.eager[i]) {  // split before launch: to the one-root-per-wave kernel
        if (l == 0) a.heavy[atomicAdd(a.n_heavy, 1ull)] = canon;
        continue;
      }
      const uint32_t s = (uint32_t)__builtin_ctzll((uint64_t)(~busy & full));
      bool derive = grouped;
      const uint64_t hk = a.pk_whandle ? a.pk_whandle[i] : ~0ull;
      if (hk != ~0ull) {  // the vertex bitmap an earlier launch kept for this root
        const uint64_t *hdr = a.pk_pool + hk;
        G.v0 = (uint32_t)hdr[0];
        G.d0 = (uint32_t)hdr[1];
        G.W0 = (uint32_t)hdr[2];
        G.xq0 = (uint32_t)hdr[3];
        G.rp = (uint32_t)hdr[4];
        G.o0 = hdr[5];
        G.rfirst = hdr[6];
        G.lh0 = hdr[7];
        G.T = hdr + kPoolHdr;
        G.cm = G.T + (uint64_t)G.W0 * G.rp;
        G.lh = G.cm + G.W0;
        derive = true;
      }
      uint32_t pruned = 0;
      const uint32_t built = derive ? pk_derive<NW, NX>(w, G, s, canon, acc, cnt, hsh, pruned)
                                    : pk_build<NW, NX>(w, scratch, s, canon, acc, cnt, hsh);
      if (built == 3 && l == 0) {  // one tree node, pruned
        ++nodes;
        words += pruned;
        if (a.root_nodes) atomicAdd(&a.root_nodes[canon], 1ull);
      }
      if (built == 2) {  // too wide for this variant: to the next launch
        if (grouped && a.wide_h) {
          if (l == 0) G.wide[nwide] = canon;
          ++nwide;
        } else if (l == 0) {
          const unsigned long long at = atomicAdd(a.n_wide, 1ull);
          a.wide[at] = canon;
          if (a.wide_h) a.wide_h[at] = ~0ull;
        }
      }
      if (built == 1) {
        if (l == 0) {  // one task (V = {0}): the root's top node
          for (uint32_t f = 0; f < C::RS; ++f) fld(f, top) = 0;
          fld(C::oV, top) = 1;
          fld(C::oMeta, top) = ((uint64_t)s << 16) | kRecRoot;
        }
        wave_fence();
        ++top;
        ++pending;
        busy |= 1u << s;
      }
    }
    if (top == 0) break;  // no records and no roots left (see the refill condition)

    // Dispatch up to 64 tasks from the top of the stack: look-lane i sees
    // record top-1-i; task lane l takes candidate k of look-record ri.
    const uint32_t nlook = top < 64 ? top : 64;
    uint32_t ci = 0;
    uint64_t lrem[NW], lmeta = 0;
#pragma unroll
    for (int q = 0; q < NW; ++q) lrem[q] = 0;
    if (l < nlook) {
      const uint32_t r = top - 1 - l;
      lmeta = fld(C::oMeta, r);
      const uint32_t first = (uint32_t)(lmeta >> 32);
#pragma unroll
      for (int q = 0; q < NW; ++q) {
        const uint32_t lo = 64u * q;
        const uint64_t keep = first <= lo ? ~0ull : (first >= lo + 64 ? 0ull : ~0ull << (first - lo));
        lrem[q] = fld(C::oV + q, r) & keep;
        ci += (uint32_t)__popcll(lrem[q]);
      }
    }
    const uint32_t lslot = (uint32_t)(lmeta >> 16) & 0xFF;
    const uint32_t incl = wave_scan_add_u32(ci), excl = incl - ci;
    uint32_t T = __builtin_amdgcn_readlane(incl, 63);  // tasks in the top 64 records: min(pending, 64) or more
    T = T < 64 ? T : 64u;
    if (top > a.pk_thr) {  // deep stack: expand the top record only
      const uint32_t c0 = __builtin_amdgcn_readlane(ci, 0);
      T = T < c0 ? T : c0;
    }
    const bool lfull = l < nlook && incl <= T;
    const bool lpart = l < nlook && excl < T && incl > T;
    const uint32_t nfull = (uint32_t)__popcll(__ballot(lfull));
    map32[l] = 0;
    wave_fence();
    if (l < nlook && excl < T) map32[excl] = ((l + 1) << 8) | excl;
    wave_fence();
    const uint32_t mv = wave_scan_max_u32(l < T ? map32[l] : 0u);
    const bool act = l < T;
    const uint32_t ri = act ? (mv >> 8) - 1 : 0u, k = act ? l - (mv & 0xFF) : 0u;

    // The task's parent record.
    uint64_t R[NW], P[NW], XP[NW], V[NW], XX[NX], key = 0;
    uint32_t meta = 0, t = 0;
#pragma unroll
    for (int q = 0; q < NW; ++q) R[q] = P[q] = XP[q] = V[q] = 0;
#pragma unroll
    for (int j = 0; j < NX; ++j) XX[j] = 0;
    if (act) {
      const uint32_t r = top - 1 - ri;
      const uint64_t m64 = fld(C::oMeta, r);
      const uint32_t first = (uint32_t)(m64 >> 32);
      uint64_t rem[NW];
#pragma unroll
      for (int q = 0; q < NW; ++q) {
        if (lists) R[q] = fld(C::oR + q, r);
        P[q] = fld(C::oP + q, r);
        XP[q] = fld(C::oX + q, r);
        V[q] = fld(C::oV + q, r);
        const uint32_t lo = 64u * q;
        rem[q] = V[q] & (first <= lo ? ~0ull : (first >= lo + 64 ? 0ull : ~0ull << (first - lo)));
      }
#pragma unroll
      for (int j = 0; j < NX; ++j) XX[j] = fld(C::oXX + j, r);
      key = fld(C::oKey, r);
      meta = (uint32_t)m64;
      // t = the k-th remaining candidate
      uint32_t kk = k, base = 0;
#pragma unroll
      for (int q = 0; q < NW - 1; ++q) {
        const uint32_t cq = (uint32_t)__popcll(rem[q]);
        if (base == (uint32_t)q * 64 && kk >= cq) {
          kk -= cq;
          base = (uint32_t)(q + 1) * 64;
        }
      }
      uint64_t rw = rem[0];
#pragma unroll
      for (int q = 1; q < NW; ++q)
        if (base == (uint32_t)q * 64) rw = rem[q];
      t = base + select64(rw, kk);
    }
    // A child's P | XP lies within its parent's P | XP: when no parent of
    // this pass reaches the last word (a root's top node may), the tasks
    // skip that word of every row they scan.
    const bool wlast = NW == 1 || __ballot(act && ((meta & kRecRoot) || (P[NW - 1] | XP[NW - 1]) != 0)) != 0;
    wave_fence();  // every parent is read before records are rewritten below
    if (lpart) {   // the partly taken record keeps its candidates from the (T - excl)-th on
      uint32_t kk = T - excl, base = 0;
#pragma unroll
      for (int q = 0; q < NW - 1; ++q) {
        const uint32_t cq = (uint32_t)__popcll(lrem[q]);
        if (base == (uint32_t)q * 64 && kk >= cq) {
          kk -= cq;
          base = (uint32_t)(q + 1) * 64;
        }
      }
      uint64_t rw = lrem[0];
#pragma unroll
      for (int q = 1; q < NW; ++q)
        if (base == (uint32_t)q * 64) rw = lrem[q];
      const uint32_t first = base + select64(rw, kk);
      fld(C::oMeta, top - 1 - l) = (lmeta & 0xFFFFFFFFull) | ((uint64_t)first << 32);
    }
    if (lfull) __hip_atomic_fetch_sub(w.m(mRefs, lslot), (uint64_t)1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    top -= nfull;

    // The task: one tree node per lane.
    const uint32_t s = (meta >> 16) & 0xFF;
    bool branch = false;
    if (act) {
      const lds64 *sl = w.slot(s);
      const uint64_t px = *w.m(mPX, s);
      const uint32_t p = (uint32_t)px, xq = (uint32_t)(px >> 32);
      const uint32_t Wr = (p + 63) / 64;  // the root's own width (accounting as the other kernel)
      uint32_t rsize;
      if (meta & kRecRoot) {  // the root's top node: R = {}, P = P0, XP = {}, XX = X0'
#pragma unroll
        for (int q = 0; q < NW; ++q) {
          R[q] = 0;
          P[q] = low_mask(p > 64u * q ? p - 64u * q : 0u);
          XP[q] = 0;
        }
#pragma unroll
        for (int j = 0; j < NX; ++j) XX[j] = low_mask(xq > 64u * j ? xq - 64u * j : 0u);
        key = *w.m(mKey, s);
        rsize = 0;
      } else {
        // Child of candidate t: the parent's candidates below t have moved
        // from P to XP, as when the sequential search reaches t.
        const uint32_t jt = t >> 6, bt = t & 63;
        uint32_t xpop = 0;
#pragma unroll
        for (int q = 0; q < NW; ++q) {
          const uint64_t below = (uint32_t)q < jt ? V[q] : ((uint32_t)q == jt ? V[q] & low_mask(bt) : 0ull);
          const uint64_t rt = sl[C::sProw + t * NW + q];
          P[q] = P[q] & ~below & rt;
          XP[q] = (XP[q] | below) & rt;
          R[q] |= (uint32_t)q == jt ? 1ull << bt : 0ull;
        }
#pragma unroll
        for (int j = 0; j < NX; ++j) {
          xpop += (uint32_t)__popcll(XX[j]);
          XX[j] &= sl[C::sXcol + t * NX + j];
        }
        key += sl[C::sG + t];
        rsize = (meta & 0xFFFF) + 1;
        words += Wr;
        // Aggressive node generation: every vertex of P' adjacent to all of
        // P' \ {itself} moves to R; XP' and XX' keep what is adjacent to it.
        uint32_t psz = 0;
#pragma unroll
        for (int q = 0; q < NW; ++q) psz += (uint32_t)__popcll(P[q]);
        bool anym = false;
        if (psz > 1) {
          words += (uint64_t)psz * Wr;
          uint64_t mk[NW], xacc[NW];
#pragma unroll
          for (int q = 0; q < NW; ++q) {
            mk[q] = 0;
            xacc[q] = ~0ull;
          }
#pragma unroll
          for (int q = 0; q < NW; ++q) {
            // two rows per pass, both loads in flight before either is used
            for (uint64_t m = P[q]; m;) {
              const uint64_t b0 = m & (~m + 1);
              m ^= b0;
              const uint64_t b1 = m & (~m + 1);
              m ^= b1;
              const uint32_t i0 = 64u * q + (uint32_t)__builtin_ctzll(b0);
              const uint32_t i1 = b1 ? 64u * q + (uint32_t)__builtin_ctzll(b1) : i0;
              uint64_t r0[NW], r1[NW];
              uint32_t c0 = 0, c1 = 0;
#pragma unroll
              for (int z = 0; z < NW; ++z) {
                r0[z] = z + 1 < NW || wlast ? (uint64_t)sl[C::sProw + i0 * NW + z] : 0ull;
                r1[z] = z + 1 < NW || wlast ? (uint64_t)sl[C::sProw + i1 * NW + z] : 0ull;
              }
#pragma unroll
              for (int z = 0; z < NW; ++z) {
                if (z + 1 < NW || wlast) {
                  c0 += (uint32_t)__popcll(P[z] & ~r0[z]);
                  c1 += (uint32_t)__popcll(P[z] & ~r1[z]);
                }
              }
              if (c0 == 1) {
                mk[q] |= b0;
#pragma unroll
                for (int z = 0; z < NW; ++z)
                  if (z + 1 < NW || wlast) xacc[z] &= r0[z];
              }
              if (b1 && c1 == 1) {
                mk[q] |= b1;
#pragma unroll
                for (int z = 0; z < NW; ++z)
                  if (z + 1 < NW || wlast) xacc[z] &= r1[z];
              }
            }
          }
          uint32_t nm = 0;
#pragma unroll
          for (int q = 0; q < NW; ++q) nm += (uint32_t)__popcll(mk[q]);
          if (nm) {
            anym = true;
#pragma unroll
            for (int q = 0; q < NW; ++q) {
              R[q] |= mk[q];
              P[q] &= ~mk[q];
              XP[q] &= xacc[q];
              for (uint64_t m = mk[q]; m; m &= m - 1) {
                const uint32_t i = 64u * q + (uint32_t)__builtin_ctzll(m);
#pragma unroll
                for (int j = 0; j < NX; ++j) XX[j] &= sl[C::sXcol + i * NX + j];
                key += sl[C::sG + i];
              }
            }
            rsize += nm;
            words += (uint64_t)nm * Wr;
          }
        }
        words += (uint64_t)xpop * (anym ? Wr + 1 : 1);
      }
      // Evaluate the node.
      uint32_t psz = 0, cw = 0, xw = 0;
      uint64_t xany = 0;
#pragma unroll
      for (int q = 0; q < NW; ++q) {
        psz += (uint32_t)__popcll(P[q]);
        cw += (uint32_t)__popcll(P[q] | XP[q]);
        xany |= XP[q];
      }
#pragma unroll
      for (int j = 0; j < NX; ++j) {
        xw += (uint32_t)__popcll(XX[j]);
        xany |= XX[j];
      }
      ++nodes;
      if (a.root_nodes) __hip_atomic_fetch_add(w.m(mNodes, s), (uint64_t)1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
      if (psz == 0) {
        if (xany == 0) pk_emit<NW, NX, LISTS>(w, s, (uint32_t)*w.m(mBase, s) + rsize, key, R, cnt, hsh);
      } else {
        words += (uint64_t)(cw + xw) * Wr;
        // pivot: (|P' \ N(u)|, row) argmin; rows of P0 first, then X0'
        uint32_t best = ~0u;
        bool zero = false;
        // Rows two per pass (both loads in flight first); the minimum does
        // not depend on the order rows are visited in.
        auto scan = [&](uint64_t m, uint32_t rbase, uint32_t kbase) {
          while (m && !zero) {
            const uint64_t b0 = m & (~m + 1);
            m ^= b0;
            const uint64_t b1 = m & (~m + 1);
            m ^= b1;
            const uint32_t u0 = (uint32_t)__builtin_ctzll(b0);
            const uint32_t u1 = b1 ? (uint32_t)__builtin_ctzll(b1) : u0;
            uint64_t r0[NW], r1[NW];
#pragma unroll
            for (int z = 0; z < NW; ++z) {
              r0[z] = z + 1 < NW || wlast ? (uint64_t)sl[rbase + u0 * NW + z] : 0ull;
              r1[z] = z + 1 < NW || wlast ? (uint64_t)sl[rbase + u1 * NW + z] : 0ull;
            }
            uint32_t c0 = 0, c1 = 0;
#pragma unroll
            for (int z = 0; z < NW; ++z) {
              if (z + 1 < NW || wlast) {
                c0 += (uint32_t)__popcll(P[z] & ~r0[z]);
                c1 += (uint32_t)__popcll(P[z] & ~r1[z]);
              }
            }
            const uint32_t k0 = (c0 << 16) | (kbase + u0), k1 = (c1 << 16) | (kbase + u1);
            best = k0 < best ? k0 : best;
            best = k1 < best ? k1 : best;
            zero = c0 == 0 || c1 == 0;
          }
        };
#pragma unroll
        for (int q = 0; q < NW; ++q) scan(P[q] | XP[q], C::sProw + 64u * q * NW, 64u * q);
#pragma unroll
        for (int j = 0; j < NX; ++j) scan(XX[j], C::sXrow + 64u * j * NW, p + 64u * j);
        if (!zero) {  // else an excluded row covers P': no maximal clique below
          words += Wr;
          const uint32_t piv = best & 0xFFFF;
          const lds64 *pr = piv < p ? sl + C::sProw + piv * NW : sl + C::sXrow + (piv - p) * NW;
#pragma unroll
          for (int q = 0; q < NW; ++q) V[q] = P[q] & ~(uint64_t)pr[q];
          branch = true;
        }
      }
      meta = (meta & 0x00FF0000u) | rsize;
    }

    // Push the branching nodes' records (the children of the deepest
    // records highest) and account.
    uint64_t bm = __ballot(branch);
    if (top + (uint32_t)__popcll(bm) > a.pk_cap) {  // unreachable by the bound above; never write past the arena
      if (l == 0) atomicOr(a.err, kErrPackStack);
      bm = 0;
      branch = false;
    }
    const uint32_t npush = (uint32_t)__popcll(bm);
    uint32_t vc = 0;
    if (branch) {
      const uint32_t r = top + npush - 1 - rank_below(bm);
#pragma unroll
      for (int q = 0; q < NW; ++q) {
        if (lists) fld(C::oR + q, r) = R[q];
        fld(C::oP + q, r) = P[q];
        fld(C::oX + q, r) = XP[q];
        fld(C::oV + q, r) = V[q];
        vc += (uint32_t)__popcll(V[q]);
      }
#pragma unroll
      for (int j = 0; j < NX; ++j) fld(C::oXX + j, r) = XX[j];
      fld(C::oKey, r) = key;
      fld(C::oMeta, r) = meta;  // first = 0: every candidate remains
      __hip_atomic_fetch_add(w.m(mRefs, s), (uint64_t)1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    }
    const uint32_t vsum = __builtin_amdgcn_readlane(wave_scan_add_u32(vc), 63);
    top += npush;
    pending = pending - T + vsum;
    wave_fence();
    // Slots whose last record is consumed are done.
    const bool done = l < S && ((busy >> l) & 1) && *w.m(mRefs, l) == 0;
    const uint64_t dm = __ballot(done);
    if (done) pk_finish<NW, NX>(w, l);
    busy &= ~(uint32_t)dm;
  }

  // Exit: the wave's sums.
  const uint64_t wc = wave_sum64(cnt), wh = wave_sum64(hsh), ww = wave_sum64(words), wn = wave_sum64(nodes);
  acc.build += 4 * wave_sum64(acc.scan);
  if (l == 0) {
    if (wc) atomicAdd(a.count, (unsigned long long)wc);
    if (wh) atomicAdd(a.hash, (unsigned long long)wh);
    if (wn) atomicAdd(a.total_nodes, (unsigned long long)wn);
    if (ww) atomicAdd(a.search_words, (unsigned long long)ww);
    if (acc.build) atomicAdd(a.build_bytes, acc.build);
  }
  wave_fence();
  for (uint32_t i = l; i < kLdsHist && i < a.hist_len; i += 64)
    if (w.L.hist[i]) atomicAdd(&a.hist[i], (unsigned long long)w.L.hist[i]);
}

// One kernel per variant and occupancy target (WPE waves per SIMD: the
// register budget 512 / WPE VGPRs).
template <int NW, int NX, int WPE, bool LISTS>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(WPE, 8))) void pack_kernel(KArgs a) {
  pack_main<NW, NX, LISTS>(a);
}

template <int NW, int NX>
hipError_t launch_variant(const KArgs &a, uint32_t waves, uint32_t wpc, size_t lds, hipStream_t stream) {
  // Runs that write clique lists carry R in every record: one instantiation
  // (4 waves per SIMD) for them, so the others hold no R at all.
  if (a.list)
    hipLaunchKernelGGL((pack_kernel<NW, NX, 4, true>), dim3(waves), dim3(64), lds, stream, a);
  else if (wpc <= 16)
    hipLaunchKernelGGL((pack_kernel<NW, NX, 4, false>), dim3(waves), dim3(64), lds, stream, a);
  else if (wpc <= 20)
    hipLaunchKernelGGL((pack_kernel<NW, NX, 5, false>), dim3(waves), dim3(64), lds, stream, a);
  else if (wpc <= 24)
    hipLaunchKernelGGL((pack_kernel<NW, NX, 6, false>), dim3(waves), dim3(64), lds, stream, a);
  else
    hipLaunchKernelGGL((pack_kernel<NW, NX, 8, false>), dim3(waves), dim3(64), lds, stream, a);
  return hipGetLastError();
}

}  // namespace

hipError_t launch_pack(const KArgs &a, int variant, uint32_t waves, uint32_t waves_per_cu, hipStream_t stream) {
  const bool batch = a.root_graph != nullptr;
  if (variant == 0) return launch_variant<1, 1>(a, waves, waves_per_cu, pack_lds_bytes(1, 1, a.pk_slots, batch), stream);
  return launch_variant<2, 2>(a, waves, waves_per_cu, pack_lds_bytes(2, 2, a.pk_slots, batch), stream);
}

}  // namespace rdmce
