/*
 * CPU oracle - TEST INFRASTRUCTURE ONLY (see oracle.h for scope, the
 * reference lines each part restates and what pins it).
 */
#define _GNU_SOURCE
#include "oracle.h"

#include <errno.h>
#include <pthread.h>
#include <stdatomic.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include <unistd.h>

uint64_t orc_splitmix64(uint64_t x) {
  uint64_t z = x + 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

static double now_s(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return ts.tv_sec + ts.tv_nsec * 1e-9;
}

static int cmp_u64(const void *a, const void *b) {
  uint64_t x = *(const uint64_t *)a, y = *(const uint64_t *)b;
  return x < y ? -1 : x > y;
}
static int cmp_u32(const void *a, const void *b) {
  uint32_t x = *(const uint32_t *)a, y = *(const uint32_t *)b;
  return x < y ? -1 : x > y;
}

/* ------------------------------------------------------------------ */
/* Graph construction: graph.cc:64-152                                */
/* ------------------------------------------------------------------ */
static int build_from_keys(uint64_t *keys, uint64_t cnt, orc_graph **out) {
  qsort(keys, cnt, sizeof(uint64_t), cmp_u64); /* graph.cc:109 */
  uint64_t m = 0;
  for (uint64_t i = 0; i < cnt; ++i) /* graph.cc:110-111 */
    if (i == 0 || keys[i] != keys[i - 1]) keys[m++] = keys[i];
  /* used ids in ascending order become 0..n-1 (graph.cc:123-130) */
  uint32_t *ends = malloc(sizeof(uint32_t) * (2 * m + 1));
  for (uint64_t i = 0; i < m; ++i) {
    ends[2 * i] = (uint32_t)(keys[i] >> 32);
    ends[2 * i + 1] = (uint32_t)keys[i];
  }
  qsort(ends, 2 * m, sizeof(uint32_t), cmp_u32);
  uint64_t n = 0;
  for (uint64_t i = 0; i < 2 * m; ++i)
    if (i == 0 || ends[i] != ends[i - 1]) ends[n++] = ends[i];
  orc_graph *g = calloc(1, sizeof(orc_graph));
  g->n = n;
  g->m = m;
  g->labels = malloc(sizeof(uint32_t) * (n + 1));
  memcpy(g->labels, ends, sizeof(uint32_t) * n);
  free(ends);
  g->rowoff = calloc(n + 1, sizeof(uint64_t));
  g->col = malloc(sizeof(uint32_t) * (2 * m + 1));
  uint32_t *a = malloc(sizeof(uint32_t) * (m + 1)), *b = malloc(sizeof(uint32_t) * (m + 1));
  for (uint64_t i = 0; i < m; ++i) {
    uint32_t x = (uint32_t)(keys[i] >> 32), y = (uint32_t)keys[i];
    uint32_t *px = bsearch(&x, g->labels, n, sizeof(uint32_t), cmp_u32);
    uint32_t *py = bsearch(&y, g->labels, n, sizeof(uint32_t), cmp_u32);
    a[i] = (uint32_t)(px - g->labels);
    b[i] = (uint32_t)(py - g->labels);
    g->rowoff[a[i] + 1]++;
    g->rowoff[b[i] + 1]++;
  }
  for (uint64_t v = 0; v < n; ++v) g->rowoff[v + 1] += g->rowoff[v];
  uint64_t *cur = malloc(sizeof(uint64_t) * (n + 1));
  memcpy(cur, g->rowoff, sizeof(uint64_t) * (n + 1));
  /* Pushing both endpoints in sorted-pair order leaves every row sorted
   * (graph.cc:135-145). */
  for (uint64_t i = 0; i < m; ++i) {
    g->col[cur[a[i]]++] = b[i];
    g->col[cur[b[i]]++] = a[i];
  }
  free(cur);
  free(a);
  free(b);
  g->max_degree = 0;
  for (uint64_t v = 0; v < n; ++v) {
    uint64_t d = g->rowoff[v + 1] - g->rowoff[v];
    if (d > g->max_degree) g->max_degree = d;
  }
  g->degeneracy = g->max_degree; /* graph.cc:150 */
  *out = g;
  return 0;
}

int orc_graph_from_pairs(const uint32_t *u, const uint32_t *v, uint64_t count, orc_graph **out) {
  uint64_t *keys = malloc(sizeof(uint64_t) * (count + 1));
  uint64_t k = 0;
  for (uint64_t i = 0; i < count; ++i) {
    uint32_t x = u[i], y = v[i];
    if (x == y) continue; /* graph.cc:88-89 */
    if (x > y) { uint32_t t = x; x = y; y = t; } /* graph.cc:90-91 */
    keys[k++] = ((uint64_t)x << 32) | y;
  }
  int rc = build_from_keys(keys, k, out);
  free(keys);
  return rc;
}

/* One `fin >> x` into uint32_t (libstdc++ semantics, see graph.cc:84):
 * whitespace skipped, optional sign, decimal digits, '-' negates modulo
 * 2^32, magnitude > 2^32-1 fails.  Returns 1 on success. */
static int extract_u32(const char *s, size_t len, size_t *pos, uint32_t *out) {
  size_t p = *pos;
  while (p < len && (s[p] == ' ' || s[p] == '\t' || s[p] == '\n' || s[p] == '\r' || s[p] == '\v' || s[p] == '\f'))
    ++p;
  int neg = 0;
  if (p < len && (s[p] == '-' || s[p] == '+')) neg = s[p++] == '-';
  size_t q = p;
  while (q < len && s[q] >= '0' && s[q] <= '9') ++q;
  *pos = q;
  if (q == p) return 0;
  char tmp[64];
  size_t nd = q - p;
  while (nd > 1 && s[p] == '0') { ++p; --nd; } /* leading zeros */
  if (nd > 20) return 0;
  memcpy(tmp, s + p, nd);
  tmp[nd] = 0;
  errno = 0;
  unsigned long long val = strtoull(tmp, NULL, 10);
  if (errno == ERANGE || val > 0xFFFFFFFFull) return 0;
  *out = neg ? (uint32_t)(0u - (uint32_t)val) : (uint32_t)val;
  return 1;
}

int orc_graph_load_coo(const char *path, orc_graph **out) {
  FILE *f = fopen(path, "rb");
  if (!f) return 1;
  fseek(f, 0, SEEK_END);
  long sz = ftell(f);
  fseek(f, 0, SEEK_SET);
  char *buf = malloc((size_t)sz + 1);
  size_t len = fread(buf, 1, (size_t)sz, f);
  fclose(f);
  uint64_t cap = 1024, k = 0;
  uint64_t *keys = malloc(sizeof(uint64_t) * cap);
  size_t pos = 0;
  while (pos < len) { /* graph.cc:83-104 */
    uint32_t x, y;
    int ok = extract_u32(buf, len, &pos, &x);
    if (ok) ok = extract_u32(buf, len, &pos, &y);
    if (ok && x != y) {
      if (x > y) { uint32_t t = x; x = y; y = t; }
      if (k == cap) keys = realloc(keys, sizeof(uint64_t) * (cap *= 2));
      keys[k++] = ((uint64_t)x << 32) | y;
    }
    while (pos < len && buf[pos] != '\n') ++pos; /* ignore rest of the line */
    if (pos < len) ++pos;
  }
  free(buf);
  int rc = build_from_keys(keys, k, out);
  free(keys);
  return rc;
}

int orc_graph_from_csr(uint64_t n, const uint64_t *rowoff, const uint32_t *col, const uint32_t *labels,
                       uint64_t degeneracy, orc_graph **out) {
  orc_graph *g = calloc(1, sizeof(orc_graph));
  g->n = n;
  g->m = rowoff[n] / 2;
  g->rowoff = malloc(sizeof(uint64_t) * (n + 1));
  memcpy(g->rowoff, rowoff, sizeof(uint64_t) * (n + 1));
  g->col = malloc(sizeof(uint32_t) * (rowoff[n] + 1));
  memcpy(g->col, col, sizeof(uint32_t) * rowoff[n]);
  g->labels = malloc(sizeof(uint32_t) * (n + 1));
  for (uint64_t v = 0; v < n; ++v) g->labels[v] = labels ? labels[v] : (uint32_t)v;
  g->max_degree = 0;
  for (uint64_t v = 0; v < n; ++v) {
    uint64_t d = rowoff[v + 1] - rowoff[v];
    if (d > g->max_degree) g->max_degree = d;
  }
  g->degeneracy = degeneracy ? degeneracy : g->max_degree;
  *out = g;
  return 0;
}

void orc_graph_free(orc_graph *g) {
  if (!g) return;
  free(g->rowoff);
  free(g->col);
  free(g->labels);
  free(g);
}

/* graph.cc:281-306: level k = 1, 2, ...: append every vertex whose
 * remaining degree equals k (ascending id), then process the appended
 * vertices FIFO, decrementing their neighbours and appending any that reach
 * k.  Then graph.cc:2-25: renumber by position, re-sort every row. */
void orc_graph_order_degeneracy(orc_graph *g) {
  uint64_t n = g->n;
  int64_t *deg = malloc(sizeof(int64_t) * (n + 1));
  uint32_t *order = malloc(sizeof(uint32_t) * (n + 1));
  for (uint64_t v = 0; v < n; ++v) deg[v] = (int64_t)(g->rowoff[v + 1] - g->rowoff[v]);
  uint64_t len = 0, level = 0;
  while (len < n) {
    uint64_t cur = len;
    ++level;
    for (uint64_t v = 0; v < n; ++v)
      if (deg[v] == (int64_t)level) order[len++] = (uint32_t)v;
    while (cur < len) {
      uint32_t v = order[cur++];
      for (uint64_t i = g->rowoff[v]; i < g->rowoff[v + 1]; ++i)
        if (--deg[g->col[i]] == (int64_t)level) order[len++] = g->col[i];
    }
  }
  uint32_t *rev = malloc(sizeof(uint32_t) * (n + 1));
  for (uint64_t i = 0; i < n; ++i) rev[order[i]] = (uint32_t)i;
  uint64_t *ro = malloc(sizeof(uint64_t) * (n + 1));
  uint32_t *co = malloc(sizeof(uint32_t) * (g->rowoff[n] + 1));
  uint32_t *lb = malloc(sizeof(uint32_t) * (n + 1));
  ro[0] = 0;
  for (uint64_t i = 0; i < n; ++i) {
    uint32_t old = order[i];
    uint64_t d = g->rowoff[old + 1] - g->rowoff[old];
    for (uint64_t k = 0; k < d; ++k) co[ro[i] + k] = rev[g->col[g->rowoff[old] + k]];
    qsort(co + ro[i], d, sizeof(uint32_t), cmp_u32);
    ro[i + 1] = ro[i] + d;
    lb[i] = g->labels[old];
  }
  free(g->rowoff);
  free(g->col);
  free(g->labels);
  g->rowoff = ro;
  g->col = co;
  g->labels = lb;
  g->degeneracy = level;
  free(deg);
  free(order);
  free(rev);
}

/* ------------------------------------------------------------------ */
/* MCE                                                                 */
/* ------------------------------------------------------------------ */
typedef struct {
  const orc_graph *g;
  int mode;
  uint64_t *outoff; /* mode 2: prefix of out-degrees, n + 1 */
  const uint64_t *ids; /* explicit canonical root ids, or NULL for [root_begin, root_end) */
  uint64_t root_begin, root_end;
  _Atomic uint64_t next;
  int want_nodes;
  uint64_t *root_nodes;
  uint64_t list_cap;
} shared_t;

typedef struct {
  shared_t *sh;
  uint64_t count, hash, nodes;
  uint64_t hist[ORC_HIST_MAX];
  /* per-root working storage */
  uint32_t *C, *rowv;
  uint64_t Ccap, rowcap;
  uint64_t *bm, bmcap;
  uint64_t *stk, stkcap;
  uint64_t *gP, gPcap;
  /* current root */
  uint64_t W, rows, p;
  uint32_t base_size;
  uint64_t base_key;
  uint32_t base_lab[2];
  uint64_t root_nodes;
  /* clique list */
  uint32_t *list;
  uint64_t list_len, list_cap;
  int overflow;
} thr_t;

#define GROW(ptr, cap, need, type)                                  \
  do {                                                              \
    if ((need) > (cap)) {                                           \
      (cap) = (need) * 2;                                           \
      (ptr) = (type *)realloc((ptr), sizeof(type) * (cap));         \
    }                                                               \
  } while (0)

static inline int popc(uint64_t x) { return __builtin_popcountll(x); }

static void emit(thr_t *t, const uint64_t *R) {
  uint32_t size = t->base_size;
  uint64_t key = t->base_key;
  for (uint64_t w = 0; w < t->W; ++w) {
    uint64_t r = R[w];
    size += (uint32_t)popc(r);
    while (r) {
      int b = __builtin_ctzll(r);
      key += t->gP[w * 64 + b];
      r &= r - 1;
    }
  }
  t->count++;
  t->hash += orc_splitmix64(key);
  t->hist[size < ORC_HIST_MAX ? size : ORC_HIST_MAX - 1]++;
  if (t->sh->list_cap) {
    if (t->list_len + size + 1 > t->sh->list_cap) {
      t->overflow = 1;
      return;
    }
    GROW(t->list, t->list_cap, t->list_len + size + 1, uint32_t);
    uint32_t *dst = t->list + t->list_len;
    dst[0] = size;
    uint32_t k = 1;
    for (uint32_t i = 0; i < t->base_size; ++i) dst[k++] = t->base_lab[i];
    for (uint64_t w = 0; w < t->W; ++w) {
      uint64_t r = R[w];
      while (r) {
        int b = __builtin_ctzll(r);
        dst[k++] = t->sh->g->labels[t->C[w * 64 + b]];
        r &= r - 1;
      }
    }
    qsort(dst + 1, size, sizeof(uint32_t), cmp_u32);
    t->list_len += size + 1;
  }
}

static inline const uint64_t *row(thr_t *t, uint64_t i) { return t->bm + i * t->W; }

static inline int subset(const uint64_t *R, const uint64_t *rw, uint64_t W) {
  for (uint64_t w = 0; w < W; ++w)
    if (R[w] & ~rw[w]) return 0;
  return 1;
}

/* One search-tree node (BkPivotBitBalance.cuh:229-382 semantics).  R and P
 * live at stack level `lvl`; children are written to level lvl+1. */
static void node(thr_t *t, uint64_t lvl) {
  const uint64_t W = t->W;
  uint64_t *R = t->stk + lvl * 3 * W, *P = R + W, *V = P + W;
  t->root_nodes++;
  int pempty = 1;
  for (uint64_t w = 0; w < W; ++w)
    if (P[w]) { pempty = 0; break; }
  if (pempty) { /* leaf: maximal iff no row contains R (:276-293) */
    for (uint64_t i = 0; i < t->rows; ++i)
      if (subset(R, row(t, i), W)) return;
    emit(t, R);
    return;
  }
  /* pivot: argmin |P \ N(u)| over rows u with R subset of N(u), lowest row
   * on ties (:229-257) */
  uint64_t best = UINT64_MAX, piv = 0;
  for (uint64_t i = 0; i < t->rows && best != 0; ++i) {
    const uint64_t *rw = row(t, i);
    if (!subset(R, rw, W)) continue;
    uint64_t c = 0;
    for (uint64_t w = 0; w < W; ++w) c += (uint64_t)popc(P[w] & ~rw[w]);
    if (c < best) { best = c; piv = i; }
  }
  if (best == 0) return;
  const uint64_t *pr = row(t, piv);
  for (uint64_t w = 0; w < W; ++w) V[w] = P[w] & ~pr[w];
  uint64_t *R2 = t->stk + (lvl + 1) * 3 * W, *P2 = R2 + W;
  for (uint64_t w = 0; w < W; ++w) {
    while (V[w]) { /* children in ascending order (:318-366) */
      int b = __builtin_ctzll(V[w]);
      uint64_t c = w * 64 + b;
      V[w] &= V[w] - 1;
      P[w] &= ~(1ull << b);
      const uint64_t *cr = row(t, c);
      uint64_t psz = 0;
      for (uint64_t k = 0; k < W; ++k) {
        R2[k] = R[k];
        P2[k] = P[k] & cr[k];
        psz += (uint64_t)popc(P2[k]);
      }
      R2[w] |= 1ull << b;
      if (psz > 1) { /* aggressive node generation (:368-382) */
        for (uint64_t k = 0; k < W; ++k) {
          uint64_t cand = P2[k], moved = 0;
          while (cand) {
            int bb = __builtin_ctzll(cand);
            cand &= cand - 1;
            const uint64_t *ir = row(t, k * 64 + bb);
            uint64_t nn = 0;
            for (uint64_t q = 0; q < W; ++q) nn += (uint64_t)popc(P2[q] & ~ir[q]);
            if (nn == 1) moved |= 1ull << bb;
          }
          /* universal vertices stay universal after removal: moving the
           * word's set at once equals the reference's chunked moves */
          R2[k] |= moved;
          P2[k] &= ~moved;
        }
      }
      node(t, lvl + 1);
    }
  }
}

static inline uint64_t lower_bound_u32(const uint32_t *a, uint64_t lo, uint64_t hi, uint32_t key) {
  while (lo < hi) {
    uint64_t mid = (lo + hi) / 2;
    if (a[mid] < key) lo = mid + 1; else hi = mid;
  }
  return lo;
}

/* Intersection of two sorted lists; binary search of the shorter list into
 * the longer when their lengths differ a lot, else a linear merge. */
static uint64_t intersect(const uint32_t *a, uint64_t na, const uint32_t *b, uint64_t nb, uint32_t *out) {
  if (na > nb) { const uint32_t *t = a; a = b; b = t; uint64_t tn = na; na = nb; nb = tn; }
  uint64_t k = 0;
  if (na * 16 < nb) {
    uint64_t lo = 0;
    for (uint64_t i = 0; i < na && lo < nb; ++i) {
      lo = lower_bound_u32(b, lo, nb, a[i]);
      if (lo < nb && b[lo] == a[i]) out[k++] = a[i];
    }
    return k;
  }
  uint64_t i = 0, j = 0;
  while (i < na && j < nb) {
    if (a[i] < b[j]) ++i;
    else if (a[i] > b[j]) ++j;
    else { out[k++] = a[i]; ++i; ++j; }
  }
  return k;
}

/* Calls back on every index b of P (sorted, size p) with P[b] in N(u). */
#define FOR_EACH_P_IN_NU(g, u, P, p, stop, BODY)                                   \
  do {                                                                           \
    const uint32_t *nu_ = (g)->col + (g)->rowoff[u];                             \
    uint64_t du_ = (g)->rowoff[(u) + 1] - (g)->rowoff[u];                        \
    uint64_t a_ = lower_bound_u32(nu_, 0, du_, (P)[0]);                          \
    if ((p) * 16 < du_ - a_) {                                                   \
      for (uint64_t b = 0; b < (p) && a_ < du_ && !(stop); ++b) {                \
        a_ = lower_bound_u32(nu_, a_, du_, (P)[b]);                              \
        if (a_ < du_ && nu_[a_] == (P)[b]) { BODY; }                             \
      }                                                                          \
    } else {                                                                     \
      uint64_t b = 0;                                                            \
      while (a_ < du_ && b < (p) && !(stop)) {                                   \
        if (nu_[a_] < (P)[b]) ++a_;                                              \
        else if (nu_[a_] > (P)[b]) ++b;                                          \
        else { BODY; ++a_; ++b; }                                                \
      }                                                                          \
    }                                                                            \
  } while (0)

/* Solves one root given X (ascending, size x) and P (ascending, size p)
 * stored contiguously in t->C as [P | X]. */
static void solve_sub(thr_t *t, uint64_t p, uint64_t x) {
  const orc_graph *g = t->sh->g;
  if (p == 0) { /* :535-538 */
    if (x == 0) {
      uint64_t zero = 0;
      t->W = 1;
      emit(t, &zero);
    }
    return;
  }
  const uint64_t W = (p + 63) / 64;
  uint32_t *P = t->C, *X = t->C + p;
  /* X rows without a neighbour in P are dropped (graph_gpu.cuh:90-106) */
  GROW(t->rowv, t->rowcap, p + x + 1, uint32_t);
  memcpy(t->rowv, P, sizeof(uint32_t) * p);
  uint64_t rows = p;
  for (uint64_t i = 0; i < x; ++i) {
    int hit = 0;
    FOR_EACH_P_IN_NU(g, X[i], P, p, hit, hit = 1);
    if (hit) t->rowv[rows++] = X[i];
  }
  t->W = W;
  t->rows = rows;
  t->p = p;
  GROW(t->bm, t->bmcap, rows * W + 1, uint64_t);
  memset(t->bm, 0, sizeof(uint64_t) * rows * W);
  for (uint64_t i = 0; i < rows; ++i) {
    uint64_t *rw = t->bm + i * W;
    FOR_EACH_P_IN_NU(g, t->rowv[i], P, p, 0, rw[b >> 6] |= 1ull << (b & 63));
  }
  GROW(t->gP, t->gPcap, p + 1, uint64_t);
  for (uint64_t j = 0; j < p; ++j) t->gP[j] = orc_splitmix64(g->labels[P[j]]);
  GROW(t->stk, t->stkcap, (p + 2) * 3 * W, uint64_t);
  uint64_t *R0 = t->stk, *P0 = R0 + W;
  for (uint64_t w = 0; w < W; ++w) {
    R0[w] = 0;
    uint64_t bits = p - w * 64;
    P0[w] = bits >= 64 ? ~0ull : ((1ull << bits) - 1);
  }
  node(t, 0);
}

static void solve_root(thr_t *t, uint64_t r) {
  const orc_graph *g = t->sh->g;
  t->root_nodes = 0;
  if (t->sh->mode == 2) {
    /* root r = (v0, k-th out-neighbour of v0) in (v0, v1) order */
    uint64_t lo = 0, hi = g->n;
    while (hi - lo > 1) {
      uint64_t mid = (lo + hi) / 2;
      if (t->sh->outoff[mid] <= r) lo = mid; else hi = mid;
    }
    while (lo + 1 < g->n && t->sh->outoff[lo + 1] <= r) ++lo;
    uint32_t v0 = (uint32_t)lo;
    const uint32_t *n0 = g->col + g->rowoff[v0];
    uint64_t d0 = g->rowoff[v0 + 1] - g->rowoff[v0];
    uint64_t first_out = lower_bound_u32(n0, 0, d0, v0 + 1);
    uint32_t v1 = n0[first_out + (r - t->sh->outoff[v0])];
    const uint32_t *n1 = g->col + g->rowoff[v1];
    uint64_t d1 = g->rowoff[v1 + 1] - g->rowoff[v1];
    GROW(t->C, t->Ccap, (d0 < d1 ? d0 : d1) * 2 + 2, uint32_t);
    uint32_t *tmp = t->C + (d0 < d1 ? d0 : d1) + 1;
    uint64_t c = intersect(n0, d0, n1, d1, tmp);
    uint64_t x = lower_bound_u32(tmp, 0, c, v1); /* X below v1, P above (:462-469) */
    uint64_t p = c - x;
    uint32_t *xs = malloc(sizeof(uint32_t) * (x + 1));
    memcpy(xs, tmp, sizeof(uint32_t) * x);
    memmove(t->C, tmp + x, sizeof(uint32_t) * p);
    memcpy(t->C + p, xs, sizeof(uint32_t) * x);
    free(xs);
    t->base_size = 2;
    t->base_key = orc_splitmix64(g->labels[v0]) + orc_splitmix64(g->labels[v1]);
    t->base_lab[0] = g->labels[v0];
    t->base_lab[1] = g->labels[v1];
    solve_sub(t, p, x);
  } else {
    uint32_t v = (uint32_t)r;
    const uint32_t *nv = g->col + g->rowoff[v];
    uint64_t d = g->rowoff[v + 1] - g->rowoff[v];
    uint64_t x = 0;
    while (x < d && nv[x] < v) ++x;
    uint64_t p = d - x;
    GROW(t->C, t->Ccap, d + 1, uint32_t);
    memcpy(t->C, nv + x, sizeof(uint32_t) * p);
    memcpy(t->C + p, nv, sizeof(uint32_t) * x);
    t->base_size = 1;
    t->base_key = orc_splitmix64(g->labels[v]);
    t->base_lab[0] = g->labels[v];
    solve_sub(t, p, x);
  }
  t->nodes += t->root_nodes;
  if (t->sh->want_nodes && !t->sh->ids) t->sh->root_nodes[r - t->sh->root_begin] = t->root_nodes;
}

/* Per-root algorithmic build terms of the edge roots [begin, end) (the
 * survey's measurement formula, with the roots built as solve_root and
 * solve_sub build them): deg = d(v0) + d(v1), pdeg = sum over u in P0 of
 * d(u), rows_w = (p + x') * ceil(p / 64), x' the rows of X with a neighbour
 * in P0.  B_build = 4 deg + 4 pdeg + 8 rows_w.  Test and measurement
 * infrastructure; not part of the enumeration. */
typedef struct {
  const orc_graph *g;
  const uint64_t *outoff; /* canonical id of the first edge root of each vertex, n + 1 */
  uint64_t begin, end;
  _Atomic uint64_t next_v;
  pthread_mutex_t mu;
  uint64_t deg, pdeg, rows_w, roots;
} terms_t;

static void *terms_worker(void *arg) {
  terms_t *T = arg;
  const orc_graph *g = T->g;
  uint64_t deg = 0, pdeg = 0, rows_w = 0, roots = 0, cap = 0;
  uint32_t *buf = NULL;
  for (;;) {
    uint64_t vb = atomic_fetch_add(&T->next_v, 256);
    if (vb >= g->n) break;
    uint64_t ve = vb + 256 < g->n ? vb + 256 : g->n;
    for (uint64_t v0 = vb; v0 < ve; ++v0) {
      const uint32_t *n0 = g->col + g->rowoff[v0];
      const uint64_t d0 = g->rowoff[v0 + 1] - g->rowoff[v0];
      const uint64_t first_out = lower_bound_u32(n0, 0, d0, (uint32_t)v0 + 1);
      for (uint64_t k = first_out; k < d0; ++k) {
        const uint64_t r = T->outoff[v0] + (k - first_out);
        if (r < T->begin || r >= T->end) continue;
        const uint32_t v1 = n0[k];
        const uint32_t *n1 = g->col + g->rowoff[v1];
        const uint64_t d1 = g->rowoff[v1 + 1] - g->rowoff[v1];
        const uint64_t mn = d0 < d1 ? d0 : d1;
        if (cap < mn + 1) {
          cap = 2 * (mn + 1);
          buf = realloc(buf, sizeof(uint32_t) * cap);
        }
        const uint64_t c = intersect(n0, d0, n1, d1, buf);
        const uint64_t x = lower_bound_u32(buf, 0, c, v1), p = c - x;
        const uint32_t *P = buf + x;
        uint64_t xq = 0;
        for (uint64_t i = 0; i < x && p; ++i) {
          int hit = 0;
          FOR_EACH_P_IN_NU(g, buf[i], P, p, hit, hit = 1);
          xq += hit;
        }
        uint64_t sd = 0;
        for (uint64_t j = 0; j < p; ++j) sd += g->rowoff[P[j] + 1] - g->rowoff[P[j]];
        deg += d0 + d1;
        pdeg += sd;
        rows_w += (p + xq) * ((p + 63) / 64);
        ++roots;
      }
    }
  }
  free(buf);
  pthread_mutex_lock(&T->mu);
  T->deg += deg;
  T->pdeg += pdeg;
  T->rows_w += rows_w;
  T->roots += roots;
  pthread_mutex_unlock(&T->mu);
  return NULL;
}

int orc_build_terms(const orc_graph *g, uint64_t root_begin, uint64_t root_end, int threads, uint64_t out[4]) {
  terms_t T;
  memset(&T, 0, sizeof(T));
  T.g = g;
  T.begin = root_begin;
  T.end = root_end;
  uint64_t *oo = malloc(sizeof(uint64_t) * (g->n + 1));
  if (!oo) return ENOMEM;
  oo[0] = 0;
  for (uint64_t v = 0; v < g->n; ++v) {
    const uint64_t d = g->rowoff[v + 1] - g->rowoff[v];
    oo[v + 1] = oo[v] + (d - lower_bound_u32(g->col + g->rowoff[v], 0, d, (uint32_t)v + 1));
  }
  T.outoff = oo;
  atomic_init(&T.next_v, 0);
  pthread_mutex_init(&T.mu, NULL);
  if (threads <= 0) threads = (int)sysconf(_SC_NPROCESSORS_ONLN);
  pthread_t *th = malloc(sizeof(pthread_t) * (size_t)threads);
  for (int i = 0; i < threads; ++i) pthread_create(&th[i], NULL, terms_worker, &T);
  for (int i = 0; i < threads; ++i) pthread_join(th[i], NULL);
  free(th);
  free(oo);
  pthread_mutex_destroy(&T.mu);
  out[0] = T.roots;
  out[1] = T.deg;
  out[2] = T.pdeg;
  out[3] = T.rows_w;
  return 0;
}

static void *worker(void *arg) {
  thr_t *t = arg;
  shared_t *sh = t->sh;
  for (;;) {
    uint64_t b = atomic_fetch_add(&sh->next, 64);
    if (b >= sh->root_end) break;
    uint64_t e = b + 64 < sh->root_end ? b + 64 : sh->root_end;
    for (uint64_t r = b; r < e; ++r) solve_root(t, sh->ids ? sh->ids[r] : r);
  }
  return NULL;
}

uint64_t orc_num_roots(const orc_graph *g, int root_mode) {
  if (root_mode != 2) return g->n;
  return g->m;
}

static int mce_impl(const orc_graph *g, int root_mode, int threads, int want_root_nodes, uint64_t list_cap,
                    uint64_t root_begin, uint64_t root_end, const uint64_t *ids, orc_result **out);

int orc_mce_range(const orc_graph *g, int root_mode, int threads, int want_root_nodes, uint64_t list_cap,
                  uint64_t root_begin, uint64_t root_end, orc_result **out) {
  return mce_impl(g, root_mode, threads, want_root_nodes, list_cap, root_begin, root_end, NULL, out);
}

static int mce_impl(const orc_graph *g, int root_mode, int threads, int want_root_nodes, uint64_t list_cap,
                    uint64_t root_begin, uint64_t root_end, const uint64_t *ids, orc_result **out) {
  double t0 = now_s();
  shared_t sh;
  memset(&sh, 0, sizeof(sh));
  sh.g = g;
  sh.mode = root_mode == 1 ? 1 : 2;
  uint64_t nroots = ids ? root_end : orc_num_roots(g, sh.mode);
  sh.ids = ids;
  if (ids) want_root_nodes = 0;
  if (root_end > nroots) root_end = nroots;
  if (root_begin > root_end) root_begin = root_end;
  sh.root_begin = root_begin;
  sh.root_end = root_end;
  atomic_init(&sh.next, root_begin);
  sh.want_nodes = want_root_nodes;
  sh.list_cap = list_cap;
  if (sh.mode == 2) {
    sh.outoff = malloc(sizeof(uint64_t) * (g->n + 1));
    sh.outoff[0] = 0;
    for (uint64_t v = 0; v < g->n; ++v) {
      uint64_t cnt = 0;
      for (uint64_t i = g->rowoff[v]; i < g->rowoff[v + 1]; ++i) cnt += g->col[i] > v;
      sh.outoff[v + 1] = sh.outoff[v] + cnt;
    }
  }
  orc_result *res = calloc(1, sizeof(orc_result));
  res->n_roots = root_end - root_begin;
  if (want_root_nodes) {
    res->root_nodes = calloc(res->n_roots + 1, sizeof(uint64_t));
    sh.root_nodes = res->root_nodes;
  }
  if (threads <= 0) threads = (int)sysconf(_SC_NPROCESSORS_ONLN);
  if (threads < 1) threads = 1;
  thr_t *ts = calloc((size_t)threads, sizeof(thr_t));
  pthread_t *th = calloc((size_t)threads, sizeof(pthread_t));
  for (int i = 0; i < threads; ++i) {
    ts[i].sh = &sh;
    pthread_create(&th[i], NULL, worker, &ts[i]);
  }
  for (int i = 0; i < threads; ++i) pthread_join(th[i], NULL);
  uint64_t total_list = 0;
  for (int i = 0; i < threads; ++i) total_list += ts[i].list_len;
  if (list_cap) {
    res->list = malloc(sizeof(uint32_t) * (total_list + 1));
    res->list_cap = list_cap;
  }
  for (int i = 0; i < threads; ++i) {
    thr_t *t = &ts[i];
    res->count += t->count;
    res->hash += t->hash;
    res->total_nodes += t->nodes;
    for (int k = 0; k < ORC_HIST_MAX; ++k) res->hist[k] += t->hist[k];
    if (list_cap && t->list_len) {
      memcpy(res->list + res->list_len, t->list, sizeof(uint32_t) * t->list_len);
      res->list_len += t->list_len;
    }
    res->overflow |= t->overflow;
    free(t->C);
    free(t->rowv);
    free(t->bm);
    free(t->stk);
    free(t->gP);
    free(t->list);
  }
  free(ts);
  free(th);
  free(sh.outoff);
  res->seconds = now_s() - t0;
  *out = res;
  return 0;
}

int orc_mce_roots(const orc_graph *g, int root_mode, int threads, const uint64_t *ids, uint64_t n_ids,
                  orc_result **out) {
  uint64_t nroots = orc_num_roots(g, root_mode == 1 ? 1 : 2);
  for (uint64_t i = 0; i < n_ids; ++i)
    if (ids[i] >= nroots) return 1;
  return mce_impl(g, root_mode, threads, 0, 0, 0, n_ids, ids, out);
}

int orc_mce(const orc_graph *g, int root_mode, int threads, int want_root_nodes, uint64_t list_cap,
            orc_result **out) {
  return orc_mce_range(g, root_mode, threads, want_root_nodes, list_cap, 0, UINT64_MAX, out);
}

/* ------------------------------------------------------------------ */
/* Textbook Tomita (list based), vertex roots in the given order        */
/* ------------------------------------------------------------------ */
typedef struct {
  const orc_graph *g;
  orc_result *res;
  uint32_t *Rv;
  uint32_t rsz;
} tom_t;

static int adjacent(const orc_graph *g, uint32_t a, uint32_t b) {
  const uint32_t *na = g->col + g->rowoff[a];
  uint64_t lo = 0, hi = g->rowoff[a + 1] - g->rowoff[a];
  while (lo < hi) {
    uint64_t mid = (lo + hi) / 2;
    if (na[mid] < b) lo = mid + 1; else hi = mid;
  }
  return lo < g->rowoff[a + 1] - g->rowoff[a] && na[lo] == b;
}

static void tomita(tom_t *T, uint32_t *P, uint64_t np, uint32_t *X, uint64_t nx) {
  if (np == 0 && nx == 0) {
    uint64_t key = 0;
    for (uint32_t i = 0; i < T->rsz; ++i) key += orc_splitmix64(T->g->labels[T->Rv[i]]);
    T->res->count++;
    T->res->hash += orc_splitmix64(key);
    T->res->hist[T->rsz < ORC_HIST_MAX ? T->rsz : ORC_HIST_MAX - 1]++;
    return;
  }
  if (np == 0) return;
  /* pivot maximising |P ^ N(u)| over P u X */
  uint32_t piv = P[0];
  int64_t best = -1;
  for (uint64_t s = 0; s < np + nx; ++s) {
    uint32_t u = s < np ? P[s] : X[s - np];
    int64_t c = 0;
    for (uint64_t i = 0; i < np; ++i) c += adjacent(T->g, u, P[i]);
    if (c > best) { best = c; piv = u; }
  }
  uint32_t *cand = malloc(sizeof(uint32_t) * (np + 1));
  uint64_t nc = 0;
  for (uint64_t i = 0; i < np; ++i)
    if (!adjacent(T->g, piv, P[i])) cand[nc++] = P[i];
  uint32_t *Pw = malloc(sizeof(uint32_t) * (np + 1));
  uint32_t *Xw = malloc(sizeof(uint32_t) * (nx + np + 1));
  memcpy(Pw, P, sizeof(uint32_t) * np);
  memcpy(Xw, X, sizeof(uint32_t) * nx);
  uint64_t npw = np, nxw = nx;
  uint32_t *P2 = malloc(sizeof(uint32_t) * (np + 1)), *X2 = malloc(sizeof(uint32_t) * (nx + np + 1));
  for (uint64_t ci = 0; ci < nc; ++ci) {
    uint32_t v = cand[ci];
    uint64_t np2 = 0, nx2 = 0;
    for (uint64_t i = 0; i < npw; ++i)
      if (adjacent(T->g, v, Pw[i])) P2[np2++] = Pw[i];
    for (uint64_t i = 0; i < nxw; ++i)
      if (adjacent(T->g, v, Xw[i])) X2[nx2++] = Xw[i];
    T->Rv[T->rsz++] = v;
    tomita(T, P2, np2, X2, nx2);
    T->rsz--;
    for (uint64_t i = 0; i < npw; ++i)
      if (Pw[i] == v) { Pw[i] = Pw[--npw]; break; }
    Xw[nxw++] = v;
  }
  free(cand);
  free(Pw);
  free(Xw);
  free(P2);
  free(X2);
}

int orc_mce_tomita(const orc_graph *g, orc_result **out) {
  double t0 = now_s();
  orc_result *res = calloc(1, sizeof(orc_result));
  tom_t T;
  T.g = g;
  T.res = res;
  T.Rv = malloc(sizeof(uint32_t) * (g->n + 2));
  for (uint64_t v = 0; v < g->n; ++v) {
    const uint32_t *nv = g->col + g->rowoff[v];
    uint64_t d = g->rowoff[v + 1] - g->rowoff[v];
    uint32_t *P = malloc(sizeof(uint32_t) * (d + 1)), *X = malloc(sizeof(uint32_t) * (d + 1));
    uint64_t np = 0, nx = 0;
    for (uint64_t i = 0; i < d; ++i) {
      if (nv[i] > v) P[np++] = nv[i]; else X[nx++] = nv[i];
    }
    T.Rv[0] = (uint32_t)v;
    T.rsz = 1;
    tomita(&T, P, np, X, nx);
    free(P);
    free(X);
  }
  free(T.Rv);
  res->seconds = now_s() - t0;
  *out = res;
  return 0;
}

void orc_result_free(orc_result *r) {
  if (!r) return;
  free(r->root_nodes);
  free(r->list);
  free(r);
}
