/*
 * CPU oracle for the maximal-clique-enumeration hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py load this library, and only to check or time
 * the HIP path.  The product (rdmce_amd/) never links or calls it.
 *
 * It restates, in plain C, the reference algorithm for this path:
 *   - COO text loading semantics          RDMCE/src/graph.cc:64-152
 *   - degeneracy peel order               RDMCE/src/graph.cc:281-306, 2-25
 *   - CSR                                 RDMCE/src/graph.cc:216-236
 *   - edge roots (LEVEL_2): C = N(a)^N(b), X/P split at b
 *                                         RDMCE/src/kernels/BkPivotBitBalance.cuh:434-470
 *   - vertex roots (LEVEL_1)              BkPivotBitBalance.cuh:414-427
 *   - bitmap rows P then X, X rows without a P neighbour dropped
 *                                         RDMCE/include/graph_gpu.cuh:59-107
 *   - bitmap BK with pivot, leaf test, aggressive node generation (EXPAND_R)
 *                                         BkPivotBitBalance.cuh:229-382
 * with one deliberate difference: pivot ties are broken by the lowest row
 * index (the reference breaks them by warp lane, which depends on the warp
 * width), so that the search tree - and the per-root node count - is a
 * function of the graph alone.
 *
 * Outputs beyond the reference (which only counts): size histogram, an
 * order-independent clique hash, optional clique list, per-root node counts.
 *
 * Clique hash:  g(v) = splitmix64(label(v)),  key(K) = sum_{v in K} g(v),
 *               H = sum_K splitmix64(key(K))   (all mod 2^64)
 * where splitmix64(x) is the SplitMix64 output function (Steele et al. 2014).
 *
 * Pinned by: networkx find_cliques goldens, brute force over all subsets
 * (n <= 16), closed forms (K_n, cycles, stars, Moon-Moser, cocktail party,
 * complete multipartite), the reference's own CPU solver built from its
 * sources (oracle/_ref, vertex-rooted Tomita counts and the reference loader
 * and peel order), and the list-based Tomita in this file.
 */
#ifndef RDMCE_ORACLE_H
#define RDMCE_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_graph {
  uint64_t n, m, max_degree, degeneracy;
  uint64_t *rowoff; /* n + 1 */
  uint32_t *col;    /* 2m */
  uint32_t *labels; /* n */
} orc_graph;

#define ORC_HIST_MAX 4096

typedef struct orc_result {
  uint64_t count;
  uint64_t hash;
  uint64_t hist[ORC_HIST_MAX];
  uint64_t total_nodes;
  uint64_t n_roots;
  uint64_t *root_nodes; /* per canonical root, NULL unless requested */
  uint64_t list_len;    /* u32 entries in list */
  uint32_t *list;       /* [k, labels...]* each clique sorted ascending */
  uint64_t list_cap;
  int overflow;         /* list capacity exceeded */
  double seconds;
} orc_result;

/* Graph construction (reference loader semantics). 0 = ok. */
int orc_graph_from_pairs(const uint32_t *u, const uint32_t *v, uint64_t count, orc_graph **out);
int orc_graph_load_coo(const char *path, orc_graph **out);
int orc_graph_from_csr(uint64_t n, const uint64_t *rowoff, const uint32_t *col, const uint32_t *labels,
                       uint64_t degeneracy, orc_graph **out);
/* Reference degeneracy peel (graph.cc:281-306) + relabel. */
void orc_graph_order_degeneracy(orc_graph *g);
void orc_graph_free(orc_graph *g);

/* root_mode: 2 = edge roots (default), 1 = vertex roots.
 * threads <= 0: all online CPUs.  list_cap > 0 collects the clique list. */
int orc_mce(const orc_graph *g, int root_mode, int threads, int want_root_nodes, uint64_t list_cap,
            orc_result **out);
/* Same, restricted to canonical roots [root_begin, root_end) (timing samples). */
int orc_mce_range(const orc_graph *g, int root_mode, int threads, int want_root_nodes, uint64_t list_cap,
                  uint64_t root_begin, uint64_t root_end, orc_result **out);
/* Same over an explicit list of canonical root ids (timing samples). */
int orc_mce_roots(const orc_graph *g, int root_mode, int threads, const uint64_t *ids, uint64_t n_ids,
                  orc_result **out);
/* Textbook list-based Tomita pivoting (max |P ^ N(u)|), vertex-rooted in
 * degeneracy order; count + histogram + hash only.  Independent cross-check. */
int orc_mce_tomita(const orc_graph *g, orc_result **out);
void orc_result_free(orc_result *r);

uint64_t orc_splitmix64(uint64_t x);
uint64_t orc_num_roots(const orc_graph *g, int root_mode);
/* Algorithmic build terms of the edge roots [root_begin, root_end): out =
 * {roots, sum d(v0) + d(v1), sum over P0 members of their degree, sum of
 * rows * ceil(p / 64)} (B_build = 4 out[1] + 4 out[2] + 8 out[3]). */
int orc_build_terms(const orc_graph *g, uint64_t root_begin, uint64_t root_end, int threads, uint64_t out[4]);

#ifdef __cplusplus
}
#endif
#endif
