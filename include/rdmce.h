/*
 * rdmce: maximal clique enumeration on AMD Instinct MI355X (gfx950).
 *
 * C ABI of librdmce.so.  Plain pointers and sizes only; every call returns an
 * mce_status (0 = MCE_OK) and never exits the process - the message of the
 * last failure on the calling thread is returned by mce_last_error().
 *
 * Ownership: inputs are caller-owned and only read during the call; graphs,
 * plans and results are library-owned and released with their *_free call.
 * Device memory is allocated and released by the library.  One solve per
 * device at a time; calls on different devices may run concurrently.
 *
 * Each entry point names the reference interface it replaces
 * (paths relative to the reference repository root).
 */
#ifndef RDMCE_H
#define RDMCE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int32_t mce_status;
#define MCE_OK 0
#define MCE_ERR_ARG 1          /* invalid argument */
#define MCE_ERR_IO 2           /* file could not be opened / read / written */
#define MCE_ERR_NOMEM 3        /* host or device allocation failed */
#define MCE_ERR_DEVICE 4       /* HIP runtime error, or no device */
#define MCE_ERR_UNSUPPORTED 5  /* input beyond a documented limit */
#define MCE_ERR_INTERNAL 6

/* Vertex orders: OrderType, RDMCE/include/graph.h:4-10 */
#define MCE_ORDER_UNCHANGED 0
#define MCE_ORDER_ASC 1
#define MCE_ORDER_DES 2
#define MCE_ORDER_DEG 3

/* Root granularity: LEVEL_2 (edge roots, reference default) or LEVEL_1
 * (vertex roots), RDMCE/include/mce_gpu.cuh:4 and
 * RDMCE/src/kernels/BkPivotBitBalance.cuh:414-470 */
#define MCE_ROOT_VERTEX 1
#define MCE_ROOT_EDGE 2

/* Root schedules (dispatch order of roots; never changes results). */
#define MCE_SCHED_PLAIN 0    /* reverse edge-slot order, BkPivotBitBalance.cuh:429-431 */
#define MCE_SCHED_CUSTOM 1   /* caller-supplied list of distinct canonical root ids
                                (mce_options.roots): all roots in any order, or a
                                subset - results then cover the listed roots only */

const char *mce_last_error(void);
const char *mce_version(void);

/* ------------------------------------------------------------------------ */
/* Host graph  (class Graph, RDMCE/include/graph.h:12-52)                    */
/* ------------------------------------------------------------------------ */
typedef struct mce_graph mce_graph;

typedef struct mce_graph_info {
  uint64_t n;            /* GetNumVertices  graph.h:39 */
  uint64_t m;            /* GetNumEdges     graph.h:40 (undirected) */
  uint64_t max_degree;   /* GetMaxDegree    graph.h:41 */
  uint64_t degeneracy;   /* GetDegeneracy   graph.h:42 */
} mce_graph_info;

/* Graph::LoadFromFile, RDMCE/src/graph.cc:64-152 (COO text: "u v" per line,
 * unparseable lines skipped, self loops dropped, duplicates merged, used ids
 * relabelled densely in ascending order). */
mce_status mce_graph_load_coo_text(const char *path, mce_graph **out);
/* Graph::LoadFromBin, RDMCE/src/graph.cc:178-213. */
mce_status mce_graph_load_bin(const char *path, mce_graph **out);
/* Graph::StoreIntoBin, RDMCE/src/graph.cc:154-176.  Writes exactly `path`
 * (mce_graph_bin_path gives the reference's extension-replacing name). */
mce_status mce_graph_store_bin(const mce_graph *g, const char *path);
mce_status mce_graph_bin_path(const char *input_path, char *buf, uint64_t buf_len);
/* Reference loader semantics applied to in-memory endpoint pairs. */
mce_status mce_graph_from_pairs(const uint32_t *u, const uint32_t *v, uint64_t count, mce_graph **out);
/* Copies a CSR (rows sorted ascending, symmetric, no self loops).  labels may
 * be NULL (identity).  degeneracy 0 = use max degree. */
mce_status mce_graph_from_csr(uint64_t n, const uint64_t *rowoff, const uint32_t *col, const uint32_t *labels,
                              uint64_t degeneracy, mce_graph **out);
/* Graph::SortByOrder, RDMCE/src/graph.cc:241-306. */
mce_status mce_graph_order(mce_graph *g, int32_t order);
mce_status mce_graph_get_info(const mce_graph *g, mce_graph_info *info);
/* Graph::GetRowOffset/GetColIdx (graph.h:43-44) plus labels; pointers stay
 * valid until the graph is modified or freed.  Offsets are 64-bit. */
mce_status mce_graph_get_csr(const mce_graph *g, const uint64_t **rowoff, const uint32_t **col,
                             const uint32_t **labels);
/* Core number per vertex (computed on first use). */
mce_status mce_graph_get_core(mce_graph *g, const uint32_t **core);
mce_status mce_graph_write_coo_text(const mce_graph *g, const char *path);
void mce_graph_free(mce_graph *g);

/* ------------------------------------------------------------------------ */
/* Synthetic generators (no reference counterpart; deterministic per seed)   */
/* ------------------------------------------------------------------------ */
mce_status mce_gen_er(uint32_t n, double p, uint64_t seed, mce_graph **out);
mce_status mce_gen_rmat(int32_t scale, uint64_t edge_factor, double a, double b, double c, uint64_t seed,
                        mce_graph **out);
mce_status mce_gen_chung_lu_planted(uint64_t n, double gamma, double mean_degree, uint32_t n_comm, uint32_t cmin,
                                    uint32_t cmax, double p_in, uint64_t seed, mce_graph **out);
mce_status mce_gen_batch_params(uint64_t seed, uint64_t index, uint32_t nmin, uint32_t nmax, double pmin,
                                double pmax, uint32_t *n, double *p, uint64_t *edge_seed);

/* ------------------------------------------------------------------------ */
/* Solver                                                                    */
/* ------------------------------------------------------------------------ */
typedef struct mce_options {
  int32_t device;            /* HIP device ordinal (MceGpuSolver(g, device_id), mce_gpu.cu:9-12) */
  int32_t root_mode;         /* MCE_ROOT_EDGE (default) or MCE_ROOT_VERTEX */
  int32_t waves_per_cu;      /* resident wavefronts per CU (1..32; 0 = 16), WARP_PER_SM analogue; also picks the
                                kernel compiled for that occupancy (<=16, <=20, <=24, <=32) */
  int32_t share_bound;       /* TASK_SHARE_BOUND: |P| at which a level is exposed; 0 = 24; <0 = off */
  int32_t want_list;         /* 1 = return the clique list */
  int32_t want_root_stats;   /* 1 = return per-root node counts and wave cycles */
  int32_t schedule;          /* MCE_SCHED_* */
  int32_t lds_tier;          /* 1 = stage small subproblems in LDS (default), 0 = HBM only */
  uint64_t list_capacity;    /* u32 entries reserved for the clique list; 0 = 1<<24 */
  uint64_t slot_bytes;       /* per-wave HBM scratch; 0 = sized automatically */
  const uint64_t *roots;     /* MCE_SCHED_CUSTOM: canonical root ids (indices into the
                                list of mce_plan_roots) in dispatch order */
  uint64_t n_roots;
  int32_t use_graph;         /* 1 = replay the run (resets + enumeration kernel) as a
                                captured hipGraph */
  uint32_t lds_bytes;        /* LDS per wave for the LDS tier; 0 = CU LDS / waves_per_cu */
  const uint8_t *eager;      /* optional, one flag per dispatched root: expose its
                                top level at once (a pre-launch split) */
  int32_t word_path;         /* roots with |P0|, |X0'| <= 64 or <= 128 search with their bitmap rows in
                                registers: 0 = on (default), -1 = off (every root in the bitmap tiers) */
  int32_t pack;              /* packed search (no reference counterpart: the reference runs one root per
                                warp, BkPivotBitBalance.cuh:429-470): 1 = roots with |P0|, |X0'| <= 64, then
                                <= 128, are searched many per wavefront, one tree node per lane, and only the
                                wider ones by the one-root-per-wavefront kernel; 0 = off (default, the plain
                                path).  Never changes a result. */
  uint32_t pack_slots;       /* subproblems resident per wavefront in the packed search (1..8; 0 = 4) */
  uint32_t pack_waves_per_cu;/* wavefronts per CU of the packed search (1..32; 0 = 16) */
  const uint64_t *runs;      /* optional, packed search over a custom edge-root list: run r is
                                roots[runs[r] .. runs[r + 1]), all with one first vertex, whose
                                bitmap the run builds once; runs[0] = 0, runs[n_runs] = n_roots.
                                NULL: maximal runs of consecutive entries sharing a first vertex */
  uint64_t n_runs;
  uint64_t reserved[2];
} mce_options;

void mce_options_init(mce_options *opt);

typedef struct mce_result {
  uint64_t count;            /* number of maximal cliques (MC numbers, mce_gpu.cu:20) */
  uint64_t hash;             /* sum over cliques of splitmix64(sum of splitmix64(label)) */
  uint32_t hist_len;         /* hist[k] = cliques of size k, k < hist_len */
  uint64_t *hist;
  uint64_t list_len;         /* u32 entries: [k, label_1 .. label_k]* (unsorted) */
  uint32_t *list;
  uint64_t n_roots;          /* canonical roots (edge: m, vertex: n) */
  uint64_t *root_nodes;      /* per canonical root: search-tree nodes (want_root_stats) */
  uint64_t *root_cycles;     /* per canonical root: wave cycles spent (want_root_stats) */
  uint64_t total_nodes;
  uint64_t deferred_roots;   /* roots re-run in the large-scratch pass */
  uint64_t shared_tasks;     /* subtrees handed between wavefronts */
  double kernel_ms;          /* device time of the enumeration kernels (HIP events) */
  double total_ms;           /* wall time of the call */
  int32_t passes;
  int32_t waves;             /* wavefronts launched */
  uint64_t search_bytes;     /* bitmap bytes read by the search (8 per 64-bit word) */
  uint64_t build_bytes;      /* bytes moved building subproblems: both root adjacency
                                lists, the candidate list and the bitmap */
  uint64_t published_levels; /* levels exposed to other wavefronts */
  uint64_t owner_tasks;      /* tasks of exposed levels that their publisher ran itself */
  uint64_t reserved[4];
} mce_result;

/* BkSolverWrapper(Graph&, device_id), BkPivotBitBalance.cuh:611-678 and
 * MceGpuSolver::Solve, mce_gpu.cu:14-22: upload, enumerate, free. */
mce_status mce_solve(const mce_graph *g, const mce_options *opt, mce_result **out);
/* Same on a caller CSR (must be in an order where every vertex has at most
 * `degeneracy` higher-ordered neighbours when degeneracy != 0). */
mce_status mce_solve_csr(uint64_t n, const uint64_t *rowoff, const uint32_t *col, const uint32_t *labels,
                         uint64_t degeneracy, const mce_options *opt, mce_result **out);
void mce_result_free(mce_result *r);

/* Device-resident plan: the CSR, root list and scratch stay in HBM, so that
 * mce_plan_run measures enumeration only (inputs resident). */
typedef struct mce_plan mce_plan;
mce_status mce_plan_create(const mce_graph *g, const mce_options *opt, mce_plan **out);
mce_status mce_plan_run(mce_plan *p, mce_result **out);
/* Canonical root descriptors (edge roots: v0 << 32 | v1 with v0 < v1 in the
 * graph's order, sorted ascending; vertex roots: v).  Host copy. */
mce_status mce_plan_roots(const mce_graph *g, int32_t root_mode, uint64_t *out, uint64_t cap, uint64_t *count);
void mce_plan_free(mce_plan *p);

/* ------------------------------------------------------------------------ */
/* Batches: many independent graphs enumerated by one persistent launch      */
/* (no reference counterpart: the reference solves one graph per launch,     */
/* BkPivotBitBalance.cuh:661)                                                */
/* ------------------------------------------------------------------------ */
typedef struct mce_batch mce_batch;
typedef struct mce_batch_result {
  uint64_t n_graphs;
  uint32_t hist_stride;      /* hist[g * hist_stride + k] = cliques of size k in graph g */
  uint64_t *count;           /* per graph */
  uint64_t *hash;            /* per graph */
  uint64_t *hist;
  uint64_t total_count;
  uint64_t total_nodes;
  double kernel_ms;          /* device time of the launch (graph replay when captured) */
  double total_ms;
  int32_t graph_replay;      /* 1 when the run was a hipGraph replay */
  int32_t passes;
  uint64_t reserved[6];
} mce_batch_result;

/* Concatenates `count` graphs (each in the order it should be searched in)
 * into one device-resident plan. */
mce_status mce_batch_create(const mce_graph *const *graphs, uint64_t count, const mce_options *opt,
                            mce_batch **out);
mce_status mce_batch_run(mce_batch *b, mce_batch_result **out);
void mce_batch_result_free(mce_batch_result *r);
void mce_batch_free(mce_batch *b);

/* ------------------------------------------------------------------------ */
/* Autotuner (replaces the reference's compile-time WARP_PER_BLOCK,          */
/* BLOCK_PER_SM and TASK_SHARE_BOUND variants, RDMCE/CMakeLists.txt:47-65,    */
/* scripts/compile.sh:53-80)                                                 */
/* ------------------------------------------------------------------------ */
typedef struct mce_tune_result {
  int32_t waves_per_cu;
  int32_t share_bound;
  int32_t lds_tier;
  int32_t trials;
  double best_ms;            /* kernel time of the winning setting on the sample */
  double base_ms;            /* kernel time of the default setting on the sample */
  uint32_t pack_slots;       /* packed search: subproblems resident per wavefront (the LDS tile width) */
  uint32_t pack_waves_per_cu;/* packed search: wavefronts per CU */
  uint64_t reserved[3];
} mce_tune_result;

/* Times the plan's kernels on a deterministic sample of its roots (every
 * `stride`-th run of edge roots sharing a first vertex, whole - or every
 * `stride`-th root without runs; 1 = all, the faithful choice: a sparse
 * sample changes the load balance the settings are about) for a coordinate
 * search.  One-root-per-wavefront plans: waves per CU in {8, 12, 16, 20, 24,
 * 28, 32} with the LDS tier on, then the sharing bound in {16, 24, 48, off}
 * at the best occupancy, then waves per CU again when the bound changed, then
 * the LDS tier off.  Packed plans (mce_options.pack): subproblems per
 * wavefront in {1, 2, 4, 8}, then packed wavefronts per CU in {12, 16, 20,
 * 24}, then the sharing bound of the wide roots' kernel in {24, off}.
 * Checks that every setting gives the same clique count and hash, and keeps
 * the fastest for later mce_plan_run calls. */
mce_status mce_plan_autotune(mce_plan *p, uint64_t stride, mce_tune_result *out);
mce_status mce_batch_autotune(mce_batch *b, uint64_t stride, mce_tune_result *out);

/* ------------------------------------------------------------------------ */
/* Learned cost model (no reference counterpart; the nearest reference      */
/* behaviour is its fixed dispatch of high-core roots first,                 */
/* BkPivotBitBalance.cuh:429-431).  An FP32 MLP (8 -> 32 -> 32 -> 1, ReLU)   */
/* predicting log(1 + cost) of each edge root; trained and evaluated by the   */
/* library's HIP kernels.  Used only to order roots and flag eager splits,  */
/* so it never changes a result.                                             */
/* ------------------------------------------------------------------------ */
#define MCE_NFEAT 8
typedef struct mce_cost_model mce_cost_model;

/* Per canonical edge root (order of mce_plan_roots), MCE_NFEAT floats:
 * log1p(deg v0), log1p(deg v1), core v0 / degeneracy, core v1 / degeneracy,
 * log1p(outdeg v0), log1p(outdeg v1), log1p(t), log1p(t)^2 with
 * t = |N+(v1) & N(v0)| (the size of the root's candidate set).  Computed on
 * the GPU.  out == NULL only reports the root count. */
mce_status mce_root_features(mce_graph *g, int32_t device, float *out, uint64_t cap, uint64_t *n_roots);
uint32_t mce_cost_model_num_params(void);     /* trainable + 2 * MCE_NFEAT normalisation */
uint32_t mce_cost_model_num_trainable(void);
mce_status mce_cost_model_create(uint64_t seed, mce_cost_model **out);
mce_status mce_cost_model_get_params(const mce_cost_model *m, float *out, uint64_t cap);
mce_status mce_cost_model_set_params(mce_cost_model *m, const float *in, uint64_t n);
/* Sets the input standardisation (per-feature mean and 1/std) from X. */
mce_status mce_cost_model_fit_normalizer(mce_cost_model *m, const float *X, uint64_t n);
/* Mean squared error over n samples and its gradient w.r.t. the trainable
 * parameters (deterministic reduction order). */
mce_status mce_cost_model_gradient(const mce_cost_model *m, int32_t device, const float *X, const float *Y,
                                   uint64_t n, float *grad, float *loss);
/* Minibatch Adam; returns the last epoch's mean squared error. */
mce_status mce_cost_model_train(mce_cost_model *m, int32_t device, const float *X, const float *Y, uint64_t n,
                                uint32_t epochs, uint32_t batch, float lr, uint64_t seed, float *final_loss);
mce_status mce_cost_model_predict(const mce_cost_model *m, int32_t device, const float *X, uint64_t n, float *out);
void mce_cost_model_free(mce_cost_model *m);

/* ------------------------------------------------------------------------ */
/* Enumeration that allocates nothing (no counterpart in the reference).      */
/* For callers that cannot allocate device memory: code under HIP graph       */
/* capture, tight serving loops, harnesses that forbid hipMalloc.  Every      */
/* scratch buffer lives in one static device array of                          */
/* MCE_NOALLOC_SCRATCH_BYTES per device, held until the process ends; calls    */
/* that overlap on one device race on it.                                      */
/* ------------------------------------------------------------------------ */
#define MCE_NOALLOC_SCRATCH_BYTES (64ull << 20)
#define MCE_NOALLOC_ERR_SCRATCH (1ull << 31)
/* mce_solve_csr (the one-root-per-wavefront kernel, reference dispatch order,
 * default settings) on a symmetric CSR already in device memory: rowoff
 * u64[n + 1], col u32[nnz] (rows ascending, no self loops; degeneracy order
 * keeps per-root scratch small), label_hash u64[n] (splitmix64 of each
 * vertex's original id, for the clique hash).  Enqueued on `stream` (a
 * hipStream_t; NULL = the null stream) with no host synchronisation; when
 * it has run, device memory `out` (caller-owned, 4 + hist_len u64) holds
 * out[0] = maximal cliques, out[1] = clique hash, out[2] = search-tree nodes,
 * out[3] = 0, or error bits when the result is incomplete (1: a root with more
 * than 4096 candidates; 2: a clique larger than hist_len - 1, counted in the
 * last entry; MCE_NOALLOC_ERR_SCRATCH: a root needed more scratch than the
 * array holds), out[4 + k] = cliques of size k.  Returns MCE_ERR_UNSUPPORTED
 * when the graph's derived arrays do not fit the array (about n + m < 2.5M). */
mce_status mce_solve_device_csr_noalloc(int32_t device, uint64_t n, uint64_t nnz, const uint64_t *rowoff,
                                        const uint32_t *col, const uint64_t *label_hash, uint32_t hist_len,
                                        void *stream, uint64_t *out);

int32_t mce_device_count(void);
mce_status mce_device_sync(int32_t device);
/* Measurement helper (no reference counterpart): times `iters` launches of
 * a streaming copy kernel over `bytes` (read `bytes`, write `bytes`), used to
 * calibrate hardware traffic counters against a known byte count. */
mce_status mce_copy_probe(int32_t device, uint64_t bytes, int32_t iters, double *ms_per_copy);

#ifdef __cplusplus
}
#endif
#endif /* RDMCE_H */
