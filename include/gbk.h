/*
 * gbk.h -- C ABI of the MI355X (gfx950) graph-kernel library libgbk.so.
 *
 * Drop-in boundary for GraphBrew's hot path: the builder's relabel step, the
 * five GAPBS kernels (seven binaries) and the timing harness.  Host arrays
 * are plain pointers + sizes; no C++ or torch types cross this boundary and
 * no exception escapes it.  Every call returns GBK_OK (0) or a negative
 * GBK_E* code; gbk_last_error() gives the message.
 *
 * Reference interface each entry point replaces (paths relative to the
 * reference root; gapbs/ = bench/include/external/gapbs/):
 *   gbk_graph_upload       CSRGraph(num_nodes, index, neighs[, inv...])  gapbs/graph.h:310-330
 *   gbk_relabel            BuilderBase::RelabelByMapping                 gapbs/builder.h:1046-1206
 *   gbk_pr (GS)            PageRankPullGS                                bench/src/pr.cc:40-94
 *   gbk_pr (SPMV)          PageRankPull                                  bench/src/pr_spmv.cc:40-83
 *   gbk_bfs                DOBFS                                         bench/src/bfs.cc:164-269
 *   gbk_cc (AFFOREST)      Afforest                                      bench/src/cc.cc:153-264
 *   gbk_cc (SV)            ShiloachVishkin                               bench/src/cc_sv.cc:64-136
 *   gbk_sssp               DeltaStep                                     bench/src/sssp.cc:160-283
 *   gbk_bc                 Brandes                                       bench/src/bc.cc:135-237
 * Results are written to caller-owned host buffers after the device work;
 * the stats structs carry the device time of the kernel region (hipEvents on
 * the context stream), which is what the harness prints as kernel time.
 *
 * Threading: one context = one HIP stream; a context is not thread-safe.
 * Graphs belong to the context that created them.
 */
#ifndef GBK_H_
#define GBK_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GBK_OK 0
#define GBK_EINVAL -1      /* bad argument (null pointer, size, source id) */
#define GBK_EPERM_MAP -2   /* mapping is not a permutation of [0, n) */
#define GBK_EHIP -3        /* HIP runtime error */
#define GBK_ENOMEM -4      /* device allocation failed */
#define GBK_ESTATE -5      /* graph lacks what the call needs (e.g. weights) */

typedef struct gbk_ctx gbk_ctx;
typedef struct gbk_graph gbk_graph;

enum gbk_pr_variant { GBK_PR_GS = 0, GBK_PR_SPMV = 1 };
enum gbk_cc_variant { GBK_CC_AFFOREST = 0, GBK_CC_SV = 1 };
/* Kernels the per-launch timer can follow (parameter "ktime"). */
enum gbk_ktime_kernel {
  GBK_KT_NONE = 0,
  GBK_KT_PR_ITER = 1,     /* one pull iteration of PR (GS or SpMV) */
  GBK_KT_SSSP_RELAX = 2,  /* one relaxation round: light rows + heavy rows */
  GBK_KT_BFS_STEP = 3,    /* one top-down or bottom-up BFS step */
  GBK_KT_BC_LEVEL = 4,    /* one forward or backward Brandes level */
  GBK_KT_LAST = 4
};

typedef struct {
  double device_ms;        /* device time of the whole call's kernels */
  int iterations;          /* executed iterations */
  int converged_iter;      /* first iteration with error < tol, -1 if none */
  double final_error;      /* L1 change of the last iteration (pr.cc:62-76) */
} gbk_pr_stats;

typedef struct {
  double device_ms;
  int steps_td, steps_bu;  /* top-down / bottom-up steps executed */
  int64_t reached;         /* vertices with parent >= 0 */
  int64_t edges_reached;   /* sum of out-degrees of reached vertices (m_r) */
} gbk_bfs_stats;

typedef struct {
  double device_ms;
  int iterations;          /* SV hook rounds, or Afforest neighbour rounds */
  int32_t skipped_label;   /* Afforest's sampled largest component label */
} gbk_cc_stats;

typedef struct {
  double device_ms;
  int64_t buckets;         /* buckets settled */
  int64_t relax_rounds;    /* frontier relaxation launches that did work */
  int64_t reached;
  int64_t edges_reached;
} gbk_sssp_stats;

typedef struct {
  double device_ms;
  int64_t edges_reached;   /* sum over sources of m_r */
  int max_depth;
} gbk_bc_stats;

typedef struct {
  double device_ms;
  int64_t n, m;
} gbk_relabel_stats;

/* ---- context --------------------------------------------------------- */
int gbk_ctx_create(int device, gbk_ctx **out);
int gbk_ctx_destroy(gbk_ctx *ctx);
const char *gbk_last_error(const gbk_ctx *ctx);
/* hipStream_t of the context, for callers that time or order work on it. */
void *gbk_ctx_stream(gbk_ctx *ctx);
int gbk_ctx_sync(gbk_ctx *ctx);
/* Tunables: "heavy_degree", "bu_group", "bfs_alpha", "bfs_beta",
 * "xcd_swizzle", "relax_grid", "graphs" (1: SSSP rounds and PR iterations
 * are captured once per call as a hipGraph and replayed; 0: launched one by
 * one); "batch_grid" (workgroups of the source-batch relax launches: 0 launches
 * as many as the GPU keeps resident, else 8..16384 in multiples of 8);
 * "pr_blocks" (row blocks of a Gauss-Seidel PageRank sweep, run in order: 0
 * picks one per ~4096 row tiles, at most 2; 1 is one launch updating in
 * place); "pr_inplace" (1: the blocks update contributions in place rather than
 * publishing them block by block); "pr_unroll" (64-edge chunks a PageRank wave keeps in flight: 1, 2,
 * 4, 8, or 0 for 8 in block sweeps, 4 in segmented sweeps and 1 otherwise); "pr_segment"
 * (PageRank's segmented sweep, which walks in-edges by source segment so gathers stay in an XCD's
 * L2: 0 (default) on graphs above 2^19 vertices, -1 never, 6..26 the smallest segment width 2^k
 * tried, widened until at most 32 segments remain); "bc_pull" (BC batch forward levels: 0
 * top-down only, 1 (default) bottom-up from in-neighbours when the frontier's out-edges outweigh
 * those of the vertices some source has not reached, 2 bottom-up at every level); "bc_pull_ratio"
 * (1..1024, default 1: how many such edges one frontier edge weighs in that choice); "sssp_wide"
 * (1, default: SSSP batches of 17..32 sources run on 64-lane rows; 0: 32-lane rows); "sssp_pull"
 * (one-source SSSP's pass over a settled bucket's heavy edges: 0 pushed from the bucket's
 * vertices, 1 (default) pulled into the unsettled vertices when their edges are fewer, 2 always
 * pulled); "sssp_heavy" (one-source SSSP: rows with more edges to walk than this, default 256, go to
 * a kernel that spreads them over all workgroups); "sssp_tail" (1, default: after a pulled heavy pass
 * the vertices left are relaxed Bellman-Ford style, every edge of each improved vertex per round,
 * to the end; 0: buckets to the end); "msbfs_group"
 * (lanes per vertex in the BFS batch's bottom-up step: 1, 2, 4, ..., 64; default 32); and
 * "ktime" (a gbk_ktime_kernel) selecting the kernel whose
 * launches are timed one by one with HIP events (loops then run eagerly).  Returns
 * GBK_EINVAL for an unknown name or bad value.  No reference counterpart:
 * the reference's kernels have no launch shape to tune. */
int gbk_ctx_set_param(gbk_ctx *ctx, const char *name, int64_t value);
int gbk_ctx_get_param(gbk_ctx *ctx, const char *name, int64_t *value);
/* Reads and resets the per-launch timer: summed launch time, launch count,
 * and, for GBK_KT_SSSP_RELAX only, the non-stale frontier rows and their
 * edges those launches scanned (counted by separate kernels outside the
 * timed event pairs; 0 otherwise).  No reference counterpart (the
 * reference's Timer, timer.h, times whole trials). */
int gbk_ktime_read(gbk_ctx *ctx, double *total_ms, int64_t *launches,
                   int64_t *rows, int64_t *edges);

/* ---- graphs ------------------------------------------------------------ */
/* Copies a host CSR to the device.  n vertices; off has n+1 entries with
 * off[0] == 0; rows sorted.  wgt may be NULL (unweighted).  For directed
 * graphs the in-CSR (in_off/in_nbr[/in_wgt]) is required; for undirected
 * graphs pass NULL and the out arrays serve as in arrays.  org_ids (length
 * n, new id -> original id) may be NULL for the identity.  Offsets that do
 * not start at 0 or that decrease, and neighbour ids outside [0, n), return
 * GBK_EINVAL before anything is copied (checked on the host). */
int gbk_graph_upload(gbk_ctx *ctx, int64_t n, int directed,
                     const int64_t *off, const int32_t *nbr, const int32_t *wgt,
                     const int64_t *in_off, const int32_t *in_nbr,
                     const int32_t *in_wgt, const int32_t *org_ids,
                     gbk_graph **out);
/* Builds the squished CSR of an edge list on the device: MakeGraphFromEL +
 * MakeCSR + SquishGraph (gapbs/builder.h:186-318, 444-493).  n < 0 takes
 * max id + 1 (FindMaxNodeID); symmetrize adds both directions; rows are
 * sorted by (neighbour, weight), duplicates keep the smallest weight, self
 * loops are dropped unless keep_self.  Directed graphs get their in-CSR.
 * device_ms (may be NULL) receives the device time after the upload. */
int gbk_graph_build(gbk_ctx *ctx, int64_t n, int64_t m_el, const int32_t *u,
                    const int32_t *v, const int32_t *w, int symmetrize,
                    int keep_self, gbk_graph **out, double *device_ms);
int gbk_graph_info(const gbk_graph *g, int64_t *n, int64_t *m, int *directed,
                   int *weighted);
/* Copies the device CSR back.  Any pointer may be NULL to skip that array. */
int gbk_graph_download(gbk_ctx *ctx, const gbk_graph *g, int64_t *off,
                       int32_t *nbr, int32_t *wgt, int64_t *in_off,
                       int32_t *in_nbr, int32_t *in_wgt, int32_t *org_ids);
int gbk_graph_free(gbk_ctx *ctx, gbk_graph *g);

/* ---- relabel (RelabelByMapping) --------------------------------------- */
/* new_ids[old] = new (host array, n entries).  Builds a new device graph
 * whose rows are the old rows moved to their new ids, neighbours remapped
 * and each row sorted by (neighbour, weight); org_ids are composed as in
 * update_org_ids (gapbs/graph.h:535-564).  A non-permutation returns
 * GBK_EPERM_MAP and creates nothing. */
int gbk_relabel(gbk_ctx *ctx, const gbk_graph *in, const int32_t *new_ids,
                gbk_graph **out, gbk_relabel_stats *stats);
/* Mapping producers on the device: algo 0 original (reorder_basic.h:65-79),
 * 1 random (reorder_basic.h:104-135: the SplitMix64 Fisher-Yates shuffle,
 * run as deterministic reservations, the same permutation), and the
 * orderings whose result is a stable sort by an out-degree key: 2 sort
 * (reorder_basic.h:182-258), 5 DBG (reorder_hub.h:267-345), 6 HubSortDBG
 * (reorder_hub.h:380-502), 7 HubClusterDBG (reorder_hub.h:505-582); other
 * orderings return GBK_EINVAL (the hub placements of 3 and 4 move vertices
 * one after another and stay on the host).  new_ids (host, n) receives the mapping,
 * equal to gbh_mapping_*; device_ms (may be NULL) its device time. */
int gbk_mapping(gbk_ctx *ctx, const gbk_graph *g, int algo, int32_t *new_ids, double *device_ms);
/* gbk_mapping then gbk_relabel without the host round trip of the mapping
 * (MakeGraph's mapping + RelabelByMapping, gapbs/builder.h:869).  new_ids
 * (host, may be NULL) receives a copy of the mapping; mapping_ms (may be
 * NULL) the mapping's device time; stats the relabel's. */
int gbk_relabel_by(gbk_ctx *ctx, const gbk_graph *in, int algo, gbk_graph **out,
                   gbk_relabel_stats *stats, int32_t *new_ids, double *mapping_ms);

/* ---- kernels ----------------------------------------------------------- */
/* scores: host float[n] (may be NULL to leave the result on the device). */
int gbk_pr(gbk_ctx *ctx, const gbk_graph *g, int variant, int max_iters,
           double tol, int fixed_work, float *scores, gbk_pr_stats *stats);
/* PageRank without any allocation: gbk_pr's kernels with their workspace
 * in one fixed device array per GPU (gbk_pr_noalloc_max_bytes(), 64 MiB:
 * about 8 bytes per vertex, so n up to ~8M), held until the process ends.
 * No reference counterpart: for callers that cannot allocate (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * It calls no hipMalloc/hipFree or allocating library routine, does not
 * synchronise and only launches kernels on the context stream, so it may be
 * captured.  scores_dev: device float[n] (caller-owned) receiving the
 * scores; report_dev: device double[3] or NULL receiving iterations, the
 * last L1 change and the first converged iteration (-1 if none).  Results
 * are ready once the context stream has reached the call.  A graph whose
 * workspace exceeds the array returns GBK_EINVAL.  Every call on one GPU
 * shares the array: calls that overlap on one GPU (from two contexts or
 * streams) race; each call initialises what it reads. */
int gbk_pr_noalloc(gbk_ctx *ctx, const gbk_graph *g, int variant, int max_iters, double tol,
                   int fixed_work, float *scores_dev, double *report_dev);
int64_t gbk_pr_noalloc_max_bytes(void);
/* parent: host int32[n], -1 for unreached; alpha/beta <= 0 use defaults. */
int gbk_bfs(gbk_ctx *ctx, const gbk_graph *g, int32_t source, int alpha,
            int beta, int32_t *parent, gbk_bfs_stats *stats);
/* comp: host int32[n]; labels are the smallest vertex id of each component. */
int gbk_cc(gbk_ctx *ctx, const gbk_graph *g, int variant, int32_t *comp,
           gbk_cc_stats *stats);
/* dist: host int32[n], INT32_MAX/2 for unreached (sssp.cc:67). */
int gbk_sssp(gbk_ctx *ctx, const gbk_graph *g, int32_t source, int32_t delta,
             int32_t *dist, gbk_sssp_stats *stats);
/* scores: host float[n], normalised by the maximum (bc.cc:224-230). */
int gbk_bc(gbk_ctx *ctx, const gbk_graph *g, const int32_t *sources,
           int num_sources, float *scores, gbk_bc_stats *stats);

/* ---- multi-source batches ---------------------------------------------- */
/* Many sources in one traversal; every source's result equals the single-
 * source call's (BFS depths, SSSP distances).  No reference counterpart as a
 * single call: the reference runs one source per trial (the trial loop of
 * BenchmarkKernel, gapbs/benchmark.h:450, around the BFSBound / SSSPBound
 * lambdas of bench/src/bfs.cc:285-287 and sssp.cc:342-344); these replace
 * that loop for callers that want the batch's throughput.  Sources are
 * processed 64 at a time. */
typedef struct {
  double device_ms;
  int64_t rounds;         /* BFS levels or SSSP relaxation rounds, summed */
  int64_t buckets;        /* SSSP buckets opened (0 for BFS) */
  int64_t edges_reached;  /* sum over sources of the out-degrees of reached vertices */
  int64_t rows_scanned;   /* SSSP: queued rows relaxed over all rounds; BC: frontier entries (0 for BFS) */
  int64_t edges_scanned;  /* SSSP / BC: out-edges walked, each once for all sources of a chunk */
} gbk_batch_stats;

/* depth: host int32[num_sources * n] (source-major, -1 unreached) or NULL;
 * reached / edges_reached: host int64[num_sources] or NULL. */
int gbk_bfs_batch(gbk_ctx *ctx, const gbk_graph *g, const int32_t *sources,
                  int num_sources, int32_t *depth, int64_t *reached,
                  int64_t *edges_reached, gbk_batch_stats *stats);
/* dist: host int32[num_sources * n] (source-major, INT32_MAX/2 unreached)
 * or NULL. */
int gbk_sssp_batch(gbk_ctx *ctx, const gbk_graph *g, const int32_t *sources,
                   int num_sources, int32_t delta, int32_t *dist,
                   int64_t *reached, int64_t *edges_reached,
                   gbk_batch_stats *stats);
/* Brandes BC over all sources (bc.cc:134-231 runs them one per iteration of
 * its source loop): up to 64 sources share each forward level and each
 * backward level; scores (host float[n] or NULL) are summed over the sources
 * and divided by their maximum as gbk_bc's, equal to them within float
 * rounding.  stats: rounds = BFS levels, rows_scanned = union-frontier
 * entries, edges_scanned = edges walked forward and backward. */
int gbk_bc_batch(gbk_ctx *ctx, const gbk_graph *g, const int32_t *sources,
                 int num_sources, float *scores, gbk_batch_stats *stats);

/* ---- many small graphs in one launch ------------------------------------ */
/* No reference counterpart as a call (the reference runs one graph per
 * process).  gbk_graph_union concatenates graphs (all directed or all
 * undirected, all weighted or not) into one device graph whose vertex ids
 * are shifted by each part's base; vbase (host, count + 1 entries) receives
 * the bases and the total. */
int gbk_graph_union(gbk_ctx *ctx, const gbk_graph *const *graphs, int count,
                    int64_t *vbase, gbk_graph **out);
/* CC of every part at once; comp (host int32[n] of the union) holds each
 * part's own labels (smallest id of the component within the part). */
int gbk_cc_many(gbk_ctx *ctx, const gbk_graph *g, const int64_t *vbase, int count,
                int variant, int32_t *comp, gbk_cc_stats *stats);
/* PR of every part at once: each part has its own base score, L1 change and
 * stopping iteration (pr.cc:77-88); scores (host float[n] of the union),
 * iters (host int32[count]) per part. */
int gbk_pr_many(gbk_ctx *ctx, const gbk_graph *g, const int64_t *vbase, int count,
                int variant, int max_iters, double tol, int fixed_work, float *scores,
                int32_t *iters, gbk_pr_stats *stats);

/* ---- learned ordering selector ------------------------------------------ */
/* No reference counterpart as a call: the reference's selector computes its
 * features on the CPU (ComputeTier0SampledGraphFeatures, reorder/
 * reorder_types.h:6066-6195) and applies a frozen threshold rule
 * (reorder/reorder_adaptive.h:302-333).  These compute graph statistics on
 * the device and train / run a small fp32 MLP (in -> 64 -> 64 -> 1, ReLU)
 * with this library's own kernels. */
#define GBK_NUM_FEATURES 12
/* Writes GBK_NUM_FEATURES floats: log2 n, log2 m, mean degree, log2(1 +
 * max degree), degree coefficient of variation, share of edges owned by
 * rows above the mean degree, sampled local clustering coefficient (1024
 * vertices, first 64 neighbours), mean log2(1 + |u - v|) / log2 n over all
 * edges, share of edges with |u - v| < 64 and < 4096, log2 of the working
 * set over 4 MiB (one XCD's L2) and over 256 MiB (Infinity Cache).  Returns
 * GBK_NUM_FEATURES, or a negative error code. */
int gbk_graph_features(gbk_ctx *ctx, const gbk_graph *g, float *features, int count);

/* The adaptive selector's Tier-0 features (ComputeTier0SampledGraphFeatures,
 * reorder/reorder_types.h:6066-6195) computed on the device: feat receives 7
 * doubles, avg_degree, degree_cv, hub_concentration, normalized_edge_span,
 * window_neighbor_overlap, packing_factor, internal_density (the order of
 * gbh_tier0_features).  sample_size 0 takes the reference's default
 * max(1024, min(sqrt(n), 8192)); at most 8192.  Equal to the host
 * restatement bit for bit except the edge span (a few ulps). */
int gbk_tier0_features(gbk_ctx *ctx, const gbk_graph *g, int64_t sample_size, uint64_t seed, double *feat);

typedef struct gbk_mlp gbk_mlp;
/* in_dim in [1, 48]; Glorot-uniform weights from 'seed', zero biases. */
int gbk_mlp_create(gbk_ctx *ctx, int in_dim, uint64_t seed, gbk_mlp **out);
int gbk_mlp_destroy(gbk_ctx *ctx, gbk_mlp *m);
/* Parameters as one flat array: W1[64][in], b1[64], W2[64][64], b2[64],
 * w3[64], b3. */
int64_t gbk_mlp_num_params(const gbk_mlp *m);
int gbk_mlp_get_params(gbk_ctx *ctx, const gbk_mlp *m, float *params);
int gbk_mlp_set_params(gbk_ctx *ctx, gbk_mlp *m, const float *params);
/* X: rows x in_dim (row-major host floats); pred: rows. */
int gbk_mlp_forward(gbk_ctx *ctx, const gbk_mlp *m, const float *X, int64_t rows, float *pred);
/* Mean squared error over the rows and its gradient w.r.t. the parameters
 * (grad may be NULL). */
int gbk_mlp_grad(gbk_ctx *ctx, gbk_mlp *m, const float *X, const float *y, int64_t rows,
                 float *grad, double *loss);
/* Adam (beta 0.9 / 0.999, eps 1e-8) over shuffled minibatches; optional
 * mean loss per epoch. */
int gbk_mlp_train(gbk_ctx *ctx, gbk_mlp *m, const float *X, const float *y, int64_t rows,
                  int epochs, int batch, float lr, uint64_t seed, double *loss_per_epoch);

#ifdef __cplusplus
}
#endif

#endif /* GBK_H_ */
