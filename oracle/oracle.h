/*
 * oracle.h -- CPU restatement of the GraphBrew/GAPBS graph-kernel path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (graphbrew_amd/, the
 * bench drivers, include/gbk.h) may include, link or call this library.
 * It is loaded by tests/ (as the checker), by __graft_entry__.smoke() (as the
 * checker) and by bench.py's cpu_baseline leg (as the timed CPU baseline).
 *
 * Every function restates one reference function; the citation next to each
 * declaration names the file:line it follows (paths relative to the
 * reference root, gapbs/ = bench/include/external/gapbs/).
 *
 * Conventions: CSR offsets are int64 (SGOffset, gapbs/graph.h:215), vertex
 * ids and weights are int32 (gapbs/benchmark.h:37-40).  Weighted graphs keep
 * weights in a separate array parallel to the neighbour array.  A NULL
 * weight pointer means "unweighted".  Functions return 0 on success and a
 * negative value on invalid input unless documented otherwise.
 */
#ifndef GBK_ORACLE_H_
#define GBK_ORACLE_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- A0 generator (gapbs/generator.h:89-186, util.h:47) ---------------- */
/* Writes (1<<scale)*degree edges.  uniform=0: R-MAT with PermuteIDs,
 * uniform=1: uniform random (no permutation). */
int orc_generate_el(int scale, int degree, int uniform, int32_t *u, int32_t *v);
/* InsertWeights: overwrites w[0..m) with U[1,255] drawn per 2^18 block. */
int orc_insert_weights(int64_t m, int32_t *w);
/* First 'count' raw outputs of std::mt19937 seeded with 'seed'. */
void orc_mt19937(uint32_t seed, int64_t count, uint32_t *out);

/* ---- A1 CSR build + squish (gapbs/builder.h:186-318, 444-493) ---------- */
/* max node id over an edge list (FindMaxNodeID, builder.h:160-171). */
int32_t orc_max_node_id(int64_t m, const int32_t *u, const int32_t *v);
/* Builds one squished CSR from an edge list.  symmetrize: both directions
 * (builder.h:444-461); transpose (directed only): build the in-CSR.  Output
 * arrays must hold n+1 offsets and up to (symmetrize ? 2m : m) entries.
 * Returns the number of squished edges (>= 0) or < 0 on error. */
int64_t orc_build_csr(int64_t n, int64_t m, const int32_t *u, const int32_t *v,
                      const int32_t *w, int symmetrize, int transpose,
                      int keep_self, int64_t *off_out, int32_t *nbr_out,
                      int32_t *wgt_out);

/* ---- A2 relabel (gapbs/builder.h:1046-1206) ----------------------------- */
/* Relabels one CSR (call twice, out and in, for a directed graph). */
int orc_relabel(int64_t n, const int64_t *off, const int32_t *nbr,
                const int32_t *wgt, const int32_t *new_ids, int64_t *off_out,
                int32_t *nbr_out, int32_t *wgt_out);
/* update_org_ids (gapbs/graph.h:535-564): org_new[new_ids[i]] = org_old[i]. */
int orc_update_org_ids(int64_t n, const int32_t *org_old,
                       const int32_t *new_ids, int32_t *org_new);

/* ---- A4 mapping helpers (builder.h:1398-1411, 2042-2081) --------------- */
uint64_t orc_mapping_fingerprint(int64_t n, const int32_t *new_ids);
/* 0 = permutation, -1 = out-of-range id, -2 = duplicate id. */
int orc_validate_permutation(int64_t n, const int32_t *ids);
/* -o 1: SplitMix64 Fisher-Yates, seed 0 (reorder_basic.h:20-37, 104-135). */
int orc_mapping_random(int64_t n, int32_t *new_ids);
/* -o 2: stable sort by degree, descending (reorder_basic.h:213-258). */
int orc_mapping_sort(int64_t n, const int64_t *off, int32_t *new_ids);
/* -o 5: degree-based grouping (reorder_hub.h:267-345). */
int orc_mapping_dbg(int64_t n, const int64_t *off, int32_t *new_ids);

/* ---- A16 SourcePicker (gapbs/benchmark.h:115-156, 301-309) ------------- */
/* Writes 'count' (original, internal) source pairs.  org_ids may be NULL
 * (identity).  Returns 0, or -1 when every vertex is isolated. */
int orc_source_picker(int64_t n, const int64_t *off, const int32_t *org_ids,
                      int count, int repeats, int32_t *orig_out,
                      int32_t *internal_out);

/* ---- A5/A6 PageRank (bench/src/pr.cc:40-94, pr_spmv.cc:40-83) ---------- */
/* Pull over the in-CSR; out-degrees from out_off.  threads<=1 is serial
 * (deterministic Gauss-Seidel order).  *iters_out = executed iterations. */
int orc_pr_gs(int64_t n, const int64_t *in_off, const int32_t *in_nbr,
              const int64_t *out_off, int max_iters, double tol, int fixed,
              int threads, float *scores, int *iters_out, double *err_out);
int orc_pr_spmv(int64_t n, const int64_t *in_off, const int32_t *in_nbr,
                const int64_t *out_off, int max_iters, double tol, int fixed,
                int threads, float *scores, int *iters_out, double *err_out);
/* PRVerifier's push-iteration residual (pr.cc:117-134). */
double orc_pr_residual(int64_t n, const int64_t *out_off,
                       const int32_t *out_nbr, const float *scores);
/* FP64 fixed point of the same update (power iteration to 'tol' L1). */
int orc_pr_fp64(int64_t n, const int64_t *in_off, const int32_t *in_nbr,
                const int64_t *out_off, int max_iters, double tol,
                double *scores);

/* ---- A7/A8 BFS (bench/src/bfs.cc:54-269, graphbrew/bfs_common.h:31-97) - */
int orc_bfs_do(int64_t n, const int64_t *out_off, const int32_t *out_nbr,
               const int64_t *in_off, const int32_t *in_nbr, int32_t source,
               int alpha, int beta, int threads, int32_t *parent);
/* Serial BFS depths over out-edges (-1 unreached). */
int orc_bfs_depths(int64_t n, const int64_t *out_off, const int32_t *out_nbr,
                   int32_t source, int32_t *depth);
/* BFSVerifier: 1 = PASS, 0 = FAIL. */
int orc_bfs_verify(int64_t n, const int64_t *out_off, const int32_t *out_nbr,
                   const int64_t *in_off, const int32_t *in_nbr,
                   int32_t source, const int32_t *parent);

/* ---- A9-A11 connected components (cc.cc:82-264, cc_sv.cc:64-136) ------- */
int orc_cc_afforest(int64_t n, const int64_t *out_off, const int32_t *out_nbr,
                    const int64_t *in_off, const int32_t *in_nbr, int directed,
                    int rounds, int threads, int32_t *comp);
int orc_cc_sv(int64_t n, const int64_t *out_off, const int32_t *out_nbr,
              int threads, int32_t *comp);
/* CCVerifier (cc.cc:290-330): 1 = PASS, 0 = FAIL. */
int orc_cc_verify(int64_t n, const int64_t *out_off, const int32_t *out_nbr,
                  const int64_t *in_off, const int32_t *in_nbr, int directed,
                  const int32_t *comp);
/* Afforest's SampleFrequentElement draw indices (cc.cc:127-150). */
int orc_afforest_sample_indices(int64_t n, int num_samples, int32_t *idx);

/* ---- A12/A13 SSSP (sssp.cc:77-324) ------------------------------------- */
int orc_sssp_delta(int64_t n, const int64_t *off, const int32_t *nbr,
                   const int32_t *wgt, int32_t source, int32_t delta,
                   int threads, int32_t *dist);
int orc_dijkstra(int64_t n, const int64_t *off, const int32_t *nbr,
                 const int32_t *wgt, int32_t source, int32_t *dist);
/* 16-byte fingerprints as (xor, sum) pairs (sssp.cc:92-127). */
void orc_weight_checksum(int64_t n, const int64_t *off, const int32_t *nbr,
                         const int32_t *wgt, const int32_t *org_ids,
                         uint64_t out[2]);
void orc_distance_fingerprint(int64_t n, const int32_t *dist,
                              const int32_t *org_ids, uint64_t out[2]);

/* ---- A14/A15 Brandes BC (bc.cc:71-326) --------------------------------- */
/* Parallel Brandes port with the edge-indexed successor bitmap. */
int orc_bc(int64_t n, const int64_t *off, const int32_t *nbr,
           const int32_t *sources, int num_sources, int threads,
           float *scores);
/* BCVerifier's serial recomputation (unnormalised=0 -> normalised). */
int orc_bc_serial(int64_t n, const int64_t *off, const int32_t *nbr,
                  const int32_t *sources, int num_sources, float *scores);
/* BCScoreWithinTolerance (bc.cc:52-62): 1 when within. */
int orc_bc_within_tolerance(float expected, float actual);

/* ---- A18 Tier-0 features + frozen rule (reorder_types.h:1226-1320,
 *      6066-6195; reorder_adaptive.h:302-333) ------------------------------ */
/* feat[0..6]: avg_degree, degree_cv, hub_concentration,
 * normalized_edge_span, window_neighbor_overlap, packing_factor,
 * internal_density. */
int orc_tier0_features(int64_t n, int64_t m_directed, const int64_t *off,
                       const int32_t *nbr, int64_t sample_size,
                       uint64_t sample_seed, double feat[7]);
/* ModeledPropertyBytes for BenchmarkType codes 1..8 (reorder_types.h:1214). */
int64_t orc_modeled_property_bytes(int bench, int64_t n, int64_t m_directed);
/* ShouldUseAllKernelLowReuseRule: 1 = GraphBrew arm, 0 = fallback. */
int orc_frozen_rule(int bench, double reuse, int64_t n, double avg_degree,
                    double degree_cv, double wsr);

#ifdef __cplusplus
}
#endif

#endif  // GBK_ORACLE_H_
