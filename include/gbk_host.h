/*
 * gbk_host.h -- C ABI of libgbkhost.so, the host-side substrate of the
 * bench drivers: synthetic graph generation, edge-list readers, mapping
 * producers and files, source picking and the Tier-0 feature rule.  Pure
 * C++17/OpenMP on the CPU; the device work lives in libgbk.so (gbk.h).
 *
 * Reference interface each entry point follows (gapbs/ =
 * bench/include/external/gapbs/, reorder/ = bench/include/graphbrew/reorder/):
 *   gbh_generate_el / gbh_insert_weights  Generator::GenerateEL, InsertWeights  gapbs/generator.h:89-186
 *   gbh_read_el                           Reader::ReadInEL / ReadInWEL           gapbs/reader.h:91-185
 *   gbh_mapping_*                         GenerateOriginal/Random/SortMapping    reorder/reorder_basic.h:65-258
 *   gbh_mapping_fingerprint               BuilderBase::MappingFingerprint        gapbs/builder.h:1398-1411
 *   gbh_validate_permutation              ValidatePermutationOrThrow             gapbs/builder.h:2042-2081
 *   gbh_mapping_read / gbh_mapping_write  MAP (-o 13) files, WriteLabels         gapbs/builder.h:2125-2235, gapbs/writer.h:380-412
 *   gbh_sg_info / gbh_sg_read / gbh_sg_write  ReadSerializedGraph, WriteSerializedGraph gapbs/reader.h:445-578, gapbs/writer.h:336-377
 *   gbh_source_picker                     SourcePicker::GenerateConsistentSources gapbs/benchmark.h:115-156
 *   gbh_tier0_features / gbh_frozen_rule  ComputeTier0SampledGraphFeatures,
 *                                         ShouldUseAllKernelLowReuseRule         reorder/reorder_types.h:6066-6195, reorder/reorder_adaptive.h:302-333
 *   gbh_budgeted_rule / gbh_llc_bytes     ShouldUseBudgetedLeidenGorder,
 *                                         GetLLCSizeBytes                        reorder/reorder_adaptive.h:240-257, reorder/reorder_types.h:6030-6038
 * All calls return 0 on success and a negative code on bad input.
 */
#ifndef GBK_HOST_H_
#define GBK_HOST_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* (1<<scale)*degree edges into u, v (caller-allocated). */
int gbh_generate_el(int scale, int degree, int uniform, int32_t *u, int32_t *v);
int gbh_insert_weights(int64_t m, int32_t *w);
/* Reads a .el (weighted=0) or .wel (weighted=1) file.  First call with
 * u == NULL returns the edge count in *m; second call fills the arrays. */
int gbh_read_el(const char *path, int weighted, int64_t *m, int32_t *u, int32_t *v, int32_t *w);

/* Serialized graphs: ".sg" (int32 neighbours) and ".wsg" ((v, w) int32
 * pairs).  A one-byte directed flag, int64 m, int64 n, the out-CSR (int64
 * offsets[n+1], m neighbours), for a directed graph the in-CSR, then the n
 * int32 original ids.  gbh_sg_info reads the header (weighted = the suffix
 * is .wsg).  gbh_sg_read fills caller arrays sized from it (in_* only for a
 * directed graph): with wgt == NULL a .wsg file's weights are dropped; with
 * wgt given for a .sg file they are generated from the original ids,
 * weight_scheme 0 "unit" (1) or 1 "hash" (reader.h:48-59).  Returns -3 for
 * a truncated or inconsistent file.  gbh_sg_write writes .wsg when wgt is
 * given, else .sg (the path's suffix must agree). */
int gbh_sg_info(const char *path, int *directed, int64_t *n, int64_t *m, int *weighted);
int gbh_sg_read(const char *path, int weight_scheme, int64_t *off, int32_t *nbr, int32_t *wgt, int64_t *in_off,
                int32_t *in_nbr, int32_t *in_wgt, int32_t *org_ids);
int gbh_sg_write(const char *path, int directed, int64_t n, int64_t m, const int64_t *off, const int32_t *nbr,
                 const int32_t *wgt, const int64_t *in_off, const int32_t *in_nbr, const int32_t *in_wgt,
                 const int32_t *org_ids);

int gbh_mapping_identity(int64_t n, int32_t *new_ids);
int gbh_mapping_random(int64_t n, int32_t *new_ids);
/* use_out_degree=1 sorts by out-degree (off), else by in-degree (in_off). */
int gbh_mapping_sort(int64_t n, const int64_t *off, int32_t *new_ids);
/* -o 5: degree-based grouping (reorder/reorder_hub.h:267-345). */
int gbh_mapping_dbg(int64_t n, const int64_t *off, int32_t *new_ids);
/* Hub mappings of reorder/reorder_hub.h (degree > integer average m/n):
 * HubSort (-o 3, :125-199), HubCluster (-o 4, :203-245), HubSortDBG
 * (-o 6, :380-502), HubClusterDBG (-o 7, :505-582). */
int gbh_mapping_hubsort(int64_t n, const int64_t *off, int32_t *new_ids);
int gbh_mapping_hubcluster(int64_t n, const int64_t *off, int32_t *new_ids);
int gbh_mapping_hubsortdbg(int64_t n, const int64_t *off, int32_t *new_ids);
int gbh_mapping_hubclusterdbg(int64_t n, const int64_t *off, int32_t *new_ids);
/* Parameterised R-MAT (quadrant probabilities a, b, c, 1-a-b-c; 2^scale
 * vertices; 'edges' edges, self loops included) from a counter-based
 * seeded stream, optionally permuted; for the selector's graph families.
 * No reference counterpart (the reference's generator has fixed
 * probabilities and seed, gapbs/generator.h:89-186). */
int gbh_generate_rmat(int scale, int64_t edges, double a, double b, double c, uint64_t seed,
                      int permute, int32_t *u, int32_t *v);
/* Road-like grid of rows x cols vertices (4-neighbour: right and lower
 * edges), each edge's head rewired to a uniform random vertex with
 * probability 'rewire', weights U[1, max_weight] into w (may be NULL);
 * counter-based seeded stream.  With u == NULL returns the edge count
 * rows (cols - 1) + (rows - 1) cols; else fills the arrays and returns it;
 * -1 on bad arguments.  No reference counterpart (the reference's road
 * graphs are downloaded files). */
int64_t gbh_generate_lattice(int64_t rows, int64_t cols, double rewire, uint64_t seed, int max_weight,
                             int32_t *u, int32_t *v, int32_t *w);
uint64_t gbh_mapping_fingerprint(int64_t n, const int32_t *new_ids);
/* 0 ok, -1 out-of-range id, -2 duplicate id. */
int gbh_validate_permutation(int64_t n, const int32_t *ids);
/* Mapping files: ".so" = binary int32 new_ids, ".lo" = one id per line.
 * Read: first call with ids == NULL returns the length in *n. */
int gbh_mapping_read(const char *path, int64_t *n, int32_t *ids);
int gbh_mapping_write(const char *path, int64_t n, const int32_t *ids);

int gbh_source_picker(int64_t n, const int64_t *off, const int32_t *org_ids, int count,
                      int repeats, int32_t *orig_out, int32_t *internal_out);
/* Afforest's sampled indices (default-seeded mt19937, 1024 draws). */
int gbh_afforest_samples(int64_t n, int count, int32_t *idx);

/* feat: avg_degree, degree_cv, hub_concentration, normalized_edge_span,
 * window_neighbor_overlap, packing_factor, internal_density. */
int gbh_tier0_features(int64_t n, const int64_t *off, const int32_t *nbr, int64_t sample_size,
                       uint64_t seed, double feat[7]);
int64_t gbh_modeled_property_bytes(int bench, int64_t n, int64_t m);
int gbh_frozen_rule(int bench, double reuse, int64_t n, double avg_degree, double degree_cv,
                    double wsr);
/* ShouldUseBudgetedLeidenGorder (reorder/reorder_adaptive.h:240-257). */
int gbh_budgeted_rule(int bench, double reuse, int64_t n, double avg_degree, double degree_cv,
                      double hub_concentration);
/* GetLLCSizeBytes (reorder/reorder_types.h:6030-6038): host L3, else L2, else 30 MB. */
int64_t gbh_llc_bytes(void);

/* ---- built-in verifiers (gapbs/benchmark.h:475-485) and fingerprints ---- */
/* BFSVerifier (graphbrew/bfs_common.h:31-97): 1 = PASS. */
int gbh_verify_bfs(int64_t n, const int64_t *out_off, const int32_t *out_nbr,
                   const int64_t *in_off, const int32_t *in_nbr, int32_t source,
                   const int32_t *parent);
/* CCVerifier (bench/src/cc.cc:290-330): 1 = PASS. */
int gbh_verify_cc(int64_t n, const int64_t *out_off, const int32_t *out_nbr,
                  const int64_t *in_off, const int32_t *in_nbr, int directed,
                  const int32_t *comp);
/* SSSPVerifier (bench/src/sssp.cc:294-324): number of mismatches (0 = PASS). */
int64_t gbh_verify_sssp(int64_t n, const int64_t *off, const int32_t *nbr,
                        const int32_t *wgt, int32_t source, const int32_t *dist);
/* BCVerifier (bench/src/bc.cc:256-326): 1 = PASS. */
int gbh_verify_bc(int64_t n, const int64_t *off, const int32_t *nbr,
                  const int32_t *sources, int num_sources, const float *scores,
                  double *max_abs_delta);
/* PRVerifier's residual (bench/src/pr.cc:117-134). */
double gbh_pr_residual(int64_t n, const int64_t *out_off, const int32_t *out_nbr,
                       const float *scores);
/* WeightChecksum / DistanceFingerprint (bench/src/sssp.cc:92-127). */
void gbh_weight_checksum(int64_t n, const int64_t *off, const int32_t *nbr,
                         const int32_t *wgt, const int32_t *org_ids, uint64_t out[2]);
void gbh_distance_fingerprint(int64_t n, const int32_t *dist, const int32_t *org_ids,
                              uint64_t out[2]);

#ifdef __cplusplus
}
#endif

#endif /* GBK_HOST_H_ */
