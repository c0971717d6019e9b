// ref_glue.cc -- C entry points over the reference's own header-only code.
//
// TEST INFRASTRUCTURE ONLY.  oracle/Makefile compiles this file against the
// reference's headers where they lie (bench/include/... under the reference
// root) into oracle/_ref/libref.so.  The tests and tests/golden/
// make_ref_vectors.py call it to pin the oracle (oracle.cpp) and to produce
// the mappings of the reference's CPU reorderers (RCM, Leiden-Gorder8) that
// the GPU path consumes as mapping files; the product never loads it.
// The frozen rule's header and the kernels' .cc files are not built: they
// reach Boost (gapbs/builder.h:73-81, external/corder/vec2d.h:5-6).
//
// Only standard-library headers are added before the reference's own (its
// reorder_types.h uses std::ifstream without including <fstream>).  Nothing
// here restates an algorithm: each function builds the reference's CSRGraph
// from plain arrays and calls the reference function named in its comment.
#include <fstream>
#include <iomanip>
#include <iostream>
#include <sstream>

#include "generator.h"
#include "reorder/reorder_types.h"
#include "reorder/reorder_basic.h"
#include "reorder/reorder_hub.h"
#include "reorder/reorder_rcm.h"
#include "reorder/reorder_graphbrew.h"

namespace {

using WN = NodeWeight<int32_t, int32_t>;

template <typename D>
D make_dest(int32_t v, const int32_t *w, int64_t k) {
  if constexpr (std::is_same_v<D, int32_t>) {
    (void)w;
    (void)k;
    return v;
  } else {
    return WN(v, w[k]);
  }
}

template <typename D>
D **index_of(int64_t n, const int64_t *off, const int32_t *nbr, const int32_t *w) {
  pvector<SGOffset> offsets(n + 1);
  for (int64_t i = 0; i <= n; i++) offsets[i] = off[i];
  D *neighs = new D[off[n] > 0 ? off[n] : 1];
  for (int64_t k = 0; k < off[n]; k++) neighs[k] = make_dest<D>(nbr[k], w, k);
  return CSRGraph<int32_t, D>::GenIndex(offsets, neighs);
}

// CSRGraph (gapbs/graph.h:293-309) over copies of the caller's arrays; the
// graph owns them and frees them with delete[] (graph.h:268-285).
template <typename D>
CSRGraph<int32_t, D> make_graph(int64_t n, int directed, const int64_t *off, const int32_t *nbr,
                                const int32_t *w, const int64_t *in_off, const int32_t *in_nbr,
                                const int32_t *in_w) {
  D **out = index_of<D>(n, off, nbr, w);
  if (!directed) return CSRGraph<int32_t, D>(n, out, out[0]);
  D **in = index_of<D>(n, in_off, in_nbr, in_w);
  return CSRGraph<int32_t, D>(n, out, out[0], in, in[0]);
}

template <typename D>
void read_csr(const CSRGraph<int32_t, D> &g, bool in, int64_t *off, int32_t *nbr, int32_t *w) {
  int64_t k = 0;
  for (int64_t u = 0; u < g.num_nodes(); u++) {
    off[u] = k;
    auto row = in ? g.in_neigh(static_cast<int32_t>(u)) : g.out_neigh(static_cast<int32_t>(u));
    for (auto d : row) {
      if constexpr (std::is_same_v<D, int32_t>) {
        nbr[k] = d;
      } else {
        nbr[k] = d.v;
        if (w) w[k] = d.w;
      }
      k++;
    }
  }
  off[g.num_nodes()] = k;
}

}  // namespace

extern "C" {

// Generator<int32>(scale, degree).GenerateEL(uniform) (gapbs/generator.h:89-166).
int ref_generate_el(int scale, int degree, int uniform, int32_t *u, int32_t *v) {
  Generator<int32_t> gen(scale, degree);
  auto el = gen.GenerateEL(uniform != 0);
  for (size_t e = 0; e < el.size(); e++) {
    u[e] = el[e].u;
    v[e] = el[e].v;
  }
  return 0;
}

// Generator::InsertWeights over an edge list of m weighted edges
// (gapbs/generator.h:171-186); the weights depend only on the edge index.
int ref_insert_weights(int64_t m, int32_t *w) {
  pvector<EdgePair<int32_t, WN>> el(m);
  Generator<int32_t, WN>::InsertWeights(el);
  for (int64_t e = 0; e < m; e++) w[e] = el[e].v.w;
  return 0;
}

// RelabelByMappingStandalone (reorder/reorder_types.h:711-860), the
// builder-free copy of BuilderBase::RelabelByMapping.  Outputs are sized
// like the inputs; in_* only for a directed graph, w/out_w only weighted.
int ref_relabel(int64_t n, int directed, const int64_t *off, const int32_t *nbr, const int32_t *w,
                const int64_t *in_off, const int32_t *in_nbr, const int32_t *in_w, const int32_t *new_ids,
                int64_t *out_off, int32_t *out_nbr, int32_t *out_w, int64_t *out_in_off,
                int32_t *out_in_nbr, int32_t *out_in_w) {
  pvector<int32_t> ids(n);
  for (int64_t i = 0; i < n; i++) ids[i] = new_ids[i];
  if (w) {
    auto g = make_graph<WN>(n, directed, off, nbr, w, in_off, in_nbr, in_w);
    auto r = RelabelByMappingStandalone(g, ids);
    read_csr(r, false, out_off, out_nbr, out_w);
    if (directed) read_csr(r, true, out_in_off, out_in_nbr, out_in_w);
  } else {
    auto g = make_graph<int32_t>(n, directed, off, nbr, nullptr, in_off, in_nbr, nullptr);
    auto r = RelabelByMappingStandalone(g, ids);
    read_csr(r, false, out_off, out_nbr, nullptr);
    if (directed) read_csr(r, true, out_in_off, out_in_nbr, nullptr);
  }
  return 0;
}

// The mapping of ordering 'algo' as MakeGraph's GenerateMapping dispatches it
// with the default out-degree flag (gapbs/builder.h:1426-1449, 1530-1558;
// command_line.h:53): 1 random (reorder_basic.h:104), 2 sort (:213), 3
// HubSort, 4 HubCluster, 5 DBG, 6 HubSortDBG, 7 HubClusterDBG
// (reorder_hub.h:125-505), 11 RCM in its "bnf" variant (reorder_rcm.h:559).
int ref_mapping(int algo, int64_t n, int directed, const int64_t *off, const int32_t *nbr,
                const int64_t *in_off, const int32_t *in_nbr, int32_t *new_ids) {
  auto g = make_graph<int32_t>(n, directed, off, nbr, nullptr, in_off, in_nbr, nullptr);
  pvector<int32_t> ids(n, -1);
  const bool out = true;
  switch (algo) {
    case 0: GenerateOriginalMapping(g, ids); break;
    case 1: GenerateRandomMapping(g, ids); break;
    case 2: GenerateSortMapping(g, ids, out); break;
    case 3: GenerateHubSortMapping(g, ids, out); break;
    case 4: GenerateHubClusterMapping(g, ids, out); break;
    case 5: GenerateDBGMapping(g, ids, out); break;
    case 6: GenerateHubSortDBGMapping(g, ids, out); break;
    case 7: GenerateHubClusterDBGMapping(g, ids, out); break;
    case 11: GenerateRCMBNFOrderMapping<int32_t, int32_t, int32_t, true>(g, ids, std::string()); break;
    default: return -1;
  }
  for (int64_t i = 0; i < n; i++) new_ids[i] = ids[i];
  return 0;
}

// The GraphBrew pipeline for "-o 12:<options>" (options ':'-separated, e.g.
// "leiden:compose:sg_none:comm_size_desc:intra_gorder:gw8") as
// GenerateGraphBrewMappingUnified runs it for a non-layer ordering
// (gapbs/builder.h:3669-3783): auto resolution, the CLI parser, then
// generateGraphBrewMapping (reorder_graphbrew.h:7837).  Returns -2 for an
// option string the delegated path does not cover (Rabbit, layer orders).
int ref_graphbrew_mapping(const char *options, int64_t n, int directed, const int64_t *off,
                          const int32_t *nbr, const int64_t *in_off, const int32_t *in_nbr,
                          int32_t *new_ids) {
  auto g = make_graph<int32_t>(n, directed, off, nbr, nullptr, in_off, in_nbr, nullptr);
  std::vector<std::string> opts;
  std::string s(options ? options : ""), tok;
  std::stringstream ss(s);
  while (std::getline(ss, tok, ':')) opts.push_back(tok);
  const double res = computeAutoResolution<int32_t, int32_t>(g);
  graphbrew::GraphBrewConfig config = graphbrew::parseGraphBrewCliConfig(opts, res);
  if (config.algorithm == graphbrew::GraphBrewAlgorithm::RABBIT_ORDER) return -2;
  if (config.ordering == graphbrew::OrderingStrategy::LAYER) return -2;
  if (config.communityOrder == graphbrew::CommunityOrder::CapacityRuns) return -2;
  pvector<int32_t> ids(n, -1);
  graphbrew::generateGraphBrewMapping<uint32_t>(g, ids, config);
  for (int64_t i = 0; i < n; i++) new_ids[i] = ids[i];
  return 0;
}

// ComputeTier0SampledGraphFeatures (reorder/reorder_types.h:6066-6195):
// avg_degree, degree_cv, hub_concentration, normalized edge span, window
// neighbour overlap, packing factor, internal density.
int ref_tier0_features(int64_t n, int directed, const int64_t *off, const int32_t *nbr,
                       const int64_t *in_off, const int32_t *in_nbr, int64_t sample_size,
                       uint64_t seed, double feat[7]) {
  auto g = make_graph<int32_t>(n, directed, off, nbr, nullptr, in_off, in_nbr, nullptr);
  auto f = ComputeTier0SampledGraphFeatures(g, static_cast<size_t>(sample_size), seed);
  feat[0] = f.avg_degree;
  feat[1] = f.degree_variance;
  feat[2] = f.hub_concentration;
  feat[3] = f.avg_reuse_distance;
  feat[4] = f.window_neighbor_overlap;
  feat[5] = f.packing_factor;
  feat[6] = f.internal_density;
  return 0;
}

// ModeledPropertyBytes (reorder/reorder_types.h:1226-1259); -1 where the
// reference throws (BENCH_GENERIC).
int64_t ref_modeled_property_bytes(int bench, int64_t n, int64_t m) {
  try {
    return static_cast<int64_t>(
        ModeledPropertyBytes(static_cast<BenchmarkType>(bench), static_cast<uint64_t>(n),
                             static_cast<uint64_t>(m)));
  } catch (const std::exception &) {
    return -1;
  }
}

// ShouldUseAllKernelLowReuseRule (reorder/reorder_adaptive.h:302-333) is not
// built here: reorder_adaptive.h needs reorder.h, which reaches Boost through
// external/corder/vec2d.h:5-6.  The rule is pinned by the reference's own
// evidence records instead (docs/allkernel-lowreuse-evidence.json).

}  // extern "C"
