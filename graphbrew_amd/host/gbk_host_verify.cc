// gbk_host_verify.cc -- the bench binaries' built-in verifiers and the
// ordering-invariant fingerprints, on the host (libgbkhost.so).  These run
// after a trial with -v, exactly where the reference runs its serial checks
// (gapbs/benchmark.h:475-485).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <queue>
#include <unordered_map>
#include <utility>
#include <vector>

#include "gbk_host.h"

namespace {

uint64_t mix64(uint64_t x) {  // bench/src/sssp.cc:77-82
  x += 0x9e3779b97f4a7c15ULL;
  x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ULL;
  x = (x ^ (x >> 27)) * 0x94d049bb133111ebULL;
  return x ^ (x >> 31);
}

std::vector<int32_t> bfs_levels(int64_t n, const int64_t *off, const int32_t *nbr, int32_t s) {
  std::vector<int32_t> depth(n, -1);
  std::vector<int32_t> frontier{s};
  depth[s] = 0;
  for (size_t head = 0; head < frontier.size(); head++) {
    const int32_t x = frontier[head];
    for (int64_t k = off[x]; k < off[x + 1]; k++) {
      const int32_t y = nbr[k];
      if (depth[y] < 0) {
        depth[y] = depth[x] + 1;
        frontier.push_back(y);
      }
    }
  }
  return depth;
}

}  // namespace

extern "C" {

// BFSVerifier (graphbrew/bfs_common.h:31-97): serial depths; every reached
// vertex's parent must be an in-neighbour one level up; reachability equal.
int gbh_verify_bfs(int64_t n, const int64_t *out_off, const int32_t *out_nbr, const int64_t *in_off,
                   const int32_t *in_nbr, int32_t source, const int32_t *parent) {
  if (source < 0 || source >= n) return 0;
  const auto depth = bfs_levels(n, out_off, out_nbr, source);
  for (int64_t v = 0; v < n; v++) {
    const bool reached = depth[v] != -1;
    if (reached && parent[v] != -1) {
      if (v == source) {
        if (parent[v] != v || depth[v] != 0) return 0;
        continue;
      }
      bool ok = false;
      for (int64_t k = in_off[v]; k < in_off[v + 1] && !ok; k++) {
        if (in_nbr[k] != parent[v]) continue;
        if (depth[in_nbr[k]] != depth[v] - 1) return 0;
        ok = true;
      }
      if (!ok) return 0;
    } else if (depth[v] != parent[v]) {
      return 0;
    }
  }
  return 1;
}

// CCVerifier (bench/src/cc.cc:290-330): a search from one vertex of each
// label never meets another label, and every vertex is visited.
int gbh_verify_cc(int64_t n, const int64_t *out_off, const int32_t *out_nbr, const int64_t *in_off,
                  const int32_t *in_nbr, int directed, const int32_t *comp) {
  std::unordered_map<int32_t, int32_t> seed;
  for (int64_t v = 0; v < n; v++) seed[comp[v]] = static_cast<int32_t>(v);
  std::vector<uint8_t> seen(n, 0);
  std::vector<int32_t> stack;
  for (const auto &kv : seed) {
    stack.assign(1, kv.second);
    seen[kv.second] = 1;
    while (!stack.empty()) {
      const int32_t x = stack.back();
      stack.pop_back();
      for (int pass = 0; pass < (directed ? 2 : 1); pass++) {
        const int64_t *o = pass ? in_off : out_off;
        const int32_t *a = pass ? in_nbr : out_nbr;
        for (int64_t k = o[x]; k < o[x + 1]; k++) {
          const int32_t y = a[k];
          if (comp[y] != kv.first) return 0;
          if (!seen[y]) {
            seen[y] = 1;
            stack.push_back(y);
          }
        }
      }
    }
  }
  for (int64_t v = 0; v < n; v++)
    if (!seen[v]) return 0;
  return 1;
}

// SSSPVerifier (bench/src/sssp.cc:294-324): serial Dijkstra, exact equality.
// Returns the number of mismatching vertices (0 = PASS).
int64_t gbh_verify_sssp(int64_t n, const int64_t *off, const int32_t *nbr, const int32_t *wgt,
                        int32_t source, const int32_t *dist) {
  const int32_t inf = INT32_MAX / 2;
  std::vector<int32_t> best(n, inf);
  using Item = std::pair<int32_t, int32_t>;
  std::priority_queue<Item, std::vector<Item>, std::greater<Item>> pq;
  best[source] = 0;
  pq.push({0, source});
  while (!pq.empty()) {
    const auto [d, x] = pq.top();
    pq.pop();
    if (d != best[x]) continue;
    for (int64_t k = off[x]; k < off[x + 1]; k++) {
      const int32_t nd = d + wgt[k];
      if (nd < best[nbr[k]]) {
        best[nbr[k]] = nd;
        pq.push({nd, nbr[k]});
      }
    }
  }
  int64_t bad = 0;
  for (int64_t v = 0; v < n; v++) bad += best[v] != dist[v];
  return bad;
}

// BCVerifier (bench/src/bc.cc:256-326): serial Brandes from the same sources,
// normalised, each score within 1e-5 abs + 1e-5 rel.  max_abs_delta (may be
// NULL) receives the largest absolute difference.  Returns 1 = PASS.
int gbh_verify_bc(int64_t n, const int64_t *off, const int32_t *nbr, const int32_t *sources,
                  int num_sources, const float *scores, double *max_abs_delta) {
  std::vector<float> ref(n, 0.0f);
  std::vector<int> depth(n);
  std::vector<double> paths(n);
  std::vector<float> dep(n);
  std::vector<int32_t> order;
  for (int it = 0; it < num_sources; it++) {
    const int32_t s = sources[it];
    std::fill(depth.begin(), depth.end(), -1);
    std::fill(paths.begin(), paths.end(), 0.0);
    std::fill(dep.begin(), dep.end(), 0.0f);
    order.assign(1, s);
    depth[s] = 0;
    paths[s] = 1;
    for (size_t h = 0; h < order.size(); h++) {
      const int32_t x = order[h];
      for (int64_t k = off[x]; k < off[x + 1]; k++) {
        const int32_t y = nbr[k];
        if (depth[y] == -1) {
          depth[y] = depth[x] + 1;
          order.push_back(y);
        }
        if (depth[y] == depth[x] + 1) paths[y] += paths[x];
      }
    }
    // farthest first: the BFS order is non-decreasing in depth
    for (size_t h = order.size(); h-- > 0;) {
      const int32_t x = order[h];
      for (int64_t k = off[x]; k < off[x + 1]; k++) {
        const int32_t y = nbr[k];
        if (depth[y] == depth[x] + 1) dep[x] += (paths[x] / paths[y]) * (1 + dep[y]);
      }
      ref[x] += dep[x];
    }
  }
  const float top = n ? *std::max_element(ref.begin(), ref.end()) : 0.0f;
  int ok = 1;
  double worst = 0;
  for (int64_t v = 0; v < n; v++) {
    const float e = ref[v] / top, a = scores[v];
    const float diff = e > a ? e - a : a - e;
    const float scale = e > a ? e : a;
    worst = std::max(worst, static_cast<double>(diff));
    if (!(diff <= 1e-5f + 1e-5f * scale)) ok = 0;
  }
  if (max_abs_delta) *max_abs_delta = worst;
  return ok;
}

// PRVerifier's residual (bench/src/pr.cc:117-134): one serial push sweep.
double gbh_pr_residual(int64_t n, const int64_t *out_off, const int32_t *out_nbr,
                       const float *scores) {
  const float damp = 0.85f;
  const float base = (1.0f - damp) / n;
  std::vector<float> in(n, 0.0f);
  for (int64_t u = 0; u < n; u++) {
    const int64_t d = out_off[u + 1] - out_off[u];
    const float c = d ? scores[u] / d : 0.0f;
    for (int64_t k = out_off[u]; k < out_off[u + 1]; k++) in[out_nbr[k]] += c;
  }
  double err = 0;
  for (int64_t v = 0; v < n; v++) err += std::fabs(base + damp * in[v] - scores[v]);
  return err;
}

// WeightChecksum and DistanceFingerprint (bench/src/sssp.cc:92-127):
// order-invariant hashes over original ids.  out = {xor, sum}.
void gbh_weight_checksum(int64_t n, const int64_t *off, const int32_t *nbr, const int32_t *wgt,
                         const int32_t *org_ids, uint64_t out[2]) {
  uint64_t x = 0, s = 0;
  for (int64_t u = 0; u < n; u++) {
    const uint64_t a = mix64(static_cast<uint32_t>(org_ids ? org_ids[u] : u));
    for (int64_t k = off[u]; k < off[u + 1]; k++) {
      const uint64_t b = static_cast<uint32_t>(org_ids ? org_ids[nbr[k]] : nbr[k]);
      const uint64_t h = mix64(a ^ (mix64(b) << 1) ^ (mix64(static_cast<uint32_t>(wgt[k])) << 7));
      x ^= h;
      s += h;
    }
  }
  out[0] = x;
  out[1] = s;
}

void gbh_distance_fingerprint(int64_t n, const int32_t *dist, const int32_t *org_ids,
                              uint64_t out[2]) {
  uint64_t x = 0, s = 0;
  for (int64_t v = 0; v < n; v++) {
    const uint64_t o = static_cast<uint32_t>(org_ids ? org_ids[v] : v);
    const uint64_t h = mix64(mix64(o) ^ (mix64(static_cast<uint32_t>(dist[v])) << 1));
    x ^= h;
    s += h;
  }
  out[0] = x;
  out[1] = s;
}

}  // extern "C"
