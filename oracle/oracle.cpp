// oracle.cpp -- CPU restatement of the GraphBrew/GAPBS graph-kernel path.
//
// TEST INFRASTRUCTURE ONLY: the checker for tests/, __graft_entry__.smoke()
// and the timed CPU baseline of bench.py.  The product never links it.
//
// Each function cites the reference function it restates (paths relative to
// the reference root; gapbs/ = bench/include/external/gapbs/).  The kernels
// keep the reference's OpenMP structure so that, run on all host cores, they
// time the same algorithm the reference binaries run; with threads <= 1 they
// run serially, which makes the schedule-dependent ones (PR Gauss-Seidel, BFS
// parents) deterministic for the checker.
//
// Parity pins (see tests/test_oracle.py): mt19937 10000th-output KAT (C++
// standard), the reference's random-mapping KAT (2,6,0,3,4,1,5,7) from
// scripts/test/test_lightweight_baselines.py:76, the relabel fixtures of
// bench/tests/test_relabel_weights.cc:23-134, the tiny graphs under
// scripts/test/data/, closed forms, invariants and brute force.

#include "oracle.h"

#include <omp.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <numeric>
#include <queue>
#include <random>
#include <unordered_map>
#include <unordered_set>
#include <utility>
#include <vector>

namespace {

const int64_t kRandSeed = 27491095;        // gapbs/util.h:47
const int64_t kGenBlock = int64_t(1) << 18;  // gapbs/generator.h:192

struct ThreadScope {
  // Runs a region with a fixed OpenMP team size and restores the old one.
  int saved;
  explicit ThreadScope(int threads) : saved(omp_get_max_threads()) {
    omp_set_num_threads(threads < 1 ? 1 : threads);
  }
  ~ThreadScope() { omp_set_num_threads(saved); }
};

// UniDist (gapbs/generator.h:32-62): maps a 32-bit engine onto [0, max].
class UniDist32 {
 public:
  UniDist32(uint32_t max_value, std::mt19937 &rng) : rng_(rng) {
    no_mod_ = rng_.max() == max_value;
    mod_ = max_value + 1u;
    uint32_t rem = static_cast<uint32_t>(rng_.max() % mod_);
    cutoff_ = (rem == mod_ - 1) ? 0u
                                : static_cast<uint32_t>(rng_.max() - rem);
  }
  uint32_t operator()() {
    uint32_t r = static_cast<uint32_t>(rng_());
    if (no_mod_) return r;
    if (cutoff_ != 0) {
      while (r >= cutoff_) r = static_cast<uint32_t>(rng_());
    }
    return r % mod_;
  }

 private:
  std::mt19937 &rng_;
  bool no_mod_;
  uint32_t mod_;
  uint32_t cutoff_;
};

inline bool CasI32(int32_t *p, int32_t expect, int32_t desired) {
  return __sync_bool_compare_and_swap(p, expect, desired);
}
inline int32_t LoadI32(const int32_t *p) {
  return __atomic_load_n(p, __ATOMIC_RELAXED);
}
inline void StoreI32(int32_t *p, int32_t v) {
  __atomic_store_n(p, v, __ATOMIC_RELAXED);
}

uint64_t Mix64(uint64_t value) {  // bench/src/sssp.cc:77-82
  value += 0x9e3779b97f4a7c15ULL;
  value = (value ^ (value >> 30)) * 0xbf58476d1ce4e5b9ULL;
  value = (value ^ (value >> 27)) * 0x94d049bb133111ebULL;
  return value ^ (value >> 31);
}

inline int64_t Deg(const int64_t *off, int64_t u) { return off[u + 1] - off[u]; }

// Bitmap with atomic set (gapbs/bitmap.h:21-76).
struct Bits {
  std::vector<uint64_t> w;
  explicit Bits(int64_t nbits) : w((nbits + 63) / 64, 0) {}
  void reset() { std::fill(w.begin(), w.end(), 0); }
  bool get(int64_t i) const { return (w[i >> 6] >> (i & 63)) & 1; }
  void set_atomic(int64_t i) {
    __sync_fetch_and_or(&w[i >> 6], uint64_t(1) << (i & 63));
  }
};

}  // namespace

extern "C" {

// ---------------------------------------------------------------------------
// A0: generator (gapbs/generator.h)
// ---------------------------------------------------------------------------

void orc_mt19937(uint32_t seed, int64_t count, uint32_t *out) {
  std::mt19937 rng(seed);
  for (int64_t i = 0; i < count; i++) out[i] = static_cast<uint32_t>(rng());
}

// MakeRMatEL + PermuteIDs (generator.h:89-99, 119-153) and MakeUniformEL
// (generator.h:101-117).  Per-2^18-edge-block reseeding makes the output
// independent of the thread count.
int orc_generate_el(int scale, int degree, int uniform, int32_t *u,
                    int32_t *v) {
  if (scale < 0 || scale > 30 || degree <= 0) return -1;
  const int64_t n = int64_t(1) << scale;
  const int64_t m = n * degree;
  if (uniform) {
#pragma omp parallel
    {
      std::mt19937 rng;
      UniDist32 udist(static_cast<uint32_t>(n - 1), rng);
#pragma omp for
      for (int64_t block = 0; block < m; block += kGenBlock) {
        rng.seed(static_cast<uint32_t>(kRandSeed + block / kGenBlock));
        const int64_t end = std::min(block + kGenBlock, m);
        for (int64_t e = block; e < end; e++) {
          // Edge(udist(), udist()) (generator.h:112): GCC evaluates the
          // arguments right to left, so the destination is drawn first.
          v[e] = static_cast<int32_t>(udist());
          u[e] = static_cast<int32_t>(udist());
        }
      }
    }
    return 0;
  }
  const uint32_t maxu = std::numeric_limits<uint32_t>::max();
  const uint32_t A = 0.57 * maxu, B = 0.19 * maxu, C = 0.19 * maxu;
#pragma omp parallel
  {
    std::mt19937 rng;
#pragma omp for
    for (int64_t block = 0; block < m; block += kGenBlock) {
      rng.seed(static_cast<uint32_t>(kRandSeed + block / kGenBlock));
      const int64_t end = std::min(block + kGenBlock, m);
      for (int64_t e = block; e < end; e++) {
        int32_t src = 0, dst = 0;
        for (int depth = 0; depth < scale; depth++) {
          uint32_t r = static_cast<uint32_t>(rng());
          src <<= 1;
          dst <<= 1;
          if (r < A + B) {
            if (r > A) dst++;
          } else {
            src++;
            if (r > A + B + C) dst++;
          }
        }
        u[e] = src;
        v[e] = dst;
      }
    }
  }
  // PermuteIDs: one std::shuffle of the identity with mt19937(kRandSeed).
  std::vector<int32_t> perm(n);
  std::iota(perm.begin(), perm.end(), 0);
  std::mt19937 prng(static_cast<uint32_t>(kRandSeed));
  std::shuffle(perm.begin(), perm.end(), prng);
#pragma omp parallel for
  for (int64_t e = 0; e < m; e++) {
    u[e] = perm[u[e]];
    v[e] = perm[v[e]];
  }
  return 0;
}

// InsertWeights (generator.h:171-186): U[1,255] via UniDist(254).
int orc_insert_weights(int64_t m, int32_t *w) {
#pragma omp parallel
  {
    std::mt19937 rng;
    UniDist32 udist(254u, rng);
#pragma omp for
    for (int64_t block = 0; block < m; block += kGenBlock) {
      rng.seed(static_cast<uint32_t>(kRandSeed + block / kGenBlock));
      const int64_t end = std::min(block + kGenBlock, m);
      for (int64_t e = block; e < end; e++)
        w[e] = static_cast<int32_t>(udist() + 1);
    }
  }
  return 0;
}

// ---------------------------------------------------------------------------
// A1: CSR build + squish (gapbs/builder.h)
// ---------------------------------------------------------------------------

int32_t orc_max_node_id(int64_t m, const int32_t *u, const int32_t *v) {
  int32_t mx = 0;  // FindMaxNodeID starts from 0 (builder.h:162)
  for (int64_t e = 0; e < m; e++) mx = std::max(mx, std::max(u[e], v[e]));
  return mx;
}

// MakeCSR (builder.h:444-461) followed by SquishCSR (builder.h:254-290):
// per-row sort by (v, w), unique on v (keeps the smallest weight), and
// removal of self loops unless keep_self.
int64_t orc_build_csr(int64_t n, int64_t m, const int32_t *u, const int32_t *v,
                      const int32_t *w, int symmetrize, int transpose,
                      int keep_self, int64_t *off_out, int32_t *nbr_out,
                      int32_t *wgt_out) {
  if (n < 0) return -1;
  std::vector<int64_t> deg(n + 1, 0);
  for (int64_t e = 0; e < m; e++) {
    if (u[e] < 0 || u[e] >= n || v[e] < 0 || v[e] >= n) return -2;
    if (symmetrize || !transpose) deg[u[e]]++;
    if (symmetrize || transpose) deg[v[e]]++;
  }
  std::vector<int64_t> start(n + 1, 0);
  for (int64_t i = 0; i < n; i++) start[i + 1] = start[i] + deg[i];
  std::vector<std::pair<int32_t, int32_t>> rows(start[n]);
  std::vector<int64_t> fill(start.begin(), start.end() - 1);
  for (int64_t e = 0; e < m; e++) {
    const int32_t we = w ? w[e] : 1;
    if (symmetrize || !transpose) rows[fill[u[e]]++] = {v[e], we};
    if (symmetrize || transpose) rows[fill[v[e]]++] = {u[e], we};
  }
  int64_t total = 0;
  off_out[0] = 0;
  for (int64_t r = 0; r < n; r++) {
    auto b = rows.begin() + start[r], e = rows.begin() + start[r + 1];
    if (w)
      std::sort(b, e);  // NodeWeight::operator< is (v, w) (graph.h:136-140)
    else
      std::sort(b, e, [](const auto &x, const auto &y) { return x.first < y.first; });
    int32_t last = -1;
    bool have = false;
    for (auto it = b; it != e; ++it) {
      if (have && it->first == last) continue;  // unique on v (graph.h:143)
      have = true;
      last = it->first;
      if (!keep_self && it->first == r) continue;
      nbr_out[total] = it->first;
      if (wgt_out) wgt_out[total] = it->second;
      total++;
    }
    off_out[r + 1] = total;
  }
  return total;
}

// ---------------------------------------------------------------------------
// A2/A3: relabel (builder.h:1046-1206) and org_ids (graph.h:535-566)
// ---------------------------------------------------------------------------

int orc_validate_permutation(int64_t n, const int32_t *ids) {
  std::vector<uint8_t> seen(n, 0);  // builder.h:2060-2081
  for (int64_t i = 0; i < n; i++) {
    const int64_t x = ids[i];
    if (x < 0 || x >= n) return -1;
    if (seen[x]) return -2;
    seen[x] = 1;
  }
  return 0;
}

int orc_relabel(int64_t n, const int64_t *off, const int32_t *nbr,
                const int32_t *wgt, const int32_t *new_ids, int64_t *off_out,
                int32_t *nbr_out, int32_t *wgt_out) {
  if (orc_validate_permutation(n, new_ids) != 0) return -1;
  std::vector<int64_t> deg(n, 0);
  for (int64_t u = 0; u < n; u++) deg[new_ids[u]] = Deg(off, u);
  off_out[0] = 0;
  for (int64_t r = 0; r < n; r++) off_out[r + 1] = off_out[r] + deg[r];
#pragma omp parallel for schedule(dynamic, 1024)
  for (int64_t u = 0; u < n; u++) {
    const int64_t dst = off_out[new_ids[u]];
    const int64_t d = Deg(off, u);
    if (wgt) {
      std::vector<std::pair<int32_t, int32_t>> row(d);
      for (int64_t k = 0; k < d; k++)
        row[k] = {new_ids[nbr[off[u] + k]], wgt[off[u] + k]};
      std::sort(row.begin(), row.end());
      for (int64_t k = 0; k < d; k++) {
        nbr_out[dst + k] = row[k].first;
        wgt_out[dst + k] = row[k].second;
      }
    } else {
      for (int64_t k = 0; k < d; k++) nbr_out[dst + k] = new_ids[nbr[off[u] + k]];
      std::sort(nbr_out + dst, nbr_out + dst + d);
    }
  }
  return 0;
}

int orc_update_org_ids(int64_t n, const int32_t *org_old,
                       const int32_t *new_ids, int32_t *org_new) {
  for (int64_t i = 0; i < n; i++) {
    if (new_ids[i] < 0 || new_ids[i] >= n) return -1;
    org_new[new_ids[i]] = org_old ? org_old[i] : static_cast<int32_t>(i);
  }
  return 0;
}

// ---------------------------------------------------------------------------
// A4 + mapping producers
// ---------------------------------------------------------------------------

// MappingFingerprint: FNV-1a over (old, new) pairs (builder.h:1398-1411).
uint64_t orc_mapping_fingerprint(int64_t n, const int32_t *new_ids) {
  uint64_t h = 1469598103934665603ULL;
  for (int64_t i = 0; i < n; i++) {
    h ^= static_cast<uint64_t>(i);
    h *= 1099511628211ULL;
    h ^= static_cast<uint64_t>(static_cast<int64_t>(new_ids[i]));
    h *= 1099511628211ULL;
  }
  return h;
}

// GenerateRandomMapping (reorder_basic.h:20-37, 104-135).
int orc_mapping_random(int64_t n, int32_t *new_ids) {
  std::vector<int32_t> perm(n);
  std::iota(perm.begin(), perm.end(), 0);
  uint64_t state = 0;
  auto splitmix = [&state]() {
    uint64_t z = (state += 0x9e3779b97f4a7c15ULL);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
  };
  for (uint64_t remaining = static_cast<uint64_t>(n); remaining > 1; --remaining) {
    const uint64_t threshold = (0 - remaining) % remaining;
    uint64_t x;
    do {
      x = splitmix();
    } while (x < threshold);
    std::swap(perm[remaining - 1], perm[x % remaining]);
  }
  for (int64_t i = 0; i < n; i++) new_ids[i] = perm[i];
  return 0;
}

// GenerateSortMapping, out-degree, descending, stable (reorder_basic.h:213-258).
int orc_mapping_sort(int64_t n, const int64_t *off, int32_t *new_ids) {
  std::vector<std::pair<int64_t, int32_t>> p(n);
  for (int64_t v = 0; v < n; v++) p[v] = {Deg(off, v), static_cast<int32_t>(v)};
  std::stable_sort(p.begin(), p.end(),
                   [](const auto &a, const auto &b) { return a.first > b.first; });
  for (int64_t i = 0; i < n; i++) new_ids[p[i].second] = static_cast<int32_t>(i);
  return 0;
}

// GenerateDBGMapping (reorder_hub.h:267-345) restated with the reference's
// thread-local buckets and per-(bucket, thread) start offsets, serial.
int orc_mapping_dbg(int64_t n, const int64_t *off, int32_t *new_ids) {
  if (n == 0) return 0;
  const uint64_t avg = static_cast<uint64_t>(off[n] - off[0]) / n;
  const uint64_t thr[8] = {avg / 2, avg, avg * 2, avg * 4, avg * 8, avg * 16, avg * 32,
                           std::numeric_limits<uint64_t>::max()};
  const int T = 3;  // any team size gives the same answer with a static schedule
  std::vector<std::vector<std::vector<int32_t>>> local(T, std::vector<std::vector<int32_t>>(8));
  const int64_t chunk = (n + T - 1) / T;
  for (int t = 0; t < T; t++)
    for (int64_t i = t * chunk; i < std::min<int64_t>(n, (t + 1) * chunk); i++) {
      const uint64_t d = static_cast<uint64_t>(Deg(off, i));
      for (int j = 0; j < 8; j++)
        if (d <= thr[j]) {
          local[t][j].push_back(static_cast<int32_t>(i));
          break;
        }
    }
  int64_t k = 0;
  std::vector<std::vector<int64_t>> start(T, std::vector<int64_t>(8));
  for (int j = 7; j >= 0; j--)
    for (int t = 0; t < T; t++) {
      start[t][j] = k;
      k += local[t][j].size();
    }
  for (int t = 0; t < T; t++)
    for (int j = 7; j >= 0; j--) {
      int64_t p = start[t][j];
      for (int32_t v : local[t][j]) new_ids[v] = static_cast<int32_t>(p++);
    }
  return 0;
}

// ---------------------------------------------------------------------------
// A16: SourcePicker::GenerateConsistentSources (benchmark.h:115-156)
// ---------------------------------------------------------------------------

int orc_source_picker(int64_t n, const int64_t *off, const int32_t *org_ids,
                      int count, int repeats, int32_t *orig_out,
                      int32_t *internal_out) {
  if (repeats <= 0 || count < 0) return -2;
  std::vector<int32_t> internal_of(n);
  for (int64_t i = 0; i < n; i++) {
    const int32_t o = org_ids ? org_ids[i] : static_cast<int32_t>(i);
    if (o < 0 || o >= n) return -3;
    internal_of[o] = static_cast<int32_t>(i);
  }
  std::vector<int32_t> active;
  for (int64_t o = 0; o < n; o++)
    if (Deg(off, internal_of[o]) > 0) active.push_back(static_cast<int32_t>(o));
  if (active.empty()) return -1;
  std::mt19937_64 rng(kRandSeed);
  std::shuffle(active.begin(), active.end(), rng);
  const int unique_count = (count + repeats - 1) / repeats;
  int k = 0;
  for (int i = 0; i < unique_count && k < count; i++) {
    const int32_t s = active[i % active.size()];
    for (int r = 0; r < repeats && k < count; r++, k++) {
      orig_out[k] = s;
      internal_out[k] = internal_of[s];
    }
  }
  return 0;
}

// ---------------------------------------------------------------------------
// A5/A6: PageRank
// ---------------------------------------------------------------------------

// PageRankPullGS (bench/src/pr.cc:40-94).
int orc_pr_gs(int64_t n, const int64_t *in_off, const int32_t *in_nbr,
              const int64_t *out_off, int max_iters, double tol, int fixed,
              int threads, float *scores, int *iters_out, double *err_out) {
  ThreadScope ts(threads);
  const float kDamp = 0.85f;
  const float init_score = 1.0f / n;
  const float base_score = (1.0f - kDamp) / n;
  std::vector<float> contrib(n);
  for (int64_t v = 0; v < n; v++) {
    scores[v] = init_score;
    const int64_t d = Deg(out_off, v);
    contrib[v] = d ? init_score / d : 0.0f;
  }
  int executed = 0;
  double final_error = 0;
  for (int iter = 0; iter < max_iters; iter++) {
    double error = 0;
#pragma omp parallel for reduction(+ : error) schedule(dynamic, 16384)
    for (int64_t u = 0; u < n; u++) {
      float incoming = 0;
      for (int64_t k = in_off[u]; k < in_off[u + 1]; k++) {
        float c;
        __atomic_load(&contrib[in_nbr[k]], &c, __ATOMIC_RELAXED);
        incoming += c;
      }
      const float old_score = scores[u];
      scores[u] = base_score + kDamp * incoming;
      error += std::fabs(scores[u] - old_score);
      const int64_t d = Deg(out_off, u);
      const float c = d ? scores[u] / d : 0.0f;
      __atomic_store(&contrib[u], &c, __ATOMIC_RELAXED);
    }
    executed = iter + 1;
    final_error = error;
    if (!fixed && error < tol) break;
  }
  if (iters_out) *iters_out = executed;
  if (err_out) *err_out = final_error;
  return 0;
}

// PageRankPull, Jacobi (bench/src/pr_spmv.cc:40-83).
int orc_pr_spmv(int64_t n, const int64_t *in_off, const int32_t *in_nbr,
                const int64_t *out_off, int max_iters, double tol, int fixed,
                int threads, float *scores, int *iters_out, double *err_out) {
  ThreadScope ts(threads);
  const float kDamp = 0.85f;
  const float init_score = 1.0f / n;
  const float base_score = (1.0f - kDamp) / n;
  std::vector<float> contrib(n);
  for (int64_t v = 0; v < n; v++) scores[v] = init_score;
  int executed = 0;
  double final_error = 0;
  for (int iter = 0; iter < max_iters; iter++) {
    double error = 0;
#pragma omp parallel for
    for (int64_t v = 0; v < n; v++) {
      const int64_t d = Deg(out_off, v);
      contrib[v] = d ? scores[v] / d : 0.0f;
    }
#pragma omp parallel for reduction(+ : error) schedule(dynamic, 16384)
    for (int64_t u = 0; u < n; u++) {
      float incoming = 0;
      for (int64_t k = in_off[u]; k < in_off[u + 1]; k++) incoming += contrib[in_nbr[k]];
      const float old_score = scores[u];
      scores[u] = base_score + kDamp * incoming;
      error += std::fabs(scores[u] - old_score);
    }
    executed = iter + 1;
    final_error = error;
    if (!fixed && error < tol) break;
  }
  if (iters_out) *iters_out = executed;
  if (err_out) *err_out = final_error;
  return 0;
}

// PRVerifier's serial push iteration (pr.cc:117-134).
double orc_pr_residual(int64_t n, const int64_t *out_off,
                       const int32_t *out_nbr, const float *scores) {
  const float kDamp = 0.85f;
  const float base_score = (1.0f - kDamp) / n;
  std::vector<float> incoming(n, 0.0f);
  for (int64_t u = 0; u < n; u++) {
    const int64_t d = Deg(out_off, u);
    const float c = d ? scores[u] / d : 0.0f;
    for (int64_t k = out_off[u]; k < out_off[u + 1]; k++) incoming[out_nbr[k]] += c;
  }
  double error = 0;
  for (int64_t v = 0; v < n; v++) error += std::fabs(base_score + kDamp * incoming[v] - scores[v]);
  return error;
}

// The same pull update in FP64, iterated to an L1 change below tol: the
// fixed point every float run is compared against.
int orc_pr_fp64(int64_t n, const int64_t *in_off, const int32_t *in_nbr,
                const int64_t *out_off, int max_iters, double tol,
                double *scores) {
  const double d = 0.85;
  std::vector<double> contrib(n), next(n);
  for (int64_t v = 0; v < n; v++) scores[v] = 1.0 / n;
  for (int iter = 0; iter < max_iters; iter++) {
#pragma omp parallel for
    for (int64_t v = 0; v < n; v++) {
      const int64_t dv = Deg(out_off, v);
      contrib[v] = dv ? scores[v] / dv : 0.0;
    }
    double change = 0;
#pragma omp parallel for reduction(+ : change)
    for (int64_t u = 0; u < n; u++) {
      double s = 0;
      for (int64_t k = in_off[u]; k < in_off[u + 1]; k++) s += contrib[in_nbr[k]];
      next[u] = (1.0 - d) / n + d * s;
      change += std::fabs(next[u] - scores[u]);
    }
    std::copy(next.begin(), next.end(), scores);
    if (change < tol) break;
  }
  return 0;
}

// ---------------------------------------------------------------------------
// A7/A8: direction-optimizing BFS (bench/src/bfs.cc)
// ---------------------------------------------------------------------------

int orc_bfs_do(int64_t n, const int64_t *out_off, const int32_t *out_nbr,
               const int64_t *in_off, const int32_t *in_nbr, int32_t source,
               int alpha, int beta, int threads, int32_t *parent) {
  if (source < 0 || source >= n) return -1;
  ThreadScope ts(threads);
  // InitParent (bfs.cc:156-162)
#pragma omp parallel for
  for (int64_t v = 0; v < n; v++) {
    const int64_t d = Deg(out_off, v);
    parent[v] = d != 0 ? static_cast<int32_t>(-d) : -1;
  }
  parent[source] = source;
  std::vector<int32_t> queue(n > 0 ? n : 1);
  int64_t q_begin = 0, q_end = 1, q_tail = 1;  // SlidingQueue window + tail
  queue[0] = source;
  Bits curr(n), front(n);
  int64_t edges_to_check = out_off[n] - out_off[0];
  int64_t scout_count = Deg(out_off, source);
  while (q_begin != q_end) {
    if (scout_count > edges_to_check / alpha) {
      // QueueToBitmap (bfs.cc:134-140)
      front.reset();
#pragma omp parallel for
      for (int64_t i = q_begin; i < q_end; i++) front.set_atomic(queue[i]);
      int64_t awake_count = q_end - q_begin, old_awake_count;
      q_begin = q_end;
      do {  // BUStep (bfs.cc:54-89)
        old_awake_count = awake_count;
        awake_count = 0;
        curr.reset();
#pragma omp parallel for reduction(+ : awake_count) schedule(dynamic, 1024)
        for (int64_t u = 0; u < n; u++) {
          if (parent[u] < 0) {
            for (int64_t k = in_off[u]; k < in_off[u + 1]; k++) {
              const int32_t v = in_nbr[k];
              if (front.get(v)) {
                parent[u] = v;
                awake_count++;
                curr.set_atomic(u);
                break;
              }
            }
          }
        }
        std::swap(front.w, curr.w);
      } while ((awake_count >= old_awake_count) || (awake_count > n / beta));
      // BitmapToQueue (bfs.cc:142-154)
      q_tail = q_end;
      for (int64_t v = 0; v < n; v++)
        if (front.get(v)) queue[q_tail++] = static_cast<int32_t>(v);
      q_begin = q_end;
      q_end = q_tail;
      scout_count = 1;
    } else {
      edges_to_check -= scout_count;
      scout_count = 0;
      int64_t sc = 0;
      // TDStep (bfs.cc:92-131) with thread-local queue buffers
#pragma omp parallel reduction(+ : sc)
      {
        std::vector<int32_t> local;
#pragma omp for nowait
        for (int64_t i = q_begin; i < q_end; i++) {
          const int32_t u = queue[i];
          for (int64_t k = out_off[u]; k < out_off[u + 1]; k++) {
            const int32_t v = out_nbr[k];
            const int32_t cur = parent[v];
            if (cur < 0 && CasI32(&parent[v], cur, u)) {
              local.push_back(v);
              sc += -cur;
            }
          }
        }
        const int64_t at = __sync_fetch_and_add(&q_tail, (int64_t)local.size());
        std::copy(local.begin(), local.end(), queue.begin() + at);
      }
      scout_count = sc;
      q_begin = q_end;
      q_end = q_tail;
    }
  }
#pragma omp parallel for
  for (int64_t v = 0; v < n; v++)
    if (parent[v] < -1) parent[v] = -1;
  return 0;
}

int orc_bfs_depths(int64_t n, const int64_t *out_off, const int32_t *out_nbr,
                   int32_t source, int32_t *depth) {
  if (source < 0 || source >= n) return -1;
  std::fill(depth, depth + n, -1);
  std::vector<int32_t> q;
  q.reserve(n);
  depth[source] = 0;
  q.push_back(source);
  for (size_t i = 0; i < q.size(); i++) {
    const int32_t u = q[i];
    for (int64_t k = out_off[u]; k < out_off[u + 1]; k++) {
      const int32_t v = out_nbr[k];
      if (depth[v] == -1) {
        depth[v] = depth[u] + 1;
        q.push_back(v);
      }
    }
  }
  return 0;
}

// BFSVerifier (graphbrew/bfs_common.h:31-97).
int orc_bfs_verify(int64_t n, const int64_t *out_off, const int32_t *out_nbr,
                   const int64_t *in_off, const int32_t *in_nbr,
                   int32_t source, const int32_t *parent) {
  std::vector<int32_t> depth(n);
  if (orc_bfs_depths(n, out_off, out_nbr, source, depth.data()) != 0) return 0;
  for (int64_t v = 0; v < n; v++) {
    if (depth[v] != -1 && parent[v] != -1) {
      if (v == source) {
        if (parent[v] != v || depth[v] != 0) return 0;
        continue;
      }
      bool found = false;
      for (int64_t k = in_off[v]; k < in_off[v + 1]; k++) {
        const int32_t p = in_nbr[k];
        if (p != parent[v]) continue;
        if (depth[p] != depth[v] - 1) return 0;
        found = true;
        break;
      }
      if (!found) return 0;
    } else if (depth[v] != parent[v]) {
      return 0;
    }
  }
  return 1;
}

// ---------------------------------------------------------------------------
// A9-A11: connected components
// ---------------------------------------------------------------------------

namespace {
// Link (cc.cc:82-96).
void Link(int32_t u, int32_t v, int32_t *comp) {
  int32_t p1 = LoadI32(&comp[u]);
  int32_t p2 = LoadI32(&comp[v]);
  while (p1 != p2) {
    const int32_t high = p1 > p2 ? p1 : p2;
    const int32_t low = p1 + (p2 - high);
    const int32_t p_high = LoadI32(&comp[high]);
    if ((p_high == low) || (p_high == high && CasI32(&comp[high], high, low))) break;
    p1 = LoadI32(&comp[LoadI32(&comp[high])]);
    p2 = LoadI32(&comp[low]);
  }
}
// Compress (cc.cc:100-124).
void Compress(int64_t n, int32_t *comp) {
#pragma omp parallel for schedule(dynamic, 16384)
  for (int64_t v = 0; v < n; v++) {
    int32_t p = LoadI32(&comp[v]);
    int32_t gp = LoadI32(&comp[p]);
    while (p != gp) {
      StoreI32(&comp[v], gp);
      p = gp;
      gp = LoadI32(&comp[p]);
    }
  }
}
}  // namespace

// SampleFrequentElement's draw indices: default-seeded std::mt19937 through
// std::uniform_int_distribution<NodeID>(0, n-1) (cc.cc:127-135).
int orc_afforest_sample_indices(int64_t n, int num_samples, int32_t *idx) {
  std::mt19937 gen;
  std::uniform_int_distribution<int32_t> dist(0, static_cast<int32_t>(n - 1));
  for (int i = 0; i < num_samples; i++) idx[i] = dist(gen);
  return 0;
}

// Afforest (cc.cc:153-264).
int orc_cc_afforest(int64_t n, const int64_t *out_off, const int32_t *out_nbr,
                    const int64_t *in_off, const int32_t *in_nbr, int directed,
                    int rounds, int threads, int32_t *comp) {
  ThreadScope ts(threads);
#pragma omp parallel for
  for (int64_t v = 0; v < n; v++) comp[v] = static_cast<int32_t>(v);
  if (n == 0) return 0;
  for (int r = 0; r < rounds; r++) {
#pragma omp parallel for schedule(dynamic, 16384)
    for (int64_t u = 0; u < n; u++) {
      if (out_off[u] + r < out_off[u + 1])
        Link(static_cast<int32_t>(u), out_nbr[out_off[u] + r], comp);
    }
    Compress(n, comp);
  }
  // SampleFrequentElement (cc.cc:127-150)
  const int kSamples = 1024;
  std::vector<int32_t> idx(kSamples);
  orc_afforest_sample_indices(n, kSamples, idx.data());
  std::unordered_map<int32_t, int> counts(32);
  for (int i = 0; i < kSamples; i++) counts[comp[idx[i]]]++;
  auto best = std::max_element(
      counts.begin(), counts.end(),
      [](const auto &a, const auto &b) { return a.second < b.second; });
  const int32_t c = best->first;
#pragma omp parallel for schedule(dynamic, 16384)
  for (int64_t u = 0; u < n; u++) {
    if (LoadI32(&comp[u]) == c) continue;
    for (int64_t k = out_off[u] + rounds; k < out_off[u + 1]; k++)
      Link(static_cast<int32_t>(u), out_nbr[k], comp);
    if (directed)
      for (int64_t k = in_off[u]; k < in_off[u + 1]; k++)
        Link(static_cast<int32_t>(u), in_nbr[k], comp);
  }
  Compress(n, comp);
  return 0;
}

// ShiloachVishkin (cc_sv.cc:64-136).
int orc_cc_sv(int64_t n, const int64_t *out_off, const int32_t *out_nbr,
              int threads, int32_t *comp) {
  ThreadScope ts(threads);
  for (int64_t v = 0; v < n; v++) comp[v] = static_cast<int32_t>(v);
  bool change = true;
  while (change) {
    change = false;
#pragma omp parallel for
    for (int64_t u = 0; u < n; u++) {
      for (int64_t k = out_off[u]; k < out_off[u + 1]; k++) {
        const int32_t cu = LoadI32(&comp[u]);
        const int32_t cv = LoadI32(&comp[out_nbr[k]]);
        if (cu == cv) continue;
        const int32_t high = cu > cv ? cu : cv;
        const int32_t low = cu + (cv - high);
        if (high == LoadI32(&comp[high]) && CasI32(&comp[high], high, low)) {
          __atomic_store_n(&change, true, __ATOMIC_RELAXED);
        }
      }
    }
    Compress(n, comp);
  }
  return 0;
}

// CCVerifier (cc.cc:290-330).
int orc_cc_verify(int64_t n, const int64_t *out_off, const int32_t *out_nbr,
                  const int64_t *in_off, const int32_t *in_nbr, int directed,
                  const int32_t *comp) {
  std::unordered_map<int32_t, int32_t> label_to_source;
  for (int64_t v = 0; v < n; v++) label_to_source[comp[v]] = static_cast<int32_t>(v);
  std::vector<uint8_t> visited(n, 0);
  std::vector<int32_t> frontier;
  frontier.reserve(n);
  for (const auto &ls : label_to_source) {
    const int32_t label = ls.first;
    frontier.clear();
    frontier.push_back(ls.second);
    visited[ls.second] = 1;
    for (size_t i = 0; i < frontier.size(); i++) {
      const int32_t u = frontier[i];
      for (int64_t k = out_off[u]; k < out_off[u + 1]; k++) {
        const int32_t v = out_nbr[k];
        if (comp[v] != label) return 0;
        if (!visited[v]) {
          visited[v] = 1;
          frontier.push_back(v);
        }
      }
      if (directed) {
        for (int64_t k = in_off[u]; k < in_off[u + 1]; k++) {
          const int32_t v = in_nbr[k];
          if (comp[v] != label) return 0;
          if (!visited[v]) {
            visited[v] = 1;
            frontier.push_back(v);
          }
        }
      }
    }
  }
  for (int64_t v = 0; v < n; v++)
    if (!visited[v]) return 0;
  return 1;
}

// ---------------------------------------------------------------------------
// A12/A13: SSSP
// ---------------------------------------------------------------------------

// DeltaStep with thread-local bins and bucket fusion (sssp.cc:129-283).
int orc_sssp_delta(int64_t n, const int64_t *off, const int32_t *nbr,
                   const int32_t *wgt, int32_t source, int32_t delta,
                   int threads, int32_t *dist) {
  if (source < 0 || source >= n || delta <= 0) return -1;
  ThreadScope ts(threads);
  const int32_t kDistInf = std::numeric_limits<int32_t>::max() / 2;
  const size_t kMaxBin = std::numeric_limits<size_t>::max() / 2;
  const size_t kBinSizeThreshold = 1000;
  for (int64_t v = 0; v < n; v++) dist[v] = kDistInf;
  dist[source] = 0;
  std::vector<int32_t> frontier(std::max<int64_t>(off[n] - off[0], 1));
  size_t shared_indexes[2] = {0, kMaxBin};
  size_t frontier_tails[2] = {1, 0};
  frontier[0] = source;
  auto relax = [&](int32_t u, std::vector<std::vector<int32_t>> &bins) {
    for (int64_t k = off[u]; k < off[u + 1]; k++) {
      const int32_t v = nbr[k];
      int32_t old_dist = dist[v];
      const int32_t new_dist = dist[u] + wgt[k];
      while (new_dist < old_dist) {
        if (CasI32(&dist[v], old_dist, new_dist)) {
          const size_t b = new_dist / delta;
          if (b >= bins.size()) bins.resize(b + 1);
          bins[b].push_back(v);
          break;
        }
        old_dist = dist[v];
      }
    }
  };
#pragma omp parallel
  {
    std::vector<std::vector<int32_t>> bins;
    size_t iter = 0;
    while (shared_indexes[iter & 1] != kMaxBin) {
      size_t &curr_bin = shared_indexes[iter & 1];
      size_t &next_bin = shared_indexes[(iter + 1) & 1];
      size_t &curr_tail = frontier_tails[iter & 1];
      size_t &next_tail = frontier_tails[(iter + 1) & 1];
#pragma omp for nowait schedule(dynamic, 64)
      for (size_t i = 0; i < curr_tail; i++) {
        const int32_t u = frontier[i];
        if (dist[u] >= delta * static_cast<int32_t>(curr_bin)) relax(u, bins);
      }
      while (curr_bin < bins.size() && !bins[curr_bin].empty() &&
             bins[curr_bin].size() < kBinSizeThreshold) {
        std::vector<int32_t> copy = bins[curr_bin];
        bins[curr_bin].resize(0);
        for (int32_t u : copy) relax(u, bins);
      }
      for (size_t i = curr_bin; i < bins.size(); i++) {
        if (!bins[i].empty()) {
#pragma omp critical
          next_bin = std::min(next_bin, i);
          break;
        }
      }
#pragma omp barrier
#pragma omp single nowait
      {
        curr_bin = kMaxBin;
        curr_tail = 0;
      }
      if (next_bin < bins.size()) {
        const size_t at = __sync_fetch_and_add(&next_tail, bins[next_bin].size());
        std::copy(bins[next_bin].begin(), bins[next_bin].end(), frontier.begin() + at);
        bins[next_bin].resize(0);
      }
      iter++;
#pragma omp barrier
    }
  }
  return 0;
}

// SSSPVerifier's serial Dijkstra (sssp.cc:294-324).
int orc_dijkstra(int64_t n, const int64_t *off, const int32_t *nbr,
                 const int32_t *wgt, int32_t source, int32_t *dist) {
  if (source < 0 || source >= n) return -1;
  const int32_t kDistInf = std::numeric_limits<int32_t>::max() / 2;
  for (int64_t v = 0; v < n; v++) dist[v] = kDistInf;
  dist[source] = 0;
  using WN = std::pair<int32_t, int32_t>;
  std::priority_queue<WN, std::vector<WN>, std::greater<WN>> mq;
  mq.push({0, source});
  while (!mq.empty()) {
    const int32_t td = mq.top().first, u = mq.top().second;
    mq.pop();
    if (td == dist[u]) {
      for (int64_t k = off[u]; k < off[u + 1]; k++) {
        if (td + wgt[k] < dist[nbr[k]]) {
          dist[nbr[k]] = td + wgt[k];
          mq.push({td + wgt[k], nbr[k]});
        }
      }
    }
  }
  return 0;
}

// WeightChecksum (sssp.cc:92-111).
void orc_weight_checksum(int64_t n, const int64_t *off, const int32_t *nbr,
                         const int32_t *wgt, const int32_t *org_ids,
                         uint64_t out[2]) {
  uint64_t x = 0, s = 0;
  for (int64_t u = 0; u < n; u++) {
    const uint64_t ou = static_cast<uint32_t>(org_ids ? org_ids[u] : u);
    for (int64_t k = off[u]; k < off[u + 1]; k++) {
      const uint64_t ov = static_cast<uint32_t>(org_ids ? org_ids[nbr[k]] : nbr[k]);
      const uint64_t h = Mix64(Mix64(ou) ^ (Mix64(ov) << 1) ^
                               (Mix64(static_cast<uint32_t>(wgt[k])) << 7));
      x ^= h;
      s += h;
    }
  }
  out[0] = x;
  out[1] = s;
}

// DistanceFingerprint (sssp.cc:113-127).
void orc_distance_fingerprint(int64_t n, const int32_t *dist,
                              const int32_t *org_ids, uint64_t out[2]) {
  uint64_t x = 0, s = 0;
  for (int64_t v = 0; v < n; v++) {
    const uint64_t o = static_cast<uint32_t>(org_ids ? org_ids[v] : v);
    const uint64_t d = static_cast<uint32_t>(dist[v]);
    const uint64_t h = Mix64(Mix64(o) ^ (Mix64(d) << 1));
    x ^= h;
    s += h;
  }
  out[0] = x;
  out[1] = s;
}

// ---------------------------------------------------------------------------
// A14/A15: Brandes BC
// ---------------------------------------------------------------------------

// Brandes + PBFS (bc.cc:71-237): level-synchronous forward pass with double
// path counts and an edge-indexed successor bitmap, then a deepest-first
// backward pass; scores normalised by their maximum.
// Score is float for the reference's arithmetic; double gives the exact
// yardstick that float accumulation over a hub's successors drifts from.
extern "C++" {
template <class Score>
int bc_brandes(int64_t n, const int64_t *off, const int32_t *nbr,
                      const int32_t *sources, int num_sources, int threads,
                      Score *scores) {
  ThreadScope ts(threads);
  const int64_t m = off[n] - off[0];
  std::fill(scores, scores + n, Score(0));
  std::vector<double> sigma(n);
  std::vector<int32_t> depths(n);
  std::vector<Score> deltas(n);
  std::vector<int32_t> queue(std::max<int64_t>(n, 1));
  Bits succ(m);
  for (int it = 0; it < num_sources; it++) {
    const int32_t source = sources[it];
    if (source < 0 || source >= n) return -1;
    std::fill(sigma.begin(), sigma.end(), 0.0);
    std::fill(depths.begin(), depths.end(), -1);
    succ.reset();
    std::vector<int64_t> depth_index;
    depths[source] = 0;
    sigma[source] = 1;
    queue[0] = source;
    int64_t q_begin = 0, q_end = 1, tail = 1;
    depth_index.push_back(0);
    int32_t depth = 0;
    while (q_begin != q_end) {
      depth++;
#pragma omp parallel
      {
        std::vector<int32_t> local;
#pragma omp for schedule(dynamic, 64) nowait
        for (int64_t i = q_begin; i < q_end; i++) {
          const int32_t u = queue[i];
          for (int64_t k = off[u]; k < off[u + 1]; k++) {
            const int32_t v = nbr[k];
            if (depths[v] == -1 && CasI32(&depths[v], -1, depth)) local.push_back(v);
            if (depths[v] == depth) {
              succ.set_atomic(k - off[0]);
#pragma omp atomic
              sigma[v] += sigma[u];
            }
          }
        }
        const int64_t at = __sync_fetch_and_add(&tail, (int64_t)local.size());
        std::copy(local.begin(), local.end(), queue.begin() + at);
      }
      depth_index.push_back(q_end);
      q_begin = q_end;
      q_end = tail;
    }
    depth_index.push_back(q_end);
    std::fill(deltas.begin(), deltas.end(), Score(0));
    for (int d = static_cast<int>(depth_index.size()) - 2; d >= 0; d--) {
#pragma omp parallel for schedule(dynamic, 64)
      for (int64_t i = depth_index[d]; i < depth_index[d + 1]; i++) {
        const int32_t u = queue[i];
        Score delta_u = 0;
        for (int64_t k = off[u]; k < off[u + 1]; k++) {
          if (succ.get(k - off[0])) {
            const int32_t v = nbr[k];
            delta_u += (sigma[u] / sigma[v]) * (1 + deltas[v]);
          }
        }
        deltas[u] = delta_u;
        scores[u] += delta_u;
      }
    }
  }
  Score biggest = 0;
  for (int64_t v = 0; v < n; v++) biggest = std::max(biggest, scores[v]);
  for (int64_t v = 0; v < n; v++) scores[v] = scores[v] / biggest;
  return 0;
}
}  // extern "C++"

int orc_bc(int64_t n, const int64_t *off, const int32_t *nbr,
           const int32_t *sources, int num_sources, int threads,
           float *scores) {
  return bc_brandes<float>(n, off, nbr, sources, num_sources, threads, scores);
}

// The same traversal with double deltas and scores.
int orc_bc_fp64(int64_t n, const int64_t *off, const int32_t *nbr,
                const int32_t *sources, int num_sources, int threads,
                double *scores) {
  return bc_brandes<double>(n, off, nbr, sources, num_sources, threads, scores);
}

// BCVerifier's serial recomputation (bc.cc:256-300).
int orc_bc_serial(int64_t n, const int64_t *off, const int32_t *nbr,
                  const int32_t *sources, int num_sources, float *scores) {
  std::fill(scores, scores + n, 0.0f);
  for (int it = 0; it < num_sources; it++) {
    const int32_t source = sources[it];
    if (source < 0 || source >= n) return -1;
    std::vector<int> depths(n, -1);
    depths[source] = 0;
    std::vector<double> sigma(n, 0);
    sigma[source] = 1;
    std::vector<int32_t> to_visit;
    to_visit.reserve(n);
    to_visit.push_back(source);
    for (size_t i = 0; i < to_visit.size(); i++) {
      const int32_t u = to_visit[i];
      for (int64_t k = off[u]; k < off[u + 1]; k++) {
        const int32_t v = nbr[k];
        if (depths[v] == -1) {
          depths[v] = depths[u] + 1;
          to_visit.push_back(v);
        }
        if (depths[v] == depths[u] + 1) sigma[v] += sigma[u];
      }
    }
    std::vector<std::vector<int32_t>> at_depth;
    for (int64_t v = 0; v < n; v++) {
      if (depths[v] != -1) {
        if (depths[v] >= static_cast<int>(at_depth.size())) at_depth.resize(depths[v] + 1);
        at_depth[depths[v]].push_back(static_cast<int32_t>(v));
      }
    }
    std::vector<float> deltas(n, 0.0f);
    for (int d = static_cast<int>(at_depth.size()) - 1; d >= 0; d--) {
      for (int32_t u : at_depth[d]) {
        for (int64_t k = off[u]; k < off[u + 1]; k++) {
          const int32_t v = nbr[k];
          if (depths[v] == depths[u] + 1)
            deltas[u] += (sigma[u] / sigma[v]) * (1 + deltas[v]);
        }
        scores[u] += deltas[u];
      }
    }
  }
  const float biggest = n > 0 ? *std::max_element(scores, scores + n) : 0.0f;
  for (int64_t v = 0; v < n; v++) scores[v] = scores[v] / biggest;
  return 0;
}

int orc_bc_within_tolerance(float expected, float actual) {
  const float kAbs = 1e-5f, kRel = 1e-5f;  // bc.cc:52-62
  const float delta = expected > actual ? expected - actual : actual - expected;
  const float scale = expected > actual ? expected : actual;
  return delta <= kAbs + kRel * scale;
}

// ---------------------------------------------------------------------------
// A18: Tier-0 sampled features + frozen rule
// ---------------------------------------------------------------------------

// ComputeTier0SampledGraphFeatures (reorder_types.h:6066-6195).
int orc_tier0_features(int64_t n, int64_t m_directed, const int64_t *off,
                       const int32_t *nbr, int64_t sample_size,
                       uint64_t sample_seed, double feat[7]) {
  (void)m_directed;
  for (int i = 0; i < 7; i++) feat[i] = 0;
  if (n <= 0) return 0;
  if (sample_size == 0)
    sample_size = std::max<int64_t>(
        1024, std::min<int64_t>(static_cast<int64_t>(std::sqrt(static_cast<double>(n))), 8192));
  sample_size = std::min<int64_t>(sample_size, n);
  std::vector<std::pair<int64_t, int64_t>> dn;
  dn.reserve(sample_size);
  double degree_sum = 0;
  const uint64_t offset = (sample_seed * 0x9E3779B97F4A7C15ULL) % static_cast<uint64_t>(n);
  for (int64_t i = 0; i < sample_size; i++) {
    const uint64_t base = ((2 * static_cast<uint64_t>(i) + 1) * static_cast<uint64_t>(n)) /
                          (2 * static_cast<uint64_t>(sample_size));
    const int64_t node = static_cast<int64_t>((base + offset) % static_cast<uint64_t>(n));
    const int64_t d = Deg(off, node);
    dn.emplace_back(d, node);
    degree_sum += d;
  }
  const double avg = degree_sum / sample_size;
  double var_sum = 0;
  for (const auto &p : dn) var_sum += (p.first - avg) * (p.first - avg);
  const double variance = sample_size > 1 ? var_sum / (sample_size - 1) : 0.0;
  const double cv = avg > 0 ? std::sqrt(variance) / avg : 0.0;
  std::sort(dn.begin(), dn.end(), [](const auto &l, const auto &r) {
    if (l.first != r.first) return l.first > r.first;
    return l.second < r.second;
  });
  const int64_t hub_count = std::max<int64_t>(1, sample_size / 10);
  double hub_sum = 0;
  std::unordered_set<int64_t> hubs;
  for (int64_t i = 0; i < hub_count; i++) {
    hub_sum += dn[i].first;
    hubs.insert(dn[i].second);
  }
  const int64_t kPerLine = 64 / static_cast<int64_t>(sizeof(double));
  std::unordered_set<int64_t> lines;
  for (int64_t h : hubs) lines.insert(h / kPerLine);
  double span_sum = 0;
  int64_t sampled = 0, window = 0;
  const double denom = static_cast<double>(std::max<int64_t>(1, n - 1));
  for (const auto &p : dn) {
    const int64_t node = p.second, node_line = node / kPerLine;
    for (int64_t k = off[node]; k < off[node + 1]; k++) {
      const int64_t d = nbr[k];
      span_sum += static_cast<double>(std::llabs(d - node)) / denom;
      sampled++;
      if (std::llabs(d / kPerLine - node_line) <= 1) window++;
    }
  }
  feat[0] = avg;
  feat[1] = cv;
  feat[2] = degree_sum > 0 ? hub_sum / degree_sum : 0.0;
  feat[3] = sampled > 0 ? span_sum / sampled : 0.0;
  feat[4] = sampled > 0 ? static_cast<double>(window) / sampled : 0.0;
  feat[5] = lines.empty() ? 0.0 : static_cast<double>(hubs.size()) / lines.size();
  feat[6] = n > 1 ? avg / (n - 1) : 0.0;
  return 0;
}

// ModeledPropertyBytes (reorder_types.h:1226-1259); bench codes follow
// BenchmarkType (reorder_types.h:1214-1224): 1 PR, 2 BFS, 3 CC, 4 SSSP,
// 5 BC, 6 TC, 7 PR_SPMV, 8 CC_SV.
int64_t orc_modeled_property_bytes(int bench, int64_t n, int64_t m) {
  const int64_t bitmap = (n + 7) / 8;
  switch (bench) {
    case 1: case 7: return n * 2 * 4;
    case 2: return n * 2 * 4 + 2 * bitmap;
    case 3: case 8: return n * 4 + bitmap;
    case 4: return n * 4 + m * 4;
    case 5: return n * (4 + 8 + 2 * 4) + (m + 7) / 8;
    case 6: return n * 4;
    default: return -1;
  }
}

// ShouldUseAllKernelLowReuseRule (reorder_adaptive.h:302-333).
int orc_frozen_rule(int bench, double reuse, int64_t n, double avg_degree,
                    double degree_cv, double wsr) {
  const bool supported = bench == 1 || bench == 7 || bench == 2 || bench == 3 ||
                         bench == 8 || bench == 5 || bench == 4;
  return supported && reuse <= 2.0 && n >= 1000 && wsr <= 3.2 &&
         ((degree_cv <= 2.68 && (avg_degree <= 60.0 || wsr <= 0.82)) ||
          degree_cv > 8.0);
}

}  // extern "C"
