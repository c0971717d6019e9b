// gbk_host.cc -- host substrate of the bench drivers (libgbkhost.so).
//
// Deterministic pieces that must reproduce the reference bit for bit (the
// generator streams, the random mapping, the source order) use the same
// engines and the same draw order as the reference; everything else is
// organised for the host side of a GPU pipeline (block-parallel generation,
// buffered file parsing).
#include "gbk_host.h"

#include <omp.h>
#include <unistd.h>

#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <limits>
#include <numeric>
#include <random>
#include <string>
#include <unordered_set>
#include <vector>

namespace {

constexpr uint32_t kSeed = 27491095u;            // gapbs/util.h:47 (kRandSeed)
constexpr int64_t kEdgesPerStream = int64_t(1) << 18;  // generator.h:192

// Draws in [0, bound) from a 32-bit engine the way the reference's UniDist
// does (generator.h:32-62): reject the top partial bucket, then reduce.
struct BoundedDraw {
  uint32_t bound;
  uint32_t reject_from;  // 0 = never reject
  explicit BoundedDraw(uint32_t b) : bound(b) {
    const uint64_t range = 4294967296ull;
    const uint32_t tail = static_cast<uint32_t>((range - 1) % b);
    reject_from = (tail == b - 1) ? 0u : static_cast<uint32_t>(4294967295u - tail);
  }
  uint32_t operator()(std::mt19937 &rng) const {
    uint32_t x = static_cast<uint32_t>(rng());
    if (bound == 0) return x;  // full 32-bit range
    if (reject_from)
      while (x >= reject_from) x = static_cast<uint32_t>(rng());
    return x % bound;
  }
};

// One R-MAT edge: 'scale' quadrant choices with a=.57, b=c=.19, d=.05
// expressed as 32-bit thresholds (generator.h:120-121, 131-145).
struct RmatQuadrants {
  uint32_t a, ab, abc;
  RmatQuadrants() {
    const uint32_t top = std::numeric_limits<uint32_t>::max();
    const uint32_t A = 0.57 * top, B = 0.19 * top, C = 0.19 * top;
    a = A;
    ab = A + B;
    abc = A + B + C;
  }
  void edge(std::mt19937 &rng, int scale, int32_t &src, int32_t &dst) const {
    int32_t s = 0, d = 0;
    for (int level = 0; level < scale; level++) {
      const uint32_t r = static_cast<uint32_t>(rng());
      const int32_t row_bit = r >= ab;
      const int32_t col_bit = row_bit ? (r > abc) : (r > a);
      s = (s << 1) | row_bit;
      d = (d << 1) | col_bit;
    }
    src = s;
    dst = d;
  }
};

inline uint64_t splitmix_next(uint64_t &state) {
  uint64_t z = (state += 0x9e3779b97f4a7c15ULL);
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}

bool ends_with(const std::string &s, const char *suf) {
  const size_t k = std::strlen(suf);
  return s.size() >= k && s.compare(s.size() - k, k, suf) == 0;
}

}  // namespace

extern "C" {

int gbh_generate_el(int scale, int degree, int uniform, int32_t *u, int32_t *v) {
  if (scale < 0 || scale > 30 || degree <= 0 || !u || !v) return -1;
  const int64_t n = int64_t(1) << scale;
  const int64_t m = n * degree;
  const int64_t streams = (m + kEdgesPerStream - 1) / kEdgesPerStream;
  if (uniform) {
    // MakeUniformEL (generator.h:101-117): no id permutation afterwards.
    const BoundedDraw draw(static_cast<uint32_t>(n));
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t sidx = 0; sidx < streams; sidx++) {
      std::mt19937 rng(kSeed + static_cast<uint32_t>(sidx));
      const int64_t e1 = std::min(m, (sidx + 1) * kEdgesPerStream);
      for (int64_t e = sidx * kEdgesPerStream; e < e1; e++) {
        // Edge(udist(), udist()) (generator.h:112): the reference's GCC build
        // evaluates the constructor arguments right to left, so the
        // destination takes the first draw of each edge.
        v[e] = static_cast<int32_t>(draw(rng));
        u[e] = static_cast<int32_t>(draw(rng));
      }
    }
    return 0;
  }
  const RmatQuadrants q;
#pragma omp parallel for schedule(dynamic, 1)
  for (int64_t sidx = 0; sidx < streams; sidx++) {
    std::mt19937 rng(kSeed + static_cast<uint32_t>(sidx));
    const int64_t e1 = std::min(m, (sidx + 1) * kEdgesPerStream);
    for (int64_t e = sidx * kEdgesPerStream; e < e1; e++) q.edge(rng, scale, u[e], v[e]);
  }
  // PermuteIDs (generator.h:89-99): one shuffle of the identity.
  std::vector<int32_t> relabel(n);
  std::iota(relabel.begin(), relabel.end(), 0);
  std::mt19937 prng(kSeed);
  std::shuffle(relabel.begin(), relabel.end(), prng);
#pragma omp parallel for schedule(static)
  for (int64_t e = 0; e < m; e++) {
    u[e] = relabel[u[e]];
    v[e] = relabel[v[e]];
  }
  return 0;
}

int gbh_insert_weights(int64_t m, int32_t *w) {
  if (m < 0 || (m > 0 && !w)) return -1;
  const BoundedDraw draw(255u);  // U[0,254] + 1 (generator.h:171-186)
  const int64_t streams = (m + kEdgesPerStream - 1) / kEdgesPerStream;
#pragma omp parallel for schedule(dynamic, 1)
  for (int64_t sidx = 0; sidx < streams; sidx++) {
    std::mt19937 rng(kSeed + static_cast<uint32_t>(sidx));
    const int64_t e1 = std::min(m, (sidx + 1) * kEdgesPerStream);
    for (int64_t e = sidx * kEdgesPerStream; e < e1; e++) w[e] = static_cast<int32_t>(draw(rng) + 1);
  }
  return 0;
}

// ReadInEL / ReadInWEL (reader.h:91-185): skips empty lines and lines that
// start with '%' or '#'; a line contributes an edge when it starts with two
// (three) integers.
int gbh_read_el(const char *path, int weighted, int64_t *m, int32_t *u, int32_t *v, int32_t *w) {
  if (!path || !m) return -1;
  FILE *f = std::fopen(path, "r");
  if (!f) return -2;
  int64_t count = 0;
  char line[4096];
  while (std::fgets(line, sizeof(line), f)) {
    if (line[0] == '\0' || line[0] == '\n' || line[0] == '\r' || line[0] == '%' || line[0] == '#')
      continue;
    char *p = line, *end = nullptr;
    errno = 0;
    const long a = std::strtol(p, &end, 10);
    if (end == p) continue;
    p = end;
    const long b = std::strtol(p, &end, 10);
    if (end == p) continue;
    p = end;
    long c = 1;
    if (weighted) {
      c = std::strtol(p, &end, 10);
      if (end == p) continue;
    }
    if (u) {
      if (count >= *m) break;
      u[count] = static_cast<int32_t>(a);
      v[count] = static_cast<int32_t>(b);
      if (w) w[count] = static_cast<int32_t>(c);
    }
    count++;
  }
  std::fclose(f);
  *m = count;
  return 0;
}

// ------------------------------------------------------ serialized graphs
// Layout of .sg / .wsg files (gapbs/writer.h:336-377, gapbs/reader.h:445-578):
// a one-byte directed flag, int64 edge count m, int64 vertex count n, the
// out-CSR (int64 offsets[n+1], then m neighbours: int32 ids in .sg, (v, w)
// int32 pairs in .wsg), for a directed graph the in-CSR the same way, and
// last the n int32 original ids.
namespace {

struct SgHeader {
  bool directed = false;
  int64_t m = 0, n = 0;
  bool weighted = false;
};

bool sg_suffix(const std::string &p, bool *weighted) {
  if (ends_with(p, ".wsg")) {
    *weighted = true;
    return true;
  }
  if (ends_with(p, ".sg")) {
    *weighted = false;
    return true;
  }
  return false;
}

// Edge weights generated for a .sg file read as weighted (reader.h:48-59):
// scheme 0 "unit" gives 1; scheme 1 "hash" mixes the original ids (ordered
// for an undirected graph) with SplitMix64's finalizer.
int32_t sg_generated_weight(int32_t src, int32_t dst, bool directed, int scheme) {
  if (scheme == 0) return 1;
  uint64_t a = static_cast<uint32_t>(src), b = static_cast<uint32_t>(dst);
  if (!directed && b < a) std::swap(a, b);
  uint64_t z = ((a << 32) ^ b ^ 0x27491095ULL) + 0x9e3779b97f4a7c15ULL;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  z ^= z >> 31;
  return static_cast<int32_t>(z % 255ULL + 1ULL);
}

}  // namespace

int gbh_sg_info(const char *path, int *directed, int64_t *n, int64_t *m, int *weighted) {
  if (!path || !directed || !n || !m || !weighted) return -1;
  SgHeader h;
  if (!sg_suffix(path, &h.weighted)) return -1;
  std::ifstream f(path, std::ios::binary);
  if (!f) return -2;
  char d = 0;
  f.read(&d, 1);
  f.read(reinterpret_cast<char *>(&h.m), 8);
  f.read(reinterpret_cast<char *>(&h.n), 8);
  if (!f || h.m < 0 || h.n < 0 || h.n > INT32_MAX) return -3;
  *directed = d != 0;
  *n = h.n;
  *m = h.m;
  *weighted = h.weighted;
  return 0;
}

int gbh_sg_read(const char *path, int weight_scheme, int64_t *off, int32_t *nbr, int32_t *wgt, int64_t *in_off,
                int32_t *in_nbr, int32_t *in_wgt, int32_t *org_ids) {
  int directed = 0, file_weighted = 0;
  int64_t n = 0, m = 0;
  if (int rc = gbh_sg_info(path, &directed, &n, &m, &file_weighted)) return rc;
  if (!off || !nbr || !org_ids || (weight_scheme != 0 && weight_scheme != 1)) return -1;
  if (directed && (!in_off || !in_nbr || (wgt && !in_wgt))) return -1;
  std::ifstream f(path, std::ios::binary);
  if (!f) return -2;
  const std::streamoff header = 1 + 8 + 8;
  const std::streamoff entry = file_weighted ? 8 : 4;
  const std::streamoff csr_bytes = (n + 1) * 8 + m * entry;
  f.seekg(0, std::ios::end);
  const std::streamoff size = f.tellg();
  if (size != header + csr_bytes * (directed ? 2 : 1) + n * 4) return -3;  // truncated or padded
  f.seekg(header + csr_bytes * (directed ? 2 : 1));
  f.read(reinterpret_cast<char *>(org_ids), n * 4);
  f.seekg(header);
  std::vector<int32_t> pairs;
  auto read_csr = [&](int64_t *o, int32_t *nb, int32_t *w, bool in_csr) -> int {
    f.read(reinterpret_cast<char *>(o), (n + 1) * 8);
    if (!f || o[0] != 0 || o[n] != m) return -3;
    for (int64_t v = 0; v < n; v++)
      if (o[v + 1] < o[v]) return -3;
    if (file_weighted) {
      pairs.resize(2 * static_cast<size_t>(m));
      f.read(reinterpret_cast<char *>(pairs.data()), m * 8);
      for (int64_t e = 0; e < m; e++) {
        nb[e] = pairs[2 * e];
        if (w) w[e] = pairs[2 * e + 1];
      }
    } else {
      f.read(reinterpret_cast<char *>(nb), m * 4);
      if (w) {
        // out-CSR rows are sources, in-CSR rows destinations (reader.h:505-564)
#pragma omp parallel for schedule(dynamic, 1024)
        for (int64_t v = 0; v < n; v++)
          for (int64_t e = o[v]; e < o[v + 1]; e++) {
            const int32_t a = org_ids[v], b = nb[e] >= 0 && nb[e] < n ? org_ids[nb[e]] : 0;
            w[e] = in_csr ? sg_generated_weight(b, a, true, weight_scheme)
                          : sg_generated_weight(a, b, directed != 0, weight_scheme);
          }
      }
    }
    if (!f) return -3;
    for (int64_t e = 0; e < m; e++)
      if (nb[e] < 0 || nb[e] >= n) return -3;
    return 0;
  };
  if (int rc = read_csr(off, nbr, wgt, false)) return rc;
  if (directed)
    if (int rc = read_csr(in_off, in_nbr, wgt ? in_wgt : nullptr, true)) return rc;
  return 0;
}

int gbh_sg_write(const char *path, int directed, int64_t n, int64_t m, const int64_t *off, const int32_t *nbr,
                 const int32_t *wgt, const int64_t *in_off, const int32_t *in_nbr, const int32_t *in_wgt,
                 const int32_t *org_ids) {
  bool weighted = false;
  if (!path || !sg_suffix(path, &weighted) || n < 0 || m < 0 || n > INT32_MAX || !off || !nbr || !org_ids)
    return -1;
  if (weighted != (wgt != nullptr) || (directed && (!in_off || !in_nbr || (weighted && !in_wgt)))) return -1;
  std::ofstream f(path, std::ios::binary);
  if (!f) return -2;
  const char d = directed ? 1 : 0;
  f.write(&d, 1);
  f.write(reinterpret_cast<const char *>(&m), 8);
  f.write(reinterpret_cast<const char *>(&n), 8);
  std::vector<int32_t> pairs;
  auto write_csr = [&](const int64_t *o, const int32_t *nb, const int32_t *w) {
    f.write(reinterpret_cast<const char *>(o), (n + 1) * 8);
    if (!w) {
      f.write(reinterpret_cast<const char *>(nb), m * 4);
      return;
    }
    pairs.resize(2 * static_cast<size_t>(m));
    for (int64_t e = 0; e < m; e++) {
      pairs[2 * e] = nb[e];
      pairs[2 * e + 1] = w[e];
    }
    f.write(reinterpret_cast<const char *>(pairs.data()), m * 8);
  };
  write_csr(off, nbr, wgt);
  if (directed) write_csr(in_off, in_nbr, in_wgt);
  f.write(reinterpret_cast<const char *>(org_ids), n * 4);
  return f ? 0 : -3;
}

int gbh_mapping_identity(int64_t n, int32_t *new_ids) {
  if (n < 0 || (n > 0 && !new_ids)) return -1;
#pragma omp parallel for
  for (int64_t i = 0; i < n; i++) new_ids[i] = static_cast<int32_t>(i);
  return 0;
}

// GenerateRandomMapping (reorder_basic.h:104-135): Fisher-Yates from the
// back with SplitMix64 (state 0) and rejection of the biased low range.
int gbh_mapping_random(int64_t n, int32_t *new_ids) {
  if (n < 0 || (n > 0 && !new_ids)) return -1;
  for (int64_t i = 0; i < n; i++) new_ids[i] = static_cast<int32_t>(i);
  uint64_t state = 0;
  for (uint64_t left = static_cast<uint64_t>(n); left > 1; --left) {
    const uint64_t floor_reject = (0 - left) % left;
    uint64_t r = splitmix_next(state);
    while (r < floor_reject) r = splitmix_next(state);
    std::swap(new_ids[left - 1], new_ids[r % left]);
  }
  return 0;
}

}  // extern "C"

namespace {

// Stable bucket numbering in parallel: vertices get new ids bucket by bucket
// (bucket B-1 first, down to 0), in id order inside a bucket -- the order a
// sequential pass produces.  Each thread counts a contiguous id range, a
// prefix over (bucket, thread) gives every thread its first id per bucket.
template <class BucketOf>
void stable_buckets_desc(int64_t n, int64_t buckets, BucketOf bucket_of, int32_t *new_ids) {
  const int T = std::max(1, omp_get_max_threads());
  std::vector<int64_t> cnt(static_cast<size_t>(T) * buckets, 0);
  const int64_t chunk = (n + T - 1) / T;
#pragma omp parallel num_threads(T)
  {
    const int t = omp_get_thread_num();
    int64_t *c = cnt.data() + static_cast<size_t>(t) * buckets;
    for (int64_t v = t * chunk; v < std::min(n, (t + 1) * chunk); v++) c[bucket_of(v)]++;
  }
  int64_t run = 0;
  for (int64_t b = buckets - 1; b >= 0; b--)
    for (int t = 0; t < T; t++) {
      int64_t &c = cnt[static_cast<size_t>(t) * buckets + b];
      const int64_t x = c;
      c = run;
      run += x;
    }
#pragma omp parallel num_threads(T)
  {
    const int t = omp_get_thread_num();
    int64_t *c = cnt.data() + static_cast<size_t>(t) * buckets;
    for (int64_t v = t * chunk; v < std::min(n, (t + 1) * chunk); v++)
      new_ids[v] = static_cast<int32_t>(c[bucket_of(v)]++);
  }
}

}  // namespace

extern "C" {

// GenerateSortMapping (reorder_basic.h:213-258): stable, degree descending
// -- a stable counting sort on the degree when the degree range is small
// enough for per-thread counters, a stable comparison sort otherwise.
int gbh_mapping_sort(int64_t n, const int64_t *off, int32_t *new_ids) {
  if (n < 0 || (n > 0 && (!off || !new_ids))) return -1;
  if (n == 0) return 0;
  int64_t max_deg = 0;
#pragma omp parallel for reduction(max : max_deg)
  for (int64_t v = 0; v < n; v++) max_deg = std::max<int64_t>(max_deg, off[v + 1] - off[v]);
  if ((max_deg + 1) * std::max(1, omp_get_max_threads()) <= (int64_t{1} << 26)) {
    stable_buckets_desc(n, max_deg + 1, [off](int64_t v) { return off[v + 1] - off[v]; }, new_ids);
    return 0;
  }
  std::vector<int32_t> order(n);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [off](int32_t a, int32_t b) {
    return (off[a + 1] - off[a]) > (off[b + 1] - off[b]);
  });
#pragma omp parallel for
  for (int64_t i = 0; i < n; i++) new_ids[order[i]] = static_cast<int32_t>(i);
  return 0;
}

// GenerateDBGMapping (reorder_hub.h:267-345): degree-based grouping into 8
// buckets with thresholds avg/2, avg, 2avg, ..., 32avg, inf; the highest
// bucket gets the lowest ids; inside a bucket vertices keep id order (the
// reference's static schedule hands each thread a contiguous id range).
int gbh_mapping_dbg(int64_t n, const int64_t *off, int32_t *new_ids) {
  if (n < 0 || (n > 0 && (!off || !new_ids))) return -1;
  if (n == 0) return 0;
  const uint64_t avg = static_cast<uint64_t>(off[n] - off[0]) / static_cast<uint64_t>(n);
  const uint64_t limit[8] = {avg / 2, avg, avg * 2, avg * 4, avg * 8, avg * 16, avg * 32,
                             std::numeric_limits<uint64_t>::max()};
  auto bucket_of = [&](int64_t v) {
    const uint64_t d = static_cast<uint64_t>(off[v + 1] - off[v]);
    int b = 0;
    while (d > limit[b]) b++;
    return static_cast<int64_t>(b);
  };
  stable_buckets_desc(n, 8, bucket_of, new_ids);
  return 0;
}

// MappingFingerprint (builder.h:1398-1411).  The offset basis is the one the
// reference writes, 1469598103934665603.
uint64_t gbh_mapping_fingerprint(int64_t n, const int32_t *new_ids) {
  uint64_t h = 1469598103934665603ULL;
  const uint64_t prime = 1099511628211ULL;
  for (int64_t i = 0; i < n; i++) {
    h = (h ^ static_cast<uint64_t>(i)) * prime;
    h = (h ^ static_cast<uint64_t>(static_cast<int64_t>(new_ids[i]))) * prime;
  }
  return h;
}

int gbh_validate_permutation(int64_t n, const int32_t *ids) {
  if (n < 0 || (n > 0 && !ids)) return -3;
  std::vector<uint64_t> seen((n + 63) / 64, 0);
  for (int64_t i = 0; i < n; i++) {
    const int64_t x = ids[i];
    if (x < 0 || x >= n) return -1;
    uint64_t &word = seen[x >> 6];
    const uint64_t bit = uint64_t(1) << (x & 63);
    if (word & bit) return -2;
    word |= bit;
  }
  return 0;
}

int gbh_mapping_read(const char *path, int64_t *n, int32_t *ids) {
  if (!path || !n) return -1;
  const std::string p(path);
  if (ends_with(p, ".so")) {
    std::ifstream f(p, std::ios::binary | std::ios::ate);
    if (!f) return -2;
    const int64_t bytes = f.tellg();
    if (bytes % 4) return -3;
    if (!ids) {
      *n = bytes / 4;
      return 0;
    }
    f.seekg(0);
    f.read(reinterpret_cast<char *>(ids), std::min<int64_t>(bytes, *n * 4));
    return f ? 0 : -3;
  }
  std::ifstream f(p);
  if (!f) return -2;
  int64_t count = 0;
  long long x;
  while (f >> x) {
    if (ids) {
      if (count >= *n) break;
      ids[count] = static_cast<int32_t>(x);
    }
    count++;
  }
  *n = count;
  return 0;
}

int gbh_mapping_write(const char *path, int64_t n, const int32_t *ids) {
  if (!path || n < 0 || (n > 0 && !ids)) return -1;
  const std::string p(path);
  if (ends_with(p, ".so")) {
    std::ofstream f(p, std::ios::binary);
    if (!f) return -2;
    f.write(reinterpret_cast<const char *>(ids), n * 4);
    return f ? 0 : -3;
  }
  std::ofstream f(p);
  if (!f) return -2;
  for (int64_t i = 0; i < n; i++) f << ids[i] << '\n';
  return f ? 0 : -3;
}

// SourcePicker::GenerateConsistentSources (benchmark.h:115-156): original
// ids of non-isolated vertices, shuffled by mt19937_64(kRandSeed), grouped
// into runs of 'repeats', resolved to internal ids through org_ids.
int gbh_source_picker(int64_t n, const int64_t *off, const int32_t *org_ids, int count,
                      int repeats, int32_t *orig_out, int32_t *internal_out) {
  if (repeats <= 0 || count < 0 || !off) return -2;
  std::vector<int32_t> internal(n);
  for (int64_t i = 0; i < n; i++) {
    const int64_t o = org_ids ? org_ids[i] : i;
    if (o < 0 || o >= n) return -3;
    internal[o] = static_cast<int32_t>(i);
  }
  std::vector<int32_t> candidates;
  candidates.reserve(n);
  for (int64_t o = 0; o < n; o++) {
    const int32_t x = internal[o];
    if (off[x + 1] > off[x]) candidates.push_back(static_cast<int32_t>(o));
  }
  if (candidates.empty()) return -1;
  std::mt19937_64 rng(static_cast<uint64_t>(kSeed));
  std::shuffle(candidates.begin(), candidates.end(), rng);
  const int groups = (count + repeats - 1) / repeats;
  int k = 0;
  for (int gi = 0; gi < groups; gi++) {
    const int32_t o = candidates[gi % candidates.size()];
    for (int r = 0; r < repeats && k < count; r++, k++) {
      orig_out[k] = o;
      internal_out[k] = internal[o];
    }
  }
  return 0;
}

int gbh_afforest_samples(int64_t n, int count, int32_t *idx) {
  if (n <= 0 || count < 0 || (count > 0 && !idx)) return -1;
  std::mt19937 gen;
  std::uniform_int_distribution<int32_t> pick(0, static_cast<int32_t>(n - 1));
  for (int i = 0; i < count; i++) idx[i] = pick(gen);
  return 0;
}

// ComputeTier0SampledGraphFeatures (reorder_types.h:6066-6195).
int gbh_tier0_features(int64_t n, const int64_t *off, const int32_t *nbr, int64_t sample_size,
                       uint64_t seed, double feat[7]) {
  if (!feat) return -1;
  std::fill(feat, feat + 7, 0.0);
  if (n <= 0) return 0;
  if (sample_size == 0) {
    const int64_t root = static_cast<int64_t>(std::sqrt(static_cast<double>(n)));
    sample_size = std::max<int64_t>(1024, std::min<int64_t>(root, 8192));
  }
  sample_size = std::min(sample_size, n);
  const uint64_t shift = (seed * 0x9E3779B97F4A7C15ULL) % static_cast<uint64_t>(n);
  struct Sample {
    int64_t degree, node;
  };
  std::vector<Sample> picks(sample_size);
  double deg_total = 0;
  for (int64_t i = 0; i < sample_size; i++) {
    const uint64_t mid = ((2 * static_cast<uint64_t>(i) + 1) * static_cast<uint64_t>(n)) /
                         (2 * static_cast<uint64_t>(sample_size));
    const int64_t node = static_cast<int64_t>((mid + shift) % static_cast<uint64_t>(n));
    picks[i] = {off[node + 1] - off[node], node};
    deg_total += static_cast<double>(picks[i].degree);
  }
  const double mean = deg_total / sample_size;
  double ss = 0;
  for (const auto &p : picks) ss += (p.degree - mean) * (p.degree - mean);
  const double sd = sample_size > 1 ? std::sqrt(ss / (sample_size - 1)) : 0.0;
  std::sort(picks.begin(), picks.end(), [](const Sample &a, const Sample &b) {
    return a.degree != b.degree ? a.degree > b.degree : a.node < b.node;
  });
  const int64_t hubs = std::max<int64_t>(1, sample_size / 10);
  double hub_total = 0;
  std::unordered_set<int64_t> hub_nodes, hub_lines;
  for (int64_t i = 0; i < hubs; i++) {
    hub_total += static_cast<double>(picks[i].degree);
    hub_nodes.insert(picks[i].node);
  }
  constexpr int64_t kLine = 64 / sizeof(double);
  for (int64_t h : hub_nodes) hub_lines.insert(h / kLine);
  double span = 0;
  int64_t edges = 0, near = 0;
  const double denom = static_cast<double>(std::max<int64_t>(1, n - 1));
  for (const auto &p : picks) {
    for (int64_t k = off[p.node]; k < off[p.node + 1]; k++) {
      const int64_t d = nbr[k];
      span += static_cast<double>(std::llabs(d - p.node)) / denom;
      edges++;
      if (std::llabs(d / kLine - p.node / kLine) <= 1) near++;
    }
  }
  feat[0] = mean;
  feat[1] = mean > 0 ? sd / mean : 0.0;
  feat[2] = deg_total > 0 ? hub_total / deg_total : 0.0;
  feat[3] = edges ? span / edges : 0.0;
  feat[4] = edges ? static_cast<double>(near) / edges : 0.0;
  feat[5] = hub_lines.empty() ? 0.0 : static_cast<double>(hub_nodes.size()) / hub_lines.size();
  feat[6] = n > 1 ? mean / (n - 1) : 0.0;
  return 0;
}

// ModeledPropertyBytes (reorder_types.h:1226-1259); bench codes follow the
// reference's BenchmarkType: 1 PR, 2 BFS, 3 CC, 4 SSSP, 5 BC, 6 TC,
// 7 PR_SPMV, 8 CC_SV.
int64_t gbh_modeled_property_bytes(int bench, int64_t n, int64_t m) {
  const int64_t bits = (n + 7) / 8;
  if (bench == 1 || bench == 7) return 8 * n;
  if (bench == 2) return 8 * n + 2 * bits;
  if (bench == 3 || bench == 8) return 4 * n + bits;
  if (bench == 4) return 4 * n + 4 * m;
  if (bench == 5) return 20 * n + (m + 7) / 8;
  if (bench == 6) return 4 * n;
  return -1;
}

// ShouldUseAllKernelLowReuseRule (reorder_adaptive.h:302-333).
int gbh_frozen_rule(int bench, double reuse, int64_t n, double avg_degree, double degree_cv,
                    double wsr) {
  const bool kernel_ok = bench == 1 || bench == 2 || bench == 3 || bench == 4 || bench == 5 ||
                         bench == 7 || bench == 8;
  if (!kernel_ok || reuse > 2.0 || n < 1000 || wsr > 3.2) return 0;
  if (degree_cv > 8.0) return 1;
  return degree_cv <= 2.68 && (avg_degree <= 60.0 || wsr <= 0.82);
}

// ShouldUseBudgetedLeidenGorder (reorder_adaptive.h:240-257): PageRank
// (either variant) used once on a dense, moderately skewed graph whose hubs
// hold at least 30% of the sampled degree.
int gbh_budgeted_rule(int bench, double reuse, int64_t n, double avg_degree, double degree_cv,
                      double hub_concentration) {
  const bool pr = bench == 1 || bench == 7;
  return pr && reuse <= 1.0 && n >= 1000 && avg_degree >= 45.0 && degree_cv < 3.0 && hub_concentration >= 0.3;
}

// GetLLCSizeBytes (reorder_types.h:6030-6038): the host's L3, else its L2,
// else 30 MB.  The frozen rules' working-set ratio divides by it.
int64_t gbh_llc_bytes(void) {
  long llc = sysconf(_SC_LEVEL3_CACHE_SIZE);
  if (llc > 0) return llc;
  llc = sysconf(_SC_LEVEL2_CACHE_SIZE);
  if (llc > 0) return llc;
  return int64_t(30) * 1024 * 1024;
}

}  // extern "C"

// ---- hub-based mappings (reorder/reorder_hub.h) ---------------------------
// Hubs are vertices whose degree exceeds the integer average m / n.
namespace {

int64_t hub_avg(int64_t n, const int64_t *off) {
  return n > 0 ? (off[n] - off[0]) / n : 0;
}

// Places hubs[j] at new id j; every other vertex keeps its own id unless
// that id is taken by a hub, in which case it moves into a position a hub
// vacated (AssignHubOrderPreservingSourceIds, reorder_hub.h:36-86): a
// non-hub at position p < h whose slot went to a hub takes the old id of the
// hub that now sits at p when that hub came from beyond h; the rest fill the
// remaining vacated positions in order.
int place_hubs(int64_t n, const std::vector<int32_t> &hubs, int32_t *new_ids) {
  const int64_t h = static_cast<int64_t>(hubs.size());
  std::vector<uint8_t> hub(n, 0);
  std::fill(new_ids, new_ids + n, -1);
  for (int64_t j = 0; j < h; j++) {
    hub[hubs[j]] = 1;
    new_ids[hubs[j]] = static_cast<int32_t>(j);
  }
  std::vector<int32_t> spill;
  for (int64_t v = h; v < n; v++) {
    if (!hub[v]) {
      new_ids[v] = static_cast<int32_t>(v);
      continue;
    }
    const int32_t p = new_ids[v];  // the hub's new position, below h
    if (!hub[p] && new_ids[p] == -1) new_ids[p] = static_cast<int32_t>(v);
    else spill.push_back(static_cast<int32_t>(v));
  }
  size_t next = 0;
  for (int64_t v = 0; v < h; v++) {
    if (new_ids[v] != -1) continue;
    if (next >= spill.size()) return -2;
    new_ids[v] = spill[next++];
  }
  return next == spill.size() ? 0 : -2;
}

// Hubs ordered by degree, then id, both descending (std::greater on
// (degree, id) pairs, reorder_hub.h:151-153).
std::vector<int32_t> hubs_by_degree(int64_t n, const int64_t *off) {
  const int64_t avg = hub_avg(n, off);
  std::vector<int32_t> hubs;
  for (int64_t v = 0; v < n; v++)
    if (off[v + 1] - off[v] > avg) hubs.push_back(static_cast<int32_t>(v));
  std::sort(hubs.begin(), hubs.end(), [off](int32_t a, int32_t b) {
    const int64_t da = off[a + 1] - off[a], db = off[b + 1] - off[b];
    return da != db ? da > db : a > b;
  });
  return hubs;
}

std::vector<int32_t> hubs_by_id(int64_t n, const int64_t *off) {
  const int64_t avg = hub_avg(n, off);
  std::vector<int32_t> hubs;
  for (int64_t v = 0; v < n; v++)
    if (off[v + 1] - off[v] > avg) hubs.push_back(static_cast<int32_t>(v));
  return hubs;
}


// Hubs first in the given order, then every other vertex in id order
// (the DBG variants, reorder_hub.h:380-502 and 505-582).
void hubs_then_rest(int64_t n, const std::vector<int32_t> &hubs, int32_t *new_ids) {
  std::vector<uint8_t> hub(n, 0);
  for (int32_t v : hubs) hub[v] = 1;
  // non-hubs keep id order after the hubs: a stable two-bucket numbering
  stable_buckets_desc(n, 2, [&hub](int64_t v) { return static_cast<int64_t>(hub[v]); }, new_ids);
  int32_t next = 0;
  for (int32_t v : hubs) new_ids[v] = next++;
}

}  // namespace

extern "C" {

int gbh_mapping_hubsort(int64_t n, const int64_t *off, int32_t *new_ids) {
  if (n < 0 || (n > 0 && (!off || !new_ids))) return -1;
  return place_hubs(n, hubs_by_degree(n, off), new_ids);
}

int gbh_mapping_hubcluster(int64_t n, const int64_t *off, int32_t *new_ids) {
  if (n < 0 || (n > 0 && (!off || !new_ids))) return -1;
  return place_hubs(n, hubs_by_id(n, off), new_ids);
}

int gbh_mapping_hubsortdbg(int64_t n, const int64_t *off, int32_t *new_ids) {
  if (n < 0 || (n > 0 && (!off || !new_ids))) return -1;
  hubs_then_rest(n, hubs_by_degree(n, off), new_ids);
  return 0;
}

int gbh_mapping_hubclusterdbg(int64_t n, const int64_t *off, int32_t *new_ids) {
  if (n < 0 || (n > 0 && (!off || !new_ids))) return -1;
  hubs_then_rest(n, hubs_by_id(n, off), new_ids);
  return 0;
}

}  // extern "C"

// ---- parameterised R-MAT for the selector's corpus ------------------------
// Not the reference generator (that one is gbh_generate_el): R-MAT after
// Chakrabarti et al. with quadrant probabilities a, b, c, 1-a-b-c, driven by
// a counter-based stream (SplitMix64 of seed and edge index) so that edges
// are generated in parallel and the output does not depend on the thread
// count; ids are optionally relabelled by a seeded Fisher-Yates permutation.
extern "C" int gbh_generate_rmat(int scale, int64_t edges, double a, double b, double c, uint64_t seed,
                                 int permute, int32_t *u, int32_t *v) {
  if (scale < 1 || scale > 30 || edges < 0 || !u || !v || a < 0 || b < 0 || c < 0 || a + b + c > 1.0)
    return -1;
  const double ab = a + b, abc = a + b + c;
#pragma omp parallel for schedule(static)
  for (int64_t e = 0; e < edges; e++) {
    uint64_t state = seed * 0x9E3779B97F4A7C15ULL ^ (static_cast<uint64_t>(e) * 0xD1B54A32D192ED03ULL);
    uint32_t x = 0, y = 0;
    for (int bit = 0; bit < scale; bit++) {
      const double r = static_cast<double>(splitmix_next(state) >> 11) * (1.0 / 9007199254740992.0);
      const uint32_t down = r >= ab, right = (r >= a && r < ab) || r >= abc;
      x |= down << bit;
      y |= right << bit;
    }
    u[e] = static_cast<int32_t>(x);
    v[e] = static_cast<int32_t>(y);
  }
  if (permute) {
    const int64_t n = int64_t{1} << scale;
    std::vector<int32_t> p(n);
    std::iota(p.begin(), p.end(), 0);
    uint64_t state = seed ^ 0x5851F42D4C957F2DULL;
    for (int64_t left = n; left > 1; --left) std::swap(p[left - 1], p[splitmix_next(state) % left]);
#pragma omp parallel for schedule(static)
    for (int64_t e = 0; e < edges; e++) {
      u[e] = p[u[e]];
      v[e] = p[v[e]];
    }
  }
  return 0;
}

// A road-like grid: rows x cols vertices (id = r * cols + c), each joined to
// its right and lower neighbour; with probability 'rewire' an edge's head is
// replaced by a uniform random vertex (self loops and duplicates are left to
// the CSR build's squish).  Weights, when w is given, are uniform in
// [1, max_weight].  Every edge draws from a counter-based stream (SplitMix64
// of the seed and the edge index), so the output does not depend on the
// thread count.  Edge count: rows (cols - 1) + (rows - 1) cols.  No
// reference counterpart: the road-network graphs of the reference's
// experiments are downloaded files; this is their synthetic stand-in.
extern "C" int64_t gbh_generate_lattice(int64_t rows, int64_t cols, double rewire, uint64_t seed, int max_weight,
                                        int32_t *u, int32_t *v, int32_t *w) {
  if (rows < 1 || cols < 1 || rows * cols > INT32_MAX || rewire < 0 || rewire > 1 || max_weight < 1) return -1;
  const int64_t horiz = rows * (cols - 1), total = horiz + (rows - 1) * cols;
  if (!u || !v) return total;
  const uint64_t n = static_cast<uint64_t>(rows * cols);
  const uint64_t cut = static_cast<uint64_t>(rewire * 18446744073709551615.0);
#pragma omp parallel for schedule(static)
  for (int64_t e = 0; e < total; e++) {
    int64_t a, b;
    if (e < horiz) {
      const int64_t r = e / (cols - 1), c = e % (cols - 1);
      a = r * cols + c;
      b = a + 1;
    } else {
      const int64_t k = e - horiz;
      a = k;
      b = k + cols;
    }
    uint64_t state = seed * 0x9E3779B97F4A7C15ULL ^ (static_cast<uint64_t>(e) * 0xD1B54A32D192ED03ULL);
    if (rewire > 0 && splitmix_next(state) <= cut) b = static_cast<int64_t>(splitmix_next(state) % n);
    u[e] = static_cast<int32_t>(a);
    v[e] = static_cast<int32_t>(b);
    if (w) w[e] = static_cast<int32_t>(1 + splitmix_next(state) % static_cast<uint64_t>(max_weight));
  }
  return total;
}
