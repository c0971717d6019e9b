// gbk_many.hip -- many small graphs in one launch: PR and CC over the
// disjoint union of a batch of graphs.
//
// The reference runs one graph per process (bench/src/pr.cc, cc.cc main).
// For batches of small graphs a launch per graph leaves the GPU idle; here
// the graphs are concatenated into one CSR (vertex and edge ids shifted by
// each graph's base) and one pass of every kernel covers all of them.
//   CC   components of a disjoint union are the components of its parts and
//        the smallest id of a component shifts with its graph's base, so
//        running CC on the union and subtracting the base gives each graph's
//        labels bit for bit.
//   PR   each vertex uses its own graph's base score (1 - d) / n_g; the L1
//        change is summed per graph (wave-segmented by graph id, one atomic
//        per run of a graph's rows) and a graph whose change falls below the
//        tolerance stops updating, exactly when a run on that graph alone
//        would stop (pr.cc:77-88).
#include <algorithm>
#include <vector>

#include "gbk_expand.h"

int gbk_graph_alloc(gbk_ctx *ctx, int64_t n, int64_t m, int64_t m_in, int directed, int weighted,
                    gbk_graph **out);
int gbk_graph_finalize(gbk_ctx *ctx, gbk_graph *g);
int gbk_cc_device(gbk_ctx *ctx, const gbk_graph *g, int variant, int32_t *comp, gbk_cc_stats *stats);

namespace gbk {

__global__ void union_rows(const int64_t *off, int64_t n, int64_t vbase, int64_t ebase, int64_t *off_out,
                           const int32_t *org, int32_t *org_out) {
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t v = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; v < n; v += stride) {
    off_out[vbase + v] = off[v] - off[0] + ebase;
    org_out[vbase + v] = static_cast<int32_t>(vbase + org[v]);
  }
}

__global__ void union_edges(const int32_t *nbr, const int32_t *wgt, int64_t m, int64_t vbase, int64_t ebase,
                            int32_t *nbr_out, int32_t *wgt_out) {
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; e < m; e += stride) {
    nbr_out[ebase + e] = static_cast<int32_t>(nbr[e] + vbase);
    if (wgt) wgt_out[ebase + e] = wgt[e];
  }
}

__global__ void store_i64(int64_t *p, int64_t v) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *p = v;
}

// gid[v] = graph of vertex v (vbase has count + 1 entries).
__global__ void graph_ids(const int64_t *vbase, int count, int64_t n, int32_t *gid) {
  const int64_t v = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (v >= n) return;
  int lo = 0, hi = count;  // last graph with vbase[g] <= v
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (vbase[mid] <= v) lo = mid;
    else hi = mid;
  }
  gid[v] = lo;
}

__global__ void shift_labels(int32_t *comp, const int32_t *gid, const int64_t *vbase, int64_t n) {
  const int64_t v = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (v < n) comp[v] = static_cast<int32_t>(comp[v] - vbase[gid[v]]);
}

struct ManyState {
  int done;
  int iters;
  int remaining;
  int pad;
};

__global__ void many_init(const int64_t *out_off, const int32_t *gid, const float *gbase_init, int64_t n,
                          float *scores, float *contrib) {
  const int64_t v = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (v < n) {
    const int64_t d = out_off[v + 1] - out_off[v];
    const float init = gbase_init[gid[v]];
    scores[v] = init;
    contrib[v] = d ? init / static_cast<float>(d) : 0.0f;
  }
}

__global__ void many_graph_init(const int64_t *vbase, int count, float damp, float *gbase, float *ginit,
                                double *gerr, int *gdone, int *giters, ManyState *st) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g < count) {
    const float ng = static_cast<float>(vbase[g + 1] - vbase[g]);
    ginit[g] = 1.0f / ng;
    gbase[g] = (1.0f - damp) / ng;
    gerr[g] = 0.0;
    gdone[g] = 0;
    giters[g] = 0;
  }
  if (g == 0) {
    st->done = 0;
    st->iters = 0;
    st->remaining = count;
  }
}

// Pull iteration over the union (same units as the one-graph kernel: heavy
// chunks into hacc, then one wave per tile).  Rows of finished graphs keep
// their scores.
//
// Gauss-Seidel (GS, contributions updated in place) walks the tiles in the
// order 'order' gives per XCD (xoff[x] .. xoff[x + 1] for the workgroups
// with blockIdx.x % 8 == x, which the GPU dispatches to XCD x): every part
// lives on one XCD, so its in-place updates land in that XCD's L2 where its
// other waves read them (agent-scope loads, past the CU's L1), and its
// tiles come in order of their position inside the part, so the waves in
// flight sweep each part front to back as the reference's few-thread sweep
// does.  Spread over all XCDs and all at once, a small part's rows read
// almost only last sweep's values and its L1 change shrinks far more slowly
// than the reference's (measured on the bench's 256 R-MAT parts: 20 parts
// still above tol after 20 sweeps and 8 failing the reference's verifier;
// with this order 3 and 0).  Jacobi (GS false) reads the previous sweep only
// and walks the tiles in index order.
template <bool GS>
__device__ __forceinline__ float read_contrib(const float *c, int32_t v) {
  if constexpr (GS)
    return __hip_atomic_load(const_cast<float *>(c) + v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  else
    return c[v];
}

template <bool GS>
__global__ void __launch_bounds__(kBlock)
many_pull_iter(RowTiles rt, const int32_t *in_nbr, const int64_t *out_off, const int32_t *gid,
               const float *gbase, const int *gdone, float *scores, const float *contrib_cur,
               float *contrib_next, double *gerr, float *hacc, const ManyState *st, float damp,
               const int32_t *order, const int32_t *xoff) {
  if (*reinterpret_cast<const volatile int *>(&st->done)) return;
  __shared__ float acc[kBlock];
  __shared__ float hred[kWavesPerBlock];
  const int lane = lane_id();
  const int wv = threadIdx.x / kWave;
  const int64_t chunks = rt.nheavy ? rt.hprefix[rt.nheavy] : 0;
  for (int64_t gc = blockIdx.x; gc < chunks; gc += gridDim.x) {
    int64_t h, first, last;
    heavy_chunk_of(rt, gc, h, first, last);
    const int64_t s = rt.off[rt.heavy[h]];
    float part = 0;
    for (int64_t k = first + threadIdx.x; k < last; k += blockDim.x)
      part += read_contrib<GS>(contrib_cur, in_nbr[s + k]);
    const float sum = block_sum(part, hred);
    if (threadIdx.x == 0) atomicAdd(&hacc[h], sum);
  }
  auto tile = [&](int64_t t) {
    const int64_t u = rt.tiles[t] + lane;
    const bool valid = u < rt.tiles[t + 1];
    int64_t start = 0, len = 0;
    int g = -1;
    bool light = false;
    if (valid) {
      g = gid[u];
      start = rt.off[u];
      len = rt.off[u + 1] - start;
      light = len <= rt.heavy_deg && !gdone[g];
    }
    if (!light) len = 0;
    acc[threadIdx.x] = 0.0f;
    const int64_t incl = wave_incl_scan(len);
    const int64_t excl = incl - len;
    const int64_t total = __shfl(incl, kWave - 1);
    for (int64_t c = 0; c < total; c += kWave) {
      const int64_t e = c + lane;
      int lo = 0;
#pragma unroll
      for (int step = 32; step >= 1; step >>= 1) {
        const int probe = lo + step;
        const int64_t x = __shfl(excl, probe & (kWave - 1));
        if (probe < kWave && x <= e) lo = probe;
      }
      const int64_t ostart = __shfl(start, lo);
      const int64_t oexcl = __shfl(excl, lo);
      const bool act = e < total;
      const float val = act ? read_contrib<GS>(contrib_cur, in_nbr[ostart + (e - oexcl)]) : 0.0f;
      bool tail;
      const float run = seg_run_sum(val, act ? lo : kWave, tail);
      if (act && tail) acc[wv * kWave + lo] += run;
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "workgroup");
    __builtin_amdgcn_wave_barrier();
    double err = 0.0;
    if (light) {
      const float old = scores[u];
      const float nw = gbase[g] + damp * acc[threadIdx.x];
      scores[u] = nw;
      err = fabs(static_cast<double>(nw - old));
      const int64_t d = out_off[u + 1] - out_off[u];
      contrib_next[u] = d ? nw / static_cast<float>(d) : 0.0f;
    }
    // rows of a tile are consecutive, so graph ids form runs across lanes;
    // the run's last lane adds the run (whether or not its own row was light)
    bool tail;
    const double run = seg_run_sum(err, valid ? g : -1 - lane, tail);
    if (valid && tail && run != 0.0) atomicAdd(&gerr[g], run);
    __builtin_amdgcn_wave_barrier();
  };
  if (order) {
    // the launch has a multiple of 8 workgroups: blockIdx.x % 8 is the XCD
    const int x = static_cast<int>(blockIdx.x & 7);
    const int64_t xwaves = static_cast<int64_t>(gridDim.x >> 3) * kWavesPerBlock;
    for (int64_t i = xoff[x] + static_cast<int64_t>(blockIdx.x >> 3) * kWavesPerBlock + wv; i < xoff[x + 1];
         i += xwaves)
      tile(order[i]);
  } else {
    const int64_t waves = static_cast<int64_t>(gridDim.x) * kWavesPerBlock;
    for (int64_t t = static_cast<int64_t>(blockIdx.x) * kWavesPerBlock + wv; t < rt.ntiles; t += waves) tile(t);
  }
}

// One workgroup: heavy rows, then each unfinished graph's convergence test.
__global__ void many_finish(const int32_t *heavy, int64_t nheavy, float *hacc, const int32_t *gid,
                            const float *gbase, int *gdone, int *giters, double *gerr, int count, int iter,
                            double tol, int fixed, const int64_t *out_off, float *scores, float *contrib_next,
                            float damp, ManyState *st) {
  if (*reinterpret_cast<volatile int *>(&st->done)) return;
  for (int64_t h = threadIdx.x; h < nheavy; h += blockDim.x) {
    const int32_t u = heavy[h];
    const int g = gid[u];
    const float sum = hacc[h];
    hacc[h] = 0.0f;
    if (gdone[g]) continue;
    const float old = scores[u];
    const float nw = gbase[g] + damp * sum;
    scores[u] = nw;
    atomicAdd(&gerr[g], fabs(static_cast<double>(nw - old)));
    const int64_t d = out_off[u + 1] - out_off[u];
    contrib_next[u] = d ? nw / static_cast<float>(d) : 0.0f;
  }
  __threadfence();
  __syncthreads();
  __shared__ int left[1024 / kWave];
  int mine = 0;
  for (int g = threadIdx.x; g < count; g += blockDim.x) {
    if (gdone[g]) continue;
    const double e = __hip_atomic_load(&gerr[g], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    gerr[g] = 0.0;
    giters[g] = iter + 1;
    if (!fixed && e < tol) gdone[g] = 1;
    else mine++;
  }
  mine = wave_sum(mine);
  if (lane_id() == 0) left[threadIdx.x / kWave] = mine;
  __syncthreads();
  if (threadIdx.x == 0) {
    int t = 0;
    for (int i = 0; i < static_cast<int>(blockDim.x) / kWave; i++) t += left[i];
    st->remaining = t;
    st->iters = iter + 1;
    if (t == 0) st->done = 1;
  }
}

}  // namespace gbk

using namespace gbk;

extern "C" int gbk_graph_union(gbk_ctx *ctx, const gbk_graph *const *graphs, int count, int64_t *vbase,
                               gbk_graph **out) {
  if (!ctx || !graphs || count <= 0 || !out) return set_error(ctx, GBK_EINVAL, "union: bad arguments");
  int64_t n = 0, m = 0;
  for (int i = 0; i < count; i++) {
    const gbk_graph *g = graphs[i];
    if (!g || g->directed != graphs[0]->directed || g->weighted != graphs[0]->weighted)
      return set_error(ctx, GBK_EINVAL, "union: graphs must all be directed or all undirected, and alike in weights");
    n += g->n;
    m += g->m;
  }
  if (n >= INT32_MAX) return set_error(ctx, GBK_EINVAL, "union: more than 2^31 vertices");
  const int directed = graphs[0]->directed, weighted = graphs[0]->weighted;
  gbk_graph *u = nullptr;
  if (int rc = gbk_graph_alloc(ctx, n, m, m, directed, weighted, &u)) return rc;
  hipStream_t s = ctx->stream;
  int64_t vb = 0, eb = 0;
  for (int i = 0; i < count; i++) {
    const gbk_graph *g = graphs[i];
    if (vbase) vbase[i] = vb;
    if (g->n > 0)
      union_rows<<<grid_for(g->n), kBlock, 0, s>>>(g->off, g->n, vb, eb, u->off, g->org_ids, u->org_ids);
    if (g->m > 0)
      union_edges<<<grid_for(g->m), kBlock, 0, s>>>(g->nbr, g->wgt, g->m, vb, eb, u->nbr, u->wgt);
    if (directed) {
      if (g->n > 0)
        union_rows<<<grid_for(g->n), kBlock, 0, s>>>(g->in_off, g->n, vb, eb, u->in_off, g->org_ids, u->org_ids);
      if (g->m > 0)
        union_edges<<<grid_for(g->m), kBlock, 0, s>>>(g->in_nbr, g->in_wgt, g->m, vb, eb, u->in_nbr, u->in_wgt);
    }
    vb += g->n;
    eb += g->m;
  }
  if (vbase) vbase[count] = vb;
  store_i64<<<1, 1, 0, s>>>(u->off + n, m);
  if (directed) store_i64<<<1, 1, 0, s>>>(u->in_off + n, m);
  GBK_LAUNCH_CHECK(ctx, "graph union");
  if (int rc = gbk_graph_finalize(ctx, u)) {
    gbk_graph_free(ctx, u);
    return rc;
  }
  *out = u;
  return GBK_OK;
}

namespace {

// Tile order of the union's Gauss-Seidel sweep (see many_pull_iter): each
// tile belongs to the part of its first row; parts go to the 8 XCDs largest
// first onto the XCD with the fewest tiles; an XCD's tiles are ordered by
// their position inside their part (then part, then index).  order gets
// every tile once, xoff the 9 XCD offsets.
void xcd_tile_order(const std::vector<int32_t> &tiles, const int64_t *vbase, int count, std::vector<int32_t> &order,
                    std::vector<int32_t> &xoff) {
  const int64_t ntiles = static_cast<int64_t>(tiles.size()) - 1;
  std::vector<int32_t> part(ntiles);
  std::vector<int64_t> first(count, -1), cnt(count, 0);
  int p = 0;
  for (int64_t t = 0; t < ntiles; t++) {
    while (p + 1 < count && vbase[p + 1] <= tiles[t]) p++;
    part[t] = p;
    if (first[p] < 0) first[p] = t;
    cnt[p]++;
  }
  std::vector<int> by_size(count);
  for (int i = 0; i < count; i++) by_size[i] = i;
  std::stable_sort(by_size.begin(), by_size.end(), [&](int a, int b) { return cnt[a] > cnt[b]; });
  std::vector<int> xcd_of(count, 0);
  int64_t load[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (int q : by_size) {
    int best = 0;
    for (int x = 1; x < 8; x++)
      if (load[x] < load[best]) best = x;
    xcd_of[q] = best;
    load[best] += cnt[q];
  }
  struct Key {
    int64_t pos;
    int32_t part, t;
  };
  std::vector<std::vector<Key>> per(8);
  for (int64_t t = 0; t < ntiles; t++) {
    const int q = part[t];
    per[xcd_of[q]].push_back({((t - first[q]) << 24) / cnt[q], q, static_cast<int32_t>(t)});
  }
  order.clear();
  xoff.assign(9, 0);
  for (int x = 0; x < 8; x++) {
    std::sort(per[x].begin(), per[x].end(), [](const Key &a, const Key &b) {
      return a.pos != b.pos ? a.pos < b.pos : (a.part != b.part ? a.part < b.part : a.t < b.t);
    });
    for (const Key &k : per[x]) order.push_back(k.t);
    xoff[x + 1] = static_cast<int32_t>(order.size());
  }
}

int many_setup(gbk_ctx *ctx, const gbk_graph *g, const int64_t *vbase, int count, int64_t **dvb, int32_t **gid) {
  if (!ctx || !g || !vbase || count <= 0) return set_error(ctx, GBK_EINVAL, "many: bad arguments");
  if (vbase[0] != 0 || vbase[count] != g->n) return set_error(ctx, GBK_EINVAL, "many: vertex bases do not cover the graph");
  for (int i = 0; i < count; i++)
    if (vbase[i + 1] <= vbase[i]) return set_error(ctx, GBK_EINVAL, "many: every graph needs a vertex");
  *dvb = static_cast<int64_t *>(scratch(ctx, kS8, sizeof(int64_t) * (count + 1)));
  *gid = static_cast<int32_t *>(scratch(ctx, kS9, sizeof(int32_t) * (g->n + 1)));
  if (!*dvb || !*gid) return set_error(ctx, GBK_ENOMEM, "many scratch");
  GBK_HIP(ctx, hipMemcpyAsync(*dvb, vbase, sizeof(int64_t) * (count + 1), hipMemcpyHostToDevice, ctx->stream));
  graph_ids<<<grid_for(g->n), kBlock, 0, ctx->stream>>>(*dvb, count, g->n, *gid);
  GBK_LAUNCH_CHECK(ctx, "graph ids");
  return GBK_OK;
}

}  // namespace

extern "C" int gbk_cc_many(gbk_ctx *ctx, const gbk_graph *g, const int64_t *vbase, int count, int variant,
                           int32_t *comp_out, gbk_cc_stats *stats) {
  int64_t *dvb = nullptr;
  int32_t *gid = nullptr;
  if (int rc = many_setup(ctx, g, vbase, count, &dvb, &gid)) return rc;
  if (variant != GBK_CC_AFFOREST && variant != GBK_CC_SV) return set_error(ctx, GBK_EINVAL, "cc many: variant");
  auto *comp = static_cast<int32_t *>(scratch(ctx, kS10, sizeof(int32_t) * (g->n + 1)));
  if (!comp) return set_error(ctx, GBK_ENOMEM, "cc many comp");
  if (int rc = gbk_cc_device(ctx, g, variant, comp, stats)) return rc;
  shift_labels<<<grid_for(g->n), kBlock, 0, ctx->stream>>>(comp, gid, dvb, g->n);
  GBK_LAUNCH_CHECK(ctx, "shift labels");
  if (comp_out) {
    GBK_HIP(ctx, hipMemcpyAsync(comp_out, comp, sizeof(int32_t) * g->n, hipMemcpyDeviceToHost, ctx->stream));
  }
  GBK_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return GBK_OK;
}

extern "C" int gbk_pr_many(gbk_ctx *ctx, const gbk_graph *g, const int64_t *vbase, int count, int variant,
                           int max_iters, double tol, int fixed_work, float *scores_out, int32_t *iters_out,
                           gbk_pr_stats *stats) {
  int64_t *dvb = nullptr;
  int32_t *gid = nullptr;
  if (int rc = many_setup(ctx, g, vbase, count, &dvb, &gid)) return rc;
  if (max_iters <= 0 || (variant != GBK_PR_GS && variant != GBK_PR_SPMV))
    return set_error(ctx, GBK_EINVAL, "pr many: bad arguments");
  const int64_t n = g->n;
  hipStream_t s = ctx->stream;
  const RowTiles rt = row_tiles(g, true);
  const unsigned grid = rows_grid(rt);
  auto *scores = static_cast<float *>(scratch(ctx, kS0, sizeof(float) * n));
  auto *c0 = static_cast<float *>(scratch(ctx, kS1, sizeof(float) * n));
  auto *c1 = static_cast<float *>(scratch(ctx, kS2, sizeof(float) * n));
  auto *st = static_cast<ManyState *>(scratch(ctx, kS3, sizeof(ManyState) + 64));
  auto *hacc = static_cast<float *>(scratch(ctx, kS6, sizeof(float) * (rt.nheavy + 1)));
  // per graph: base, initial score, error, done flag, iterations
  char *pg = static_cast<char *>(scratch(ctx, kS7, (sizeof(float) * 2 + sizeof(double) + sizeof(int) * 2) * (count + 1) + 256));
  if (!scores || !c0 || !c1 || !st || !hacc || !pg) return set_error(ctx, GBK_ENOMEM, "pr many scratch");
  // Gauss-Seidel: the per-XCD tile order (a multiple of 8 workgroups)
  const bool gs = variant == GBK_PR_GS;
  int32_t *order = nullptr, *xoff = nullptr;
  unsigned pgrid = grid;
  if (gs && rt.ntiles > 0) {
    std::vector<int32_t> tiles(rt.ntiles + 1), ho, hx;
    GBK_HIP(ctx, hipMemcpyAsync(tiles.data(), rt.tiles, sizeof(int32_t) * (rt.ntiles + 1), hipMemcpyDeviceToHost, s));
    GBK_HIP(ctx, hipStreamSynchronize(s));
    xcd_tile_order(tiles, vbase, count, ho, hx);
    order = static_cast<int32_t *>(scratch(ctx, kS4, sizeof(int32_t) * (ho.size() + 9)));
    if (!order) return set_error(ctx, GBK_ENOMEM, "pr many order");
    xoff = order + ho.size();
    GBK_HIP(ctx, hipMemcpyAsync(order, ho.data(), sizeof(int32_t) * ho.size(), hipMemcpyHostToDevice, s));
    GBK_HIP(ctx, hipMemcpyAsync(xoff, hx.data(), sizeof(int32_t) * 9, hipMemcpyHostToDevice, s));
    pgrid = std::max(8u, (grid + 7u) / 8u * 8u);
  }
  auto *gerr = reinterpret_cast<double *>(pg);
  auto *gbase = reinterpret_cast<float *>(gerr + count + 1);
  auto *ginit = gbase + count + 1;
  auto *gdone = reinterpret_cast<int *>(ginit + count + 1);
  auto *giters = gdone + count + 1;
  const float kDamp = 0.85f;
  // all iterations are queued (one captured graph); finished graphs and a
  // finished batch make the kernels return at once
  auto iterations = [&]() {
  for (int it = 0; it < max_iters; it++) {
    float *cur = c0, *nxt = c0;
    if (gs) {
      many_pull_iter<true><<<pgrid, kBlock, 0, s>>>(rt, g->in_nbr, g->off, gid, gbase, gdone, scores, c0, c0, gerr,
                                                    hacc, st, kDamp, order, xoff);
    } else {
      cur = (it & 1) ? c1 : c0;
      nxt = (it & 1) ? c0 : c1;
      // a finished graph keeps its contributions in both buffers
      many_pull_iter<false><<<grid, kBlock, 0, s>>>(rt, g->in_nbr, g->off, gid, gbase, gdone, scores, cur, nxt,
                                                    gerr, hacc, st, kDamp, nullptr, nullptr);
    }
    // rows of a finished graph are skipped from then on; with ping-pong
    // contributions their stale entries are read only by rows of the same
    // (finished) graph, since the parts of a union share no edges
    many_finish<<<1, 1024, 0, s>>>(rt.heavy, rt.nheavy, hacc, gid, gbase, gdone, giters, gerr, count, it, tol,
                                   fixed_work, g->off, scores, nxt, kDamp, st);
  }
  };
  LoopGraph loop(ctx);
  if (int rc = loop.capture(iterations)) return rc;
  EventTimer timer(ctx);
  fill<float>(ctx, hacc, 0.0f, rt.nheavy + 1);
  many_graph_init<<<grid_for(count), kBlock, 0, s>>>(dvb, count, kDamp, gbase, ginit, gerr, gdone, giters, st);
  many_init<<<grid_for(n), kBlock, 0, s>>>(g->off, gid, ginit, n, scores, c0);
  if (int rc = loop.run(iterations)) return rc;
  GBK_LAUNCH_CHECK(ctx, "pr many");
  const double ms = timer.stop_ms();
  ManyState h{};
  GBK_HIP(ctx, hipMemcpyAsync(&h, st, sizeof(h), hipMemcpyDeviceToHost, s));
  if (scores_out) GBK_HIP(ctx, hipMemcpyAsync(scores_out, scores, sizeof(float) * n, hipMemcpyDeviceToHost, s));
  if (iters_out) GBK_HIP(ctx, hipMemcpyAsync(iters_out, giters, sizeof(int32_t) * count, hipMemcpyDeviceToHost, s));
  GBK_HIP(ctx, hipStreamSynchronize(s));
  if (stats) {
    stats->device_ms = ms;
    stats->iterations = h.iters;
    stats->converged_iter = h.done ? h.iters : -1;
    stats->final_error = 0.0;
  }
  return GBK_OK;
}
