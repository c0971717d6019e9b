// gbk_bfs.hip -- direction-optimizing BFS (bench/src/bfs.cc:54-269).
//
// Same encoding and switching rule as the reference: parent[v] < 0 means
// unvisited and holds -out_degree(v); the search goes bottom-up when the
// scout count (out-degree sum of the last top-down frontier) exceeds
// edges_to_check / alpha, and returns to top-down once the awake count
// shrinks and is at most n / beta.
//   top-down   frontier queue -> coalesced expansion (gbk_expand.h), a CAS
//              claims each newly seen vertex; winners are staged per wave in
//              LDS and flushed to the next queue in batches.
//   bottom-up  bitmap frontier; 'bu_group' lanes per unvisited vertex read
//              its in-neighbours in chunks and stop at the first frontier
//              hit (masked ballot inside the lane group).
// Counts (scout, awake) are per-workgroup partials summed by a one-block
// finish launch; the host reads one small counter block per step to apply
// the switching rule.
#include "gbk_expand.h"

namespace gbk {

struct BfsCounters {
  unsigned long long tail;   // queue tail
  unsigned long long scout;  // top-down scout count
  unsigned long long awake;  // bottom-up awake count
  int overflow;
  int pad;
};

__global__ void bfs_init(const int64_t *out_off, int64_t n, int32_t source, int32_t *parent,
                         int32_t *queue, BfsCounters *c) {
  const int64_t v = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (v < n) {
    const int64_t d = out_off[v + 1] - out_off[v];
    parent[v] = v == source ? source : (d != 0 ? static_cast<int32_t>(-d) : -1);
  }
  if (v == 0) {
    queue[0] = source;
    c->tail = 1;
    c->scout = 0;
    c->awake = 0;
    c->overflow = 0;
  }
}

struct TdEdge {
  const int32_t *nbr;
  int32_t *parent;
  WaveQueue q;
  unsigned long long scout;
  __device__ void operator()(int64_t u, int64_t k, bool active) {
    bool won = false;
    int32_t v = 0;
    if (active) {
      v = nbr[k];
      const int32_t cur = ld_agent(&parent[v]);
      if (cur < 0 && atomicCAS(&parent[v], cur, static_cast<int32_t>(u)) == cur) {
        won = true;
        scout += static_cast<unsigned long long>(-static_cast<int64_t>(cur));
      }
    }
    q.push(won, v);
  }
};

__global__ void __launch_bounds__(kBlock)
bfs_td_step(const int64_t *out_off, const int32_t *out_nbr, int32_t *parent, int32_t *queue,
            int64_t qb, int64_t qe, int64_t cap, int64_t heavy_deg, BfsCounters *c,
            unsigned long long *partial, HeavyQueue hq) {
  __shared__ int32_t seg[kWavesPerBlock * kSeg];
  TdEdge f;
  f.nbr = out_nbr;
  f.parent = parent;
  f.scout = 0;
  f.q.init(seg + (threadIdx.x / kWave) * kSeg, &c->tail, queue, cap, &c->overflow);
  auto rowlen = [&](int64_t u, int64_t &s, int64_t &l) {
    s = out_off[u];
    l = out_off[u + 1] - s;
  };
  expand_list_light(queue + qb, qe - qb, heavy_deg, rowlen, f, hq);
  f.q.flush_block();
  block_partial(f.scout, partial);
}

// Heavy rows of the frontier, spread over all workgroups.
__global__ void __launch_bounds__(kBlock)
bfs_td_heavy(const int32_t *out_nbr, int32_t *parent, int32_t *queue, int64_t cap, BfsCounters *c,
             unsigned long long *partial, HeavyQueue hq) {
  __shared__ int32_t seg[kWavesPerBlock * kSeg];
  TdEdge f;
  f.nbr = out_nbr;
  f.parent = parent;
  f.scout = 0;
  f.q.init(seg + (threadIdx.x / kWave) * kSeg, &c->tail, queue, cap, &c->overflow);
  expand_heavy(hq, f);
  f.q.flush_block();
  block_partial(f.scout, partial);
}

template <int G>
__global__ void __launch_bounds__(kBlock)
bfs_bu_step(const int64_t *in_off, const int32_t *in_nbr, int64_t n, int32_t *parent,
            const uint32_t *front, uint32_t *next, unsigned long long *partial, int swz) {
  const int lane = lane_id();
  const int glane = lane % G;
  const int gbase = lane - glane;
  const int64_t groups_per_block = kBlock / G;
  const int64_t tiles = (n + groups_per_block - 1) / groups_per_block;
  unsigned long long awake = 0;
  for (int64_t t = swizzled_block(swz); t < tiles; t += gridDim.x) {
    const int64_t u = t * groups_per_block + threadIdx.x / G;
    bool todo = false;
    int64_t s = 0, e = 0;
    if (u < n && parent[u] < 0) {
      s = in_off[u];
      e = in_off[u + 1];
      todo = s < e;
    }
    // group-uniform loop
    for (int64_t base = s; todo; base += G) {
      const int64_t k = base + glane;
      const bool hit = k < e && bit_get(front, in_nbr[k]);
      const uint64_t m = __ballot(hit);
      const uint64_t gm = (G == 64) ? m : ((m >> gbase) & ((1ull << G) - 1));
      if (gm) {
        const int first = __ffsll(static_cast<long long>(gm)) - 1;
        if (glane == first) {
          parent[u] = in_nbr[k];
          bit_set_atomic(next, u);
          awake++;
        }
        todo = false;
      } else if (base + G >= e) {
        todo = false;
      }
    }
  }
  block_partial(awake, partial);
}

__global__ void bfs_queue_to_bitmap(const int32_t *queue, int64_t qb, int64_t qe, uint32_t *bm) {
  const int64_t i = qb + static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < qe) bit_set_atomic(bm, queue[i]);
}

__global__ void __launch_bounds__(kBlock)
bfs_bitmap_to_queue(const uint32_t *bm, int64_t n, int32_t *queue, int64_t cap, BfsCounters *c) {
  __shared__ int32_t seg[kWavesPerBlock * kSeg];
  WaveQueue q;
  q.init(seg + (threadIdx.x / kWave) * kSeg, &c->tail, queue, cap, &c->overflow);
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t base = static_cast<int64_t>(blockIdx.x) * blockDim.x; base < n; base += stride) {
    const int64_t v = base + threadIdx.x;
    const bool set = v < n && bit_get(bm, v);
    q.push(set, static_cast<int32_t>(v));
  }
  q.flush_block();
}

// Writes the summed partials (one or two ranges) into a counter (one workgroup).
__global__ void bfs_store_count(const unsigned long long *partial, int count,
                                unsigned long long *dst, const unsigned long long *partial2 = nullptr,
                                int count2 = 0) {
  __shared__ unsigned long long red[1024 / kWave];
  unsigned long long s = 0;
  for (int i = threadIdx.x; i < count; i += blockDim.x) s += partial[i];
  for (int i = threadIdx.x; i < count2; i += blockDim.x) s += partial2[i];
  s = wave_sum(s);
  if (lane_id() == 0) red[threadIdx.x / kWave] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long t = 0;
    for (int i = 0; i < static_cast<int>(blockDim.x) / kWave; i++) t += red[i];
    *dst = t;
  }
}

__global__ void bfs_finish(const int64_t *out_off, int64_t n, int32_t *parent,
                           unsigned long long *p_reached, unsigned long long *p_edges) {
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  unsigned long long reached = 0, edges = 0;
  for (int64_t v = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; v < n; v += stride) {
    int32_t p = parent[v];
    if (p < -1) parent[v] = p = -1;
    if (p >= 0) {
      reached++;
      edges += static_cast<unsigned long long>(out_off[v + 1] - out_off[v]);
    }
  }
  block_partial(reached, p_reached);
  block_partial(edges, p_edges);
}

static void launch_bu(gbk_ctx *ctx, int G, const gbk_graph *g, int32_t *parent,
                      const uint32_t *front, uint32_t *next, unsigned long long *partial,
                      unsigned grid) {
  const int swz = ctx->tune.xcd_swizzle;
  hipStream_t s = ctx->stream;
  switch (G) {
    case 1: bfs_bu_step<1><<<grid, kBlock, 0, s>>>(g->in_off, g->in_nbr, g->n, parent, front, next, partial, swz); break;
    case 2: bfs_bu_step<2><<<grid, kBlock, 0, s>>>(g->in_off, g->in_nbr, g->n, parent, front, next, partial, swz); break;
    case 4: bfs_bu_step<4><<<grid, kBlock, 0, s>>>(g->in_off, g->in_nbr, g->n, parent, front, next, partial, swz); break;
    case 16: bfs_bu_step<16><<<grid, kBlock, 0, s>>>(g->in_off, g->in_nbr, g->n, parent, front, next, partial, swz); break;
    case 32: bfs_bu_step<32><<<grid, kBlock, 0, s>>>(g->in_off, g->in_nbr, g->n, parent, front, next, partial, swz); break;
    case 64: bfs_bu_step<64><<<grid, kBlock, 0, s>>>(g->in_off, g->in_nbr, g->n, parent, front, next, partial, swz); break;
    default: bfs_bu_step<8><<<grid, kBlock, 0, s>>>(g->in_off, g->in_nbr, g->n, parent, front, next, partial, swz); break;
  }
}

}  // namespace gbk

using namespace gbk;

// Runs one DO-BFS on the device; the parent array stays in 'parent'.
int gbk_bfs_device(gbk_ctx *ctx, const gbk_graph *g, int32_t source, int alpha, int beta,
                   int32_t *parent, gbk_bfs_stats *stats) {
  const int64_t n = g->n;
  hipStream_t s = ctx->stream;
  const int64_t words = (n + 31) / 32 + 1;
  const int64_t cap = n + 64;
  auto *queue = static_cast<int32_t *>(scratch(ctx, kS5, sizeof(int32_t) * cap));
  auto *bm0 = static_cast<uint32_t *>(scratch(ctx, kS6, sizeof(uint32_t) * words));
  auto *bm1 = static_cast<uint32_t *>(scratch(ctx, kS7, sizeof(uint32_t) * words));
  auto *c = static_cast<BfsCounters *>(scratch(ctx, kS8, sizeof(BfsCounters) + 64));
  auto *partial = static_cast<unsigned long long *>(scratch(ctx, kS9, sizeof(unsigned long long) * 2 * kGridCap));
  auto *fin = static_cast<unsigned long long *>(scratch(ctx, kS10, 64));
  auto *hc = static_cast<BfsCounters *>(pinned(ctx, sizeof(BfsCounters) * 2));
  HeavyQueue hq;
  if (!queue || !bm0 || !bm1 || !c || !partial || !fin || !hc || !make_heavy_queue(ctx, g, kS11, &hq))
    return set_error(ctx, GBK_ENOMEM, "bfs scratch");
  hq.overflow = &c->overflow;  // reported with the step counters
  int64_t out_deg_src = 0;
  {
    int64_t o[2];
    GBK_HIP(ctx, hipMemcpyAsync(o, g->off + source, sizeof(o), hipMemcpyDeviceToHost, s));
    GBK_HIP(ctx, hipStreamSynchronize(s));
    out_deg_src = o[1] - o[0];
  }
  const int G = ctx->tune.bu_group;
  const unsigned bu_grid = expand_grid((n + (kBlock / G) - 1) / (kBlock / G) * kBlock);
  const unsigned n_grid = expand_grid(n);
  EventTimer timer(ctx);
  bfs_init<<<grid_for(n), kBlock, 0, s>>>(g->off, n, source, parent, queue, c);
  auto read_counters = [&]() -> int {
    GBK_HIP(ctx, hipMemcpyAsync(hc, c, sizeof(BfsCounters), hipMemcpyDeviceToHost, s));
    GBK_HIP(ctx, hipStreamSynchronize(s));
    if (hc->overflow) return set_error(ctx, GBK_EHIP, "bfs: queue overflow");
    return GBK_OK;
  };
  int64_t qb = 0, qe = 1;
  int64_t edges_to_check = g->m;
  int64_t scout_count = out_deg_src;
  int steps_td = 0, steps_bu = 0;
  uint32_t *front = bm0, *curr = bm1;
  while (qe > qb) {
    if (scout_count > edges_to_check / alpha) {
      fill<uint32_t>(ctx, front, 0u, words);
      bfs_queue_to_bitmap<<<grid_for(qe - qb), kBlock, 0, s>>>(queue, qb, qe, front);
      int64_t awake_count = qe - qb, old_awake_count;
      qb = qe;
      do {
        old_awake_count = awake_count;
        fill<uint32_t>(ctx, curr, 0u, words);
        const int kt = kt_begin(ctx, GBK_KT_BFS_STEP);
        launch_bu(ctx, G, g, parent, front, curr, partial, bu_grid);
        kt_end(ctx, kt);
        bfs_store_count<<<1, 1024, 0, s>>>(partial, static_cast<int>(bu_grid), &c->awake);
        GBK_LAUNCH_CHECK(ctx, "bfs_bu_step");
        if (int rc = read_counters()) return rc;
        awake_count = static_cast<int64_t>(hc->awake);
        std::swap(front, curr);
        steps_bu++;
      } while ((awake_count >= old_awake_count) || (awake_count > n / beta));
      bfs_bitmap_to_queue<<<n_grid, kBlock, 0, s>>>(front, n, queue, cap, c);
      GBK_LAUNCH_CHECK(ctx, "bfs_bitmap_to_queue");
      if (int rc = read_counters()) return rc;
      qe = static_cast<int64_t>(hc->tail);
      scout_count = 1;
    } else {
      edges_to_check -= scout_count;
      const unsigned grid = expand_grid(qe - qb);
      fill<unsigned long long>(ctx, hq.count, 0ull, 1);
      const int kt = kt_begin(ctx, GBK_KT_BFS_STEP);
      bfs_td_step<<<grid, kBlock, 0, s>>>(g->off, g->nbr, parent, queue, qb, qe, cap,
                                          ctx->tune.heavy_degree, c, partial, hq);
      bfs_td_heavy<<<kHeavyGrid, kBlock, 0, s>>>(g->nbr, parent, queue, cap, c, partial + kGridCap, hq);
      kt_end(ctx, kt);
      bfs_store_count<<<1, 1024, 0, s>>>(partial, static_cast<int>(grid), &c->scout,
                                         partial + kGridCap, static_cast<int>(kHeavyGrid));
      GBK_LAUNCH_CHECK(ctx, "bfs_td_step");
      if (int rc = read_counters()) return rc;
      scout_count = static_cast<int64_t>(hc->scout);
      qb = qe;
      qe = static_cast<int64_t>(hc->tail);
      steps_td++;
    }
  }
  bfs_finish<<<n_grid, kBlock, 0, s>>>(g->off, n, parent, partial, partial + kGridCap);
  bfs_store_count<<<1, 1024, 0, s>>>(partial, static_cast<int>(n_grid), fin);
  bfs_store_count<<<1, 1024, 0, s>>>(partial + kGridCap, static_cast<int>(n_grid), fin + 1);
  GBK_LAUNCH_CHECK(ctx, "bfs_finish");
  const double ms = timer.stop_ms();
  unsigned long long hf[2] = {0, 0};
  GBK_HIP(ctx, hipMemcpyAsync(hf, fin, sizeof(hf), hipMemcpyDeviceToHost, s));
  GBK_HIP(ctx, hipStreamSynchronize(s));
  if (stats) {
    stats->device_ms = ms;
    stats->steps_td = steps_td;
    stats->steps_bu = steps_bu;
    stats->reached = static_cast<int64_t>(hf[0]);
    stats->edges_reached = static_cast<int64_t>(hf[1]);
  }
  return GBK_OK;
}

extern "C" int gbk_bfs(gbk_ctx *ctx, const gbk_graph *g, int32_t source, int alpha, int beta,
                       int32_t *parent_out, gbk_bfs_stats *stats) {
  if (!ctx || !g) return set_error(ctx, GBK_EINVAL, "bfs: null argument");
  if (source < 0 || source >= g->n) return set_error(ctx, GBK_EINVAL, "bfs: source out of range");
  if (alpha <= 0) alpha = ctx->tune.bfs_alpha;
  if (beta <= 0) beta = ctx->tune.bfs_beta;
  auto *parent = static_cast<int32_t *>(scratch(ctx, kS4, sizeof(int32_t) * (g->n + 1)));
  if (!parent) return set_error(ctx, GBK_ENOMEM, "bfs parent");
  int rc = gbk_bfs_device(ctx, g, source, alpha, beta, parent, stats);
  if (rc) return rc;
  if (parent_out) {
    GBK_HIP(ctx, hipMemcpyAsync(parent_out, parent, sizeof(int32_t) * g->n, hipMemcpyDeviceToHost,
                                ctx->stream));
    GBK_HIP(ctx, hipStreamSynchronize(ctx->stream));
  }
  return GBK_OK;
}
