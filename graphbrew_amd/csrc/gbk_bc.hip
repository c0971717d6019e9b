// gbk_bc.hip -- Brandes betweenness centrality (bench/src/bc.cc:71-237).
//
// Per source, as in the reference:
//   forward   level-synchronous BFS; a CAS on depth claims a vertex for the
//             next level; every edge (u,v) with depth[v] == depth[u]+1 sets
//             its bit in an edge-indexed successor bitmap (bc.cc:109, m bits)
//             and adds sigma[u] into sigma[v] (double; path counts are
//             integers < 2^53, so the atomic sum is exact in any order).
//   backward  levels deepest first; delta[u] = sum over successor edges of
//             sigma[u]/sigma[v] * (1 + delta[v]); scores[u] += delta[u].
//             Once delta[v] is final its factor (1 + delta[v]) / sigma[v]
//             (computed in double, kept as float) is all a predecessor
//             needs, so a successor edge costs one random 4-byte read
//             instead of a double and a float from two arrays.
// Scores are divided by their maximum at the end (bc.cc:224-230).
// Kernels: the forward uses the coalesced frontier expansion; the backward
// walks each level's edges wave-wise with per-row LDS accumulators; rows
// above the heavy threshold are cut into chunks spread over every wave of
// the grid, whose double partial sums a finish kernel adds per row in chunk
// order (the same sum whatever the schedule).
#include "gbk_expand.h"

namespace gbk {

struct BcCounters {
  unsigned long long tail;
  unsigned long long reach;   // sum over sources of m_r
  unsigned long long max_bits;
  int overflow;
  int pad;
};

__global__ void bc_init_source(int64_t n, int32_t source, int32_t *depth, double *sigma,
                               int32_t *queue, BcCounters *c) {
  const int64_t v = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (v < n) {
    depth[v] = v == source ? 0 : -1;
    sigma[v] = v == source ? 1.0 : 0.0;
  }
  if (v == 0) {
    queue[0] = source;
    c->tail = 1;
    c->overflow = 0;
  }
}

struct BcForwardEdge {
  const int32_t *nbr;
  int32_t *depth;
  double *sigma;
  uint32_t *succ;
  WaveQueue q;
  int64_t edge_base;
  int32_t d;
  __device__ void operator()(int64_t u, int64_t k, bool active) {
    bool claimed = false;
    int32_t v = 0;
    uint32_t bits = 0;
    if (active) {
      v = nbr[k];
      // A plain load sees either a depth fixed in an earlier level (final)
      // or -1 / d from this level; only -1 needs the CAS, whose return value
      // is the coherent depth (bc.cc:104-112 test it twice).
      int32_t x = depth[v];
      if (x == -1) {
        x = atomicCAS(&depth[v], -1, d);
        claimed = x == -1;
      }
      if (x == -1 || x == d) {
        bits = 1u << ((k - edge_base) & 31);
        unsafeAtomicAdd(&sigma[v], sigma[u]);
      }
    }
    // Successor bits: lanes holding consecutive edges share bitmap words, so
    // OR the bits over each run of lanes with the same word and let the
    // run's last lane write it (one atomic per word instead of per edge).
    const int lane = lane_id();
    const int64_t word = active ? (k - edge_base) >> 5 : -1 - lane;
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
      const uint32_t b = __shfl_up(bits, o);
      const int64_t w = __shfl_up(word, o);
      if (lane >= o && w == word) bits |= b;
    }
    // The shuffle runs on every lane: inside the short-circuit it would run
    // only on lanes 0..62, and a lane reading a disabled lane gets 0.
    const int64_t next = __shfl_down(word, 1);
    const bool tail = lane == kWave - 1 || next != word;
    if (tail && bits) atomicOr(&succ[word], bits);
    q.push(claimed, v);
  }
};

__global__ void __launch_bounds__(kBlock)
bc_forward_level(const int64_t *off, const int32_t *nbr, int32_t *depth, double *sigma,
                 uint32_t *succ, int32_t *queue, int64_t lb, int64_t le, int64_t cap, int32_t d,
                 int64_t heavy_deg, BcCounters *c, HeavyQueue hq) {
  __shared__ int32_t seg[kWavesPerBlock * kSeg];
  BcForwardEdge f;
  f.nbr = nbr;
  f.depth = depth;
  f.sigma = sigma;
  f.succ = succ;
  f.q.init(seg + (threadIdx.x / kWave) * kSeg, &c->tail, queue, cap, &c->overflow);
  f.edge_base = off[0];
  f.d = d;
  auto rowlen = [&](int64_t u, int64_t &s, int64_t &l) {
    s = off[u];
    l = off[u + 1] - s;
  };
  expand_list_light(queue + lb, le - lb, heavy_deg, rowlen, f, hq);
  f.q.flush_block();
}

// Heavy rows of the level, spread over all workgroups.
__global__ void __launch_bounds__(kBlock)
bc_forward_heavy(const int64_t *off, const int32_t *nbr, int32_t *depth, double *sigma,
                 uint32_t *succ, int32_t *queue, int64_t cap, int32_t d, BcCounters *c,
                 HeavyQueue hq) {
  __shared__ int32_t seg[kWavesPerBlock * kSeg];
  BcForwardEdge f;
  f.nbr = nbr;
  f.depth = depth;
  f.sigma = sigma;
  f.succ = succ;
  f.q.init(seg + (threadIdx.x / kWave) * kSeg, &c->tail, queue, cap, &c->overflow);
  f.edge_base = off[0];
  f.d = d;
  expand_heavy(hq, f);
  f.q.flush_block();
}

// One level of the dependency accumulation; coef[v] = (1 + delta[v]) /
// sigma[v] of the deeper levels, written here for this level's vertices.
// Rows above heavy_deg are queued for bc_backward_heavy.
__global__ void __launch_bounds__(kBlock)
bc_backward_level(const int64_t *off, const int32_t *nbr, const int32_t *list, int64_t count,
                  const double *sigma, const uint32_t *succ, float *coef, float *scores,
                  int64_t heavy_deg, HeavyQueue hq) {
  __shared__ float acc[kBlock];
  const int64_t edge_base = off[0];
  const int lane = lane_id();
  const int wv = threadIdx.x / kWave;
  for (int64_t tile = blockIdx.x; tile * kBlock < count; tile += gridDim.x) {
    acc[threadIdx.x] = 0.0f;
    __syncthreads();
    const int64_t i = tile * kBlock + threadIdx.x;
    int32_t u = 0;
    int64_t start = 0, len = 0;
    bool light = false;
    if (i < count) {
      u = list[i];
      start = off[u];
      len = off[u + 1] - start;
      if (len > heavy_deg) {
        heavy_push(hq, u, start, len);
        len = 0;
      } else {
        light = true;
      }
    }
    const double su = sigma[u];
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
      const double osig = __shfl(su, lo);
      const bool act = e < total;
      float term = 0.0f;
      if (act) {
        const int64_t k = ostart + (e - oexcl);
        if (bit_get(succ, k - edge_base)) {
          term = static_cast<float>(osig * static_cast<double>(coef[nbr[k]]));
        }
      }
      bool tail;
      const float run = seg_run_sum(term, act ? lo : kWave, tail);
      if (act && tail) acc[wv * kWave + lo] += run;
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "workgroup");
    __syncthreads();
    if (light) {
      const float du = acc[threadIdx.x];
      coef[u] = static_cast<float>((1.0 + du) / su);
      scores[u] += du;
    }
    __syncthreads();
  }
}

// Chunks of the level's heavy rows, one wave each over the whole grid:
// hpart[c] = sum over the successor edges of chunk c of sigma[u] * coef[v].
__global__ void __launch_bounds__(kBlock)
bc_backward_heavy(const int64_t *off, const int32_t *nbr, const double *sigma, const uint32_t *succ,
                  const float *coef, HeavyQueue hq, double *hpart) {
  int64_t rows, total;
  heavy_extent(hq, rows, total);
  if (rows == 0) return;
  const int64_t edge_base = off[0];
  const int lane = lane_id();
  const int64_t waves = static_cast<int64_t>(gridDim.x) * kWavesPerBlock;
  for (int64_t c = static_cast<int64_t>(blockIdx.x) * kWavesPerBlock + threadIdx.x / kWave; c < total;
       c += waves) {
    int64_t lo = 0, hi = rows;  // last row with prefix[row] <= c
    while (hi - lo > 1) {
      const int64_t mid = (lo + hi) >> 1;
      if (hq.prefix[mid] <= c) lo = mid;
      else hi = mid;
    }
    const int64_t first = hq.start[lo] + (c - hq.prefix[lo]) * hq.chunk;
    const int64_t rest = hq.start[lo] + hq.len[lo] - first;
    const int64_t end = first + (rest < hq.chunk ? rest : hq.chunk);
    const double su = sigma[hq.row[lo]];
    double part = 0.0;
    constexpr int U = 4;
    for (int64_t k0 = first; k0 < end; k0 += U * kWave) {
      int32_t v[U];
      bool on[U];
#pragma unroll
      for (int j = 0; j < U; j++) {
        const int64_t k = k0 + j * kWave + lane;
        on[j] = k < end && bit_get(succ, k - edge_base);
        v[j] = on[j] ? nbr[k] : 0;
      }
#pragma unroll
      for (int j = 0; j < U; j++)
        if (on[j]) part += su * static_cast<double>(coef[v[j]]);
    }
    part = wave_sum(part);
    if (lane == 0) hpart[c] = part;
  }
}

// One lane per heavy row: its chunks' partials added in order.
__global__ void bc_backward_heavy_finish(HeavyQueue hq, const double *sigma, const double *hpart,
                                         float *coef, float *scores) {
  int64_t rows, total;
  heavy_extent(hq, rows, total);
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t h = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; h < rows; h += stride) {
    double tot = 0.0;
    const int64_t end = h + 1 < rows ? hq.prefix[h + 1] : total;
    for (int64_t c = hq.prefix[h]; c < end; c++) tot += hpart[c];
    const int32_t u = hq.row[h];
    coef[u] = static_cast<float>((1.0 + tot) / sigma[u]);
    scores[u] += static_cast<float>(tot);
  }
}

__global__ void bc_reach_edges(const int64_t *off, const int32_t *depth, int64_t n,
                               unsigned long long *partial) {
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  unsigned long long e = 0;
  for (int64_t v = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; v < n; v += stride)
    if (depth[v] >= 0) e += static_cast<unsigned long long>(off[v + 1] - off[v]);
  block_partial(e, partial);
}

}  // namespace gbk

using namespace gbk;

// Workgroups of the heavy-row launches (chunks of hub rows spread over the
// grid): 1024 instead of the frontier kernels' 512 took 4 sources on
// Kronecker scale 22 from 25.0 to 23.0 ms.
constexpr unsigned kBcHeavyGrid = 1024;

int gbk_bc_device(gbk_ctx *ctx, const gbk_graph *g, const int32_t *sources, int num_sources,
                  float *scores, gbk_bc_stats *stats) {
  const int64_t n = g->n;
  hipStream_t s = ctx->stream;
  const int64_t words = (g->m + 31) / 32 + 1;
  auto *depth = static_cast<int32_t *>(scratch(ctx, kS5, sizeof(int32_t) * (n + 1)));
  auto *sigma = static_cast<double *>(scratch(ctx, kS6, sizeof(double) * (n + 1)));
  auto *coef = static_cast<float *>(scratch(ctx, kS7, sizeof(float) * (n + 1)));
  auto *queue = static_cast<int32_t *>(scratch(ctx, kS8, sizeof(int32_t) * (n + 64)));
  auto *succ = static_cast<uint32_t *>(scratch(ctx, kS9, sizeof(uint32_t) * words));
  auto *c = static_cast<BcCounters *>(scratch(ctx, kS10, sizeof(BcCounters) + 64));
  auto *hc = static_cast<BcCounters *>(pinned(ctx, sizeof(BcCounters)));
  if (!depth || !sigma || !coef || !queue || !succ || !c || !hc)
    return set_error(ctx, GBK_ENOMEM, "bc scratch");
  auto *partial = static_cast<unsigned long long *>(scratch(ctx, kS11, sizeof(unsigned long long) * kGridCap));
  HeavyQueue hq;
  if (!partial || !make_heavy_queue(ctx, g, kS12, &hq)) return set_error(ctx, GBK_ENOMEM, "bc partials");
  hq.overflow = &c->overflow;
  // chunks of one level's heavy rows: at most m / chunk + one per row
  auto *hpart = static_cast<double *>(scratch(ctx, kS13, sizeof(double) * (g->m / hq.chunk + hq.cap + 1)));
  if (!hpart) return set_error(ctx, GBK_ENOMEM, "bc heavy partials");
  const int64_t H = ctx->tune.heavy_degree;
  const int64_t cap = n + 64;
  const unsigned ng = expand_grid(n);
  EventTimer timer(ctx);
  fill<float>(ctx, scores, 0.0f, n);
  fill<unsigned long long>(ctx, &c->reach, 0ull, 1);
  int max_depth = 0;
  std::vector<int64_t> level;
  for (int it = 0; it < num_sources; it++) {
    const int32_t src = sources[it];
    fill<uint32_t>(ctx, succ, 0u, words);
    bc_init_source<<<grid_for(n), kBlock, 0, s>>>(n, src, depth, sigma, queue, c);
    level.assign({0, 1});
    int32_t d = 0;
    for (;;) {
      d++;
      const int64_t lb = level[level.size() - 2], le = level.back();
      fill<unsigned long long>(ctx, hq.count, 0ull, 1);
      const int kt = kt_begin(ctx, GBK_KT_BC_LEVEL);
      bc_forward_level<<<expand_grid(le - lb), kBlock, 0, s>>>(g->off, g->nbr, depth, sigma, succ,
                                                              queue, lb, le, cap, d, H, c, hq);
      bc_forward_heavy<<<kBcHeavyGrid, kBlock, 0, s>>>(g->off, g->nbr, depth, sigma, succ, queue, cap,
                                                     d, c, hq);
      kt_end(ctx, kt);
      GBK_LAUNCH_CHECK(ctx, "bc_forward_level");
      GBK_HIP(ctx, hipMemcpyAsync(hc, c, sizeof(BcCounters), hipMemcpyDeviceToHost, s));
      GBK_HIP(ctx, hipStreamSynchronize(s));
      const int64_t tail = static_cast<int64_t>(hc->tail);
      if (hc->overflow || tail > n + 1) return set_error(ctx, GBK_EHIP, "bc: queue overflow");
      if (tail == le) break;
      level.push_back(tail);
    }
    max_depth = std::max<int>(max_depth, static_cast<int>(level.size()) - 2);
    for (int lv = static_cast<int>(level.size()) - 2; lv >= 0; lv--) {
      const int64_t lb = level[lv], cnt = level[lv + 1] - lb;
      fill<unsigned long long>(ctx, hq.count, 0ull, 1);
      const int kt = kt_begin(ctx, GBK_KT_BC_LEVEL);
      bc_backward_level<<<expand_grid(cnt), kBlock, 0, s>>>(g->off, g->nbr, queue + lb, cnt, sigma,
                                                           succ, coef, scores, H, hq);
      bc_backward_heavy<<<kBcHeavyGrid, kBlock, 0, s>>>(g->off, g->nbr, sigma, succ, coef, hq, hpart);
      bc_backward_heavy_finish<<<4, kBlock, 0, s>>>(hq, sigma, hpart, coef, scores);
      kt_end(ctx, kt);
    }
    GBK_LAUNCH_CHECK(ctx, "bc_backward_heavy");
    GBK_HIP(ctx, hipMemcpyAsync(hc, c, sizeof(BcCounters), hipMemcpyDeviceToHost, s));
    GBK_HIP(ctx, hipStreamSynchronize(s));
    if (hc->overflow) return set_error(ctx, GBK_EHIP, "bc: heavy queue overflow");
    bc_reach_edges<<<ng, kBlock, 0, s>>>(g->off, depth, n, partial);
    sum_partials<unsigned long long><<<1, 1024, 0, s>>>(partial, static_cast<int>(ng), &c->reach, 1);
    GBK_LAUNCH_CHECK(ctx, "bc_backward_level");
  }
  auto *pmax = reinterpret_cast<unsigned int *>(partial);
  bc_max<<<ng, kBlock, 0, s>>>(scores, n, pmax);
  bc_normalize<<<ng, kBlock, 0, s>>>(scores, n, pmax, static_cast<int>(ng));
  GBK_LAUNCH_CHECK(ctx, "bc_normalize");
  const double ms = timer.stop_ms();
  GBK_HIP(ctx, hipMemcpyAsync(hc, c, sizeof(BcCounters), hipMemcpyDeviceToHost, s));
  GBK_HIP(ctx, hipStreamSynchronize(s));
  if (stats) {
    stats->device_ms = ms;
    stats->edges_reached = static_cast<int64_t>(hc->reach);
    stats->max_depth = max_depth;
  }
  return GBK_OK;
}

extern "C" int gbk_bc(gbk_ctx *ctx, const gbk_graph *g, const int32_t *sources, int num_sources,
                      float *scores_out, gbk_bc_stats *stats) {
  if (!ctx || !g || !sources || num_sources <= 0) return set_error(ctx, GBK_EINVAL, "bc: bad arguments");
  for (int i = 0; i < num_sources; i++)
    if (sources[i] < 0 || sources[i] >= g->n) return set_error(ctx, GBK_EINVAL, "bc: source out of range");
  auto *scores = static_cast<float *>(scratch(ctx, kS4, sizeof(float) * (g->n + 1)));
  if (!scores) return set_error(ctx, GBK_ENOMEM, "bc scores");
  int rc = gbk_bc_device(ctx, g, sources, num_sources, scores, stats);
  if (rc) return rc;
  if (scores_out) {
    GBK_HIP(ctx, hipMemcpyAsync(scores_out, scores, sizeof(float) * g->n, hipMemcpyDeviceToHost, ctx->stream));
    GBK_HIP(ctx, hipStreamSynchronize(ctx->stream));
  }
  return GBK_OK;
}
