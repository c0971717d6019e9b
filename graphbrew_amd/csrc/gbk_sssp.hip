// gbk_sssp.hip -- delta-stepping SSSP (bench/src/sssp.cc:129-283).
//
// Buckets of width delta as in the reference: a vertex is relaxed from the
// current bucket b only while dist[v] >= delta*b (stale copies are skipped,
// sssp.cc:191-192), relaxations take the minimum atomically, and vertices
// that land in bucket b are processed again in b before the next bucket
// opens (the reference's bucket fusion does the same for small bins).
//
// GPU organisation:
//   word[v]   one 64-bit word per vertex: distance in the high half, in the
//             low half the relaxation round that last queued v (0 = none).
//             A relaxation is a single 64-bit atomic minimum of
//             (candidate << 32 | tag): it lowers the distance and records
//             the round together, and the value it returns says both
//             whether the distance improved and whether v is already queued
//             for the next round -- one atomic instead of a minimum plus a
//             separate test-and-set of a queued flag.  The tag is the round
//             number only when the candidate falls in the current bucket,
//             so a later improvement into the bucket still queues v.
//   near[2]   ping-pong queues of the current bucket, staged per wave in LDS.
//   later buckets are not queued at all: when the current bucket drains, one
//             dense pass over the words finds the smallest bucket still
//             holding an unsettled vertex and a second pass queues that
//             bucket.  Two streaming passes per bucket cost far less than a
//             pending list would in random atomics (one per improvement).
// Each lane relaxes four edges per trip (independent loads and atomics in
// flight together): the path is bound by the latency of dependent random
// accesses, not by bandwidth.
// Kernels of one step: relax, then min-bucket and split (each acting only in
// its phase), each followed by a one-workgroup kernel that advances the
// device state block.  The host queues steps without waiting and checks the
// 'done' flag after 2, 4, ... 64 steps.  Distances are integers, so the
// result is exact whatever the order of relaxations.
#include <climits>

#include "gbk_expand.h"

namespace gbk {

constexpr int32_t kDistInf = INT32_MAX / 2;  // sssp.cc:67
constexpr int kRelaxBatch = 4;               // edges per lane per trip

__device__ __forceinline__ int64_t word_dist(unsigned long long w) { return static_cast<int64_t>(w >> 32); }
__device__ __forceinline__ unsigned long long make_word(int64_t d, uint32_t tag) {
  return (static_cast<unsigned long long>(d) << 32) | tag;
}

struct SsspState {
  unsigned long long near_cnt[2];
  unsigned long long next_bucket;
  long long bucket;
  long long relax_rounds;
  long long buckets;
  int cur, phase, done, overflow;
};

__global__ void sssp_init(int64_t n, int32_t source, unsigned long long *word, int32_t *near0,
                          SsspState *st) {
  const int64_t v = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (v < n) word[v] = make_word(v == source ? 0 : kDistInf, 0);
  if (v == 0) {
    near0[0] = source;
    st->near_cnt[0] = 1;
    st->near_cnt[1] = 0;
    st->next_bucket = ULLONG_MAX;
    st->bucket = 0;
    st->relax_rounds = 0;
    st->buckets = 0;
    st->cur = 0;
    st->phase = 0;
    st->done = 0;
    st->overflow = 0;
  }
}

struct RelaxEdge {
  const int32_t *nbr;
  const int32_t *wgt;
  unsigned long long *word;
  WaveQueue nearq;
  int64_t delta;
  int64_t bucket;
  uint32_t tag;  // this round's tag, never 0

  template <int U>
  __device__ __forceinline__ void batch(const int64_t (&u)[U], const int64_t (&k)[U], const bool (&act)[U]) {
    int32_t v[U];
    int64_t cand[U];
    unsigned long long wv[U];
#pragma unroll
    for (int j = 0; j < U; j++) {
      v[j] = nbr[k[j]];
      cand[j] = wgt[k[j]];
    }
#pragma unroll
    for (int j = 0; j < U; j++) {
      cand[j] += word_dist(word[u[j]]);
      wv[j] = word[v[j]];
    }
    bool push[U];
#pragma unroll
    for (int j = 0; j < U; j++) {
      push[j] = false;
      // test first: most relaxations do not improve, and a plain load can
      // only overstate dist[v] (values only decrease), never hide a win
      if (act[j] && cand[j] < word_dist(wv[j])) {
        const bool here = cand[j] / delta == bucket;
        const unsigned long long old = atomicMin(&word[v[j]], make_word(cand[j], here ? tag : 0u));
        push[j] = here && cand[j] < word_dist(old) && static_cast<uint32_t>(old) != tag;
      }
    }
#pragma unroll
    for (int j = 0; j < U; j++) nearq.push(push[j], v[j]);
  }
};

__device__ __forceinline__ void relax_setup(RelaxEdge &f, int32_t *seg, const int32_t *nbr,
                                            const int32_t *wgt, unsigned long long *word,
                                            int32_t *near0, int32_t *near1, int64_t cap,
                                            int64_t delta, SsspState *st) {
  const volatile SsspState *vs = st;
  const int cur = vs->cur;
  f.nbr = nbr;
  f.wgt = wgt;
  f.word = word;
  f.nearq.init(seg + (threadIdx.x / kWave) * kSeg, &st->near_cnt[1 - cur], cur ? near0 : near1, cap,
               &st->overflow);
  f.delta = delta;
  f.bucket = vs->bucket;
  f.tag = static_cast<uint32_t>(vs->relax_rounds + 1);
}

__global__ void __launch_bounds__(kBlock)
sssp_relax(const int64_t *off, const int32_t *nbr, const int32_t *wgt, unsigned long long *word,
           int32_t *near0, int32_t *near1, int64_t cap, int64_t delta, int64_t heavy_deg,
           SsspState *st, HeavyQueue hq) {
  const volatile SsspState *vs = st;
  if (vs->done || vs->phase != 0) return;
  __shared__ int32_t seg[kWavesPerBlock * kSeg];
  const int cur = vs->cur;
  // the tail counts past cap only when a push overflowed (reported by the host)
  const int64_t count = min(static_cast<int64_t>(vs->near_cnt[cur]), cap);
  RelaxEdge f;
  relax_setup(f, seg, nbr, wgt, word, near0, near1, cap, delta, st);
  const int64_t lower = delta * f.bucket;
  auto rowlen = [&](int64_t u, int64_t &s, int64_t &l) {
    s = off[u];
    l = word_dist(word[u]) >= lower ? off[u + 1] - s : 0;
  };
  expand_list_light_batch<kRelaxBatch>(cur ? near1 : near0, count, heavy_deg, rowlen, f, hq);
  f.nearq.flush_block();
}

// Heavy rows of the near queue, spread over all workgroups.
__global__ void __launch_bounds__(kBlock)
sssp_relax_heavy(const int32_t *nbr, const int32_t *wgt, unsigned long long *word, int32_t *near0,
                 int32_t *near1, int64_t cap, int64_t delta, SsspState *st, HeavyQueue hq) {
  const volatile SsspState *vs = st;
  if (vs->done || vs->phase != 0) return;
  __shared__ int32_t seg[kWavesPerBlock * kSeg];
  RelaxEdge f;
  relax_setup(f, seg, nbr, wgt, word, near0, near1, cap, delta, st);
  expand_heavy_batch<kRelaxBatch>(hq, f);
  f.nearq.flush_block();
}

// Work of the coming relaxation round, for the kernel timer only: rows of
// the near queue that are not stale and their edges, as per-workgroup
// partials (rows in partial[0, grid), edges in partial[grid, 2 grid)).
__global__ void sssp_count_round(const int64_t *off, const unsigned long long *word,
                                 const int32_t *near0, const int32_t *near1, int64_t delta,
                                 const SsspState *st, unsigned long long *partial) {
  const volatile SsspState *vs = st;
  unsigned long long rows = 0, edges = 0;
  if (!vs->done && vs->phase == 0) {
    const int cur = vs->cur;
    const int32_t *q = cur ? near1 : near0;
    const int64_t count = static_cast<int64_t>(vs->near_cnt[cur]);
    const int64_t lower = delta * vs->bucket;
    const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < count; i += stride) {
      const int32_t u = q[i];
      if (word_dist(word[u]) >= lower) {
        rows++;
        edges += static_cast<unsigned long long>(off[u + 1] - off[u]);
      }
    }
  }
  block_partial(rows, partial);
  block_partial(edges, partial + gridDim.x);
}

// One workgroup: closes a relaxation round (swap the near queues; when the
// next one is empty the bucket is done and the search for the next opens).
__global__ void sssp_relax_done(SsspState *st, unsigned long long *heavy_count) {
  if (threadIdx.x != 0 || st->done || st->phase != 0) return;
  *heavy_count = 0;
  const int cur = st->cur;
  st->near_cnt[cur] = 0;
  st->cur = 1 - cur;
  st->relax_rounds += 1;
  if (st->near_cnt[1 - cur] == 0) st->phase = 1;
}

// Smallest bucket above the current one that holds a reached vertex.
__global__ void sssp_min_bucket(const unsigned long long *word, int64_t n, int64_t delta,
                                const SsspState *st, unsigned long long *partial) {
  const volatile SsspState *vs = st;
  if (vs->done || vs->phase != 1) return;
  const int64_t bucket = vs->bucket;
  unsigned long long best = ULLONG_MAX;
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t v = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; v < n; v += stride) {
    const int64_t d = word_dist(word[v]);
    if (d < kDistInf) {
      const int64_t b = d / delta;
      if (b > bucket && static_cast<unsigned long long>(b) < best) best = static_cast<unsigned long long>(b);
    }
  }
  for (int o = kWave / 2; o > 0; o >>= 1) {
    const unsigned long long x = __shfl_xor(best, o);
    best = x < best ? x : best;
  }
  __shared__ unsigned long long red[kWavesPerBlock];
  if (lane_id() == 0) red[threadIdx.x / kWave] = best;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long m = red[0];
    for (int i = 1; i < kWavesPerBlock; i++) m = red[i] < m ? red[i] : m;
    partial[blockIdx.x] = m;
  }
}

__global__ void sssp_min_done(const unsigned long long *partial, int count, SsspState *st) {
  if (st->done || st->phase != 1) return;
  __shared__ unsigned long long red[1024 / kWave];
  unsigned long long m = ULLONG_MAX;
  for (int i = threadIdx.x; i < count; i += blockDim.x) m = partial[i] < m ? partial[i] : m;
  for (int o = kWave / 2; o > 0; o >>= 1) {
    const unsigned long long x = __shfl_xor(m, o);
    m = x < m ? x : m;
  }
  if (lane_id() == 0) red[threadIdx.x / kWave] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    m = red[0];
    for (int i = 1; i < static_cast<int>(blockDim.x) / kWave; i++) m = red[i] < m ? red[i] : m;
    st->next_bucket = m;
    if (m == ULLONG_MAX) st->done = 1;
    else st->phase = 2;
  }
}

// Queues every vertex of the chosen bucket (each vertex once).
__global__ void __launch_bounds__(kBlock)
sssp_split(const unsigned long long *word, int64_t n, int32_t *near0, int32_t *near1, int64_t cap,
           int64_t delta, SsspState *st) {
  const volatile SsspState *vs = st;
  if (vs->done || vs->phase != 2) return;
  __shared__ int32_t seg[kWavesPerBlock * kSeg];
  const int cur = vs->cur;
  const int64_t nb = static_cast<int64_t>(vs->next_bucket);
  WaveQueue q;
  q.init(seg + (threadIdx.x / kWave) * kSeg, &st->near_cnt[cur], cur ? near1 : near0, cap,
         &st->overflow);
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t base = static_cast<int64_t>(blockIdx.x) * blockDim.x; base < n; base += stride) {
    const int64_t v = base + threadIdx.x;
    bool take = false;
    if (v < n) {
      const int64_t d = word_dist(word[v]);
      take = d < kDistInf && d / delta == nb;
    }
    q.push(take, static_cast<int32_t>(v));
  }
  q.flush_block();
}

__global__ void sssp_split_done(SsspState *st) {
  if (threadIdx.x != 0 || st->done || st->phase != 2) return;
  st->bucket = static_cast<long long>(st->next_bucket);
  st->buckets += 1;
  st->phase = 0;
}

// Distances out of the words, plus reached vertices and their out-edges.
__global__ void sssp_finish(const int64_t *off, const unsigned long long *word, int64_t n,
                            int32_t *dist, unsigned long long *p_reached, unsigned long long *p_edges) {
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  unsigned long long r = 0, e = 0;
  for (int64_t v = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; v < n; v += stride) {
    const int32_t d = static_cast<int32_t>(word_dist(word[v]));
    dist[v] = d;
    if (d != kDistInf) {
      r++;
      e += static_cast<unsigned long long>(off[v + 1] - off[v]);
    }
  }
  block_partial(r, p_reached);
  block_partial(e, p_edges);
}

}  // namespace gbk

using namespace gbk;

int gbk_sssp_device(gbk_ctx *ctx, const gbk_graph *g, int32_t source, int32_t delta, int32_t *dist,
                    gbk_sssp_stats *stats) {
  const int64_t n = g->n;
  hipStream_t s = ctx->stream;
  const int64_t cap = n + 64;
  auto *word = static_cast<unsigned long long *>(scratch(ctx, kS5, sizeof(unsigned long long) * cap));
  auto *q = static_cast<int32_t *>(scratch(ctx, kS7, sizeof(int32_t) * cap * 2));
  auto *st = static_cast<SsspState *>(scratch(ctx, kS8, sizeof(SsspState) + 64));
  auto *fin = static_cast<unsigned long long *>(scratch(ctx, kS9, 64));
  auto *partial = static_cast<unsigned long long *>(scratch(ctx, kS10, sizeof(unsigned long long) * 2 * kGridCap));
  auto *hs = static_cast<SsspState *>(pinned(ctx, sizeof(SsspState)));
  HeavyQueue hq;
  if (!word || !q || !st || !fin || !partial || !hs || !make_heavy_queue(ctx, g, kS11, &hq))
    return set_error(ctx, GBK_ENOMEM, "sssp scratch");
  hq.overflow = &st->overflow;
  int32_t *near0 = q, *near1 = q + cap;
  const int64_t H = ctx->tune.heavy_degree;
  const unsigned relax_grid = static_cast<unsigned>(ctx->tune.relax_grid);
  const bool count = ctx->ktime == GBK_KT_SSSP_RELAX && ctx->kt_dev;
  const unsigned count_grid = static_cast<unsigned>(kKtPartials < 256 ? kKtPartials : 256);
  const unsigned scan_grid = expand_grid(n);
  // one bucket step: relaxation round, then the bucket bookkeeping (each
  // kernel returns at once unless the state's phase is its own)
  auto step = [&]() {
      if (count) {
        unsigned long long *part = ctx->kt_dev + 2;
        sssp_count_round<<<count_grid, kBlock, 0, s>>>(g->off, word, near0, near1, delta, st, part);
        sum_partials<unsigned long long><<<1, 1024, 0, s>>>(part, static_cast<int>(count_grid), ctx->kt_dev, 1);
        sum_partials<unsigned long long><<<1, 1024, 0, s>>>(part + count_grid, static_cast<int>(count_grid),
                                                          ctx->kt_dev + 1, 1);
      }
      const int kt = kt_begin(ctx, GBK_KT_SSSP_RELAX);
      sssp_relax<<<relax_grid, kBlock, 0, s>>>(g->off, g->nbr, g->wgt, word, near0, near1, cap, delta, H,
                                               st, hq);
      sssp_relax_heavy<<<kHeavyGrid, kBlock, 0, s>>>(g->nbr, g->wgt, word, near0, near1, cap, delta, st,
                                                     hq);
      kt_end(ctx, kt);
      sssp_relax_done<<<1, kWave, 0, s>>>(st, hq.count);
      sssp_min_bucket<<<scan_grid, kBlock, 0, s>>>(word, n, delta, st, partial);
      sssp_min_done<<<1, 1024, 0, s>>>(partial, static_cast<int>(scan_grid), st);
      sssp_split<<<scan_grid, kBlock, 0, s>>>(word, n, near0, near1, cap, delta, st);
      sssp_split_done<<<1, kWave, 0, s>>>(st);
  };
  // steps are queued blindly in blocks of kStepsPerGraph (a captured graph)
  // and the host reads the state after 4, 8, ... 64 of them
  constexpr int kStepsPerGraph = 4;
  auto block = [&]() {
    for (int j = 0; j < kStepsPerGraph; j++) step();
  };
  LoopGraph loop(ctx);
  if (int rc = loop.capture(block)) return rc;
  EventTimer timer(ctx);
  sssp_init<<<grid_for(n), kBlock, 0, s>>>(n, source, word, near0, st);
  int batch = kStepsPerGraph;
  int64_t launched = 0;
  for (;;) {
    for (int j = 0; j < batch; j += kStepsPerGraph)
      if (int rc = loop.run(block)) return rc;
    launched += batch;
    GBK_LAUNCH_CHECK(ctx, "sssp step");
    GBK_HIP(ctx, hipMemcpyAsync(hs, st, sizeof(SsspState), hipMemcpyDeviceToHost, s));
    GBK_HIP(ctx, hipStreamSynchronize(s));
    kt_collect(ctx);
    if (hs->overflow) return set_error(ctx, GBK_EHIP, "sssp: queue overflow");
    if (hs->done) break;
    if (batch < 64) batch *= 2;
    // every step either settles a bucket or shrinks the work of the current
    // one; more steps than buckets times rounds would mean a broken state
    if (launched > 4 * (n + 16) + 1024) return set_error(ctx, GBK_EHIP, "sssp: no convergence");
  }
  sssp_finish<<<scan_grid, kBlock, 0, s>>>(g->off, word, n, dist, partial, partial + kGridCap);
  sum_partials<unsigned long long><<<1, 1024, 0, s>>>(partial, static_cast<int>(scan_grid), fin, 0);
  sum_partials<unsigned long long><<<1, 1024, 0, s>>>(partial + kGridCap, static_cast<int>(scan_grid), fin + 1, 0);
  GBK_LAUNCH_CHECK(ctx, "sssp_finish");
  const double ms = timer.stop_ms();
  unsigned long long hf[2] = {0, 0};
  GBK_HIP(ctx, hipMemcpyAsync(hf, fin, sizeof(hf), hipMemcpyDeviceToHost, s));
  GBK_HIP(ctx, hipStreamSynchronize(s));
  if (stats) {
    stats->device_ms = ms;
    stats->buckets = hs->buckets + 1;
    stats->relax_rounds = hs->relax_rounds;
    stats->reached = static_cast<int64_t>(hf[0]);
    stats->edges_reached = static_cast<int64_t>(hf[1]);
  }
  return GBK_OK;
}

extern "C" int gbk_sssp(gbk_ctx *ctx, const gbk_graph *g, int32_t source, int32_t delta,
                        int32_t *dist_out, gbk_sssp_stats *stats) {
  if (!ctx || !g) return set_error(ctx, GBK_EINVAL, "sssp: null argument");
  if (!g->weighted) return set_error(ctx, GBK_ESTATE, "sssp: graph has no weights");
  if (source < 0 || source >= g->n) return set_error(ctx, GBK_EINVAL, "sssp: source out of range");
  if (delta <= 0) return set_error(ctx, GBK_EINVAL, "sssp: delta must be positive");
  auto *dist = static_cast<int32_t *>(scratch(ctx, kS4, sizeof(int32_t) * (g->n + 1)));
  if (!dist) return set_error(ctx, GBK_ENOMEM, "sssp dist");
  int rc = gbk_sssp_device(ctx, g, source, delta, dist, stats);
  if (rc) return rc;
  if (dist_out) {
    GBK_HIP(ctx, hipMemcpyAsync(dist_out, dist, sizeof(int32_t) * g->n, hipMemcpyDeviceToHost, ctx->stream));
    GBK_HIP(ctx, hipStreamSynchronize(ctx->stream));
  }
  return GBK_OK;
}
