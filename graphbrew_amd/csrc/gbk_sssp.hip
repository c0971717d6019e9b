// gbk_sssp.hip -- delta-stepping SSSP (bench/src/sssp.cc:129-283).
//
// Buckets of width delta as in the reference: a vertex is relaxed from the
// current bucket b only while dist[v] >= delta*b (stale copies are skipped,
// sssp.cc:191-192), relaxations use an atomic minimum, and vertices that
// land in bucket b are processed again in b before the next bucket opens
// (the reference's bucket fusion does the same for small bins).  On the GPU:
//   near[2]  ping-pong queues of the current bucket (deduplicated per round)
//   far[2]   one pending list for every later bucket (deduplicated per list)
// Three grid-wide kernels form one step -- relax the near queue, find the
// smallest pending bucket, split the pending list -- each followed by a
// one-workgroup kernel that advances the device state block; every kernel
// acts only in its phase.  The host queues steps without waiting and checks
// the 'done' flag every few steps, so the number of host round trips is
// logarithmic in the number of steps.  Queue appends are staged per wave in
// LDS (gbk_expand.h WaveQueue), so the shared tails see one atomic per
// batch of up to 512 entries.  Distances are
// integers, so the result is exact whatever the order of relaxations.
#include <climits>

#include "gbk_expand.h"

namespace gbk {

constexpr int32_t kDistInf = INT32_MAX / 2;  // sssp.cc:67

struct SsspState {
  unsigned long long near_cnt[2];
  unsigned long long far_cnt[2];
  unsigned long long next_bucket;
  long long bucket;
  long long relax_rounds;
  long long buckets;
  int cur, fcur, phase, done;
  int overflow;
  unsigned int blocks_done;
};

__global__ void sssp_init(int64_t n, int32_t source, int32_t *dist, int32_t *stamp,
                          int32_t *fstamp, int32_t *near0, SsspState *st) {
  const int64_t v = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (v < n) {
    dist[v] = v == source ? 0 : kDistInf;
    stamp[v] = -1;
    fstamp[v] = -1;
  }
  if (v == 0) {
    near0[0] = source;
    st->near_cnt[0] = 1;
    st->near_cnt[1] = 0;
    st->far_cnt[0] = 0;
    st->far_cnt[1] = 0;
    st->next_bucket = ULLONG_MAX;
    st->bucket = 0;
    st->relax_rounds = 0;
    st->buckets = 0;
    st->cur = 0;
    st->fcur = 0;
    st->phase = 0;
    st->done = 0;
    st->overflow = 0;
    st->blocks_done = 0;
  }
}

struct RelaxEdge {
  const int32_t *nbr;
  const int32_t *wgt;
  int32_t *dist;
  int32_t *stamp;
  int32_t *fstamp;
  WaveQueue nearq;
  WaveQueue farq;
  int64_t delta;
  int64_t bucket;
  int32_t near_tag;
  int32_t far_gen;
  __device__ void operator()(int64_t u, int64_t k, bool active) {
    bool to_near = false, to_far = false;
    int32_t v = 0;
    if (active) {
      v = nbr[k];
      const int32_t nd = dist[u] + wgt[k];
      // test first: most relaxations do not improve, and a plain load can
      // only overstate dist[v] (values only decrease), never hide a win
      const int32_t old = nd < dist[v] ? atomicMin(&dist[v], nd) : nd;
      if (nd < old) {
        if (nd / delta == bucket) to_near = atomicExch(&stamp[v], near_tag) != near_tag;
        else to_far = atomicExch(&fstamp[v], far_gen) != far_gen;
      }
    }
    nearq.push(to_near, v);
    farq.push(to_far, v);
  }
};

__global__ void __launch_bounds__(kBlock)
sssp_relax(const int64_t *off, const int32_t *nbr, const int32_t *wgt, int32_t *dist,
           int32_t *stamp, int32_t *fstamp, int32_t *near0, int32_t *near1, int32_t *far0,
           int32_t *far1, int64_t cap, int64_t delta, int64_t heavy_deg, SsspState *st) {
  const volatile SsspState *vs = st;
  if (vs->done || vs->phase != 0) return;
  __shared__ int32_t seg[2 * kWavesPerBlock * kSeg];
  const int cur = vs->cur, fcur = vs->fcur;
  const int64_t bucket = vs->bucket;
  const int64_t count = static_cast<int64_t>(vs->near_cnt[cur]);
  const int32_t *list = cur ? near1 : near0;
  const int w = threadIdx.x / kWave;
  RelaxEdge f;
  f.nbr = nbr;
  f.wgt = wgt;
  f.dist = dist;
  f.stamp = stamp;
  f.fstamp = fstamp;
  f.nearq.init(seg + w * kSeg, &st->near_cnt[1 - cur], cur ? near0 : near1, cap, &st->overflow);
  f.farq.init(seg + (kWavesPerBlock + w) * kSeg, &st->far_cnt[fcur], fcur ? far1 : far0, cap,
              &st->overflow);
  f.delta = delta;
  f.bucket = bucket;
  f.near_tag = static_cast<int32_t>(2 * vs->relax_rounds + 2);
  f.far_gen = static_cast<int32_t>(vs->buckets);
  const int64_t lower = delta * bucket;
  auto rowlen = [&](int64_t u, int64_t &s, int64_t &l) {
    s = off[u];
    l = dist[u] >= lower ? off[u + 1] - s : 0;
  };
  expand_list(list, count, heavy_deg, rowlen, f);
  f.nearq.flush();
  f.farq.flush();
}

// One workgroup: closes a relaxation round (swap the near queues; open the
// bucket search when the next near queue is empty).
__global__ void sssp_relax_done(SsspState *st) {
  if (threadIdx.x != 0 || st->done || st->phase != 0) return;
  const int cur = st->cur;
  st->near_cnt[cur] = 0;
  st->cur = 1 - cur;
  st->relax_rounds += 1;
  if (st->near_cnt[1 - cur] == 0) {
    st->phase = 1;
    st->next_bucket = ULLONG_MAX;
  }
}

__global__ void sssp_min_bucket(const int32_t *dist, const int32_t *far0, const int32_t *far1,
                                int64_t delta, SsspState *st, unsigned long long *partial) {
  const volatile SsspState *vs = st;
  if (vs->done || vs->phase != 1) return;
  const int fcur = vs->fcur;
  const int64_t bucket = vs->bucket;
  const int64_t count = static_cast<int64_t>(vs->far_cnt[fcur]);
  const int32_t *far = fcur ? far1 : far0;
  unsigned long long best = ULLONG_MAX;
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < count; i += stride) {
    const int64_t b = dist[far[i]] / delta;
    if (b > bucket && static_cast<unsigned long long>(b) < best) best = static_cast<unsigned long long>(b);
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
    for (int i = 1; i < static_cast<int>(blockDim.x) / kWave; i++) m = red[i] < m ? red[i] : m;
    m = red[0] < m ? red[0] : m;
    st->next_bucket = m;
    if (m == ULLONG_MAX) {
      st->done = 1;
      st->far_cnt[st->fcur] = 0;
    } else {
      st->phase = 2;
    }
  }
}

__global__ void __launch_bounds__(kBlock)
sssp_split(const int32_t *dist, int32_t *stamp, int32_t *fstamp, int32_t *near0,
           int32_t *near1, int32_t *far0, int32_t *far1, int64_t cap, int64_t delta,
           SsspState *st) {
  const volatile SsspState *vs = st;
  if (vs->done || vs->phase != 2) return;
  __shared__ int32_t seg[2 * kWavesPerBlock * kSeg];
  const int cur = vs->cur, fcur = vs->fcur;
  const int64_t nb = static_cast<int64_t>(vs->next_bucket);
  const int64_t count = static_cast<int64_t>(vs->far_cnt[fcur]);
  const int32_t *far = fcur ? far1 : far0;
  const int32_t near_tag = static_cast<int32_t>(2 * vs->relax_rounds + 1);
  const int32_t gen = static_cast<int32_t>(vs->buckets + 1);
  const int w = threadIdx.x / kWave;
  WaveQueue nq, fq;
  nq.init(seg + w * kSeg, &st->near_cnt[cur], cur ? near1 : near0, cap, &st->overflow);
  fq.init(seg + (kWavesPerBlock + w) * kSeg, &st->far_cnt[1 - fcur], fcur ? far0 : far1, cap,
          &st->overflow);
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  // uniform trip count per wave so the queue pushes can ballot
  for (int64_t base = static_cast<int64_t>(blockIdx.x) * blockDim.x; base < count; base += stride) {
    const int64_t i = base + threadIdx.x;
    bool to_near = false, to_far = false;
    int32_t v = 0;
    if (i < count) {
      v = far[i];
      const int64_t b = dist[v] / delta;
      if (b == nb) to_near = atomicExch(&stamp[v], near_tag) != near_tag;
      else if (b > nb) to_far = atomicExch(&fstamp[v], gen) != gen;
    }
    nq.push(to_near, v);
    fq.push(to_far, v);
  }
  nq.flush();
  fq.flush();
}

__global__ void sssp_split_done(SsspState *st) {
  if (threadIdx.x != 0 || st->done || st->phase != 2) return;
  st->far_cnt[st->fcur] = 0;
  st->fcur = 1 - st->fcur;
  st->bucket = static_cast<long long>(st->next_bucket);
  st->buckets += 1;
  st->phase = 0;
}

__global__ void sssp_finish(const int64_t *off, const int32_t *dist, int64_t n,
                            unsigned long long *p_reached, unsigned long long *p_edges) {
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  unsigned long long r = 0, e = 0;
  for (int64_t v = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; v < n; v += stride) {
    if (dist[v] != kDistInf) {
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
  auto *stamp = static_cast<int32_t *>(scratch(ctx, kS5, sizeof(int32_t) * cap));
  auto *fstamp = static_cast<int32_t *>(scratch(ctx, kS6, sizeof(int32_t) * cap));
  auto *q = static_cast<int32_t *>(scratch(ctx, kS7, sizeof(int32_t) * cap * 4));
  auto *st = static_cast<SsspState *>(scratch(ctx, kS8, sizeof(SsspState) + 64));
  auto *fin = static_cast<unsigned long long *>(scratch(ctx, kS9, 64));
  auto *hs = static_cast<SsspState *>(pinned(ctx, sizeof(SsspState)));
  if (!stamp || !fstamp || !q || !st || !fin || !hs) return set_error(ctx, GBK_ENOMEM, "sssp scratch");
  int32_t *near0 = q, *near1 = q + cap, *far0 = q + 2 * cap, *far1 = q + 3 * cap;
  const int64_t H = ctx->tune.heavy_degree;
  const unsigned grid = 1024;
  auto *partial = static_cast<unsigned long long *>(scratch(ctx, kS10, sizeof(unsigned long long) * 2 * kGridCap));
  if (!partial) return set_error(ctx, GBK_ENOMEM, "sssp partials");
  EventTimer timer(ctx);
  sssp_init<<<grid_for(n), kBlock, 0, s>>>(n, source, dist, stamp, fstamp, near0, st);
  int batch = 2;
  int64_t launched = 0;
  for (;;) {
    for (int j = 0; j < batch; j++) {
      sssp_relax<<<grid, kBlock, 0, s>>>(g->off, g->nbr, g->wgt, dist, stamp, fstamp, near0, near1,
                                         far0, far1, cap, delta, H, st);
      sssp_relax_done<<<1, kWave, 0, s>>>(st);
      sssp_min_bucket<<<grid, kBlock, 0, s>>>(dist, far0, far1, delta, st, partial);
      sssp_min_done<<<1, 1024, 0, s>>>(partial, static_cast<int>(grid), st);
      sssp_split<<<grid, kBlock, 0, s>>>(dist, stamp, fstamp, near0, near1, far0, far1, cap, delta, st);
      sssp_split_done<<<1, kWave, 0, s>>>(st);
    }
    launched += batch;
    GBK_LAUNCH_CHECK(ctx, "sssp step");
    GBK_HIP(ctx, hipMemcpyAsync(hs, st, sizeof(SsspState), hipMemcpyDeviceToHost, s));
    GBK_HIP(ctx, hipStreamSynchronize(s));
    if (hs->overflow) return set_error(ctx, GBK_EHIP, "sssp: queue overflow");
    if (hs->done) break;
    if (batch < 64) batch *= 2;
    if (launched > 4 * (n + 16) + 1024) return set_error(ctx, GBK_EHIP, "sssp: no convergence");
  }
  const unsigned ng = expand_grid(n);
  sssp_finish<<<ng, kBlock, 0, s>>>(g->off, dist, n, partial, partial + kGridCap);
  sum_partials<unsigned long long><<<1, 1024, 0, s>>>(partial, static_cast<int>(ng), fin, 0);
  sum_partials<unsigned long long><<<1, 1024, 0, s>>>(partial + kGridCap, static_cast<int>(ng), fin + 1, 0);
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
