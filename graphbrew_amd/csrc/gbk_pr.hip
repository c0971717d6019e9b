// gbk_pr.hip -- PageRank pull on gfx950.
//
// PageRankPullGS (bench/src/pr.cc:40-94): Gauss-Seidel, contributions are
// updated in place so later vertices of the same sweep may see new values
// (the reference's relaxed atomics allow exactly this).
// PageRankPull (bench/src/pr_spmv.cc:40-83): Jacobi; the contribution pass
// of the next iteration is fused into this iteration's store (contrib_next =
// score / out_degree), which is the same float the reference recomputes.
//
// One launch per iteration plus a one-workgroup finish launch.  The pull
// launch works on the graph's fixed, edge-balanced units: heavy in-rows (more
// than the graph's heavy threshold) in 1024-edge chunks, each summed by a
// workgroup and added into the row's accumulator, then tiles of at most 64
// rows and about 2048 edges, one per wave.  A tile's in-edges are contiguous
// in the CSR, so the wave streams them in 64-edge chunks with coalesced
// neighbour reads, maps each edge to its owner lane by a 6-step shuffle
// search over the tile's edge prefix, gathers contrib[v] and accumulates
// into the owner's LDS slot.  However the rows are ordered (degree-sorted
// orders put all hubs side by side), no unit is larger than a chunk or a
// tile.  The finish launch applies the heavy rows, adds the per-workgroup L1
// changes, decides convergence and raises a device flag that makes the
// remaining queued launches return at once -- no host synchronisation
// inside the iteration loop and no same-address atomics beyond one per
// heavy chunk.
#include "gbk_expand.h"

namespace gbk {

struct PrState {
  int done;
  int iters;
  int converged;
  int pad;
  double last_err;
};

__global__ void pr_init(const int64_t *out_off, int64_t n, float init, float *scores,
                        float *contrib, PrState *st) {
  const int64_t v = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (v < n) {
    const int64_t d = out_off[v + 1] - out_off[v];
    scores[v] = init;
    contrib[v] = d ? init / static_cast<float>(d) : 0.0f;
  }
  if (v == 0) {
    st->done = 0;
    st->iters = 0;
    st->converged = -1;
    st->last_err = 0;
  }
}

// One workgroup: applies the heavy rows (their sums were accumulated chunk
// by chunk into hacc by the pull launch), adds their L1 change to the
// per-workgroup partials of iteration 'iter' and decides convergence
// (pr.cc:77-88).
__global__ void pr_finish(const double *partial, int count, PrState *st, double *errs, int iter,
                          double tol, int fixed, const int32_t *heavy, int64_t nheavy, float *hacc,
                          const int64_t *out_off, float *scores, float *contrib_next, float base,
                          float damp) {
  if (*reinterpret_cast<volatile int *>(&st->done)) return;
  __shared__ double red[1024 / kWave];
  double s = 0;
  for (int i = threadIdx.x; i < count; i += blockDim.x) s += partial[i];
  for (int64_t h = threadIdx.x; h < nheavy; h += blockDim.x) {
    const int32_t u = heavy[h];
    const float sum = hacc[h];
    hacc[h] = 0.0f;
    const float old = scores[u];
    const float nw = base + damp * sum;
    scores[u] = nw;
    s += fabs(static_cast<double>(nw - old));
    const int64_t d = out_off[u + 1] - out_off[u];
    contrib_next[u] = d ? nw / static_cast<float>(d) : 0.0f;
  }
  s = wave_sum(s);
  if (lane_id() == 0) red[threadIdx.x / kWave] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double total = 0;
    for (int i = 0; i < static_cast<int>(blockDim.x) / kWave; i++) total += red[i];
    errs[iter] = total;
    st->last_err = total;
    st->iters = iter + 1;
    if (st->converged < 0 && total < tol) st->converged = iter + 1;
    if (!fixed && total < tol) st->done = 1;
  }
}

// Heavy in-rows chunk by chunk (workgroup sums added into hacc), then light
// rows tile by tile (one wave per tile of at most 64 rows, the owner of each
// edge found by a shuffle search over the tile's edge prefix).
template <bool GS>
__global__ void __launch_bounds__(kBlock)
pr_pull_iter(RowTiles rt, const int32_t *in_nbr, const int64_t *out_off, float *scores,
             const float *contrib_cur, float *contrib_next, const PrState *st, double *partial,
             float *hacc, float base, float damp) {
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
    for (int64_t k = first + threadIdx.x; k < last; k += blockDim.x) part += contrib_cur[in_nbr[s + k]];
    const float sum = block_sum(part, hred);
    if (threadIdx.x == 0) atomicAdd(&hacc[h], sum);
  }
  double err = 0;
  const int64_t waves = static_cast<int64_t>(gridDim.x) * kWavesPerBlock;
  for (int64_t t = static_cast<int64_t>(blockIdx.x) * kWavesPerBlock + wv; t < rt.ntiles; t += waves) {
    const int64_t u = rt.tiles[t] + lane;
    const bool valid = u < rt.tiles[t + 1];
    int64_t start = 0, len = 0;
    if (valid) {
      start = rt.off[u];
      len = rt.off[u + 1] - start;
    }
    const bool light = valid && len <= rt.heavy_deg;
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
      const float val = act ? contrib_cur[in_nbr[ostart + (e - oexcl)]] : 0.0f;
      bool tail;
      const float run = seg_run_sum(val, act ? lo : kWave, tail);
      // one writer per owner per chunk; chunks are sequential in this wave
      if (act && tail) acc[wv * kWave + lo] += run;
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "workgroup");
    __builtin_amdgcn_wave_barrier();
    if (light) {
      const float old = scores[u];
      const float nw = base + damp * acc[threadIdx.x];
      scores[u] = nw;
      err += fabs(static_cast<double>(nw - old));
      const int64_t d = out_off[u + 1] - out_off[u];
      contrib_next[u] = d ? nw / static_cast<float>(d) : 0.0f;
    }
    __builtin_amdgcn_wave_barrier();
  }
  block_partial(err, partial);
}

}  // namespace gbk

using namespace gbk;

extern "C" int gbk_pr(gbk_ctx *ctx, const gbk_graph *g, int variant, int max_iters, double tol,
                      int fixed_work, float *scores_out, gbk_pr_stats *stats) {
  if (!ctx || !g || max_iters <= 0 || (variant != GBK_PR_GS && variant != GBK_PR_SPMV))
    return set_error(ctx, GBK_EINVAL, "pr: bad arguments");
  const int64_t n = g->n;
  if (n == 0) return set_error(ctx, GBK_EINVAL, "pr: empty graph");
  hipStream_t s = ctx->stream;
  const RowTiles rt = row_tiles(g, true);
  const unsigned grid = rows_grid(rt);
  auto *scores = static_cast<float *>(scratch(ctx, kS0, sizeof(float) * n));
  auto *c0 = static_cast<float *>(scratch(ctx, kS1, sizeof(float) * n));
  auto *c1 = static_cast<float *>(scratch(ctx, kS2, sizeof(float) * n));
  auto *st = static_cast<PrState *>(scratch(ctx, kS3, sizeof(PrState)));
  auto *errs = static_cast<double *>(scratch(ctx, kS4, sizeof(double) * max_iters));
  auto *partial = static_cast<double *>(scratch(ctx, kS5, sizeof(double) * grid));
  auto *hacc = static_cast<float *>(scratch(ctx, kS6, sizeof(float) * (rt.nheavy + 1)));
  if (!scores || !c0 || !c1 || !st || !errs || !partial || !hacc)
    return set_error(ctx, GBK_ENOMEM, "pr scratch");
  const float kDamp = 0.85f;
  const float init = 1.0f / static_cast<float>(n);
  const float base = (1.0f - kDamp) / static_cast<float>(n);
  // all iterations are queued (one captured graph); after convergence the
  // kernels return at once on the state's done flag
  auto iterations = [&]() {
  for (int it = 0; it < max_iters; it++) {
    const int kt = kt_begin(ctx, GBK_KT_PR_ITER);
    float *cur = c0, *nxt = c0;
    if (variant == GBK_PR_GS) {
      pr_pull_iter<true><<<grid, kBlock, 0, s>>>(rt, g->in_nbr, g->off, scores, c0, c0, st, partial, hacc,
                                                 base, kDamp);
    } else {
      cur = (it & 1) ? c1 : c0;
      nxt = (it & 1) ? c0 : c1;
      pr_pull_iter<false><<<grid, kBlock, 0, s>>>(rt, g->in_nbr, g->off, scores, cur, nxt, st, partial,
                                                  hacc, base, kDamp);
    }
    kt_end(ctx, kt);
    pr_finish<<<1, 1024, 0, s>>>(partial, static_cast<int>(grid), st, errs, it, tol, fixed_work, rt.heavy,
                                 rt.nheavy, hacc, g->off, scores, nxt, base, kDamp);
  }
  };
  LoopGraph loop(ctx);
  if (int rc = loop.capture(iterations)) return rc;
  EventTimer timer(ctx);
  fill<float>(ctx, hacc, 0.0f, rt.nheavy + 1);
  pr_init<<<grid_for(n), kBlock, 0, s>>>(g->off, n, init, scores, c0, st);
  if (int rc = loop.run(iterations)) return rc;
  GBK_LAUNCH_CHECK(ctx, "pr_pull_iter");
  const double ms = timer.stop_ms();
  PrState h{};
  GBK_HIP(ctx, hipMemcpyAsync(&h, st, sizeof(h), hipMemcpyDeviceToHost, s));
  if (scores_out) GBK_HIP(ctx, hipMemcpyAsync(scores_out, scores, sizeof(float) * n, hipMemcpyDeviceToHost, s));
  GBK_HIP(ctx, hipStreamSynchronize(s));
  if (stats) {
    stats->device_ms = ms;
    stats->iterations = h.iters;
    stats->converged_iter = h.converged;
    stats->final_error = h.last_err;
  }
  return GBK_OK;
}
