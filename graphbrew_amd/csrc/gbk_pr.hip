// gbk_pr.hip -- PageRank pull on gfx950.
//
// PageRankPullGS (bench/src/pr.cc:40-94): Gauss-Seidel, contributions are
// updated in place so later vertices of the same sweep may see new values
// (the reference's relaxed atomics allow exactly this).
// PageRankPull (bench/src/pr_spmv.cc:40-83): Jacobi; the contribution pass
// of the next iteration is fused into this iteration's store (contrib_next =
// score / out_degree), which is the same float the reference recomputes.
//
// One launch per iteration plus a one-workgroup finish launch.  A block
// takes either a heavy row (more than heavy_degree in-edges, summed by the
// whole workgroup) or 4 windows of 64 consecutive vertices, one per wave;
// a window's in-edges are contiguous in the CSR, so the wave streams them in
// 64-edge chunks with coalesced neighbour reads, maps each edge to its
// owner lane by a 6-step shuffle search over the window's edge prefix,
// gathers contrib[v] and accumulates into the owner's LDS slot.  Blocks
// grid-stride so that only ~2K per-block partial L1 changes are written;
// the finish launch adds them, decides convergence and raises a device flag
// that makes the remaining queued launches return at once -- no host
// synchronisation inside the iteration loop and no same-address atomics.
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

// One workgroup: sums the per-workgroup L1 changes of iteration 'iter' and
// decides convergence (pr.cc:77-88).
__global__ void pr_finish(const double *partial, int count, PrState *st, double *errs, int iter,
                          double tol, int fixed) {
  if (*reinterpret_cast<volatile int *>(&st->done)) return;
  __shared__ double red[1024 / kWave];
  double s = 0;
  for (int i = threadIdx.x; i < count; i += blockDim.x) s += partial[i];
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

// Grid-stride over [heavy rows][64-row windows per wave].
template <bool GS>
__global__ void __launch_bounds__(kBlock)
pr_pull_iter(const int64_t *in_off, const int32_t *in_nbr, const int64_t *out_off, int64_t n,
             const int32_t *heavy, int64_t nheavy, int64_t heavy_deg, float *scores,
             const float *contrib_cur, float *contrib_next, const PrState *st, double *partial,
             float base, float damp) {
  if (*reinterpret_cast<const volatile int *>(&st->done)) return;
  __shared__ float acc[kBlock];
  __shared__ float hred[kWavesPerBlock];
  const int lane = lane_id();
  const int wv = threadIdx.x / kWave;
  const int64_t windows = (n + kBlock - 1) / kBlock;
  const int64_t items = nheavy + windows;
  double err = 0;
  for (int64_t b = blockIdx.x; b < items; b += gridDim.x) {
    if (b < nheavy) {
      // one heavy row per workgroup
      const int32_t u = heavy[b];
      const int64_t s = in_off[u], e = in_off[u + 1];
      float part = 0;
      for (int64_t k = s + threadIdx.x; k < e; k += blockDim.x) part += contrib_cur[in_nbr[k]];
      const float sum = block_sum(part, hred);
      if (threadIdx.x == 0) {
        const float old = scores[u];
        const float nw = base + damp * sum;
        scores[u] = nw;
        err += fabs(static_cast<double>(nw - old));
        const int64_t d = out_off[u + 1] - out_off[u];
        contrib_next[u] = d ? nw / static_cast<float>(d) : 0.0f;
      }
      continue;
    }
    const int64_t u = ((b - nheavy) * kWavesPerBlock + wv) * kWave + lane;
    const bool valid = u < n;
    int64_t start = 0, len = 0;
    if (valid) {
      start = in_off[u];
      len = in_off[u + 1] - start;
    }
    const bool light = valid && len <= heavy_deg;
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
  const int64_t windows = (n + kBlock - 1) / kBlock;
  const unsigned grid = expand_grid((windows + g->n_heavy_in) * kBlock);
  auto *scores = static_cast<float *>(scratch(ctx, kS0, sizeof(float) * n));
  auto *c0 = static_cast<float *>(scratch(ctx, kS1, sizeof(float) * n));
  auto *c1 = static_cast<float *>(scratch(ctx, kS2, sizeof(float) * n));
  auto *st = static_cast<PrState *>(scratch(ctx, kS3, sizeof(PrState)));
  auto *errs = static_cast<double *>(scratch(ctx, kS4, sizeof(double) * max_iters));
  auto *partial = static_cast<double *>(scratch(ctx, kS5, sizeof(double) * grid));
  if (!scores || !c0 || !c1 || !st || !errs || !partial) return set_error(ctx, GBK_ENOMEM, "pr scratch");
  const float kDamp = 0.85f;
  const float init = 1.0f / static_cast<float>(n);
  const float base = (1.0f - kDamp) / static_cast<float>(n);
  EventTimer timer(ctx);
  pr_init<<<grid_for(n), kBlock, 0, s>>>(g->off, n, init, scores, c0, st);
  for (int it = 0; it < max_iters; it++) {
    const int kt = kt_begin(ctx, GBK_KT_PR_ITER);
    if (variant == GBK_PR_GS) {
      pr_pull_iter<true><<<grid, kBlock, 0, s>>>(g->in_off, g->in_nbr, g->off, n, g->heavy_in,
                                                 g->n_heavy_in, ctx->tune.heavy_degree, scores, c0,
                                                 c0, st, partial, base, kDamp);
    } else {
      float *cur = (it & 1) ? c1 : c0;
      float *nxt = (it & 1) ? c0 : c1;
      pr_pull_iter<false><<<grid, kBlock, 0, s>>>(g->in_off, g->in_nbr, g->off, n, g->heavy_in,
                                                  g->n_heavy_in, ctx->tune.heavy_degree, scores, cur,
                                                  nxt, st, partial, base, kDamp);
    }
    kt_end(ctx, kt);
    pr_finish<<<1, 1024, 0, s>>>(partial, static_cast<int>(grid), st, errs, it, tol, fixed_work);
  }
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
