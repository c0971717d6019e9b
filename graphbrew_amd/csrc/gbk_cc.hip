// gbk_cc.hip -- connected components: Afforest (bench/src/cc.cc:82-264) and
// Shiloach-Vishkin (bench/src/cc_sv.cc:64-136).
//
// Both keep the reference's union rules (hook the larger label under the
// smaller one by CAS), so every label only decreases and stays inside its
// component; at convergence each component is labelled by its smallest
// vertex id, which makes the output independent of the schedule.
// Loads of comp[] that decide a CAS use agent-scope atomics, so a spinning
// Link always sees other workgroups' hooks (no stale per-CU cache line).
#include <random>
#include <unordered_map>
#include <utility>

#include "gbk_expand.h"

namespace gbk {

__device__ __forceinline__ void link(int32_t u, int32_t v, int32_t *comp) {
  int32_t p1 = ld_agent(&comp[u]);
  int32_t p2 = ld_agent(&comp[v]);
  while (p1 != p2) {
    const int32_t high = p1 > p2 ? p1 : p2;
    const int32_t low = p1 + (p2 - high);
    const int32_t p_high = ld_agent(&comp[high]);
    if (p_high == low || (p_high == high && atomicCAS(&comp[high], high, low) == high)) break;
    p1 = ld_agent(&comp[ld_agent(&comp[high])]);
    p2 = ld_agent(&comp[low]);
  }
}

__global__ void cc_init(int32_t *comp, int64_t n) {
  const int64_t v = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (v < n) comp[v] = static_cast<int32_t>(v);
}

// Compress (cc.cc:100-124): pointer jumping; each step strictly lowers
// comp[v], so the loop ends.
__global__ void cc_compress(int32_t *comp, int64_t n) {
  const int64_t v = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (v >= n) return;
  int32_t p = ld_agent(&comp[v]);
  int32_t gp = ld_agent(&comp[p]);
  while (p != gp) {
    st_agent(&comp[v], gp);
    p = gp;
    gp = ld_agent(&comp[p]);
  }
}

// Neighbour round r: link u with its r-th neighbour (cc.cc:171-187).
__global__ void cc_link_round(const int64_t *off, const int32_t *nbr, int64_t n, int r,
                              int32_t *comp) {
  const int64_t u = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (u >= n) return;
  const int64_t k = off[u] + r;
  if (k < off[u + 1]) link(static_cast<int32_t>(u), nbr[k], comp);
}

__global__ void cc_gather(const int32_t *comp, const int32_t *idx, int k, int32_t *out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < k) out[i] = comp[idx[i]];
}

struct LinkEdge {
  const int32_t *nbr;
  int32_t *comp;
  __device__ void operator()(int64_t u, int64_t k, bool active) {
    if (active) link(static_cast<int32_t>(u), nbr[k], comp);
  }
};

// Final phase (cc.cc:192-250): every vertex outside the sampled largest
// component links its remaining edges (skip = rounds for the out-CSR, 0 for
// the in-CSR of a directed graph).
__global__ void __launch_bounds__(kBlock)
cc_final_link(RowTiles rt, const int32_t *nbr, int64_t skip, int32_t c, int32_t *comp) {
  LinkEdge f{nbr, comp};
  auto rowlen = [&](int64_t u, int64_t &s, int64_t &l) {
    s = rt.off[u] + skip;
    l = rt.off[u + 1] - s;
    if (l < 0 || ld_agent(&comp[u]) == c) l = 0;
  };
  expand_rows(rt, rowlen, f);
}

// SV hook (cc_sv.cc:86-106) made deterministic: a sweep reads the labels
// as they stood before it ('comp', not written during the sweep) and lowers
// each offered root's entry in 'nxt' (a copy of them) to the smallest label
// offered, by atomic minimum.  The reference's plain stores let whichever
// offer lands last win, so its sweep count depends on the schedule (5 to
// 7.5 sweeps at Kronecker scale 22 on the GPU with first-offer-wins);
// taking the smallest makes every sweep a function of the previous one and
// converges in 4 there.  Labels only decrease along a hook (low < high), so
// no cycle forms, and the converged labels -- each component's smallest
// vertex id -- are the reference's.
// The labels are fully compressed between sweeps, so every label is a
// root and the reference's root test always holds.  An edge offers its
// higher label's root the lower one.  Each edge is read once: a directed
// graph stores it in one row; an undirected graph stores it in both, and it
// is taken from the row of its higher endpoint (neighbour id below the
// row's), so a sweep gathers half the neighbour labels and makes the same
// offers.  Offers to the row's own root (a neighbour's label below the
// row's) are reduced to their minimum in registers, one atomic per row run
// (in the first sweep, where every label is its vertex, all offers are of
// that kind); offers to a neighbour's root are sent edge by edge.
__device__ __forceinline__ int32_t seg_run_min(int32_t x, int32_t key, bool &tail) {
  const int lane = lane_id();
#pragma unroll
  for (int o = 1; o < kWave; o <<= 1) {
    const int32_t y = __shfl_up(x, o);
    const int32_t k = __shfl_up(key, o);
    if (lane >= o && k == key) x = y < x ? y : x;
  }
  const int32_t next = __shfl_down(key, 1);
  tail = (lane == kWave - 1) || (next != key);
  return x;
}

struct HookEdge {
  const int32_t *nbr;
  const int32_t *comp;
  int32_t *nxt;
  int directed;
  int changed;
  __device__ __forceinline__ void offer(int32_t root, int32_t low) {
    // nxt only decreases: a stale read (this XCD's L2 may hold an older
    // line) can only send an atomic that was not needed, never skip one
    // that was
    if (low < root && low < nxt[root] && atomicMin(&nxt[root], low) > low) changed = 1;
  }
  // Heavy rows: every lane of the wave is on the same row.
  __device__ void operator()(int64_t u, int64_t k, bool active) {
    const int32_t cu = comp[u];
    const int32_t v = active ? nbr[k] : 0;
    const int32_t cv = active && (directed || v < u) ? comp[v] : cu;
    if (cv > cu) offer(cv, cu);
    const int32_t m = wave_min(cv < cu ? cv : INT32_MAX);
    if (lane_id() == 0 && m < cu) offer(cu, m);
  }
  // Tiled rows: U edges per lane in flight, rows in contiguous lane runs.
  template <int U>
  __device__ __forceinline__ void batch(const int64_t (&u)[U], const int64_t (&k)[U], const bool (&act)[U]) {
    int32_t v[U], cu[U], cv[U];
#pragma unroll
    for (int j = 0; j < U; j++) v[j] = act[j] ? nbr[k[j]] : 0;
#pragma unroll
    for (int j = 0; j < U; j++) {
      cu[j] = comp[u[j]];
      cv[j] = act[j] && (directed || v[j] < u[j]) ? comp[v[j]] : cu[j];
    }
#pragma unroll
    for (int j = 0; j < U; j++) {
      if (cv[j] > cu[j]) offer(cv[j], cu[j]);
      bool tail;
      const int32_t m = seg_run_min(cv[j] < cu[j] ? cv[j] : INT32_MAX, static_cast<int32_t>(u[j]), tail);
      if (tail && m < cu[j]) offer(cu[j], m);
    }
  }
};

__global__ void __launch_bounds__(kBlock)
cc_sv_hook(RowTiles rt, const int32_t *nbr, const int32_t *comp, int32_t *nxt, int directed, int *change) {
  HookEdge f{nbr, comp, nxt, directed, 0};
  auto rowlen = [&](int64_t u, int64_t &s, int64_t &l) {
    s = rt.off[u];
    l = rt.off[u + 1] - s;
  };
  expand_rows_batch<4>(rt, rowlen, f);
  if (f.changed) *change = 1;
}

}  // namespace gbk

using namespace gbk;

// SampleFrequentElement's indices (cc.cc:127-135): default-seeded mt19937
// through uniform_int_distribution<int32>(0, n-1), 1024 draws.
static const std::vector<int32_t> &afforest_samples(int64_t n) {
  static int64_t cached_n = -1;
  static std::vector<int32_t> idx;
  if (cached_n != n) {
    std::mt19937 gen;
    std::uniform_int_distribution<int32_t> dist(0, static_cast<int32_t>(n - 1));
    idx.resize(1024);
    for (auto &x : idx) x = dist(gen);
    cached_n = n;
  }
  return idx;
}

int gbk_cc_device(gbk_ctx *ctx, const gbk_graph *g, int variant, int32_t *comp,
                  gbk_cc_stats *stats) {
  const int64_t n = g->n;
  hipStream_t s = ctx->stream;
  EventTimer timer(ctx);
  cc_init<<<grid_for(n), kBlock, 0, s>>>(comp, n);
  int iterations = 0;
  int32_t skipped = -1;
  const RowTiles out_rows = row_tiles(g, false);
  if (variant == GBK_CC_AFFOREST) {
    const int rounds = 2;
    for (int r = 0; r < rounds; r++) {
      cc_link_round<<<grid_for(n), kBlock, 0, s>>>(g->off, g->nbr, n, r, comp);
      cc_compress<<<grid_for(n), kBlock, 0, s>>>(comp, n);
    }
    iterations = rounds;
    const auto &idx = afforest_samples(n);
    auto *idx_d = static_cast<int32_t *>(scratch(ctx, kS5, sizeof(int32_t) * 2048));
    auto *hs = static_cast<int32_t *>(pinned(ctx, sizeof(int32_t) * 1024));
    if (!idx_d || !hs) return set_error(ctx, GBK_ENOMEM, "cc sample");
    GBK_HIP(ctx, hipMemcpyAsync(idx_d, idx.data(), sizeof(int32_t) * 1024, hipMemcpyHostToDevice, s));
    cc_gather<<<4, kBlock, 0, s>>>(comp, idx_d, 1024, idx_d + 1024);
    GBK_HIP(ctx, hipMemcpyAsync(hs, idx_d + 1024, sizeof(int32_t) * 1024, hipMemcpyDeviceToHost, s));
    GBK_HIP(ctx, hipStreamSynchronize(s));
    std::unordered_map<int32_t, int> counts(32);
    for (int i = 0; i < 1024; i++) counts[hs[i]]++;
    auto best = counts.begin();
    for (auto it = counts.begin(); it != counts.end(); ++it)
      if (best->second < it->second) best = it;
    skipped = best->first;
    cc_final_link<<<rows_grid(out_rows), kBlock, 0, s>>>(out_rows, g->nbr, rounds, skipped, comp);
    if (g->directed) {
      const RowTiles in_rows = row_tiles(g, true);
      cc_final_link<<<rows_grid(in_rows), kBlock, 0, s>>>(in_rows, g->in_nbr, 0, skipped, comp);
    }
    cc_compress<<<grid_for(n), kBlock, 0, s>>>(comp, n);
    GBK_LAUNCH_CHECK(ctx, "afforest");
  } else {
    auto *change = static_cast<int *>(scratch(ctx, kS5, 64));
    auto *other = static_cast<int32_t *>(scratch(ctx, kS6, sizeof(int32_t) * (n + 1)));
    auto *hc = static_cast<int *>(pinned(ctx, sizeof(int)));
    if (!change || !other || !hc) return set_error(ctx, GBK_ENOMEM, "sv flag");
    int32_t *cur = comp, *nxt = other;
    for (;;) {
      fill<int>(ctx, change, 0, 1);
      GBK_HIP(ctx, hipMemcpyAsync(nxt, cur, sizeof(int32_t) * n, hipMemcpyDeviceToDevice, s));
      cc_sv_hook<<<rows_grid(out_rows), kBlock, 0, s>>>(out_rows, g->nbr, cur, nxt, g->directed, change);
      cc_compress<<<grid_for(n), kBlock, 0, s>>>(nxt, n);
      GBK_LAUNCH_CHECK(ctx, "sv");
      std::swap(cur, nxt);
      GBK_HIP(ctx, hipMemcpyAsync(hc, change, sizeof(int), hipMemcpyDeviceToHost, s));
      GBK_HIP(ctx, hipStreamSynchronize(s));
      iterations++;
      if (!*hc) break;
      // every productive round removes at least one root
      if (iterations > n + 2) return set_error(ctx, GBK_EHIP, "sv: no convergence");
    }
    if (cur != comp) GBK_HIP(ctx, hipMemcpyAsync(comp, cur, sizeof(int32_t) * n, hipMemcpyDeviceToDevice, s));
  }
  const double ms = timer.stop_ms();
  if (stats) {
    stats->device_ms = ms;
    stats->iterations = iterations;
    stats->skipped_label = skipped;
  }
  return GBK_OK;
}

extern "C" int gbk_cc(gbk_ctx *ctx, const gbk_graph *g, int variant, int32_t *comp_out,
                      gbk_cc_stats *stats) {
  if (!ctx || !g || (variant != GBK_CC_AFFOREST && variant != GBK_CC_SV))
    return set_error(ctx, GBK_EINVAL, "cc: bad arguments");
  if (g->n == 0) return set_error(ctx, GBK_EINVAL, "cc: empty graph");
  auto *comp = static_cast<int32_t *>(scratch(ctx, kS4, sizeof(int32_t) * (g->n + 1)));
  if (!comp) return set_error(ctx, GBK_ENOMEM, "cc comp");
  int rc = gbk_cc_device(ctx, g, variant, comp, stats);
  if (rc) return rc;
  if (comp_out) {
    GBK_HIP(ctx, hipMemcpyAsync(comp_out, comp, sizeof(int32_t) * g->n, hipMemcpyDeviceToHost, ctx->stream));
    GBK_HIP(ctx, hipStreamSynchronize(ctx->stream));
  }
  return GBK_OK;
}
