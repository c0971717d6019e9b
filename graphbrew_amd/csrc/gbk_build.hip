// gbk_build.hip -- CSR construction from an edge list on the GPU
// (MakeGraphFromEL + MakeCSR + SquishGraph, gapbs/builder.h:186-318, 444-493).
//
//   count    degrees by atomic add (both endpoints when symmetrizing)
//   scan     raw offsets
//   scatter  each edge into its row via a per-row cursor (row order arbitrary)
//   sort     every row by (neighbour, weight): the relabel sorter with the
//            identity mapping
//   squish   keep the first of equal neighbours (= smallest weight) and drop
//            self loops unless keep_self; a scan of the keep flags gives
//            both the final offsets (value at each row start) and every kept
//            edge's slot
// The result equals the reference's squished CSR exactly: a row's content is
// the sorted set of its distinct neighbours whatever the scatter order.
#include "gbk_expand.h"

int gbk_graph_alloc(gbk_ctx *ctx, int64_t n, int64_t m, int64_t m_in, int directed, int weighted,
                    gbk_graph **out);
int gbk_graph_finalize(gbk_ctx *ctx, gbk_graph *g);

namespace gbk {

__global__ void build_max_id(const int32_t *u, const int32_t *v, int64_t m, int *mx, int *bad) {
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  int best = 0, neg = 0;
  for (int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; e < m; e += stride) {
    const int a = u[e], b = v[e];
    best = max(best, max(a, b));
    neg |= (a < 0) | (b < 0);
  }
  best = wave_max(best);
  if (lane_id() == 0) atomicMax(mx, best);
  if (neg) atomicOr(bad, 1);
}

__global__ void build_count(const int32_t *u, const int32_t *v, int64_t m, int sym, int transpose,
                            int64_t n, unsigned long long *deg, int *bad) {
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; e < m; e += stride) {
    const int32_t a = u[e], b = v[e];
    if (a < 0 || a >= n || b < 0 || b >= n) {
      atomicOr(bad, 1);
      continue;
    }
    if (sym || !transpose) atomicAdd(&deg[a], 1ull);
    if (sym || transpose) atomicAdd(&deg[b], 1ull);
  }
}

__global__ void build_scatter(const int32_t *u, const int32_t *v, const int32_t *w, int64_t m,
                              int sym, int transpose, int64_t n, unsigned long long *cursor,
                              int32_t *nbr, int32_t *wgt) {
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; e < m; e += stride) {
    const int32_t a = u[e], b = v[e];
    if (a < 0 || a >= n || b < 0 || b >= n) continue;
    const int32_t we = w ? w[e] : 1;
    if (sym || !transpose) {
      const unsigned long long p = atomicAdd(&cursor[a], 1ull);
      nbr[p] = b;
      if (wgt) wgt[p] = we;
    }
    if (sym || transpose) {
      const unsigned long long p = atomicAdd(&cursor[b], 1ull);
      nbr[p] = a;
      if (wgt) wgt[p] = we;
    }
  }
}

__global__ void build_iota(int32_t *a, int64_t n) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) a[i] = static_cast<int32_t>(i);
}

struct KeepEdge {
  const int32_t *nbr;
  const int64_t *off;
  int32_t *keep;
  int keep_self;
  __device__ void operator()(int64_t u, int64_t k, bool active) {
    if (!active) return;
    const int32_t x = nbr[k];
    const bool dup = k > off[u] && nbr[k - 1] == x;
    keep[k] = (!dup && (keep_self || x != u)) ? 1 : 0;
  }
};

__global__ void __launch_bounds__(kBlock)
build_keep_flags(const int64_t *off, const int32_t *nbr, int64_t n, int keep_self, int32_t *keep,
                 int64_t heavy_deg) {
  KeepEdge f{nbr, off, keep, keep_self};
  auto rowlen = [&](int64_t u, int64_t &s, int64_t &l) {
    s = off[u];
    l = off[u + 1] - s;
  };
  expand_all(n, heavy_deg, rowlen, f);
}

__global__ void build_final_offsets(const int64_t *sorted_off, const int64_t *pos, int64_t n,
                                    int64_t *off) {
  const int64_t u = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (u <= n) off[u] = pos[sorted_off[u]];
}

__global__ void build_compact(const int32_t *keep, const int64_t *pos, const int32_t *nbr,
                              const int32_t *wgt, int64_t m, int32_t *nbr_out, int32_t *wgt_out) {
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; k < m; k += stride) {
    if (keep[k]) {
      nbr_out[pos[k]] = nbr[k];
      if (wgt) wgt_out[pos[k]] = wgt[k];
    }
  }
}

// Temporary device buffers freed on scope exit.
struct DevTemp {
  std::vector<void *> ptrs;
  void *get(size_t bytes) {
    void *p = nullptr;
    if (hipMalloc(&p, bytes ? bytes : 4) != hipSuccess) return nullptr;
    ptrs.push_back(p);
    return p;
  }
  ~DevTemp() {
    for (void *p : ptrs) hipFree(p);
  }
};

// Builds one squished CSR; returns its final edge count in *m_out and the
// device arrays in (off, nbr, wgt) allocated from 'keep_alive'.
static int build_one(gbk_ctx *ctx, DevTemp &tmp, const int32_t *u, const int32_t *v,
                     const int32_t *w, int64_t m_el, int64_t n, int sym, int transpose,
                     int keep_self, const int32_t *iota, int64_t *m_out, int64_t **off_o,
                     int32_t **nbr_o, int32_t **wgt_o, DevTemp &keep_alive) {
  hipStream_t s = ctx->stream;
  const int64_t raw_m = sym ? 2 * m_el : m_el;
  auto *deg = static_cast<unsigned long long *>(tmp.get(sizeof(int64_t) * (n + 1)));
  auto *raw_off = static_cast<int64_t *>(tmp.get(sizeof(int64_t) * (n + 1)));
  auto *cursor = static_cast<unsigned long long *>(tmp.get(sizeof(int64_t) * (n + 1)));
  auto *raw_nbr = static_cast<int32_t *>(tmp.get(sizeof(int32_t) * raw_m));
  auto *raw_w = w ? static_cast<int32_t *>(tmp.get(sizeof(int32_t) * raw_m)) : nullptr;
  auto *bad = static_cast<int *>(tmp.get(64));
  if (!deg || !raw_off || !cursor || !raw_nbr || (w && !raw_w) || !bad)
    return set_error(ctx, GBK_ENOMEM, "build scratch");
  fill<unsigned long long>(ctx, deg, 0ull, n + 1);
  fill<int>(ctx, bad, 0, 1);
  const unsigned grid = grid_for(m_el, kBlock, 16384);
  build_count<<<grid, kBlock, 0, s>>>(u, v, m_el, sym, transpose, n, deg, bad);
  int rc = exclusive_scan_i64(ctx, reinterpret_cast<int64_t *>(deg), raw_off, n);
  if (rc) return rc;
  GBK_HIP(ctx, hipMemcpyAsync(cursor, raw_off, sizeof(int64_t) * n, hipMemcpyDeviceToDevice, s));
  build_scatter<<<grid, kBlock, 0, s>>>(u, v, w, m_el, sym, transpose, n, cursor, raw_nbr, raw_w);
  GBK_LAUNCH_CHECK(ctx, "build_scatter");
  int hbad = 0;
  GBK_HIP(ctx, hipMemcpyAsync(&hbad, bad, sizeof(int), hipMemcpyDeviceToHost, s));
  GBK_HIP(ctx, hipStreamSynchronize(s));
  if (hbad) return set_error(ctx, GBK_EINVAL, "build: edge endpoint outside [0, n)");
  // sort rows
  auto *srt_off = static_cast<int64_t *>(tmp.get(sizeof(int64_t) * (n + 1)));
  auto *srt_nbr = static_cast<int32_t *>(tmp.get(sizeof(int32_t) * raw_m));
  auto *srt_w = w ? static_cast<int32_t *>(tmp.get(sizeof(int32_t) * raw_m)) : nullptr;
  if (!srt_off || !srt_nbr || (w && !srt_w)) return set_error(ctx, GBK_ENOMEM, "build sort");
  rc = sort_rows(ctx, n, raw_off, raw_nbr, raw_w, iota, reinterpret_cast<int64_t *>(deg), srt_off,
                 srt_nbr, srt_w);
  if (rc) return rc;
  // squish
  auto *keep = static_cast<int32_t *>(tmp.get(sizeof(int32_t) * raw_m));
  auto *pos = static_cast<int64_t *>(tmp.get(sizeof(int64_t) * (raw_m + 1)));
  if (!keep || !pos) return set_error(ctx, GBK_ENOMEM, "build squish");
  build_keep_flags<<<expand_grid(n), kBlock, 0, s>>>(srt_off, srt_nbr, n, keep_self, keep,
                                                     ctx->tune.heavy_degree);
  rc = exclusive_scan_i32(ctx, keep, pos, raw_m);
  if (rc) return rc;
  int64_t total = 0;
  GBK_HIP(ctx, hipMemcpyAsync(&total, pos + raw_m, sizeof(int64_t), hipMemcpyDeviceToHost, s));
  GBK_HIP(ctx, hipStreamSynchronize(s));
  auto *off = static_cast<int64_t *>(keep_alive.get(sizeof(int64_t) * (n + 1)));
  auto *nbr = static_cast<int32_t *>(keep_alive.get(sizeof(int32_t) * total));
  auto *wgt = w ? static_cast<int32_t *>(keep_alive.get(sizeof(int32_t) * total)) : nullptr;
  if (!off || !nbr || (w && !wgt)) return set_error(ctx, GBK_ENOMEM, "build output");
  build_final_offsets<<<grid_for(n + 1), kBlock, 0, s>>>(srt_off, pos, n, off);
  build_compact<<<grid_for(raw_m, kBlock, 16384), kBlock, 0, s>>>(keep, pos, srt_nbr, srt_w, raw_m,
                                                                  nbr, wgt);
  GBK_LAUNCH_CHECK(ctx, "build_compact");
  GBK_HIP(ctx, hipStreamSynchronize(s));
  *m_out = total;
  *off_o = off;
  *nbr_o = nbr;
  *wgt_o = wgt;
  return GBK_OK;
}

}  // namespace gbk

using namespace gbk;

extern "C" int gbk_graph_build(gbk_ctx *ctx, int64_t n, int64_t m_el, const int32_t *u,
                               const int32_t *v, const int32_t *w, int symmetrize, int keep_self,
                               gbk_graph **out, double *device_ms) {
  if (!ctx || !out || m_el < 0 || (m_el > 0 && (!u || !v)))
    return set_error(ctx, GBK_EINVAL, "build: bad arguments");
  *out = nullptr;
  hipStream_t s = ctx->stream;
  DevTemp tmp;
  auto *du = static_cast<int32_t *>(tmp.get(sizeof(int32_t) * m_el));
  auto *dv = static_cast<int32_t *>(tmp.get(sizeof(int32_t) * m_el));
  auto *dw = w ? static_cast<int32_t *>(tmp.get(sizeof(int32_t) * m_el)) : nullptr;
  auto *mx = static_cast<int *>(tmp.get(64));
  if (!du || !dv || (w && !dw) || !mx) return set_error(ctx, GBK_ENOMEM, "build upload");
  GBK_HIP(ctx, hipMemcpyAsync(du, u, sizeof(int32_t) * m_el, hipMemcpyHostToDevice, s));
  GBK_HIP(ctx, hipMemcpyAsync(dv, v, sizeof(int32_t) * m_el, hipMemcpyHostToDevice, s));
  if (w) GBK_HIP(ctx, hipMemcpyAsync(dw, w, sizeof(int32_t) * m_el, hipMemcpyHostToDevice, s));
  EventTimer timer(ctx);
  if (n < 0) {
    // FindMaxNodeID (builder.h:160-171) + 1
    fill<int>(ctx, mx, 0, 2);
    build_max_id<<<grid_for(m_el, kBlock, 4096), kBlock, 0, s>>>(du, dv, m_el, mx, mx + 1);
    int h[2] = {0, 0};
    GBK_HIP(ctx, hipMemcpyAsync(h, mx, sizeof(h), hipMemcpyDeviceToHost, s));
    GBK_HIP(ctx, hipStreamSynchronize(s));
    if (h[1]) return set_error(ctx, GBK_EINVAL, "build: negative vertex id");
    n = static_cast<int64_t>(h[0]) + 1;
  }
  if (n > INT32_MAX) return set_error(ctx, GBK_EINVAL, "build: too many vertices");
  auto *iota = static_cast<int32_t *>(tmp.get(sizeof(int32_t) * (n + 1)));
  if (!iota) return set_error(ctx, GBK_ENOMEM, "build iota");
  build_iota<<<grid_for(n), kBlock, 0, s>>>(iota, n);
  DevTemp outbufs;
  int64_t m_out = 0, m_in = 0;
  int64_t *off = nullptr, *in_off = nullptr;
  int32_t *nbr = nullptr, *wgt = nullptr, *in_nbr = nullptr, *in_wgt = nullptr;
  int rc = build_one(ctx, tmp, du, dv, dw, m_el, n, symmetrize, 0, keep_self, iota, &m_out, &off, &nbr,
                     &wgt, outbufs);
  if (!rc && !symmetrize)
    rc = build_one(ctx, tmp, du, dv, dw, m_el, n, 0, 1, keep_self, iota, &m_in, &in_off, &in_nbr,
                   &in_wgt, outbufs);
  if (rc) return rc;
  // hand the buffers to a graph object
  auto *g = new gbk_graph();
  g->n = n;
  g->m = m_out;
  g->directed = !symmetrize;
  g->weighted = w != nullptr;
  g->off = off;
  g->nbr = nbr;
  g->wgt = wgt;
  if (symmetrize) {
    g->in_off = off;
    g->in_nbr = nbr;
    g->in_wgt = wgt;
  } else {
    g->in_off = in_off;
    g->in_nbr = in_nbr;
    g->in_wgt = in_wgt;
  }
  g->owned = outbufs.ptrs;
  outbufs.ptrs.clear();
  void *org = nullptr;
  if (hipMalloc(&org, sizeof(int32_t) * (n + 1)) != hipSuccess) {
    gbk_graph_free(ctx, g);
    return set_error(ctx, GBK_ENOMEM, "build org_ids");
  }
  g->owned.push_back(org);
  g->org_ids = static_cast<int32_t *>(org);
  build_iota<<<grid_for(n), kBlock, 0, s>>>(g->org_ids, n);
  const double ms = timer.stop_ms();
  rc = gbk_graph_finalize(ctx, g);
  if (rc) {
    gbk_graph_free(ctx, g);
    return rc;
  }
  if (device_ms) *device_ms = ms;
  *out = g;
  return GBK_OK;
}
