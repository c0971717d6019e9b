// gbk_expand.h -- edge-expansion building blocks shared by the kernels.
//
// Two shapes of work occur on this path:
//  * "all rows": every vertex's edge list (PR, CC final link, SV hook).  A
//    wave owns 64 consecutive rows, whose edges are contiguous in the CSR.
//  * "frontier": the rows named by a vertex list (TD-BFS, SSSP, BC).  A
//    wave owns 64 consecutive list entries.
// In both, lanes hold (row, first edge, length); a wave-wide prefix of the
// lengths lets the wave walk the union of the 64 edge lists in 64-edge
// chunks, each lane locating the row of its edge with a 6-step shuffle
// search.  Neighbour reads are therefore coalesced across row boundaries,
// and a wave never idles on short rows.  Rows longer than the heavy
// threshold are taken out of the wave walk and processed by the whole
// workgroup (256 lanes per row) afterwards, so one hub cannot serialise a
// wave.  Functors are called with an 'active' flag by every lane of the
// wave, so they may use wave-aggregated appends.
#pragma once

#include <algorithm>

#include "gbk_common.h"

namespace gbk {

__device__ __forceinline__ int32_t ld_agent(const int32_t *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_agent(int32_t *p, int32_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Lane whose range [excl, excl + len) holds edge e of the union (6-step
// shuffle search over the non-decreasing exclusive prefixes); call in
// uniform control flow.
__device__ __forceinline__ int walk_owner(int64_t excl, int64_t e) {
  int lo = 0;
#pragma unroll
  for (int step = 32; step >= 1; step >>= 1) {
    const int probe = lo + step;
    const int64_t x = __shfl(excl, probe & (kWave - 1));
    if (probe < kWave && x <= e) lo = probe;
  }
  return lo;
}

// Walks the union of the lanes' edge ranges.  f(row, edge_index, active).
template <class F>
__device__ __forceinline__ void wave_walk(int64_t row, int64_t start, int64_t len, F &f) {
  const int lane = lane_id();
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
    const int64_t orow = __shfl(row, lo);
    const int64_t ostart = __shfl(start, lo);
    const int64_t oexcl = __shfl(excl, lo);
    const bool active = e < total;
    f(orow, active ? ostart + (e - oexcl) : 0, active);
  }
}

// Batched walk: each lane takes U edges per trip (U consecutive 64-edge
// chunks of the union), so the functor's batch<U>() can issue U independent
// loads and atomics before it waits on any.  Inactive lanes get edge 0 and
// their row's id, both valid indices whenever the walk runs.
template <int U, class F>
__device__ __forceinline__ void wave_walk_batch(int64_t row, int64_t start, int64_t len, F &f) {
  const int lane = lane_id();
  const int64_t incl = wave_incl_scan(len);
  const int64_t excl = incl - len;
  const int64_t total = __shfl(incl, kWave - 1);
  for (int64_t c = 0; c < total; c += U * kWave) {
    int64_t rows[U], edges[U];
    bool act[U];
#pragma unroll
    for (int j = 0; j < U; j++) {
      const int64_t e = c + j * kWave + lane;
      int lo = 0;
#pragma unroll
      for (int step = 32; step >= 1; step >>= 1) {
        const int probe = lo + step;
        const int64_t x = __shfl(excl, probe & (kWave - 1));
        if (probe < kWave && x <= e) lo = probe;
      }
      act[j] = e < total;
      rows[j] = __shfl(row, lo);
      const int64_t ostart = __shfl(start, lo);
      const int64_t oexcl = __shfl(excl, lo);
      edges[j] = act[j] ? ostart + (e - oexcl) : 0;
    }
    f.template batch<U>(rows, edges, act);
  }
}

// wave_walk_batch that also hands each edge its row's value 'val' (one
// per lane, e.g. the row's distance read when the row was picked), so the
// functor needs no load per edge for it.
template <int U, class F>
__device__ __forceinline__ void wave_walk_batch_v(int64_t row, int64_t start, int64_t len, int64_t val, F &f) {
  const int lane = lane_id();
  const int64_t incl = wave_incl_scan(len);
  const int64_t excl = incl - len;
  const int64_t total = __shfl(incl, kWave - 1);
  for (int64_t c = 0; c < total; c += U * kWave) {
    int64_t rows[U], edges[U], vals[U];
    bool act[U];
#pragma unroll
    for (int j = 0; j < U; j++) {
      const int64_t e = c + j * kWave + lane;
      int lo = 0;
#pragma unroll
      for (int step = 32; step >= 1; step >>= 1) {
        const int probe = lo + step;
        const int64_t x = __shfl(excl, probe & (kWave - 1));
        if (probe < kWave && x <= e) lo = probe;
      }
      act[j] = e < total;
      rows[j] = __shfl(row, lo);
      vals[j] = __shfl(val, lo);
      const int64_t ostart = __shfl(start, lo);
      const int64_t oexcl = __shfl(excl, lo);
      edges[j] = act[j] ? ostart + (e - oexcl) : 0;
    }
    f.template batch<U>(rows, edges, vals, act);
  }
}

// Segmented inclusive sum across a 64-edge chunk whose owners ('seg', one
// per lane, non-decreasing across lanes) form contiguous runs.  Returns the
// run total in the last lane of each run (other lanes hold partial sums);
// 'tail' tells the caller which lanes end a run.  Replaces same-address LDS
// atomics, which serialise when many lanes of one wave share an owner.
template <typename T>
__device__ __forceinline__ T seg_run_sum(T x, int seg, bool &tail) {
  const int lane = lane_id();
#pragma unroll
  for (int o = 1; o < kWave; o <<= 1) {
    const T y = __shfl_up(x, o);
    const int s = __shfl_up(seg, o);
    if (lane >= o && s == seg) x += y;
  }
  const int next = __shfl_down(seg, 1);
  tail = (lane == kWave - 1) || (next != seg);
  return x;
}

// Block-cooperative walk of one heavy row: uniform trip count per wave.
template <class F>
__device__ __forceinline__ void block_walk(int64_t row, int64_t start, int64_t len, F &f) {
  for (int64_t base = 0; base < len; base += blockDim.x) {
    const int64_t e = base + threadIdx.x;
    const bool active = e < len;
    f(row, active ? start + e : 0, active);
  }
}

// Frontier expansion over list[0, count).  Each workgroup handles tiles of
// 256 entries (grid-stride); heavy rows of a tile are queued in LDS and
// walked by the whole workgroup after the light rows.
// rowlen(row, &start, &len) returns the edge range to walk for that row
// (len may be 0 to skip it).
template <class RowLen, class F>
__device__ void expand_list(const int32_t *list, int64_t count, int64_t heavy_deg, RowLen rowlen,
                            F &f) {
  __shared__ int32_t heavy_rows[kBlock];
  __shared__ int64_t heavy_start[kBlock];
  __shared__ int64_t heavy_len[kBlock];
  __shared__ int heavy_n;
  for (int64_t tile = blockIdx.x; tile * kBlock < count; tile += gridDim.x) {
    if (threadIdx.x == 0) heavy_n = 0;
    __syncthreads();
    const int64_t i = tile * kBlock + threadIdx.x;
    int64_t row = 0, start = 0, len = 0;
    if (i < count) {
      row = list[i];
      rowlen(row, start, len);
      if (len > heavy_deg) {
        const int slot = atomicAdd(&heavy_n, 1);
        heavy_rows[slot] = static_cast<int32_t>(row);
        heavy_start[slot] = start;
        heavy_len[slot] = len;
        len = 0;
      }
    }
    wave_walk(row, start, len, f);
    __syncthreads();
    const int nh = heavy_n;
    for (int h = 0; h < nh; h++) block_walk(heavy_rows[h], heavy_start[h], heavy_len[h], f);
    __syncthreads();
  }
}

// Heavy rows found in a frontier, deferred to a second kernel that spreads
// each of them over every workgroup (one hub no longer holds one workgroup
// for hundreds of microseconds while the rest of the GPU idles).
struct HeavyQueue {
  int32_t *row;
  int64_t *start;
  int64_t *len;
  int64_t *prefix;            // first chunk of each queued row (cap)
  unsigned long long *count;  // rows << 32 | chunks queued; reset to 0 before the light kernel
  int64_t cap;
  int *overflow;
  int64_t chunk;              // edges of a row per chunk
};

constexpr int64_t kHeavyChunk = 1024;
constexpr unsigned kHeavyGrid = 512;

// Device storage for a HeavyQueue in scratch slot 'slot': at most
// m / heavy_deg rows can exceed heavy_deg (the context's heavy degree unless
// 'heavy_deg' is given).  Returns false on allocation failure.  The count and
// overflow words are cleared on the stream.
inline bool make_heavy_queue(gbk_ctx *ctx, const gbk_graph *g, int slot, HeavyQueue *hq, int64_t heavy_deg = 0) {
  const int64_t H = heavy_deg > 0 ? heavy_deg : ctx->tune.heavy_degree;
  const int64_t cap = std::min<int64_t>(g->n, g->m / (H > 0 ? H : 1) + 1) + 1;
  // a level's chunk total is counted in 32 bits (chunks are >= 64 edges)
  if (g->m / 64 + cap >= (int64_t{1} << 32)) return false;
  const size_t bytes = 128 + cap * (4 + 8 + 8 + 8) + 64;
  char *p = static_cast<char *>(scratch(ctx, slot, bytes));
  if (!p) return false;
  hq->count = reinterpret_cast<unsigned long long *>(p);
  hq->overflow = reinterpret_cast<int *>(p + 64);
  hq->start = reinterpret_cast<int64_t *>(p + 128);
  hq->len = hq->start + cap;
  hq->prefix = hq->len + cap;
  hq->row = reinterpret_cast<int32_t *>(hq->prefix + cap + 1);
  hq->cap = cap;
  hq->chunk = kHeavyChunk;
  fill<unsigned long long>(ctx, hq->count, 0ull, 1);
  fill<int>(ctx, hq->overflow, 0, 1);
  return true;
}

// Queue a heavy row.  One atomic on 'count' reserves the row's slot (high 32
// bits) together with its chunk range (low 32 bits), so prefix[] ascends in
// queue order and the heavy kernel needs no scan pass first.  Past 'cap' the
// row is dropped and 'overflow' is raised for the host to report.
__device__ __forceinline__ void heavy_push(const HeavyQueue &hq, int32_t row, int64_t start, int64_t len) {
  const auto chunks = static_cast<unsigned long long>((len + hq.chunk - 1) / hq.chunk);
  const unsigned long long old = atomicAdd(hq.count, (1ull << 32) | chunks);
  const auto at = static_cast<int64_t>(old >> 32);
  if (at < hq.cap) {
    hq.row[at] = row;
    hq.start[at] = start;
    hq.len[at] = len;
    hq.prefix[at] = static_cast<int64_t>(old & 0xffffffffull);
  } else {
    *hq.overflow = 1;
  }
}

// Rows stored and chunks they cover (the chunks reserved by dropped rows,
// which follow every stored row's, are left out).
__device__ __forceinline__ void heavy_extent(const HeavyQueue &hq, int64_t &rows, int64_t &total) {
  const unsigned long long w = *hq.count;
  const auto pushed = static_cast<int64_t>(w >> 32);
  rows = min(pushed, hq.cap);
  total = static_cast<int64_t>(w & 0xffffffffull);
  if (pushed > hq.cap) total = hq.prefix[rows - 1] + (hq.len[rows - 1] + hq.chunk - 1) / hq.chunk;
}

// Light part of a frontier expansion: rows above heavy_deg go to 'hq'.
template <class RowLen, class F>
__device__ void expand_list_light(const int32_t *list, int64_t count, int64_t heavy_deg,
                                  RowLen rowlen, F &f, const HeavyQueue &hq) {
  for (int64_t tile = blockIdx.x; tile * kBlock < count; tile += gridDim.x) {
    const int64_t i = tile * kBlock + threadIdx.x;
    int64_t row = 0, start = 0, len = 0;
    if (i < count) {
      row = list[i];
      rowlen(row, start, len);
      if (len > heavy_deg) {
        heavy_push(hq, static_cast<int32_t>(row), start, len);
        len = 0;
      }
    }
    wave_walk(row, start, len, f);
  }
}

// expand_list_light with the batched walk.
template <int U, class RowLen, class F>
__device__ void expand_list_light_batch(const int32_t *list, int64_t count, int64_t heavy_deg,
                                        RowLen rowlen, F &f, const HeavyQueue &hq) {
  for (int64_t tile = blockIdx.x; tile * kBlock < count; tile += gridDim.x) {
    const int64_t i = tile * kBlock + threadIdx.x;
    int64_t row = 0, start = 0, len = 0;
    if (i < count) {
      row = list[i];
      rowlen(row, start, len);
      if (len > heavy_deg) {
        heavy_push(hq, static_cast<int32_t>(row), start, len);
        len = 0;
      }
    }
    wave_walk_batch<U>(row, start, len, f);
  }
}

// Heavy part: the chunks of all queued rows are numbered globally and
// workgroup b takes chunks b, b+G, b+2G, ... (each found by binary search in
// the chunk prefix), so the work spreads evenly however many rows there are.
template <class F>
__device__ void expand_heavy(const HeavyQueue &hq, F &f) {
  int64_t rows, total;
  heavy_extent(hq, rows, total);
  if (rows == 0) return;
  for (int64_t gc = blockIdx.x; gc < total; gc += gridDim.x) {
    int64_t lo = 0, hi = rows;  // last row with prefix[row] <= gc
    while (hi - lo > 1) {
      const int64_t mid = (lo + hi) >> 1;
      if (hq.prefix[mid] <= gc) lo = mid;
      else hi = mid;
    }
    const int64_t len = hq.len[lo], start = hq.start[lo], row = hq.row[lo];
    const int64_t first = (gc - hq.prefix[lo]) * hq.chunk;
    const int64_t last = min(len, first + hq.chunk);
    for (int64_t base = first; base < last; base += blockDim.x) {
      const int64_t e = base + threadIdx.x;
      const bool active = e < last;
      f(row, active ? start + e : 0, active);
    }
  }
}

// expand_heavy with U edges per lane per trip (rows of the batch are the
// same row; edges past the chunk end are inactive and read edge 0).
template <int U, class F>
__device__ void expand_heavy_batch(const HeavyQueue &hq, F &f) {
  int64_t rows, total;
  heavy_extent(hq, rows, total);
  if (rows == 0) return;
  for (int64_t gc = blockIdx.x; gc < total; gc += gridDim.x) {
    int64_t lo = 0, hi = rows;  // last row with prefix[row] <= gc
    while (hi - lo > 1) {
      const int64_t mid = (lo + hi) >> 1;
      if (hq.prefix[mid] <= gc) lo = mid;
      else hi = mid;
    }
    const int64_t len = hq.len[lo], start = hq.start[lo], row = hq.row[lo];
    const int64_t first = (gc - hq.prefix[lo]) * hq.chunk;
    const int64_t last = min(len, first + hq.chunk);
    for (int64_t base = first; base < last; base += U * static_cast<int64_t>(blockDim.x)) {
      int64_t rws[U], edges[U];
      bool act[U];
#pragma unroll
      for (int j = 0; j < U; j++) {
        const int64_t e = base + j * static_cast<int64_t>(blockDim.x) + threadIdx.x;
        act[j] = e < last;
        rws[j] = row;
        edges[j] = act[j] ? start + e : 0;
      }
      f.template batch<U>(rws, edges, act);
    }
  }
}

// expand_heavy_batch handing each edge its row's value, f.row_value(row),
// read once per chunk.
template <int U, class F>
__device__ void expand_heavy_batch_v(const HeavyQueue &hq, F &f) {
  int64_t rows, total;
  heavy_extent(hq, rows, total);
  if (rows == 0) return;
  for (int64_t gc = blockIdx.x; gc < total; gc += gridDim.x) {
    int64_t lo = 0, hi = rows;  // last row with prefix[row] <= gc
    while (hi - lo > 1) {
      const int64_t mid = (lo + hi) >> 1;
      if (hq.prefix[mid] <= gc) lo = mid;
      else hi = mid;
    }
    const int64_t len = hq.len[lo], start = hq.start[lo], row = hq.row[lo];
    const int64_t first = (gc - hq.prefix[lo]) * hq.chunk;
    const int64_t last = min(len, first + hq.chunk);
    const int64_t val = f.row_value(row);
    for (int64_t base = first; base < last; base += U * static_cast<int64_t>(blockDim.x)) {
      int64_t rws[U], edges[U], vals[U];
      bool act[U];
#pragma unroll
      for (int j = 0; j < U; j++) {
        const int64_t e = base + j * static_cast<int64_t>(blockDim.x) + threadIdx.x;
        act[j] = e < last;
        rws[j] = row;
        vals[j] = val;
        edges[j] = act[j] ? start + e : 0;
      }
      f.template batch<U>(rws, edges, vals, act);
    }
  }
}

// ------------------------------------------------------- all rows, balanced
// Work units of an all-rows pass over one CSR, fixed per graph: heavy rows
// (degree above the graph's threshold) cut into kHeavyChunkEdges-edge chunks
// numbered by hprefix, and tiles of at most 64 rows and about 2048 light
// edges.  Workgroups take heavy chunks, then waves take tiles, so the work of
// a unit is bounded whatever the order of the rows (a window of consecutive
// high-degree rows no longer lands on one wave).
struct RowTiles {
  const int64_t *off;
  const int32_t *tiles;    // ntiles + 1 starts
  int64_t ntiles;
  const int32_t *heavy;    // heavy rows
  const int64_t *hprefix;  // nheavy + 1 chunk prefix
  int64_t nheavy;
  int64_t heavy_deg;
};

inline RowTiles row_tiles(const gbk_graph *g, bool in_csr) {
  RowTiles t;
  t.off = in_csr ? g->in_off : g->off;
  t.tiles = in_csr ? g->tiles_in : g->tiles_out;
  t.ntiles = in_csr ? g->ntiles_in : g->ntiles_out;
  t.heavy = in_csr ? g->heavy_in : g->heavy_out;
  t.hprefix = in_csr ? g->hchunk_in : g->hchunk_out;
  t.nheavy = in_csr ? g->n_heavy_in : g->n_heavy_out;
  t.heavy_deg = g->heavy_deg;
  return t;
}

// Heavy chunk gc of the row list: the row's index h in the list and the
// chunk's edge range [first, last) relative to the row's start.
__device__ __forceinline__ void heavy_chunk_of(const RowTiles &rt, int64_t gc, int64_t &h, int64_t &first,
                                               int64_t &last) {
  int64_t lo = 0, hi = rt.nheavy;  // last row with hprefix[row] <= gc
  while (hi - lo > 1) {
    const int64_t mid = (lo + hi) >> 1;
    if (rt.hprefix[mid] <= gc) lo = mid;
    else hi = mid;
  }
  h = lo;
  const int32_t u = rt.heavy[lo];
  const int64_t deg = rt.off[u + 1] - rt.off[u];
  first = (gc - rt.hprefix[lo]) * kHeavyChunkEdges;
  last = min(deg, first + kHeavyChunkEdges);
}

// Every row of the CSR: rowlen(row, &s, &l) gives the edge range to walk
// (a sub-range of the row, or l = 0 to skip it); heavy rows are walked by
// whole workgroups chunk by chunk (positions relative to s, clipped to l),
// light rows by waves tile by tile.
template <class RowLen, class F>
__device__ void expand_rows(const RowTiles &rt, RowLen rowlen, F &f) {
  const int64_t chunks = rt.nheavy ? rt.hprefix[rt.nheavy] : 0;
  for (int64_t gc = blockIdx.x; gc < chunks; gc += gridDim.x) {
    int64_t h, first, last;
    heavy_chunk_of(rt, gc, h, first, last);
    const int64_t u = rt.heavy[h];
    int64_t s, l;
    rowlen(u, s, l);
    const int64_t hi = min(last, l);
    for (int64_t base = first; base < hi; base += blockDim.x) {
      const int64_t e = base + threadIdx.x;
      const bool active = e < hi;
      f(u, active ? s + e : 0, active);
    }
  }
  const int64_t waves = static_cast<int64_t>(gridDim.x) * kWavesPerBlock;
  for (int64_t t = static_cast<int64_t>(blockIdx.x) * kWavesPerBlock + threadIdx.x / kWave; t < rt.ntiles;
       t += waves) {
    const int64_t r = rt.tiles[t] + lane_id();
    int64_t start = 0, len = 0;
    if (r < rt.tiles[t + 1]) {
      const int64_t deg = rt.off[r + 1] - rt.off[r];
      if (deg <= rt.heavy_deg) rowlen(r, start, len);
    }
    wave_walk(r, start, len, f);
  }
}

// expand_rows with the batched walk on the light tiles (the functor's
// batch<U>() gets U edges per lane per trip; heavy chunks call it edge by
// edge as expand_rows does).
template <int U, class RowLen, class F>
__device__ void expand_rows_batch(const RowTiles &rt, RowLen rowlen, F &f) {
  const int64_t chunks = rt.nheavy ? rt.hprefix[rt.nheavy] : 0;
  for (int64_t gc = blockIdx.x; gc < chunks; gc += gridDim.x) {
    int64_t h, first, last;
    heavy_chunk_of(rt, gc, h, first, last);
    const int64_t u = rt.heavy[h];
    int64_t s, l;
    rowlen(u, s, l);
    const int64_t hi = min(last, l);
    for (int64_t base = first; base < hi; base += blockDim.x) {
      const int64_t e = base + threadIdx.x;
      const bool active = e < hi;
      f(u, active ? s + e : 0, active);
    }
  }
  const int64_t waves = static_cast<int64_t>(gridDim.x) * kWavesPerBlock;
  for (int64_t t = static_cast<int64_t>(blockIdx.x) * kWavesPerBlock + threadIdx.x / kWave; t < rt.ntiles;
       t += waves) {
    const int64_t r = rt.tiles[t] + lane_id();
    int64_t start = 0, len = 0;
    if (r < rt.tiles[t + 1]) {
      const int64_t deg = rt.off[r + 1] - rt.off[r];
      if (deg <= rt.heavy_deg) rowlen(r, start, len);
    }
    wave_walk_batch<U>(r < rt.tiles[t + 1] ? r : rt.tiles[t], start, len, f);
  }
}

// Same over all rows [0, n) (row = list index).
template <class RowLen, class F>
__device__ void expand_all(int64_t n, int64_t heavy_deg, RowLen rowlen, F &f) {
  __shared__ int32_t heavy_rows[kBlock];
  __shared__ int64_t heavy_start[kBlock];
  __shared__ int64_t heavy_len[kBlock];
  __shared__ int heavy_n;
  for (int64_t tile = blockIdx.x; tile * kBlock < n; tile += gridDim.x) {
    if (threadIdx.x == 0) heavy_n = 0;
    __syncthreads();
    const int64_t row = tile * kBlock + threadIdx.x;
    int64_t start = 0, len = 0;
    if (row < n) {
      rowlen(row, start, len);
      if (len > heavy_deg) {
        const int slot = atomicAdd(&heavy_n, 1);
        heavy_rows[slot] = static_cast<int32_t>(row);
        heavy_start[slot] = start;
        heavy_len[slot] = len;
        len = 0;
      }
    }
    wave_walk(row, start, len, f);
    __syncthreads();
    const int nh = heavy_n;
    for (int h = 0; h < nh; h++) block_walk(heavy_rows[h], heavy_start[h], heavy_len[h], f);
    __syncthreads();
  }
}

// Number of workgroups for a grid-stride expansion over 'items' entries.
// Capped so that per-workgroup epilogues (queue flushes, partial sums) stay
// few: same-address global atomics serialise in L2 at tens of ns each.
constexpr int64_t kGridCap = 2048;
inline unsigned expand_grid(int64_t items) {
  int64_t g = (items + kBlock - 1) / kBlock;
  if (g < 1) g = 1;
  if (g > kGridCap) g = kGridCap;
  return static_cast<unsigned>(g);
}

// Number of workgroups for a balanced all-rows pass.
inline unsigned rows_grid(const RowTiles &rt) {
  const int64_t units = rt.ntiles / kWavesPerBlock + 1;
  return static_cast<unsigned>(units > kGridCap ? kGridCap : units);
}

// Per-wave queue staged in LDS.  push() is called by every lane of the wave
// (uniform control flow); entries gather in the wave's LDS segment and go to
// the global queue in batches of up to kSeg with one atomic per batch, as
// contiguous coalesced stores.  flush() must be called once at the end.
constexpr int kSeg = 512;

struct WaveQueue {
  int32_t *lds;                 // this wave's kSeg-entry segment
  int cnt;                      // entries staged (wave-uniform)
  unsigned long long *tail;     // global tail counter
  int32_t *out;                 // global queue
  int64_t cap;                  // capacity of 'out'
  int *overflow;
  int segcap;                   // entries of the LDS segment (<= kSeg)

  __device__ void init(int32_t *seg, unsigned long long *t, int32_t *o, int64_t c, int *ovf,
                       int seg_entries = kSeg) {
    lds = seg;
    cnt = 0;
    tail = t;
    out = o;
    cap = c;
    overflow = ovf;
    segcap = seg_entries;
  }
  __device__ void flush() {
    if (cnt == 0) return;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    unsigned long long base = 0;
    if (lane_id() == 0) base = atomicAdd(tail, static_cast<unsigned long long>(cnt));
    base = __shfl(base, 0);
    for (int i = lane_id(); i < cnt; i += kWave) {
      const int64_t at = static_cast<int64_t>(base) + i;
      if (at < cap) out[at] = lds[i];
      else *overflow = 1;
    }
    __builtin_amdgcn_wave_barrier();
    cnt = 0;
  }
  __device__ void push(bool pred, int32_t v) {
    const uint64_t mask = __ballot(pred);
    if (mask == 0) return;
    const int k = __popcll(mask);
    if (cnt + k > segcap) flush();
    if (pred) lds[cnt + lanemask_lt_count(mask)] = v;
    cnt += k;
  }
  // Final flush of a whole workgroup, called by every thread: one tail atomic
  // per workgroup instead of one per wave (same-address atomics serialize,
  // and a resident grid ends with thousands of waves flushing at once).
  __device__ void flush_block() {
    __shared__ int wcnt[kWavesPerBlock];
    __shared__ unsigned long long bbase;
    const int w = threadIdx.x / kWave;
    __syncthreads();
    if (lane_id() == 0) wcnt[w] = cnt;
    __syncthreads();
    if (threadIdx.x == 0) {
      int tot = 0;
      for (int v = 0; v < kWavesPerBlock; v++) tot += wcnt[v];
      bbase = tot ? atomicAdd(tail, static_cast<unsigned long long>(tot)) : 0ull;
    }
    __syncthreads();
    unsigned long long base = bbase;
    for (int v = 0; v < w; v++) base += static_cast<unsigned long long>(wcnt[v]);
    for (int i = lane_id(); i < cnt; i += kWave) {
      const int64_t at = static_cast<int64_t>(base) + i;
      if (at < cap) out[at] = lds[i];
      else *overflow = 1;
    }
    cnt = 0;
  }
};

// Spread counters.  Atomics on one address from thousands of workgroups
// serialize (about 11 ns each on this GPU, so a resident grid's end-of-kernel
// adds cost tens of microseconds); a spread counter is kSpread slots
// kSpreadStride 8-byte words (256 bytes) apart, workgroup b adding into slot
// b mod kSpread, summed by one wave afterwards.
constexpr int kSpread = 32;
constexpr int kSpreadStride = 32;
constexpr size_t kSpreadBytes = sizeof(unsigned long long) * kSpread * kSpreadStride;

// Adds every lane's v into this workgroup's slot (every thread calls it).
__device__ __forceinline__ void block_add_spread(unsigned long long v, unsigned long long *slots) {
  __shared__ unsigned long long red[kWavesPerBlock];
  v = block_sum(v, red);
  if (threadIdx.x == 0 && v) atomicAdd(&slots[(blockIdx.x % kSpread) * kSpreadStride], v);
}

// Called by all lanes of one wave (in a kernel after the adding ones): the
// slots' total, and the slots cleared.
__device__ __forceinline__ unsigned long long wave_take_spread(unsigned long long *slots) {
  unsigned long long x = 0;
  const int l = lane_id();
  if (l < kSpread) {
    x = slots[l * kSpreadStride];
    slots[l * kSpreadStride] = 0;
  }
  return wave_sum(x);
}

// Adds every lane's v into *total with one atomic per workgroup (every
// thread calls it).
__device__ __forceinline__ void block_add(unsigned long long v, unsigned long long *total) {
  __shared__ unsigned long long red[kWavesPerBlock];
  v = block_sum(v, red);
  if (threadIdx.x == 0 && v) atomicAdd(total, v);
}

// Block-wide partial sum written to partial[blockIdx.x]; a one-workgroup
// finalize kernel adds the partials (no same-address atomics).
template <typename T>
__device__ __forceinline__ void block_partial(T v, T *partial) {
  __shared__ T red[kWavesPerBlock];
  const T s = block_sum(v, red);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

template <typename T>
__global__ void sum_partials(const T *partial, int count, T *out, int accumulate) {
  __shared__ T red[1024 / kWave];
  T s = 0;
  for (int i = threadIdx.x; i < count; i += blockDim.x) s += partial[i];
  s = wave_sum(s);
  const int w = threadIdx.x / kWave;
  if (lane_id() == 0) red[w] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    T t = 0;
    for (int i = 0; i < static_cast<int>(blockDim.x) / kWave; i++) t += red[i];
    *out = accumulate ? *out + t : t;
  }
}

// "Last workgroup done" helper: returns true in exactly one thread (thread 0
// of the last workgroup to arrive) after every workgroup's prior writes are
// visible.  The counter is reset by that thread.
__device__ __forceinline__ bool last_block_done(unsigned int *counter) {
  __shared__ bool is_last;
  __syncthreads();
  if (threadIdx.x == 0) {
    __threadfence();
    const unsigned prev = atomicAdd(counter, 1u);
    is_last = (prev == gridDim.x - 1);
    if (is_last) {
      __threadfence();
      *counter = 0;
    }
  }
  __syncthreads();
  return is_last && threadIdx.x == 0;
}

// BC normalisation (bc.cc:224-230): max over non-negative floats via their
// bit patterns (ordered like the values), then every score divided by it.
static __global__ void bc_max(const float *scores, int64_t n, unsigned int *partial) {
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  unsigned int x = 0;
  for (int64_t v = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; v < n; v += stride)
    x = max(x, __float_as_uint(scores[v]));
  x = wave_max(x);
  __shared__ unsigned int red[kWavesPerBlock];
  if (lane_id() == 0) red[threadIdx.x / kWave] = x;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int i = 0; i < kWavesPerBlock; i++) x = max(x, red[i]);
    partial[blockIdx.x] = x;
  }
}

static __global__ void bc_normalize(float *scores, int64_t n, const unsigned int *partial, int count) {
  __shared__ float biggest;
  if (threadIdx.x == 0) {
    unsigned int x = 0;
    for (int i = 0; i < count; i++) x = max(x, partial[i]);
    biggest = __uint_as_float(x);
  }
  __syncthreads();
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t v = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; v < n; v += stride)
    scores[v] = scores[v] / biggest;
}

}  // namespace gbk
