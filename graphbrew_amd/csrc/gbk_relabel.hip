// gbk_relabel.hip -- RelabelByMapping on the GPU (gapbs/builder.h:1046-1206).
//
// Reference: for every old row u, degrees[new_ids[u]] = deg(u); prefix sum;
// the old row is copied to row new_ids[u] with every neighbour remapped
// through new_ids; each new row is sorted by (neighbour, weight).
//
// Here: (1) one pass inverts the mapping (inv[new] = old, which also proves
// it is a permutation), writes the new degrees and composes org_ids;
// (2) device exclusive scan -> new offsets; (3) rows are binned by degree
// and each bin is gathered, remapped and sorted in one fused kernel, so the
// neighbour array is read once and the relabelled array written once:
//   deg <= 64    a group of 8/16/32/64 lanes per row, one entry per lane,
//                bitonic network over shuffles (registers only)
//   deg <= 512   one wave per row, 2/4/8 entries per lane in registers,
//                bitonic network (in-lane strides in registers, the others
//                over shuffles; no LDS, no barriers)
//   deg <= 2048  one row per workgroup, bitonic sort in LDS
//   larger       2048-element chunks sorted in LDS, then merge-path passes
//                over a compact staging buffer, then copied into place.
// Keys compare as (neighbour, weight), the order of NodeWeight::operator<
// (gapbs/graph.h:136-140).
#include <algorithm>

#include "gbk_expand.h"

int gbk_graph_alloc(gbk_ctx *ctx, int64_t n, int64_t m, int64_t m_in, int directed, int weighted,
                    gbk_graph **out);
int gbk_graph_finalize(gbk_ctx *ctx, gbk_graph *g);

namespace gbk {

constexpr int kSortCap = 2048;  // elements one wave sorts in registers (32 a lane)

__global__ void relabel_invert(const int64_t *off, const int32_t *new_ids, const int32_t *org_old,
                               int64_t n, int32_t *inv, int64_t *deg_new, int32_t *org_new,
                               int32_t *err) {
  const int64_t u = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (u >= n) return;
  const int32_t r = new_ids[u];
  if (r < 0 || r >= n) {
    atomicOr(err, 1);
    return;
  }
  if (atomicCAS(&inv[r], -1, static_cast<int32_t>(u)) != -1) {
    atomicOr(err, 2);
    return;
  }
  deg_new[r] = off[u + 1] - off[u];
  org_new[r] = org_old[u];
}

// Degree classes: <=8, <=16, <=32, <=64 (lane-group register sorts), <=128,
// <=256, <=512, <=1024, <=2048 (wave register sorts), larger (chunk sort +
// merge).  Empty rows need no work.  The rows with entries are listed class
// by class in one array, ascending within a class: a count per class and
// tile of rows, one scan of those counts, and a scatter that ranks rows by
// ballots (no global atomics; per-wave queues flushing into one counter line
// took 0.7 ms for 4M rows).
constexpr int kClasses = 10;
constexpr int kClassBig = 9;
struct RelabelLists {
  int32_t *l[kClasses];
};

__device__ __forceinline__ int bin_class(int64_t d) {
  return d == 0 ? -1 : d <= 8 ? 0 : d <= 16 ? 1 : d <= 32 ? 2 : d <= 64 ? 3 : d <= 128 ? 4
       : d <= 256 ? 5 : d <= 512 ? 6 : d <= 1024 ? 7 : d <= kSortCap ? 8 : 9;
}

// Workgroup t counts the rows of each class in rows [t * tile, (t+1) * tile):
// cnt[c * gridDim.x + t].
__global__ void __launch_bounds__(kBlock)
relabel_bin_count(const int64_t *deg_new, int64_t n, int64_t tile, uint32_t *cnt) {
  __shared__ uint32_t tot[kClasses];
  if (threadIdx.x < kClasses) tot[threadIdx.x] = 0;
  __syncthreads();
  const int64_t lo = static_cast<int64_t>(blockIdx.x) * tile;
  const int64_t hi = min(n, lo + tile);
  uint32_t mine[kClasses] = {};
  for (int64_t r = lo + threadIdx.x; r < hi; r += kBlock) {
    const int cls = bin_class(deg_new[r]);
#pragma unroll
    for (int c = 0; c < kClasses; c++) mine[c] += cls == c;
  }
#pragma unroll
  for (int c = 0; c < kClasses; c++) {
    const uint32_t w = wave_sum(mine[c]);
    if (lane_id() == 0 && w) atomicAdd(&tot[c], w);
  }
  __syncthreads();
  if (threadIdx.x < kClasses) cnt[static_cast<int64_t>(threadIdx.x) * gridDim.x + blockIdx.x] = tot[threadIdx.x];
}

// One workgroup of 1024 lanes: cnt (class-major, ntiles per class) becomes
// its exclusive scan, i.e. each (class, tile)'s first slot in the list;
// counts[c] = rows of class c.
__global__ void __launch_bounds__(1024)
relabel_bin_offsets(uint32_t *cnt, int64_t ntiles, unsigned long long *counts) {
  __shared__ uint32_t wsum[1024 / kWave];
  __shared__ uint32_t start[kClasses + 1];
  const int64_t len = ntiles * kClasses;
  const int t = threadIdx.x;
  const int64_t per = (len + 1023) / 1024;
  const int64_t b = min(len, t * per), e = min(len, b + per);
  uint32_t sum = 0;
  for (int64_t i = b; i < e; i++) sum += cnt[i];
  const uint32_t incl = wave_incl_scan(sum);
  const int w = t / kWave;
  if (lane_id() == kWave - 1) wsum[w] = incl;
  __syncthreads();
  if (w == 0) {
    uint32_t x = lane_id() < 1024 / kWave ? wsum[lane_id()] : 0;
    x = wave_incl_scan(x);
    if (lane_id() < 1024 / kWave) wsum[lane_id()] = x;
  }
  __syncthreads();
  uint32_t run = incl - sum + (w > 0 ? wsum[w - 1] : 0);
  for (int64_t i = b; i < e; i++) {
    if (i % ntiles == 0) start[i / ntiles] = run;
    const uint32_t v = cnt[i];
    cnt[i] = run;
    run += v;
  }
  if (t == 1023) start[kClasses] = run;
  __syncthreads();
  if (t < kClasses) counts[t] = start[t + 1] - start[t];
}

// Workgroup t writes the rows of its tile to their class's slots in row
// order: a wave's rank from its ballot, earlier waves' counts from LDS.
__global__ void __launch_bounds__(kBlock)
relabel_bin_scatter(const int64_t *deg_new, int64_t n, int64_t tile, const uint32_t *first, int32_t *list) {
  __shared__ uint32_t base[kClasses];
  __shared__ uint32_t wc[kWavesPerBlock][kClasses];
  if (threadIdx.x < kClasses) base[threadIdx.x] = first[static_cast<int64_t>(threadIdx.x) * gridDim.x + blockIdx.x];
  const int w = threadIdx.x / kWave;
  const int64_t lo = static_cast<int64_t>(blockIdx.x) * tile;
  const int64_t hi = min(n, lo + tile);
  for (int64_t r0 = lo; r0 < hi; r0 += kBlock) {
    const int64_t r = r0 + threadIdx.x;
    const int cls = r < hi ? bin_class(deg_new[r]) : -1;
    int rank = 0;
#pragma unroll
    for (int c = 0; c < kClasses; c++) {
      const uint64_t m = __ballot(cls == c);
      if (lane_id() == 0) wc[w][c] = static_cast<uint32_t>(__popcll(m));
      if (cls == c) rank = lanemask_lt_count(m);
    }
    __syncthreads();
    if (cls >= 0) {
      uint32_t pos = base[cls] + static_cast<uint32_t>(rank);
      for (int v = 0; v < w; v++) pos += wc[v][cls];
      list[pos] = static_cast<int32_t>(r);
    }
    __syncthreads();
    if (threadIdx.x < kClasses) {
      uint32_t add = 0;
      for (int v = 0; v < kWavesPerBlock; v++) add += wc[v][threadIdx.x];
      base[threadIdx.x] += add;
    }
    __syncthreads();
  }
}

// Rows of at most G entries: G lanes hold one entry each and run a bitonic
// network over shuffles (no LDS, no barriers).  Keys compare as (v, w).
template <bool W, int G>
__global__ void __launch_bounds__(kBlock)
relabel_group_sort(const int32_t *rows, int64_t nrows, const int32_t *inv, const int64_t *off,
                   const int32_t *nbr, const int32_t *wgt, const int32_t *new_ids,
                   const int64_t *new_off, int32_t *nbr_out, int32_t *wgt_out) {
  using K = typename std::conditional<W, uint64_t, uint32_t>::type;
  const int glane = threadIdx.x % G;
  const int64_t ri = static_cast<int64_t>(blockIdx.x) * (kBlock / G) + threadIdx.x / G;
  int64_t len = 0, src = 0, dst = 0;
  if (ri < nrows) {
    const int32_t r = rows[ri];
    const int32_t u = inv[r];
    src = off[u];
    len = off[u + 1] - src;
    dst = new_off[r];
  }
  K key = static_cast<K>(~static_cast<K>(0));
  if (glane < len) {
    const int32_t v = new_ids[nbr[src + glane]];
    if constexpr (W)
      key = (static_cast<uint64_t>(static_cast<uint32_t>(v)) << 32) |
            (static_cast<uint32_t>(wgt[src + glane]) ^ 0x80000000u);
    else
      key = static_cast<uint32_t>(v);
  }
#pragma unroll
  for (int k = 2; k <= G; k <<= 1) {
#pragma unroll
    for (int j = k >> 1; j > 0; j >>= 1) {
      const K other = __shfl_xor(key, j);
      const bool up = (glane & k) == 0;
      const bool lower = (glane & j) == 0;
      key = (lower == up) ? (other < key ? other : key) : (other > key ? other : key);
    }
  }
  if (glane < len) {
    if constexpr (W) {
      nbr_out[dst + glane] = static_cast<int32_t>(key >> 32);
      wgt_out[dst + glane] = static_cast<int32_t>(static_cast<uint32_t>(key) ^ 0x80000000u);
    } else {
      nbr_out[dst + glane] = static_cast<int32_t>(key);
    }
  }
}

// Rows of at most 64*E entries: one wave per row, E entries per lane in
// registers.  Entry i sits in lane i % 64, register i / 64 (so loads and
// stores are coalesced); the bitonic network's strides below 64 pair lanes
// (shuffles), larger strides pair registers of one lane.  Keys compare as
// (v, w) through the composite integer of RowKey.
template <bool W>
struct RowKey {
  using T = typename std::conditional<W, uint64_t, uint32_t>::type;
  __device__ static T make(int32_t v, int32_t w) {
    if constexpr (W)
      return (static_cast<uint64_t>(static_cast<uint32_t>(v)) << 32) |
             (static_cast<uint32_t>(w) ^ 0x80000000u);
    else
      return static_cast<uint32_t>(v);
  }
};

// One wave sorts up to 64*E keys: load(i) gives key i (i < len), store(i,
// key) takes the i-th smallest.  Padding keys are all ones.
template <bool W, int E, class Load, class Store>
__device__ __forceinline__ void wave_sort(int64_t len, Load load, Store store) {
  using K = typename RowKey<W>::T;
  constexpr int N = kWave * E;
  const int lane = lane_id();
  K key[E];
#pragma unroll
  for (int e = 0; e < E; e++) {
    const int64_t i = static_cast<int64_t>(e) * kWave + lane;
    key[e] = i < len ? load(i) : static_cast<K>(~static_cast<K>(0));
  }
#pragma unroll
  for (int k = 2; k <= N; k <<= 1) {
#pragma unroll
    for (int j = k >> 1; j > 0; j >>= 1) {
      if (j < kWave) {
#pragma unroll
        for (int e = 0; e < E; e++) {
          const K other = __shfl_xor(key[e], j);
          const int i = e * kWave + lane;
          const bool up = (i & k) == 0;
          const bool lower = (i & j) == 0;
          key[e] = (lower == up) ? (other < key[e] ? other : key[e]) : (other > key[e] ? other : key[e]);
        }
      } else {
        const int jj = j / kWave;
#pragma unroll
        for (int e = 0; e < E; e++) {
          if ((e & jj) == 0) {
            const int e2 = e | jj;
            const int i = e * kWave + lane;
            const bool up = (i & k) == 0;
            const K a = key[e], b = key[e2];
            if ((a > b) == up) {
              key[e] = b;
              key[e2] = a;
            }
          }
        }
      }
    }
  }
#pragma unroll
  for (int e = 0; e < E; e++) {
    const int64_t i = static_cast<int64_t>(e) * kWave + lane;
    if (i < len) store(i, key[e]);
  }
}

template <bool W, int E>
__global__ void __launch_bounds__(kBlock)
relabel_wave_sort(const int32_t *rows, int64_t nrows, const int32_t *inv, const int64_t *off,
                  const int32_t *nbr, const int32_t *wgt, const int32_t *new_ids,
                  const int64_t *new_off, int32_t *nbr_out, int32_t *wgt_out) {
  using K = typename RowKey<W>::T;
  const int64_t ri = static_cast<int64_t>(blockIdx.x) * kWavesPerBlock + threadIdx.x / kWave;
  if (ri >= nrows) return;  // wave-uniform
  const int32_t r = rows[ri];
  const int32_t u = inv[r];
  const int64_t src = off[u];
  const int64_t len = off[u + 1] - src;
  const int64_t dst = new_off[r];
  wave_sort<W, E>(
      len, [&](int64_t i) { return RowKey<W>::make(new_ids[nbr[src + i]], W ? wgt[src + i] : 0); },
      [&](int64_t i, K key) {
        if constexpr (W) {
          nbr_out[dst + i] = static_cast<int32_t>(key >> 32);
          wgt_out[dst + i] = static_cast<int32_t>(static_cast<uint32_t>(key) ^ 0x80000000u);
        } else {
          nbr_out[dst + i] = static_cast<int32_t>(key);
        }
      });
}

// Big rows: composite key so a plain integer compare gives (v, w) order.
template <bool W>
struct BigKey {
  using T = typename std::conditional<W, uint64_t, uint32_t>::type;
  __device__ static T make(int32_t v, int32_t w) {
    if constexpr (W)
      return (static_cast<uint64_t>(static_cast<uint32_t>(v)) << 32) |
             (static_cast<uint32_t>(w) ^ 0x80000000u);
    else
      return static_cast<uint32_t>(v);
  }
  __device__ static int32_t v(T k) {
    if constexpr (W) return static_cast<int32_t>(k >> 32);
    else return static_cast<int32_t>(k);
  }
  __device__ static int32_t w(T k) {
    if constexpr (W) return static_cast<int32_t>(static_cast<uint32_t>(k) ^ 0x80000000u);
    else return 1;
  }
};

// Chunk sort: chunk c covers elements [cstart, cstart+len) of big row 'row'
// (staging coordinates: row i of the big list starts at stage_off[i]); one
// wave sorts a chunk of up to kSortCap keys in registers.
template <bool W>
__global__ void __launch_bounds__(kBlock)
relabel_big_chunks(const int32_t *big_rows, const int64_t *stage_off, const int32_t *chunk_row,
                   const int64_t *chunk_start, int64_t nchunks, const int32_t *inv, const int64_t *off,
                   const int32_t *nbr, const int32_t *wgt, const int32_t *new_ids,
                   typename BigKey<W>::T *stage) {
  using K = typename BigKey<W>::T;
  const int64_t c = static_cast<int64_t>(blockIdx.x) * kWavesPerBlock + threadIdx.x / kWave;
  if (c >= nchunks) return;  // wave-uniform
  const int bi = chunk_row[c];
  const int64_t cs = chunk_start[c];
  const int32_t u = inv[big_rows[bi]];
  const int64_t first = off[u] + cs;
  const int64_t len = min(static_cast<int64_t>(kSortCap), off[u + 1] - first);
  K *out = stage + stage_off[bi] + cs;
  wave_sort<W, kSortCap / kWave>(
      len, [&](int64_t i) { return BigKey<W>::make(new_ids[nbr[first + i]], W ? wgt[first + i] : 1); },
      [&](int64_t i, K key) { out[i] = key; });
}

// One merge pass of run width 'width': element t of the staging layout is
// produced by merging its pair of runs inside its row.  Each thread makes
// kMergeItems consecutive outputs, re-searching the merge path only when it
// starts or crosses a pair boundary.
constexpr int kMergeItems = 8;

template <typename K>
__global__ void __launch_bounds__(kBlock)
relabel_merge_pass(const K *src, K *dst, const int64_t *stage_off, int nbig, int64_t total,
                   int64_t width) {
  // grid-stride over groups of kMergeItems outputs (the staged rows can hold
  // more items than one launch has threads times kMergeItems)
  const int64_t step = static_cast<int64_t>(gridDim.x) * blockDim.x * kMergeItems;
  for (int64_t t0 = (static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x) * kMergeItems; t0 < total;
       t0 += step) {
  // row containing t0
  int lo = 0, hi = nbig;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (stage_off[mid] <= t0) lo = mid;
    else hi = mid;
  }
  int row = lo;
  int64_t t = t0;
  const int64_t tend = min(t0 + kMergeItems, total);
  while (t < tend) {
    while (stage_off[row + 1] <= t) row++;
    const int64_t rs = stage_off[row], rlen = stage_off[row + 1] - rs;
    const int64_t o = t - rs;
    const int64_t pb = (o / (2 * width)) * (2 * width);
    const int64_t la = min(width, rlen - pb);
    const int64_t lb = max<int64_t>(0, min(width, rlen - pb - width));
    const K *A = src + rs + pb;
    const K *Bp = A + la;
    const int64_t diag = o - pb;
    int64_t b = max<int64_t>(0, diag - lb), e = min(diag, la);
    while (b < e) {
      const int64_t mid = (b + e) >> 1;
      if (!(Bp[diag - 1 - mid] < A[mid])) b = mid + 1;
      else e = mid;
    }
    int64_t ia = b, ib = diag - b;
    const int64_t pend = rs + pb + la + lb;  // end of this pair (staging coords)
    while (t < tend && t < pend) {
      K x;
      if (ia < la && (ib >= lb || !(Bp[ib] < A[ia]))) x = A[ia++];
      else x = Bp[ib++];
      dst[t++] = x;
    }
  }
  }
}

template <bool W>
__global__ void relabel_big_scatter(const typename BigKey<W>::T *stage0, const typename BigKey<W>::T *stage1,
                                    const int64_t *stage_off, const int32_t *big_rows, int nbig,
                                    int64_t total, const int64_t *new_off, int32_t *nbr_out,
                                    int32_t *wgt_out) {
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; t < total; t += stride) {
  int lo = 0, hi = nbig;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (stage_off[mid] <= t) lo = mid;
    else hi = mid;
  }
  // a row of length L took one merge pass per width kSortCap * 2^j below L,
  // ping-ponging from stage0, so its sorted data ends in the buffer of that
  // count's parity
  const int64_t len = stage_off[lo + 1] - stage_off[lo];
  int passes = 0;
  for (int64_t w = kSortCap; w < len; w *= 2) passes++;
  const typename BigKey<W>::T key = (passes & 1) ? stage1[t] : stage0[t];
  const int64_t dst = new_off[big_rows[lo]] + (t - stage_off[lo]);
  nbr_out[dst] = BigKey<W>::v(key);
  if (W) wgt_out[dst] = BigKey<W>::w(key);
  }
}

__global__ void gather_i64(const int64_t *src, const int32_t *idx, int64_t n, int64_t *out) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) out[i] = src[idx[i]];
}

// Relabels one CSR (out or in) of 'g' into (off_out, nbr_out, wgt_out).
template <bool W>
static int relabel_csr(gbk_ctx *ctx, int64_t n, const int64_t *off, const int32_t *nbr,
                       const int32_t *wgt, const int32_t *new_ids_d, const int32_t *inv,
                       int64_t *deg_new, int64_t *off_out, int32_t *nbr_out, int32_t *wgt_out,
                       bool degrees_ready) {
  hipStream_t s = ctx->stream;
  if (!degrees_ready) return set_error(ctx, GBK_EINVAL, "relabel: degrees not computed");
  int rc = exclusive_scan_i64(ctx, deg_new, off_out, n);
  if (rc) return rc;
  RelabelLists bins;
  int32_t *const *lists = bins.l;
  // tiles of >= 4096 rows, at most 1024 of them
  const int64_t tile = std::max<int64_t>(4096, ((n + 1023) / 1024 + kBlock - 1) / kBlock * kBlock);
  const int64_t ntiles = (n + tile - 1) / tile;
  auto *list = static_cast<int32_t *>(scratch(ctx, kS4, sizeof(int32_t) * (n + 1)));
  auto *tile_cnt = static_cast<uint32_t *>(scratch(ctx, kS5, sizeof(uint32_t) * kClasses * (ntiles + 1)));
  auto *counts_d = static_cast<unsigned long long *>(scratch(ctx, kS8, 128));
  if (!list || !tile_cnt || !counts_d) return set_error(ctx, GBK_ENOMEM, "relabel bins");
  fill<unsigned long long>(ctx, counts_d, 0ull, kClasses);
  if (n > 0) {
    relabel_bin_count<<<static_cast<unsigned>(ntiles), kBlock, 0, s>>>(deg_new, n, tile, tile_cnt);
    relabel_bin_offsets<<<1, 1024, 0, s>>>(tile_cnt, ntiles, counts_d);
    relabel_bin_scatter<<<static_cast<unsigned>(ntiles), kBlock, 0, s>>>(deg_new, n, tile, tile_cnt, list);
  }
  GBK_LAUNCH_CHECK(ctx, "relabel bins");
  unsigned long long counts[kClasses];
  GBK_HIP(ctx, hipMemcpyAsync(counts, counts_d, sizeof(counts), hipMemcpyDeviceToHost, s));
  GBK_HIP(ctx, hipStreamSynchronize(s));
  {
    int64_t at = 0;
    for (int c = 0; c < kClasses; c++) {
      bins.l[c] = list + at;
      at += static_cast<int64_t>(counts[c]);
    }
  }
  // rows of <= 64 entries: lane groups sort in registers
  const int groups[4] = {8, 16, 32, 64};
  for (int c = 0; c < 4; c++) {
    const int64_t cnt = static_cast<int64_t>(counts[c]);
    if (cnt == 0) continue;
    const int64_t blocks = (cnt + kBlock / groups[c] - 1) / (kBlock / groups[c]);
    switch (groups[c]) {
      case 8: relabel_group_sort<W, 8><<<blocks, kBlock, 0, s>>>(lists[c], cnt, inv, off, nbr, wgt, new_ids_d, off_out, nbr_out, wgt_out); break;
      case 16: relabel_group_sort<W, 16><<<blocks, kBlock, 0, s>>>(lists[c], cnt, inv, off, nbr, wgt, new_ids_d, off_out, nbr_out, wgt_out); break;
      case 32: relabel_group_sort<W, 32><<<blocks, kBlock, 0, s>>>(lists[c], cnt, inv, off, nbr, wgt, new_ids_d, off_out, nbr_out, wgt_out); break;
      default: relabel_group_sort<W, 64><<<blocks, kBlock, 0, s>>>(lists[c], cnt, inv, off, nbr, wgt, new_ids_d, off_out, nbr_out, wgt_out); break;
    }
    GBK_LAUNCH_CHECK(ctx, "relabel_group_sort");
  }
  // rows of <= 2048 entries: one wave each, in registers
  for (int c = 4; c < kClassBig; c++) {
    const int64_t cnt = static_cast<int64_t>(counts[c]);
    if (cnt == 0) continue;
    const unsigned blocks = static_cast<unsigned>((cnt + kWavesPerBlock - 1) / kWavesPerBlock);
    switch (c) {
      case 4: relabel_wave_sort<W, 2><<<blocks, kBlock, 0, s>>>(lists[c], cnt, inv, off, nbr, wgt, new_ids_d, off_out, nbr_out, wgt_out); break;
      case 5: relabel_wave_sort<W, 4><<<blocks, kBlock, 0, s>>>(lists[c], cnt, inv, off, nbr, wgt, new_ids_d, off_out, nbr_out, wgt_out); break;
      case 6: relabel_wave_sort<W, 8><<<blocks, kBlock, 0, s>>>(lists[c], cnt, inv, off, nbr, wgt, new_ids_d, off_out, nbr_out, wgt_out); break;
      case 7: relabel_wave_sort<W, 16><<<blocks, kBlock, 0, s>>>(lists[c], cnt, inv, off, nbr, wgt, new_ids_d, off_out, nbr_out, wgt_out); break;
      default: relabel_wave_sort<W, 32><<<blocks, kBlock, 0, s>>>(lists[c], cnt, inv, off, nbr, wgt, new_ids_d, off_out, nbr_out, wgt_out); break;
    }
    GBK_LAUNCH_CHECK(ctx, "relabel_wave_sort");
  }
  const int nbig = static_cast<int>(counts[kClassBig]);
  int32_t *big_list = lists[kClassBig];
  if (nbig == 0) return GBK_OK;
  // Big rows: host builds the staging layout and chunk table (few rows).
  std::vector<int32_t> big(nbig);
  std::vector<int64_t> bdeg(nbig);
  GBK_HIP(ctx, hipMemcpyAsync(big.data(), big_list, sizeof(int32_t) * nbig, hipMemcpyDeviceToHost, s));
  GBK_HIP(ctx, hipStreamSynchronize(s));
  {
    // degrees of the big rows: deg_new[row], gathered on the device
    auto *tmp = static_cast<int64_t *>(scratch(ctx, kS9, sizeof(int64_t) * nbig));
    if (!tmp) return set_error(ctx, GBK_ENOMEM, "relabel big degrees");
    gather_i64<<<grid_for(nbig), kBlock, 0, s>>>(deg_new, big_list, nbig, tmp);
    GBK_LAUNCH_CHECK(ctx, "gather_i64");
    GBK_HIP(ctx, hipMemcpyAsync(bdeg.data(), tmp, sizeof(int64_t) * nbig, hipMemcpyDeviceToHost, s));
    GBK_HIP(ctx, hipStreamSynchronize(s));
  }
  {
    // longest rows first: a merge pass of width w then covers a prefix of
    // the rows (those longer than w) instead of copying every big row again
    std::vector<int> order(nbig);
    for (int i = 0; i < nbig; i++) order[i] = i;
    std::sort(order.begin(), order.end(), [&](int a, int b) {
      return bdeg[a] != bdeg[b] ? bdeg[a] > bdeg[b] : big[a] < big[b];
    });
    std::vector<int32_t> b2(nbig);
    std::vector<int64_t> d2(nbig);
    for (int i = 0; i < nbig; i++) {
      b2[i] = big[order[i]];
      d2[i] = bdeg[order[i]];
    }
    big.swap(b2);
    bdeg.swap(d2);
  }
  std::vector<int64_t> soff(nbig + 1, 0);
  std::vector<int32_t> crow;
  std::vector<int64_t> cstart;
  int64_t maxdeg = 0;
  for (int i = 0; i < nbig; i++) {
    soff[i + 1] = soff[i] + bdeg[i];
    maxdeg = std::max(maxdeg, bdeg[i]);
    for (int64_t c = 0; c < bdeg[i]; c += kSortCap) {
      crow.push_back(i);
      cstart.push_back(c);
    }
  }
  const int64_t total = soff[nbig];
  using K = typename BigKey<W>::T;
  const size_t meta_bytes = sizeof(int32_t) * nbig + sizeof(int64_t) * (nbig + 1) +
                            sizeof(int32_t) * crow.size() + sizeof(int64_t) * cstart.size() + 64;
  char *meta = static_cast<char *>(scratch(ctx, kS9, meta_bytes));
  K *st0 = static_cast<K *>(scratch(ctx, kS10, sizeof(K) * total));
  K *st1 = static_cast<K *>(scratch(ctx, kS11, sizeof(K) * total));
  if (!meta || !st0 || !st1) return set_error(ctx, GBK_ENOMEM, "relabel staging");
  auto *big_d = reinterpret_cast<int32_t *>(meta);
  auto *soff_d = reinterpret_cast<int64_t *>(meta + ((sizeof(int32_t) * nbig + 7) / 8) * 8);
  auto *crow_d = reinterpret_cast<int32_t *>(soff_d + nbig + 1);
  auto *cstart_d = reinterpret_cast<int64_t *>(
      reinterpret_cast<char *>(crow_d) + ((sizeof(int32_t) * crow.size() + 7) / 8) * 8);
  GBK_HIP(ctx, hipMemcpyAsync(big_d, big.data(), sizeof(int32_t) * nbig, hipMemcpyHostToDevice, s));
  GBK_HIP(ctx, hipMemcpyAsync(soff_d, soff.data(), sizeof(int64_t) * (nbig + 1), hipMemcpyHostToDevice, s));
  GBK_HIP(ctx, hipMemcpyAsync(crow_d, crow.data(), sizeof(int32_t) * crow.size(), hipMemcpyHostToDevice, s));
  GBK_HIP(ctx, hipMemcpyAsync(cstart_d, cstart.data(), sizeof(int64_t) * cstart.size(), hipMemcpyHostToDevice, s));
  const int64_t nchunks = static_cast<int64_t>(crow.size());
  relabel_big_chunks<W><<<static_cast<unsigned>((nchunks + kWavesPerBlock - 1) / kWavesPerBlock), kBlock, 0, s>>>(
      big_d, soff_d, crow_d, cstart_d, nchunks, inv, off, nbr, wgt, new_ids_d, st0);
  GBK_LAUNCH_CHECK(ctx, "relabel_big_chunks");
  K *src = st0, *dst = st1;
  int active = nbig;  // rows longer than the current width: a prefix
  for (int64_t width = kSortCap; width < maxdeg; width *= 2) {
    while (active > 0 && bdeg[active - 1] <= width) active--;
    const int64_t span = soff[active];
    const int64_t threads = (span + kMergeItems - 1) / kMergeItems;
    relabel_merge_pass<K><<<grid_for(threads), kBlock, 0, s>>>(src, dst, soff_d, active, span, width);
    GBK_LAUNCH_CHECK(ctx, "relabel_merge_pass");
    std::swap(src, dst);
  }
  relabel_big_scatter<W><<<grid_for(total), kBlock, 0, s>>>(st0, st1, soff_d, big_d, nbig, total, off_out,
                                                            nbr_out, wgt_out);
  GBK_LAUNCH_CHECK(ctx, "relabel_big_scatter");
  GBK_HIP(ctx, hipStreamSynchronize(s));  // staging buffers are reused by the next CSR
  return GBK_OK;
}

__global__ void degrees_through_inverse(const int64_t *off, const int32_t *inv, int64_t n,
                                        int64_t *deg_new) {
  const int64_t r = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (r < n) {
    const int32_t u = inv[r];
    deg_new[r] = off[u + 1] - off[u];
  }
}

// Sorts every row of a CSR by (neighbour, weight) into new arrays: the
// relabel machinery with the identity mapping (used by the graph builder).
int sort_rows(gbk_ctx *ctx, int64_t n, const int64_t *off, const int32_t *nbr, const int32_t *wgt,
              const int32_t *iota, int64_t *deg, int64_t *off_out, int32_t *nbr_out,
              int32_t *wgt_out) {
  if (wgt) return relabel_csr<true>(ctx, n, off, nbr, wgt, iota, iota, deg, off_out, nbr_out, wgt_out, true);
  return relabel_csr<false>(ctx, n, off, nbr, wgt, iota, iota, deg, off_out, nbr_out, wgt_out, true);
}

}  // namespace gbk

using namespace gbk;

// ------------------------------------------------------ mapping producers
// The orderings that are a stable sort of the vertices by a key of their
// out-degree, made on the device: every vertex gets the 64-bit key
// (class << 32 | id) -- ties broken by id -- and the keys are sorted (one
// wave per 2048 keys in registers, then merge-path passes); the i-th key's
// vertex gets new id i.
//   -o 0 original: the identity (reorder_basic.h:65-79)
//   -o 2 sort: degree descending, then id (reorder_basic.h:182-258)
//   -o 5 DBG: 8 groups at m/n * {1/2, 1, 2, 4, 8, 16, 32, inf}, the highest
//        first, id order inside (reorder_hub.h:267-345)
//   -o 6 HubSortDBG: hubs (degree > integer m/n) by degree then id, both
//        descending, then the rest in id order (reorder_hub.h:380-502)
//   -o 7 HubClusterDBG: hubs then the rest, each in id order
//        (reorder_hub.h:505-582)
// Random (-o 1) is a sequential Fisher-Yates and the hub placements (-o 3, 4)
// move displaced vertices one after another: they stay on the host.
__global__ void mapping_keys(const int64_t *off, int64_t n, int algo, uint64_t avg, uint64_t *keys) {
  const int64_t v = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (v >= n) return;
  const uint64_t d = static_cast<uint64_t>(off[v + 1] - off[v]);
  const uint64_t dc = d < 0xffffffffull ? d : 0xffffffffull;
  const uint64_t id = static_cast<uint64_t>(v);
  uint64_t key = id;
  if (algo == 2) {
    key = ((0xffffffffull - dc) << 32) | id;
  } else if (algo == 5) {
    uint64_t lim = avg / 2;
    int b = 0;
    while (b < 7 && d > lim) {
      lim = b == 0 ? avg : lim * 2;
      b++;
    }
    key = (static_cast<uint64_t>(7 - b) << 32) | id;
  } else if (algo == 6) {
    key = d > avg ? ((0xffffffffull - dc) << 31) | (0x7fffffffull - id) : (1ull << 63) | id;
  } else if (algo == 7) {
    key = d > avg ? id : (1ull << 63) | id;
  }
  keys[v] = key;
}

__global__ void __launch_bounds__(kBlock)
mapping_sort_chunks(const uint64_t *keys, int64_t n, uint64_t *out) {
  const int64_t c = static_cast<int64_t>(blockIdx.x) * kWavesPerBlock + threadIdx.x / kWave;
  const int64_t first = c * kSortCap;
  if (first >= n) return;  // wave-uniform
  const int64_t len = min(static_cast<int64_t>(kSortCap), n - first);
  wave_sort<true, kSortCap / kWave>(
      len, [&](int64_t i) { return keys[first + i]; }, [&](int64_t i, uint64_t k) { out[first + i] = k; });
}

__global__ void mapping_scatter(const uint64_t *sorted, int64_t n, int algo, int32_t *new_ids) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint64_t k = sorted[i];
  const int64_t v = algo == 6 && !(k >> 63) ? static_cast<int64_t>(0x7fffffffull - (k & 0x7fffffffull))
                                            : static_cast<int64_t>(k & 0xffffffffull);
  new_ids[v] = static_cast<int32_t>(i);
}

__global__ void mapping_identity(int64_t n, int32_t *new_ids) {
  const int64_t v = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (v < n) new_ids[v] = static_cast<int32_t>(v);
}

// Mapping 'algo' of g into ids_d (device, n entries).
static int device_mapping(gbk_ctx *ctx, const gbk_graph *g, int algo, int32_t *ids_d) {
  const int64_t n = g->n;
  hipStream_t s = ctx->stream;
  if (n == 0) return GBK_OK;
  if (algo == 0) {
    mapping_identity<<<grid_for(n), kBlock, 0, s>>>(n, ids_d);
    GBK_LAUNCH_CHECK(ctx, "mapping_identity");
    return GBK_OK;
  }
  if (algo == 1) return random_mapping(ctx, n, ids_d);
  auto *k0 = static_cast<uint64_t *>(scratch(ctx, kS10, sizeof(uint64_t) * n));
  auto *k1 = static_cast<uint64_t *>(scratch(ctx, kS11, sizeof(uint64_t) * n));
  auto *so = static_cast<int64_t *>(scratch(ctx, kS9, 2 * sizeof(int64_t)));
  if (!k0 || !k1 || !so) return set_error(ctx, GBK_ENOMEM, "mapping scratch");
  const int64_t so_h[2] = {0, n};
  GBK_HIP(ctx, hipMemcpyAsync(so, so_h, sizeof(so_h), hipMemcpyHostToDevice, s));
  mapping_keys<<<grid_for(n), kBlock, 0, s>>>(g->off, n, algo, static_cast<uint64_t>(g->m / n), k1);
  const int64_t chunks = (n + kSortCap - 1) / kSortCap;
  mapping_sort_chunks<<<static_cast<unsigned>((chunks + kWavesPerBlock - 1) / kWavesPerBlock), kBlock, 0, s>>>(
      k1, n, k0);
  uint64_t *src = k0, *dst = k1;
  for (int64_t width = kSortCap; width < n; width *= 2) {
    const int64_t threads = (n + kMergeItems - 1) / kMergeItems;
    relabel_merge_pass<uint64_t><<<grid_for(threads), kBlock, 0, s>>>(src, dst, so, 1, n, width);
    std::swap(src, dst);
  }
  mapping_scatter<<<grid_for(n), kBlock, 0, s>>>(src, n, algo, ids_d);
  GBK_LAUNCH_CHECK(ctx, "mapping");
  // the host copies of so_h must outlive the copy above
  GBK_HIP(ctx, hipStreamSynchronize(s));
  return GBK_OK;
}

static bool device_mapping_supported(int algo) {
  return algo == 0 || algo == 1 || algo == 2 || algo == 5 || algo == 6 || algo == 7;
}

// Relabels 'in' by the mapping already in ids_d (device).
static int relabel_device_ids(gbk_ctx *ctx, const gbk_graph *in, int32_t *ids_d, gbk_graph **out,
                              gbk_relabel_stats *stats) {
  const int64_t n = in->n;
  hipStream_t s = ctx->stream;
  auto *inv = static_cast<int32_t *>(scratch(ctx, kS1, sizeof(int32_t) * (n + 1)));
  auto *deg_new = static_cast<int64_t *>(scratch(ctx, kS2, sizeof(int64_t) * (n + 1)));
  auto *err = static_cast<int32_t *>(scratch(ctx, kS3, 64));
  if (!inv || !deg_new || !err) return set_error(ctx, GBK_ENOMEM, "relabel scratch");
  int64_t m_in = in->m;
  if (in->directed)
    GBK_HIP(ctx, hipMemcpyAsync(&m_in, in->in_off + n, sizeof(int64_t), hipMemcpyDeviceToHost, s));
  GBK_HIP(ctx, hipStreamSynchronize(s));
  gbk_graph *g = nullptr;
  int rc = gbk_graph_alloc(ctx, n, in->m, m_in, in->directed, in->weighted, &g);
  if (rc) return rc;
  EventTimer timer(ctx);
  fill<int32_t>(ctx, inv, -1, n);
  fill<int32_t>(ctx, err, 0, 1);
  if (n > 0)
    relabel_invert<<<grid_for(n), kBlock, 0, s>>>(in->off, ids_d, in->org_ids, n, inv, deg_new,
                                                  g->org_ids, err);
  int32_t herr = 0;
  hipError_t e = hipMemcpyAsync(&herr, err, sizeof(int32_t), hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) {
    gbk_graph_free(ctx, g);
    return check_hip(ctx, e, "relabel_invert");
  }
  if (herr) {
    gbk_graph_free(ctx, g);
    return set_error(ctx, GBK_EPERM_MAP,
                     herr & 1 ? "relabel: mapping contains an out-of-range ID"
                              : "relabel: mapping contains a duplicate ID");
  }
  auto run = [&](const int64_t *off, const int32_t *nbr, const int32_t *wgt, int64_t *off_o,
                 int32_t *nbr_o, int32_t *wgt_o, bool first) -> int {
    if (!first && n > 0) degrees_through_inverse<<<grid_for(n), kBlock, 0, s>>>(off, inv, n, deg_new);
    if (in->weighted)
      return relabel_csr<true>(ctx, n, off, nbr, wgt, ids_d, inv, deg_new, off_o, nbr_o, wgt_o, true);
    return relabel_csr<false>(ctx, n, off, nbr, wgt, ids_d, inv, deg_new, off_o, nbr_o, wgt_o, true);
  };
  rc = run(in->off, in->nbr, in->wgt, g->off, g->nbr, g->wgt, true);
  if (!rc && in->directed) rc = run(in->in_off, in->in_nbr, in->in_wgt, g->in_off, g->in_nbr, g->in_wgt, false);
  const double ms = timer.stop_ms();
  if (!rc) rc = gbk_graph_finalize(ctx, g);
  if (rc) {
    gbk_graph_free(ctx, g);
    return rc;
  }
  if (stats) {
    stats->device_ms = ms;
    stats->n = n;
    stats->m = in->m;
  }
  *out = g;
  return GBK_OK;
}

extern "C" int gbk_relabel(gbk_ctx *ctx, const gbk_graph *in, const int32_t *new_ids,
                           gbk_graph **out, gbk_relabel_stats *stats) {
  if (!ctx || !in || !new_ids || !out) return set_error(ctx, GBK_EINVAL, "relabel: null argument");
  *out = nullptr;
  auto *ids_d = static_cast<int32_t *>(scratch(ctx, kS0, sizeof(int32_t) * (in->n + 1)));
  if (!ids_d) return set_error(ctx, GBK_ENOMEM, "relabel scratch");
  GBK_HIP(ctx, hipMemcpyAsync(ids_d, new_ids, sizeof(int32_t) * in->n, hipMemcpyHostToDevice, ctx->stream));
  return relabel_device_ids(ctx, in, ids_d, out, stats);
}

extern "C" int gbk_mapping(gbk_ctx *ctx, const gbk_graph *g, int algo, int32_t *new_ids, double *device_ms) {
  if (!ctx || !g || !new_ids) return set_error(ctx, GBK_EINVAL, "mapping: null argument");
  if (!device_mapping_supported(algo)) return set_error(ctx, GBK_EINVAL, "mapping: ordering not made on the device");
  auto *ids_d = static_cast<int32_t *>(scratch(ctx, kS0, sizeof(int32_t) * (g->n + 1)));
  if (!ids_d) return set_error(ctx, GBK_ENOMEM, "mapping scratch");
  EventTimer timer(ctx);
  if (int rc = device_mapping(ctx, g, algo, ids_d)) return rc;
  const double ms = timer.stop_ms();
  if (device_ms) *device_ms = ms;
  GBK_HIP(ctx, hipMemcpyAsync(new_ids, ids_d, sizeof(int32_t) * g->n, hipMemcpyDeviceToHost, ctx->stream));
  GBK_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return GBK_OK;
}

extern "C" int gbk_relabel_by(gbk_ctx *ctx, const gbk_graph *in, int algo, gbk_graph **out,
                              gbk_relabel_stats *stats, int32_t *new_ids, double *mapping_ms) {
  if (!ctx || !in || !out) return set_error(ctx, GBK_EINVAL, "relabel: null argument");
  *out = nullptr;
  if (!device_mapping_supported(algo)) return set_error(ctx, GBK_EINVAL, "relabel: ordering not made on the device");
  auto *ids_d = static_cast<int32_t *>(scratch(ctx, kS0, sizeof(int32_t) * (in->n + 1)));
  if (!ids_d) return set_error(ctx, GBK_ENOMEM, "relabel scratch");
  EventTimer timer(ctx);
  if (int rc = device_mapping(ctx, in, algo, ids_d)) return rc;
  const double ms = timer.stop_ms();
  if (mapping_ms) *mapping_ms = ms;
  if (new_ids)
    GBK_HIP(ctx, hipMemcpyAsync(new_ids, ids_d, sizeof(int32_t) * in->n, hipMemcpyDeviceToHost, ctx->stream));
  return relabel_device_ids(ctx, in, ids_d, out, stats);
}
