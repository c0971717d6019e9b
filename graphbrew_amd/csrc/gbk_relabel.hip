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
//                bitonic network over lane exchanges (registers only)
//   deg <= 1024  one wave per row, 2..16 entries per lane in registers,
//                bitonic network (in-lane strides in registers, the others
//                over lane exchanges; no LDS, no barriers)
//   larger       an MSD bucket partition of each row on its remapped keys
//                (buckets of a few hundred entries), then one wave sorts
//                each bucket in registers straight into place.
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
// <=256, <=512, <=1024 (wave register sorts), class 8 unused, larger (the
// bucket path).  Empty rows need no work.  The rows with entries are listed class
// by class in one array, ascending within a class: a count per class and
// tile of rows, one scan of those counts, and a scatter that ranks rows by
// ballots (no global atomics; per-wave queues flushing into one counter line
// took 0.7 ms for 4M rows).
constexpr int kClasses = 10;
constexpr int kClassBig = 9;
struct RelabelLists {
  int32_t *l[kClasses];
};

// Rows above kBigRow entries take the bucket path: sorting 1025-2048
// entries as one 32-entry-a-lane network cost 3x a bucketed row.
constexpr int64_t kBigRow = 1024;
__device__ __forceinline__ int bin_class(int64_t d) {
  return d == 0 ? -1 : d <= 8 ? 0 : d <= 16 ? 1 : d <= 32 ? 2 : d <= 64 ? 3 : d <= 128 ? 4
       : d <= 256 ? 5 : d <= 512 ? 6 : d <= kBigRow ? 7 : kClassBig;
}

// Workgroup t counts the rows of each class in rows [t * tile, (t+1) * tile):
// cnt[c * gridDim.x + t]; the entries of the big rows are summed into
// *big_total.
__global__ void __launch_bounds__(kBlock)
relabel_bin_count(const int64_t *deg_new, int64_t n, int64_t tile, uint32_t *cnt,
                  unsigned long long *big_total) {
  __shared__ uint32_t tot[kClasses];
  if (threadIdx.x < kClasses) tot[threadIdx.x] = 0;
  __syncthreads();
  const int64_t lo = static_cast<int64_t>(blockIdx.x) * tile;
  const int64_t hi = min(n, lo + tile);
  uint32_t mine[kClasses] = {};
  unsigned long long big = 0;
  for (int64_t r = lo + threadIdx.x; r < hi; r += kBlock) {
    const int64_t d = deg_new[r];
    const int cls = bin_class(d);
#pragma unroll
    for (int c = 0; c < kClasses; c++) mine[c] += cls == c;
    if (cls == kClassBig) big += static_cast<unsigned long long>(d);
  }
#pragma unroll
  for (int c = 0; c < kClasses; c++) {
    const uint32_t w = wave_sum(mine[c]);
    if (lane_id() == 0 && w) atomicAdd(&tot[c], w);
  }
  big = wave_sum(big);
  if (lane_id() == 0 && big) atomicAdd(big_total, big);
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

// Bitonic network over groups of G lanes, one key a lane (lane exchanges
// only: no LDS, no barriers).
template <typename K, int G>
__device__ __forceinline__ K group_net(K key, int glane) {
#pragma unroll
  for (int k = 2; k <= G; k <<= 1) {
#pragma unroll
    for (int j = k >> 1; j > 0; j >>= 1) {
      const K other = sizeof(K) == 4 ? lane_xor_rt(key, j) : __shfl_xor(key, j);
      const bool up = (glane & k) == 0;
      const bool lower = (glane & j) == 0;
      key = (lower == up) ? (other < key ? other : key) : (other > key ? other : key);
    }
  }
  return key;
}

// Rows of at most G entries: G lanes hold one entry each.  Keys compare as
// (v, w).  NW (weighted rows, ids below 2^(32 - log2 G)): the 32-bit
// composite (v << log2 G) | entry index sorts and the weight follows through
// the index; a group that finds one v twice lists its row in 'redo' for the
// 64-bit sort instead of writing it.
template <bool W, int G, bool NW = false>
__global__ void __launch_bounds__(kBlock)
relabel_group_sort(const int32_t *rows, int64_t nrows, const int32_t *inv, const int64_t *off,
                   const int32_t *nbr, const int32_t *wgt, const int32_t *new_ids,
                   const int64_t *new_off, int32_t *nbr_out, int32_t *wgt_out, int32_t *redo = nullptr,
                   unsigned long long *redo_count = nullptr) {
  using K = typename std::conditional<W, uint64_t, uint32_t>::type;
  const int glane = threadIdx.x % G;
  const int64_t ri = static_cast<int64_t>(blockIdx.x) * (kBlock / G) + threadIdx.x / G;
  int64_t len = 0, src = 0, dst = 0;
  int32_t r = 0;
  if (ri < nrows) {
    r = rows[ri];
    const int32_t u = inv[r];
    src = off[u];
    len = off[u + 1] - src;
    dst = new_off[r];
  }
  if constexpr (W && NW) {
    constexpr int IB = G == 8 ? 3 : G == 16 ? 4 : G == 32 ? 5 : 6;
    uint32_t key = ~0u;
    if (glane < len) key = (static_cast<uint32_t>(new_ids[nbr[src + glane]]) << IB) | static_cast<uint32_t>(glane);
    key = group_net<uint32_t, G>(key, glane);
    const uint32_t succ = __shfl(key, (lane_id() + 1) & (kWave - 1));
    const bool dup = glane + 1 < len && (succ >> IB) == (key >> IB);
    const uint64_t gmask = G == kWave ? ~0ull : ((1ull << G) - 1) << (lane_id() / G * G);
    if (__ballot(dup) & gmask) {
      if (glane == 0) redo[atomicAdd(redo_count, 1ull)] = r;
    } else if (glane < len) {
      nbr_out[dst + glane] = static_cast<int32_t>(key >> IB);
      wgt_out[dst + glane] = wgt[src + (key & (G - 1))];
    }
    return;
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
  key = group_net<K, G>(key, glane);
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
// Bitonic network over 64*E keys held by one wave: entry i in lane i % 64,
// register i / 64.  Strides below 64 pair lanes, larger strides pair
// registers of one lane.
template <typename K, int E>
__device__ __forceinline__ void sort_regs(K (&key)[E]) {
  constexpr int N = kWave * E;
  const int lane = lane_id();
#pragma unroll
  for (int k = 2; k <= N; k <<= 1) {
#pragma unroll
    for (int j = k >> 1; j > 0; j >>= 1) {
      if (j < kWave) {
#pragma unroll
        for (int e = 0; e < E; e++) {
          // 32-bit keys up to 16 a lane exchange through DPP / permlane
          // swaps (VALU); wider rows and 64-bit keys through ds_bpermute,
          // which needs fewer registers (measured faster there)
          const K other = sizeof(K) == 4 && E <= 16 ? lane_xor_rt(key[e], j) : __shfl_xor(key[e], j);
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
}

// One wave sorts up to 64*E keys: load(i) gives key i (i < len), store(i,
// key) takes the i-th smallest.  Padding keys are all ones.
template <bool W, int E, class Load, class Store>
__device__ __forceinline__ void wave_sort(int64_t len, Load load, Store store) {
  using K = typename RowKey<W>::T;
  const int lane = lane_id();
  K key[E];
#pragma unroll
  for (int e = 0; e < E; e++) {
    const int64_t i = static_cast<int64_t>(e) * kWave + lane;
    key[e] = i < len ? load(i) : static_cast<K>(~static_cast<K>(0));
  }
  sort_regs<K, E>(key);
#pragma unroll
  for (int e = 0; e < E; e++) {
    const int64_t i = static_cast<int64_t>(e) * kWave + lane;
    if (i < len) store(i, key[e]);
  }
}

// Weighted rows with narrow ids: a row's keys (v, w) sort as the 32-bit
// composite (v << IB) | i, i the entry's index in its row (IB bits, so
// nb + IB <= 32 for ids below 2^nb), and the weight is fetched through i
// afterwards.  This is the (v, w) order whenever the row's v are distinct,
// which squished graphs guarantee; a row holding one v twice is detected
// after the sort (its sorted neighbour has the same v) and listed for a
// 64-bit redo instead of being written.  Returns true when the wave wrote
// the row.
template <int E, class LoadV, class LoadW, class Store>
__device__ __forceinline__ bool wave_sort_narrow(int64_t len, LoadV load_v, LoadW load_w, Store store) {
  constexpr int IB = 6 + (E == 1 ? 0 : E == 2 ? 1 : E == 4 ? 2 : E == 8 ? 3 : E == 16 ? 4 : 5);
  const int lane = lane_id();
  uint32_t key[E];
#pragma unroll
  for (int e = 0; e < E; e++) {
    const int i = e * kWave + lane;
    key[e] = i < len ? (static_cast<uint32_t>(load_v(i)) << IB) | static_cast<uint32_t>(i) : ~0u;
  }
  sort_regs<uint32_t, E>(key);
  bool dup = false;
#pragma unroll
  for (int e = 0; e < E; e++) {
    const uint32_t down = __shfl(key[e], (lane + 1) & (kWave - 1));
    const uint32_t next_first = e + 1 < E ? __shfl(key[e + 1 < E ? e + 1 : e], 0) : ~0u;
    const uint32_t succ = lane == kWave - 1 ? next_first : down;
    const int i = e * kWave + lane;
    dup |= i + 1 < len && (succ >> IB) == (key[e] >> IB);
  }
  if (__ballot(dup)) return false;
#pragma unroll
  for (int e = 0; e < E; e++) {
    const int i = e * kWave + lane;
    if (i < len) store(i, static_cast<int32_t>(key[e] >> IB), load_w(static_cast<int>(key[e] & ((1u << IB) - 1))));
  }
  return true;
}

template <bool W, int E, bool NW = false>
__global__ void __launch_bounds__(kBlock)
relabel_wave_sort(const int32_t *rows, int64_t nrows, const int32_t *inv, const int64_t *off,
                  const int32_t *nbr, const int32_t *wgt, const int32_t *new_ids,
                  const int64_t *new_off, int32_t *nbr_out, int32_t *wgt_out, int32_t *redo = nullptr,
                  unsigned long long *redo_count = nullptr) {
  using K = typename RowKey<W>::T;
  const int64_t ri = static_cast<int64_t>(blockIdx.x) * kWavesPerBlock + threadIdx.x / kWave;
  if (ri >= nrows) return;  // wave-uniform
  const int32_t r = rows[ri];
  const int32_t u = inv[r];
  const int64_t src = off[u];
  const int64_t len = off[u + 1] - src;
  const int64_t dst = new_off[r];
  if constexpr (W && NW) {
    const bool wrote = wave_sort_narrow<E>(
        len, [&](int i) { return new_ids[nbr[src + i]]; }, [&](int i) { return wgt[src + i]; },
        [&](int i, int32_t v, int32_t w) {
          nbr_out[dst + i] = v;
          wgt_out[dst + i] = w;
        });
    if (!wrote && lane_id() == 0) redo[atomicAdd(redo_count, 1ull)] = r;
    return;
  }
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

// Rows a narrow sort listed (one neighbour twice): sorted again on the full
// (v, w) keys, one wave each, grid-stride over the list (rows of at most
// 1024 entries).
__global__ void __launch_bounds__(kBlock)
relabel_redo_rows(const int32_t *redo, const unsigned long long *count, const int32_t *inv, const int64_t *off,
                  const int32_t *nbr, const int32_t *wgt, const int32_t *new_ids, const int64_t *new_off,
                  int32_t *nbr_out, int32_t *wgt_out) {
  const int64_t nr = static_cast<int64_t>(*count);
  const int64_t waves = static_cast<int64_t>(gridDim.x) * kWavesPerBlock;
  for (int64_t s = static_cast<int64_t>(blockIdx.x) * kWavesPerBlock + threadIdx.x / kWave; s < nr; s += waves) {
    const int32_t r = redo[s];
    const int32_t u = inv[r];
    const int64_t src = off[u];
    const int64_t len = off[u + 1] - src;
    const int64_t dst = new_off[r];
    wave_sort<true, 16>(
        len, [&](int64_t i) { return RowKey<true>::make(new_ids[nbr[src + i]], wgt[src + i]); },
        [&](int64_t i, uint64_t key) {
          nbr_out[dst + i] = static_cast<int32_t>(key >> 32);
          wgt_out[dst + i] = static_cast<int32_t>(static_cast<uint32_t>(key) ^ 0x80000000u);
        });
  }
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

// ------------------------------------------------------------- big rows
// Rows above kBigRow entries: an MSD bucket partition on the remapped keys,
// then one wave per bucket sorts it in registers straight into place.
//   level 1  workgroups of kBigChunk entries gather the keys once into a
//            staging buffer and count them per (row, bucket): a row of d
//            entries gets 2^b buckets on the top b key bits, b chosen so a
//            bucket holds about kSegTarget entries; a device scan of the
//            counts gives each bucket its staging range, a second pass over
//            the staged keys scatters them there (order inside a bucket is
//            whatever the LDS atomics give: the bucket is sorted next)
//   level k  buckets still above kSortCap entries (keys crowded in a narrow
//            range, e.g. hub rows under a degree order) are split again on
//            their next key bits, one workgroup each, ping-ponging between
//            the two staging buffers; a bucket whose key bits are used up
//            holds one key repeated and is copied as it stands
//   sort     every bucket of at most kSortCap entries is sorted by one wave
//            (4/8/16/32 entries a lane by its size) into the output row.
// No host work per row and no merge passes: each entry is gathered once,
// written to staging twice and to the output once.
constexpr int kSegTarget = 384;                     // bucket size a level aims at
constexpr int kSegMaxBits = 12;                     // at most 4096 buckets a segment and level
constexpr int kSegMaxBuckets = 1 << kSegMaxBits;
constexpr int kBigChunk = 4096;                     // entries a level-1 workgroup stages
constexpr int kChunkPer = kBigChunk / kBlock;       // entries a lane of it holds
constexpr int kSegLists = 6;                        // sort classes 0..3, copy, overflow flag
constexpr int kSegCopy = 4;

// A bucket still to be partitioned or sorted: staging position and length;
// hi_buf = (key bits not yet partitioned) << 1 | (staging buffer holding it).
struct SegRec {
  int64_t pos;
  int32_t len;
  int32_t hi_buf;
};

template <bool W>
__device__ __forceinline__ int32_t key_v(typename RowKey<W>::T k) {
  if constexpr (W) return static_cast<int32_t>(k >> 32);
  else return static_cast<int32_t>(k);
}
template <bool W>
__device__ __forceinline__ int32_t key_w(typename RowKey<W>::T k) {
  if constexpr (W) return static_cast<int32_t>(static_cast<uint32_t>(k) ^ 0x80000000u);
  else return 0;
}

// Buckets of a segment of 'len' entries with 'hi' key bits left: about
// kSegTarget entries each, at least 2, at most 2^kSegMaxBits, at most 2^hi.
__device__ __forceinline__ int seg_bits(int64_t len, int hi) {
  const int64_t want = (len + kSegTarget - 1) / kSegTarget;
  int b = want > 1 ? 64 - __clzll(static_cast<unsigned long long>(want - 1)) : 1;
  b = max(b, 1);
  b = min(b, kSegMaxBits);
  return min(b, hi);
}

// Level-1 buckets per big row: as seg_bits, at most 2^kL1MaxBits so a
// workgroup's counts, offsets and reordered chunk fit in LDS (rows beyond
// kSegTarget * 2^kL1MaxBits entries get their big buckets split at level 2).
constexpr int kL1MaxBits = 10;
constexpr int kL1MaxBuckets = 1 << kL1MaxBits;
__device__ __forceinline__ int l1_bits(int64_t d, int key_bits) { return min(seg_bits(d, key_bits), kL1MaxBits); }

// Largest i in [0, count) with a[i] <= x (a ascending, a[0] <= x).
__device__ __forceinline__ int64_t find_le(const int64_t *a, int64_t count, int64_t x) {
  int64_t lo = 0, hi = count;
  while (hi - lo > 1) {
    const int64_t mid = (lo + hi) >> 1;
    if (a[mid] <= x) lo = mid;
    else hi = mid;
  }
  return lo;
}

// Per big row i (new id big[i]): its length, level-1 bucket count and
// kBigChunk chunk count, the inputs of big_layout's scans.
__global__ void big_row_sizes(const int32_t *big, int64_t nbig, const int64_t *deg_new, int key_bits,
                              int64_t *len, int64_t *buckets, int64_t *chunks) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= nbig) return;
  const int64_t d = deg_new[big[i]];
  len[i] = d;
  buckets[i] = int64_t{1} << l1_bits(d, key_bits);
  chunks[i] = (d + kBigChunk - 1) / kBigChunk;
}

// One workgroup of 1024 lanes turns the three arrays (nbig + 1 entries
// each) into their exclusive prefixes in place (staging offsets, level-1
// bucket offsets, chunk offsets); entry [nbig] receives each total.  Lane t
// takes a contiguous run of rows, loads it whole (independent loads) and
// carries its sums through one block scan.
__global__ void __launch_bounds__(1024)
big_layout(int64_t nbig, int64_t *stage_off, int64_t *bucket_off, int64_t *chunk_off) {
  __shared__ int64_t wsum[3][1024 / kWave];
  int64_t *arr[3] = {stage_off, bucket_off, chunk_off};
  const int t = threadIdx.x;
  const int64_t per = (nbig + 1023) / 1024;
  const int64_t b = min(nbig, t * per), e = min(nbig, b + per);
  int64_t s[3] = {0, 0, 0};
  for (int q = 0; q < 3; q++) {
#pragma unroll 8
    for (int64_t i = b; i < e; i++) s[q] += arr[q][i];
  }
  int64_t incl[3];
  for (int q = 0; q < 3; q++) incl[q] = wave_incl_scan(s[q]);
  const int w = t / kWave;
  if (lane_id() == kWave - 1)
    for (int q = 0; q < 3; q++) wsum[q][w] = incl[q];
  __syncthreads();
  if (w == 0) {
    for (int q = 0; q < 3; q++) {
      int64_t x = lane_id() < 1024 / kWave ? wsum[q][lane_id()] : 0;
      x = wave_incl_scan(x);
      if (lane_id() < 1024 / kWave) wsum[q][lane_id()] = x;
    }
  }
  __syncthreads();
  for (int q = 0; q < 3; q++) {
    int64_t run = incl[q] - s[q] + (w > 0 ? wsum[q][w - 1] : 0);
#pragma unroll 8
    for (int64_t i = b; i < e; i++) {
      const int64_t x = arr[q][i];
      arr[q][i] = run;
      run += x;
    }
    if (t == 1023) arr[q][nbig] = run;
  }
}

// Level 1, pass 1: workgroup c stages chunk c of its big row (keys gathered
// through new_ids) and adds its per-bucket counts to cnt.
template <bool W>
__global__ void __launch_bounds__(kBlock)
big_count(const int32_t *big, int64_t nbig, const int64_t *stage_off, const int64_t *bucket_off,
          const int64_t *chunk_off, const int32_t *inv, const int64_t *off, const int32_t *nbr,
          const int32_t *wgt, const int32_t *new_ids, int key_bits, typename RowKey<W>::T *stage,
          unsigned long long *cnt) {
  using K = typename RowKey<W>::T;
  __shared__ uint32_t hist[kL1MaxBuckets];
  const int64_t c = blockIdx.x;
  if (c >= chunk_off[nbig]) return;  // workgroup-uniform
  const int64_t i = find_le(chunk_off, nbig, c);
  const int32_t u = inv[big[i]];
  const int64_t d = stage_off[i + 1] - stage_off[i];
  const int b = l1_bits(d, key_bits);
  const int shift = key_bits - b;
  for (int k = threadIdx.x; k < (1 << b); k += kBlock) hist[k] = 0;
  __syncthreads();
  const int64_t cs = (c - chunk_off[i]) * kBigChunk;
  const int len = static_cast<int>(min<int64_t>(kBigChunk, d - cs));
  const int64_t src = off[u] + cs;
  K *out = stage + stage_off[i] + cs;
  int32_t v[kChunkPer], w[kChunkPer];
#pragma unroll
  for (int e = 0; e < kChunkPer; e++) {
    const int j = e * kBlock + threadIdx.x;
    v[e] = j < len ? nbr[src + j] : 0;
    w[e] = W && j < len ? wgt[src + j] : 0;
  }
#pragma unroll
  for (int e = 0; e < kChunkPer; e++) {
    const int j = e * kBlock + threadIdx.x;
    if (j < len) v[e] = new_ids[v[e]];
  }
#pragma unroll
  for (int e = 0; e < kChunkPer; e++) {
    const int j = e * kBlock + threadIdx.x;
    if (j < len) {
      const K key = RowKey<W>::make(v[e], w[e]);
      out[j] = key;
      atomicAdd(&hist[static_cast<uint32_t>(key >> shift)], 1u);
    }
  }
  __syncthreads();
  for (int k = threadIdx.x; k < (1 << b); k += kBlock)
    if (hist[k]) atomicAdd(&cnt[bucket_off[i] + k], static_cast<unsigned long long>(hist[k]));
}

// Level 1, pass 2: workgroup c moves its staged chunk into the buckets; cur
// holds each bucket's next free staging slot (starts as the scan of cnt).
// (Reordering the chunk by bucket in LDS first, so each bucket's share
// leaves as one run, measured slower: 254 -> 284 us at Kronecker scale 22.)
template <bool W>
__global__ void __launch_bounds__(kBlock)
big_scatter(const int64_t *stage_off, const int64_t *bucket_off, const int64_t *chunk_off, int64_t nbig,
            int key_bits, const typename RowKey<W>::T *stage, typename RowKey<W>::T *stage2,
            unsigned long long *cur) {
  using K = typename RowKey<W>::T;
  __shared__ uint32_t hist[kL1MaxBuckets];
  __shared__ unsigned long long base[kL1MaxBuckets];
  const int64_t c = blockIdx.x;
  if (c >= chunk_off[nbig]) return;  // workgroup-uniform
  const int64_t i = find_le(chunk_off, nbig, c);
  const int64_t d = stage_off[i + 1] - stage_off[i];
  const int b = l1_bits(d, key_bits);
  const int shift = key_bits - b;
  for (int k = threadIdx.x; k < (1 << b); k += kBlock) hist[k] = 0;
  __syncthreads();
  const int64_t cs = (c - chunk_off[i]) * kBigChunk;
  const int len = static_cast<int>(min<int64_t>(kBigChunk, d - cs));
  const K *in = stage + stage_off[i] + cs;
  K key[kChunkPer];
  uint32_t rank[kChunkPer];
#pragma unroll
  for (int e = 0; e < kChunkPer; e++) {
    const int j = e * kBlock + threadIdx.x;
    key[e] = j < len ? in[j] : K{0};
  }
#pragma unroll
  for (int e = 0; e < kChunkPer; e++) {
    const int j = e * kBlock + threadIdx.x;
    rank[e] = j < len ? atomicAdd(&hist[static_cast<uint32_t>(key[e] >> shift)], 1u) : 0u;
  }
  __syncthreads();
  for (int k = threadIdx.x; k < (1 << b); k += kBlock)
    if (hist[k]) base[k] = atomicAdd(&cur[bucket_off[i] + k], static_cast<unsigned long long>(hist[k]));
  __syncthreads();
#pragma unroll
  for (int e = 0; e < kChunkPer; e++) {
    const int j = e * kBlock + threadIdx.x;
    if (j < len) stage2[base[static_cast<uint32_t>(key[e] >> shift)] + rank[e]] = key[e];
  }
}

// Appends 'rec' to the list its length selects: sort classes of at most
// 256/512/1024/2048 entries, the copy list (key bits used up), or the next
// level's list; lists hold 'cap' records, an append past that sets
// counters[kSegLists - 1].  One atomic per wave and list.
__device__ __forceinline__ void seg_emit(bool valid, const SegRec &rec, SegRec *lists, int64_t cap,
                                         unsigned long long *counters, SegRec *next, int64_t next_cap,
                                         unsigned long long *next_count) {
  int cls = -1;
  if (valid) {
    const int hi = rec.hi_buf >> 1;
    cls = rec.len <= 256 ? 0 : rec.len <= 512 ? 1 : rec.len <= 1024 ? 2 : rec.len <= kSortCap ? 3
        : hi == 0 ? kSegCopy : kSegLists;  // kSegLists: the next level
  }
  for (int c = 0; c <= kSegLists; c++) {
    if (c == kSegLists - 1) continue;
    unsigned long long *ctr = c == kSegLists ? next_count : &counters[c];
    const int64_t slot = wave_append(ctr, cls == c);
    if (cls == c) {
      if (slot < (c == kSegLists ? next_cap : cap)) (c == kSegLists ? next : lists + c * cap)[slot] = rec;
      else atomicOr(&counters[kSegLists - 1], 1ull);
    }
  }
}

// Level 1, pass 3: bucket k's record (its staging range in buffer 1).
__global__ void big_segments(const int64_t *bucket_off, const int64_t *stage_off, int64_t nbig, int key_bits,
                             const unsigned long long *cnt, const unsigned long long *cur, SegRec *lists,
                             int64_t cap, unsigned long long *counters, SegRec *next, int64_t next_cap,
                             unsigned long long *next_count) {
  const int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const int64_t nb = bucket_off[nbig];
  SegRec rec{};
  bool valid = false;
  if (k < nb) {
    const int64_t i = find_le(bucket_off, nbig, k);
    const int64_t len = static_cast<int64_t>(cnt[k]);
    if (len > 0) {
      const int b = l1_bits(stage_off[i + 1] - stage_off[i], key_bits);
      rec.pos = static_cast<int64_t>(cur[k]) - len;
      rec.len = static_cast<int32_t>(len);
      rec.hi_buf = ((key_bits - b) << 1) | 1;
      valid = true;
    }
  }
  seg_emit(valid, rec, lists, cap, counters, next, next_cap, next_count);
}

// Level >= 2: workgroup s splits segment in[s] on its next key bits into the
// other staging buffer (LDS histogram, scan and cursors) and emits the parts.
template <bool W>
__global__ void __launch_bounds__(1024)
seg_split(const SegRec *in, typename RowKey<W>::T *buf0, typename RowKey<W>::T *buf1, SegRec *lists,
          int64_t cap, unsigned long long *counters, SegRec *next, int64_t next_cap,
          unsigned long long *next_count) {
  using K = typename RowKey<W>::T;
  __shared__ uint32_t hist[kSegMaxBuckets];
  __shared__ uint32_t start[kSegMaxBuckets];
  __shared__ uint32_t wtot[1024 / kWave];
  __shared__ K wlo[1024 / kWave], whi[1024 / kWave];
  const SegRec rec = in[blockIdx.x];
  const K *src = (rec.hi_buf & 1) ? buf1 : buf0;
  K *dst = (rec.hi_buf & 1) ? buf0 : buf1;
  // the keys agree on every bit above the highest bit where their minimum
  // and maximum differ: partition just below it
  K lo = static_cast<K>(~K{0}), hi_key = 0;
  for (int64_t j = threadIdx.x; j < rec.len; j += 1024) {
    const K key = src[rec.pos + j];
    lo = key < lo ? key : lo;
    hi_key = key > hi_key ? key : hi_key;
  }
  lo = wave_min(lo);
  hi_key = wave_max(hi_key);
  if (lane_id() == 0) {
    wlo[threadIdx.x / kWave] = lo;
    whi[threadIdx.x / kWave] = hi_key;
  }
  for (int k = threadIdx.x; k < kSegMaxBuckets; k += 1024) hist[k] = 0;
  __syncthreads();
  for (int w = 0; w < 1024 / kWave; w++) {
    lo = wlo[w] < lo ? wlo[w] : lo;
    hi_key = whi[w] > hi_key ? whi[w] : hi_key;
  }
  const K diff = lo ^ hi_key;
  int hi = 0;
  if constexpr (W) hi = diff == 0 ? 0 : 64 - __clzll(diff);
  else hi = diff == 0 ? 0 : 32 - __clz(diff);
  if (hi == 0) {  // one key repeated: copied as it stands (workgroup-uniform)
    SegRec same = rec;
    same.hi_buf = rec.hi_buf & 1;
    if (threadIdx.x < kWave) seg_emit(threadIdx.x == 0, same, lists, cap, counters, next, next_cap, next_count);
    return;
  }
  const int b = seg_bits(rec.len, hi);
  const int shift = hi - b;
  const uint32_t mask = (1u << b) - 1;
  for (int64_t j = threadIdx.x; j < rec.len; j += 1024)
    atomicAdd(&hist[static_cast<uint32_t>(src[rec.pos + j] >> shift) & mask], 1u);
  __syncthreads();
  // exclusive scan of the 4096 counts: 4 per lane
  uint32_t c4[4], s = 0;
#pragma unroll
  for (int q = 0; q < 4; q++) {
    c4[q] = hist[threadIdx.x * 4 + q];
    s += c4[q];
  }
  const uint32_t incl = wave_incl_scan(s);
  if (lane_id() == kWave - 1) wtot[threadIdx.x / kWave] = incl;
  __syncthreads();
  uint32_t run = incl - s;
  for (int w = 0; w < static_cast<int>(threadIdx.x / kWave); w++) run += wtot[w];
#pragma unroll
  for (int q = 0; q < 4; q++) {
    start[threadIdx.x * 4 + q] = run;
    run += c4[q];
  }
  __syncthreads();
  // cursors: hist becomes each bucket's next slot
#pragma unroll
  for (int q = 0; q < 4; q++) hist[threadIdx.x * 4 + q] = start[threadIdx.x * 4 + q];
  __syncthreads();
  for (int64_t j = threadIdx.x; j < rec.len; j += 1024) {
    const K key = src[rec.pos + j];
    const uint32_t p = atomicAdd(&hist[static_cast<uint32_t>(key >> shift) & mask], 1u);
    dst[rec.pos + p] = key;
  }
  __syncthreads();
  for (int k0 = 0; k0 < kSegMaxBuckets; k0 += 1024) {
    const int k = k0 + threadIdx.x;
    const uint32_t len = k < (1 << b) ? hist[k] - start[k] : 0;
    SegRec part{};
    part.pos = rec.pos + start[k];
    part.len = static_cast<int32_t>(len);
    part.hi_buf = (shift << 1) | ((rec.hi_buf & 1) ^ 1);
    seg_emit(len > 0, part, lists, cap, counters, next, next_cap, next_count);
  }
}

// Where staging position 'pos' lands in the output: its big row's offset
// plus its place in the row.
__device__ __forceinline__ int64_t seg_dst(int64_t pos, const int64_t *stage_off, int64_t nbig,
                                           const int32_t *big, const int64_t *new_off) {
  const int64_t i = find_le(stage_off, nbig, pos);
  return new_off[big[i]] + (pos - stage_off[i]);
}

// Sorts the buckets of one class (at most 64*E entries) one wave each into
// the output rows; grid-stride over the class's records.  NW: weighted
// buckets sort as 32-bit composites (see wave_sort_narrow); a bucket with one
// v twice goes to 'redo' instead.
template <bool W, int E, bool NW = false>
__global__ void __launch_bounds__(kBlock)
seg_sort(const SegRec *list, const unsigned long long *count, int64_t cap, const typename RowKey<W>::T *buf0,
         const typename RowKey<W>::T *buf1, const int64_t *stage_off, int64_t nbig, const int32_t *big,
         const int64_t *new_off, int32_t *nbr_out, int32_t *wgt_out, SegRec *redo = nullptr,
         unsigned long long *redo_count = nullptr) {
  using K = typename RowKey<W>::T;
  const int64_t nseg = min(static_cast<int64_t>(*count), cap);
  const int64_t waves = static_cast<int64_t>(gridDim.x) * kWavesPerBlock;
  for (int64_t s = static_cast<int64_t>(blockIdx.x) * kWavesPerBlock + threadIdx.x / kWave; s < nseg; s += waves) {
    const SegRec rec = list[s];
    const K *src = ((rec.hi_buf & 1) ? buf1 : buf0) + rec.pos;
    const int64_t dst = seg_dst(rec.pos, stage_off, nbig, big, new_off);
    if constexpr (W && NW) {
      const bool wrote = wave_sort_narrow<E>(
          rec.len, [&](int i) { return key_v<W>(src[i]); }, [&](int i) { return key_w<W>(src[i]); },
          [&](int i, int32_t v, int32_t w) {
            nbr_out[dst + i] = v;
            wgt_out[dst + i] = w;
          });
      if (!wrote && lane_id() == 0) redo[atomicAdd(redo_count, 1ull)] = rec;
    } else {
      wave_sort<W, E>(
          rec.len, [&](int64_t i) { return src[i]; },
          [&](int64_t i, K key) {
            nbr_out[dst + i] = key_v<W>(key);
            if constexpr (W) wgt_out[dst + i] = key_w<W>(key);
          });
    }
  }
}

// Weighted buckets a narrow sort listed (one v twice): sorted on the full
// (v, w) keys (buckets of at most kSortCap entries).
__global__ void __launch_bounds__(kBlock)
seg_redo(const SegRec *list, const unsigned long long *count, const uint64_t *buf0, const uint64_t *buf1,
         const int64_t *stage_off, int64_t nbig, const int32_t *big, const int64_t *new_off, int32_t *nbr_out,
         int32_t *wgt_out) {
  const int64_t nseg = static_cast<int64_t>(*count);
  const int64_t waves = static_cast<int64_t>(gridDim.x) * kWavesPerBlock;
  for (int64_t s = static_cast<int64_t>(blockIdx.x) * kWavesPerBlock + threadIdx.x / kWave; s < nseg; s += waves) {
    const SegRec rec = list[s];
    const uint64_t *src = ((rec.hi_buf & 1) ? buf1 : buf0) + rec.pos;
    const int64_t dst = seg_dst(rec.pos, stage_off, nbig, big, new_off);
    wave_sort<true, kSortCap / kWave>(
        rec.len, [&](int64_t i) { return src[i]; },
        [&](int64_t i, uint64_t key) {
          nbr_out[dst + i] = key_v<true>(key);
          wgt_out[dst + i] = key_w<true>(key);
        });
  }
}

// Buckets whose key bits are used up hold one key repeated: copied as they are.
template <bool W>
__global__ void __launch_bounds__(kBlock)
seg_copy(const SegRec *list, const unsigned long long *count, int64_t cap, const typename RowKey<W>::T *buf0,
         const typename RowKey<W>::T *buf1, const int64_t *stage_off, int64_t nbig, const int32_t *big,
         const int64_t *new_off, int32_t *nbr_out, int32_t *wgt_out) {
  const int64_t nseg = min(static_cast<int64_t>(*count), cap);
  for (int64_t s = blockIdx.x; s < nseg; s += gridDim.x) {
    const SegRec rec = list[s];
    const auto *src = ((rec.hi_buf & 1) ? buf1 : buf0) + rec.pos;
    const int64_t dst = seg_dst(rec.pos, stage_off, nbig, big, new_off);
    for (int64_t i = threadIdx.x; i < rec.len; i += kBlock) {
      nbr_out[dst + i] = key_v<W>(src[i]);
      if constexpr (W) wgt_out[dst + i] = key_w<W>(src[i]);
    }
  }
}

// Big rows (new ids big[0..nbig), 'total' entries): the bucket partition
// above.  The host reads the list counters once after each level (usually
// one level); a list that would overflow (keys crowded far beyond what the
// capacity allows) restarts the path with four times the capacity.
template <bool W>
static int relabel_big(gbk_ctx *ctx, int64_t n, const int64_t *off, const int32_t *nbr, const int32_t *wgt,
                       const int32_t *new_ids, const int32_t *inv, const int64_t *deg_new,
                       const int64_t *new_off, const int32_t *big, int64_t nbig, int64_t total,
                       int32_t *nbr_out, int32_t *wgt_out, bool narrow) {
  using K = typename RowKey<W>::T;
  hipStream_t s = ctx->stream;
  int nb = 1;
  while (nb < 31 && (int64_t{1} << nb) < n) nb++;  // every id is below 2^nb
  const int key_bits = nb + (W ? 32 : 0);
  const int64_t chunk_bound = total / kBigChunk + nbig + 1;
  const int64_t bucket_bound = 2 * (total / kSegTarget) + 2 * nbig + 1;
  const int64_t next_cap = total / kSortCap + 16;  // segments above kSortCap entries are disjoint
  auto *lay = static_cast<int64_t *>(scratch(ctx, kS9, sizeof(int64_t) * 3 * (nbig + 1)));
  auto *buf0 = static_cast<K *>(scratch(ctx, kS10, sizeof(K) * total));
  auto *buf1 = static_cast<K *>(scratch(ctx, kS11, sizeof(K) * total));
  auto *cnt = static_cast<unsigned long long *>(scratch(ctx, kS6, sizeof(int64_t) * (bucket_bound + 1)));
  auto *cur = static_cast<unsigned long long *>(scratch(ctx, kS7, sizeof(int64_t) * (bucket_bound + 1)));
  auto *next = static_cast<SegRec *>(scratch(ctx, kS13, sizeof(SegRec) * 2 * next_cap));
  auto *ctr = static_cast<unsigned long long *>(scratch(ctx, kS14, 16 * sizeof(unsigned long long)));
  auto *hc = static_cast<unsigned long long *>(pinned(ctx, 8 * sizeof(unsigned long long)));
  if (!lay || !buf0 || !buf1 || !cnt || !cur || !next || !ctr || !hc)
    return set_error(ctx, GBK_ENOMEM, "relabel big-row scratch");
  int64_t *stage_off = lay, *bucket_off = lay + (nbig + 1), *chunk_off = lay + 2 * (nbig + 1);
  const unsigned chunk_grid = static_cast<unsigned>(std::min<int64_t>(chunk_bound, INT32_MAX));
  for (int64_t cap = bucket_bound + bucket_bound / 2 + 4096;; cap *= 4) {
    auto *lists = static_cast<SegRec *>(scratch(ctx, kS12, sizeof(SegRec) * (kSegCopy + 1) * cap));
    if (!lists) return set_error(ctx, GBK_ENOMEM, "relabel bucket lists");
    // counters: [0..4] sort classes and copy list, [5] overflow, [6]/[7] the
    // two ping-pong lists of segments to split at the next level, [8] the
    // weighted buckets listed for a 64-bit redo
    fill<unsigned long long>(ctx, ctr, 0ull, 9);
    fill<unsigned long long>(ctx, cnt, 0ull, bucket_bound + 1);
    big_row_sizes<<<grid_for(nbig), kBlock, 0, s>>>(big, nbig, deg_new, key_bits, stage_off, bucket_off,
                                                    chunk_off);
    big_layout<<<1, 1024, 0, s>>>(nbig, stage_off, bucket_off, chunk_off);
    big_count<W><<<chunk_grid, kBlock, 0, s>>>(big, nbig, stage_off, bucket_off, chunk_off, inv, off, nbr, wgt,
                                               new_ids, key_bits, buf0, cnt);
    GBK_LAUNCH_CHECK(ctx, "big_count");
    int rc = exclusive_scan_i64(ctx, reinterpret_cast<const int64_t *>(cnt), reinterpret_cast<int64_t *>(cur),
                                bucket_bound);
    if (rc) return rc;
    big_scatter<W><<<chunk_grid, kBlock, 0, s>>>(stage_off, bucket_off, chunk_off, nbig, key_bits, buf0, buf1,
                                                 cur);
    big_segments<<<grid_for(bucket_bound), kBlock, 0, s>>>(bucket_off, stage_off, nbig, key_bits, cnt, cur,
                                                           lists, cap, ctr, next, next_cap, &ctr[6]);
    GBK_LAUNCH_CHECK(ctx, "big_segments");
    int side = 0;  // next[side * next_cap ..] holds the segments of the coming level
    bool overflow = false;
    for (;;) {
      GBK_HIP(ctx, hipMemcpyAsync(hc, ctr, 8 * sizeof(unsigned long long), hipMemcpyDeviceToHost, s));
      GBK_HIP(ctx, hipStreamSynchronize(s));
      if (hc[kSegLists - 1]) {
        overflow = true;
        break;
      }
      const int64_t nsplit = static_cast<int64_t>(hc[6 + side]);
      if (nsplit == 0) break;
      if (nsplit > next_cap) return set_error(ctx, GBK_EINVAL, "relabel: bucket split list overflow");
      fill<unsigned long long>(ctx, &ctr[6 + (side ^ 1)], 0ull, 1);
      seg_split<W><<<static_cast<unsigned>(nsplit), 1024, 0, s>>>(next + side * next_cap, buf0, buf1, lists, cap,
                                                                  ctr, next + (side ^ 1) * next_cap, next_cap,
                                                                  &ctr[6 + (side ^ 1)]);
      GBK_LAUNCH_CHECK(ctx, "seg_split");
      side ^= 1;
    }
    if (overflow) continue;
    const unsigned long long *cts = ctr;
    auto sort_grid = [&](int c) {
      const int64_t k = static_cast<int64_t>(hc[c]);
      return static_cast<unsigned>(std::max<int64_t>(1, std::min<int64_t>((k + kWavesPerBlock - 1) / kWavesPerBlock,
                                                                          8192)));
    };
    // narrow weighted sorts where nb + index bits (8..11) <= 32
    // (room for every record of the four sort lists: a bucket is listed once)
    auto *redo = static_cast<SegRec *>(scratch(ctx, kS17, sizeof(SegRec) * 4 * cap));
    if (!redo) return set_error(ctx, GBK_ENOMEM, "relabel bucket redo list");
    unsigned long long *redo_count = &ctr[8];
    bool any_narrow = false;
#define GBK_SEG(c, E, IB)                                                                                   \
  if (hc[c]) {                                                                                              \
    if (W && narrow && nb + IB <= 32) {                                                                     \
      any_narrow = true;                                                                                    \
      seg_sort<W, E, true><<<sort_grid(c), kBlock, 0, s>>>(lists + c * cap, cts + c, cap, buf0, buf1,       \
                                                           stage_off, nbig, big, new_off, nbr_out, wgt_out, \
                                                           redo, redo_count);                               \
    } else {                                                                                                \
      seg_sort<W, E><<<sort_grid(c), kBlock, 0, s>>>(lists + c * cap, cts + c, cap, buf0, buf1, stage_off,  \
                                                     nbig, big, new_off, nbr_out, wgt_out);                 \
    }                                                                                                       \
  }
    GBK_SEG(0, 4, 8)
    GBK_SEG(1, 8, 9)
    GBK_SEG(2, 16, 10)
    GBK_SEG(3, 32, 11)
#undef GBK_SEG
    if (hc[kSegCopy])
      seg_copy<W><<<static_cast<unsigned>(std::min<unsigned long long>(hc[kSegCopy], 4096)), kBlock, 0, s>>>(
          lists + kSegCopy * cap, cts + kSegCopy, cap, buf0, buf1, stage_off, nbig, big, new_off, nbr_out, wgt_out);
    GBK_LAUNCH_CHECK(ctx, "seg_sort");
    if constexpr (W) {
      if (any_narrow) {
        seg_redo<<<256, kBlock, 0, s>>>(redo, redo_count, buf0, buf1, stage_off, nbig, big, new_off, nbr_out,
                                        wgt_out);
        GBK_LAUNCH_CHECK(ctx, "seg_redo");
      }
    }
    return GBK_OK;
  }
}

// Relabels one CSR (out or in) of 'g' into (off_out, nbr_out, wgt_out).
// narrow: weighted rows may sort as 32-bit composites (rows with distinct
// neighbours, as squished graphs have; the builder's raw rows repeat
// neighbours and sort on 64-bit keys directly).
template <bool W>
static int relabel_csr(gbk_ctx *ctx, int64_t n, const int64_t *off, const int32_t *nbr,
                       const int32_t *wgt, const int32_t *new_ids_d, const int32_t *inv,
                       int64_t *deg_new, int64_t *off_out, int32_t *nbr_out, int32_t *wgt_out,
                       bool degrees_ready, bool narrow = true) {
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
  // [0..kClasses) rows per class, [kClasses] big rows' entries, [kClasses+1]
  // rows listed for a redo
  auto *counts_d = static_cast<unsigned long long *>(scratch(ctx, kS8, 128));
  if (!list || !tile_cnt || !counts_d) return set_error(ctx, GBK_ENOMEM, "relabel bins");
  fill<unsigned long long>(ctx, counts_d, 0ull, kClasses + 2);
  if (n > 0) {
    relabel_bin_count<<<static_cast<unsigned>(ntiles), kBlock, 0, s>>>(deg_new, n, tile, tile_cnt,
                                                                       counts_d + kClasses);
    relabel_bin_offsets<<<1, 1024, 0, s>>>(tile_cnt, ntiles, counts_d);
    relabel_bin_scatter<<<static_cast<unsigned>(ntiles), kBlock, 0, s>>>(deg_new, n, tile, tile_cnt, list);
  }
  GBK_LAUNCH_CHECK(ctx, "relabel bins");
  unsigned long long counts[kClasses + 1];  // rows per class, then the big rows' entries
  GBK_HIP(ctx, hipMemcpyAsync(counts, counts_d, sizeof(counts), hipMemcpyDeviceToHost, s));
  GBK_HIP(ctx, hipStreamSynchronize(s));
  {
    int64_t at = 0;
    for (int c = 0; c < kClasses; c++) {
      bins.l[c] = list + at;
      at += static_cast<int64_t>(counts[c]);
    }
  }
  // Weighted rows sort as 32-bit composites where the ids leave room for the
  // entry index (nb + index bits <= 32); rows with a repeated neighbour are
  // listed in 'redo' and sorted again on 64-bit keys after the classes.
  int nb = 1;
  while (nb < 31 && (int64_t{1} << nb) < n) nb++;
  auto fits = [&](int index_bits) { return W && narrow && nb + index_bits <= 32; };
  int32_t *redo = nullptr;
  unsigned long long *redo_count = counts_d + kClasses + 1;
  if (W) {
    redo = static_cast<int32_t *>(scratch(ctx, kS16, sizeof(int32_t) * (n + 1)));
    if (!redo) return set_error(ctx, GBK_ENOMEM, "relabel redo list");
  }
  bool any_narrow = false;
  // rows of <= 64 entries: lane groups sort in registers
  const int groups[4] = {8, 16, 32, 64};
  for (int c = 0; c < 4; c++) {
    const int64_t cnt = static_cast<int64_t>(counts[c]);
    if (cnt == 0) continue;
    const int64_t blocks = (cnt + kBlock / groups[c] - 1) / (kBlock / groups[c]);
    const bool nw = fits(3 + c);
    any_narrow |= nw;
#define GBK_GROUP(G)                                                                                        \
  if (nw)                                                                                                   \
    relabel_group_sort<W, G, true><<<blocks, kBlock, 0, s>>>(lists[c], cnt, inv, off, nbr, wgt, new_ids_d,  \
                                                             off_out, nbr_out, wgt_out, redo, redo_count);  \
  else                                                                                                      \
    relabel_group_sort<W, G><<<blocks, kBlock, 0, s>>>(lists[c], cnt, inv, off, nbr, wgt, new_ids_d,       \
                                                       off_out, nbr_out, wgt_out);
    switch (groups[c]) {
      case 8: GBK_GROUP(8) break;
      case 16: GBK_GROUP(16) break;
      case 32: GBK_GROUP(32) break;
      default: GBK_GROUP(64) break;
    }
#undef GBK_GROUP
    GBK_LAUNCH_CHECK(ctx, "relabel_group_sort");
  }
  // rows of <= 1024 entries: one wave each, in registers
  for (int c = 4; c < kClassBig; c++) {
    const int64_t cnt = static_cast<int64_t>(counts[c]);
    if (cnt == 0) continue;
    const unsigned blocks = static_cast<unsigned>((cnt + kWavesPerBlock - 1) / kWavesPerBlock);
    const bool nw = fits(3 + c);  // 64 * E entries: 7..10 index bits
    any_narrow |= nw;
#define GBK_WAVE(E)                                                                                         \
  if (nw)                                                                                                   \
    relabel_wave_sort<W, E, true><<<blocks, kBlock, 0, s>>>(lists[c], cnt, inv, off, nbr, wgt, new_ids_d,   \
                                                            off_out, nbr_out, wgt_out, redo, redo_count);   \
  else                                                                                                      \
    relabel_wave_sort<W, E><<<blocks, kBlock, 0, s>>>(lists[c], cnt, inv, off, nbr, wgt, new_ids_d,        \
                                                      off_out, nbr_out, wgt_out);
    switch (c) {
      case 4: GBK_WAVE(2) break;
      case 5: GBK_WAVE(4) break;
      case 6: GBK_WAVE(8) break;
      case 7: GBK_WAVE(16) break;
      default: GBK_WAVE(32) break;
    }
#undef GBK_WAVE
    GBK_LAUNCH_CHECK(ctx, "relabel_wave_sort");
  }
  if (any_narrow) {
    relabel_redo_rows<<<256, kBlock, 0, s>>>(redo, redo_count, inv, off, nbr, wgt, new_ids_d, off_out, nbr_out,
                                             wgt_out);
    GBK_LAUNCH_CHECK(ctx, "relabel_redo_rows");
  }
  const int64_t nbig = static_cast<int64_t>(counts[kClassBig]);
  if (nbig == 0) return GBK_OK;
  return relabel_big<W>(ctx, n, off, nbr, wgt, new_ids_d, inv, deg_new, off_out, lists[kClassBig], nbig,
                        static_cast<int64_t>(counts[kClasses]), nbr_out, wgt_out, narrow);
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
  if (wgt) return relabel_csr<true>(ctx, n, off, nbr, wgt, iota, iota, deg, off_out, nbr_out, wgt_out, true, false);
  return relabel_csr<false>(ctx, n, off, nbr, wgt, iota, iota, deg, off_out, nbr_out, wgt_out, true, false);
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
