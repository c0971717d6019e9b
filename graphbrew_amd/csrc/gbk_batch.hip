// gbk_batch.hip -- many BFS / SSSP sources in one traversal.
//
// The reference runs one source per trial (the trial loop of BenchmarkKernel,
// gapbs/benchmark.h:450, calling bench/src/bfs.cc:285-287 and sssp.cc:342-344
// once per source).  Here up to 64 sources share every edge
// scan, so that one random access to a vertex's state serves all of them:
//
//   BFS   bit-parallel: seen / frontier / next are one 64-bit word per
//         vertex, bit k standing for source k.  A top-down step ORs the
//         frontier word of a row into next[x] of each neighbour x (atomic OR
//         only when it adds bits); a bottom-up step ORs the frontier words of
//         a vertex's in-neighbours until every source has seen it.  Depths
//         equal those of separate BFS runs: bit k of the frontier after d
//         steps is exactly the set of vertices at depth d from source k.
//   SSSP  lane-parallel: the distances of one vertex to the Kp (8..64)
//         sources of a batch are stored together (vertex-major), and a group
//         of Kp lanes relaxes one edge for all sources with one coalesced
//         access.  Buckets of width delta are shared: lane k relaxes from a
//         row only while its own distance lies in the current bucket; an
//         improvement into the current bucket queues the vertex for the next
//         round (per-vertex round tag, one queue entry per vertex), one into
//         a later bucket lowers the vertex's pending-bucket word
//         pend[v] with an atomic minimum.  When the current bucket drains,
//         the next one is the minimum of pend[] (dense scan) and its
//         vertices are queued.  A queued row re-registers the later buckets
//         of its other lanes when it is visited.  Integer distances make the
//         result exact whatever the order, so every source's distances equal
//         a separate delta-stepping run.
//
// Outputs are written vertex-major during the run and transposed through
// LDS tiles at the end into one array per source, while the reached
// vertices and their out-edges are counted per source.
#include <algorithm>
#include <climits>
#include <vector>

#include "gbk_expand.h"

namespace gbk {

constexpr int kBatchMax = 64;
constexpr int kMsLevelsPerBlock = 4;   // BFS batch levels queued per host check
constexpr int kBcbLevelsPerBlock = 4;  // BC batch forward levels queued per host check
constexpr int32_t kBatchInf = INT32_MAX / 2;  // sssp.cc:67
constexpr int64_t kBatchHeavy = 256;          // rows above this are chunked over groups
constexpr int64_t kBatchChunk = 256;
// BC batch: edges in flight per trip, and the waves per SIMD its kernels'
// registers are capped for (64 lanes: fewer registers, more waves measured
// faster; narrower groups keep their registers).
template <int Kp> constexpr int kBcbUnroll = Kp == 64 ? 4 : Kp <= 8 ? 16 : 8;
template <int Kp> constexpr int kBcbWaves = Kp == 64 ? 8 : 1;
constexpr int kRelaxUnroll = 4;   // edges in flight per trip (SSSP batch: fewer registers, more waves)

inline int batch_width(int k) { return k <= 8 ? 8 : k <= 16 ? 16 : k <= 32 ? 32 : 64; }

// ------------------------------------------------------------------ shared
// A wave-uniform 64-bit value moved to scalar registers.
__device__ __forceinline__ int64_t uniform64(int64_t v) {
  const uint32_t lo = static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(static_cast<int>(v & 0xffffffffll)));
  const uint32_t hi = static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(static_cast<int>(v >> 32)));
  return static_cast<int64_t>((static_cast<uint64_t>(hi) << 32) | lo);
}

template <int Kp>
__device__ __forceinline__ int32_t group_min(int32_t v) {
#pragma unroll
  for (int o = Kp / 2; o >= 1; o >>= 1) {
    const int32_t w = __shfl_xor(v, o);
    v = w < v ? w : v;
  }
  return v;
}

template <int Kp>
__device__ __forceinline__ bool group_any(bool p) {
  const uint64_t m = __ballot(p);
  if (Kp == 64) return m != 0;
  const int shift = (lane_id() / Kp) * Kp;
  return ((m >> shift) & ((1ull << Kp) - 1)) != 0;
}

// Transposes a vertex-major [n][Kp] int32 array into K source-major arrays
// (out may be null) and counts, per source, the vertices whose value is not
// 'unreached' and the sum of their out-degrees: partial[b * 64 + k] and
// partial[(grid + b) * 64 + k] for workgroup b.
__global__ void __launch_bounds__(kBlock)
batch_collect(const int32_t *vm, int64_t n, int Kp, int K, int32_t unreached, const int64_t *off,
              int32_t *out, unsigned long long *partial) {
  __shared__ int32_t tile[64][65];
  __shared__ int64_t degs[64];
  __shared__ unsigned long long cnt[2][kBatchMax];
  if (threadIdx.x < kBatchMax) {
    cnt[0][threadIdx.x] = 0;
    cnt[1][threadIdx.x] = 0;
  }
  const int w = threadIdx.x / kWave;
  for (int64_t t0 = static_cast<int64_t>(blockIdx.x) * 64; t0 < n; t0 += static_cast<int64_t>(gridDim.x) * 64) {
    __syncthreads();
    for (int i = threadIdx.x; i < 64 * Kp; i += kBlock) {
      const int64_t v = t0 + i / Kp;
      tile[i / Kp][i % Kp] = v < n ? vm[v * Kp + i % Kp] : unreached;
    }
    if (threadIdx.x < 64) {
      const int64_t v = t0 + threadIdx.x;
      degs[threadIdx.x] = v < n ? off[v + 1] - off[v] : 0;
    }
    __syncthreads();
    // wave w handles sources w, w+4, ...; lane = vertex of the tile
    for (int k = w; k < K; k += kWavesPerBlock) {
      const int vv = lane_id();
      const int64_t v = t0 + vv;
      const int32_t val = tile[vv][k];
      const bool ok = v < n && val != unreached;
      if (v < n && out) out[static_cast<int64_t>(k) * n + v] = val;
      const unsigned long long r = wave_sum(static_cast<unsigned long long>(ok ? 1 : 0));
      const unsigned long long e = wave_sum(static_cast<unsigned long long>(ok ? degs[vv] : 0));
      if (lane_id() == 0) {
        cnt[0][k] += r;
        cnt[1][k] += e;
      }
    }
  }
  __syncthreads();
  if (threadIdx.x < kBatchMax) {
    partial[static_cast<int64_t>(blockIdx.x) * kBatchMax + threadIdx.x] = cnt[0][threadIdx.x];
    partial[static_cast<int64_t>(gridDim.x + blockIdx.x) * kBatchMax + threadIdx.x] = cnt[1][threadIdx.x];
  }
}

// The counts of batch_collect without the transposition: one streaming pass
// over the vertex-major array, 16 bytes (four sources of one vertex; Kp is a
// multiple of 4) per lane and load, four loads in flight.  The grid stride
// is a multiple of Kp, so a thread always reads the same four sources'
// columns; its counts meet the workgroup's in LDS once, at the end.
__global__ void __launch_bounds__(kBlock)
batch_count(const int32_t *vm, int64_t n, int log_kp, int32_t unreached, const int64_t *off,
            unsigned long long *partial) {
  __shared__ unsigned long long cnt[2][kBatchMax];
  if (threadIdx.x < kBatchMax) {
    cnt[0][threadIdx.x] = 0;
    cnt[1][threadIdx.x] = 0;
  }
  __syncthreads();
  constexpr int U = 4;
  unsigned long long reached[4] = {}, edges[4] = {};
  const int4 *v4 = reinterpret_cast<const int4 *>(vm);
  const int64_t total = (n << log_kp) >> 2;  // 16-byte items
  const int log_q = log_kp - 2;              // items per vertex = Kp / 4
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t i0 = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i0 < total; i0 += U * stride) {
    int4 x[U];
#pragma unroll
    for (int j = 0; j < U; j++) {
      const int64_t i = i0 + j * stride;
      x[j] = i < total ? v4[i] : make_int4(unreached, unreached, unreached, unreached);
    }
#pragma unroll
    for (int j = 0; j < U; j++) {
      const bool r[4] = {x[j].x != unreached, x[j].y != unreached, x[j].z != unreached, x[j].w != unreached};
      if (!(r[0] || r[1] || r[2] || r[3])) continue;
      const int64_t v = (i0 + j * stride) >> log_q;
      const auto deg = static_cast<unsigned long long>(off[v + 1] - off[v]);
#pragma unroll
      for (int q = 0; q < 4; q++) {
        reached[q] += r[q];
        edges[q] += r[q] ? deg : 0ull;
      }
    }
  }
  const int k0 = static_cast<int>((threadIdx.x << 2) & ((1u << log_kp) - 1));
#pragma unroll
  for (int q = 0; q < 4; q++) {
    if (reached[q]) {
      atomicAdd(&cnt[0][k0 + q], reached[q]);
      atomicAdd(&cnt[1][k0 + q], edges[q]);
    }
  }
  __syncthreads();
  if (threadIdx.x < kBatchMax) {
    partial[static_cast<int64_t>(blockIdx.x) * kBatchMax + threadIdx.x] = cnt[0][threadIdx.x];
    partial[static_cast<int64_t>(gridDim.x + blockIdx.x) * kBatchMax + threadIdx.x] = cnt[1][threadIdx.x];
  }
}

// One workgroup of 1024 lanes: sums the collect partials per source.
__global__ void batch_collect_sum(const unsigned long long *partial, int grid, unsigned long long *sums) {
  __shared__ unsigned long long red[2][16][kBatchMax];
  const int k = threadIdx.x % kBatchMax, part = threadIdx.x / kBatchMax;  // 16 parts
  unsigned long long r = 0, e = 0;
  for (int b = part; b < grid; b += 16) {
    r += partial[static_cast<int64_t>(b) * kBatchMax + k];
    e += partial[static_cast<int64_t>(grid + b) * kBatchMax + k];
  }
  red[0][part][k] = r;
  red[1][part][k] = e;
  __syncthreads();
  if (threadIdx.x < kBatchMax) {
    unsigned long long a = 0, b2 = 0;
    for (int p = 0; p < 16; p++) {
      a += red[0][p][threadIdx.x];
      b2 += red[1][p][threadIdx.x];
    }
    sums[threadIdx.x] = a;
    sums[kBatchMax + threadIdx.x] = b2;
  }
}

// ------------------------------------------------------------------- BFS
// Level state of a BFS batch, on the device: the host queues levels blind
// and reads it once per block of levels.
struct MsCounters {
  unsigned long long tail;
  unsigned long long fverts;  // vertices in the new frontier
  unsigned long long fedges;  // their out-degree sum
  int overflow;
  int done;                   // the last level found nothing: later launches return at once
  int dir;                    // this level: 0 top-down, 1 bottom-up
  int level;                  // depth this level assigns
  int levels;                 // levels executed
  int pad;
};

__global__ void msbfs_sources(const int32_t *src, int K, int Kp, unsigned long long *seen,
                              unsigned long long *front, int32_t *depth_vm, MsCounters *c) {
  const int j = threadIdx.x;
  if (j < K) {
    const int64_t s = src[j];
    atomicOr(&seen[s], 1ull << j);
    atomicOr(&front[s], 1ull << j);
    depth_vm[s * Kp + j] = 0;
  }
  if (j == 0) {
    c->tail = 0;
    c->fverts = 0;
    c->fedges = 0;
    c->overflow = 0;
    c->done = 0;
    c->dir = 0;
    c->level = 1;
    c->levels = 0;
  }
}

__device__ __forceinline__ bool ms_skip(const MsCounters *c, int dir) {
  const volatile MsCounters *v = c;
  return v->done || v->dir != dir;
}

__global__ void __launch_bounds__(kBlock)
msbfs_front_queue(const unsigned long long *front, int64_t n, int32_t *queue, int64_t cap, MsCounters *c) {
  if (ms_skip(c, 0)) return;
  __shared__ int32_t seg[kWavesPerBlock * kSeg];
  WaveQueue q;
  q.init(seg + (threadIdx.x / kWave) * kSeg, &c->tail, queue, cap, &c->overflow);
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t base = static_cast<int64_t>(blockIdx.x) * blockDim.x; base < n; base += stride) {
    const int64_t v = base + threadIdx.x;
    q.push(v < n && front[v] != 0, static_cast<int32_t>(v));
  }
  q.flush_block();
}

struct MsTdEdge {
  const int32_t *nbr;
  const unsigned long long *front;
  const unsigned long long *seen;
  unsigned long long *next;
  __device__ void operator()(int64_t r, int64_t k, bool active) {
    if (!active) return;
    const int32_t x = nbr[k];
    const unsigned long long add = front[r] & ~seen[x];
    if (add & ~next[x]) atomicOr(&next[x], add);
  }
};

__global__ void __launch_bounds__(kBlock)
msbfs_td(const int64_t *off, const int32_t *nbr, const int32_t *queue, const MsCounters *c,
         const unsigned long long *front, const unsigned long long *seen, unsigned long long *next,
         int64_t heavy_deg, HeavyQueue hq) {
  if (ms_skip(c, 0)) return;
  MsTdEdge f{nbr, front, seen, next};
  auto rowlen = [&](int64_t u, int64_t &s, int64_t &l) {
    s = off[u];
    l = off[u + 1] - s;
  };
  expand_list_light(queue, static_cast<int64_t>(c->tail), heavy_deg, rowlen, f, hq);
}

__global__ void __launch_bounds__(kBlock)
msbfs_td_heavy(const int32_t *nbr, const unsigned long long *front, const unsigned long long *seen,
               unsigned long long *next, HeavyQueue hq, const MsCounters *c) {
  if (ms_skip(c, 0)) return;
  MsTdEdge f{nbr, front, seen, next};
  expand_heavy(hq, f);
}

// Bottom-up: G lanes per vertex OR the frontier words of its in-neighbours
// until every source has reached it; only this group writes next[u].
template <int G>
__global__ void __launch_bounds__(kBlock)
msbfs_bu(const int64_t *in_off, const int32_t *in_nbr, int64_t n, const unsigned long long *seen,
         const unsigned long long *front, unsigned long long *next, unsigned long long full, int swz,
         const MsCounters *c) {
  if (ms_skip(c, 1)) return;
  const int glane = lane_id() % G;
  const int64_t groups_per_block = kBlock / G;
  const int64_t tiles = (n + groups_per_block - 1) / groups_per_block;
  for (int64_t t = swizzled_block(swz); t < tiles; t += gridDim.x) {
    const int64_t u = t * groups_per_block + threadIdx.x / G;
    unsigned long long sv = full;
    int64_t s = 0, e = 0;
    if (u < n) {
      sv = seen[u];
      s = in_off[u];
      e = in_off[u + 1];
    }
    bool todo = (sv & full) != full && s < e;
    const bool any = todo;
    unsigned long long acc = 0;
    for (int64_t base = s; todo; base += G) {  // group-uniform loop
      const int64_t k = base + glane;
      unsigned long long x = k < e ? front[in_nbr[k]] : 0ull;
#pragma unroll
      for (int o = G / 2; o >= 1; o >>= 1) x |= __shfl_xor(x, o);
      acc |= x;
      if (((acc | sv) & full) == full || base + G >= e) todo = false;
    }
    if (any && glane == 0) {
      const unsigned long long add = acc & ~sv & full;
      if (add) next[u] = add;
    }
  }
}

__global__ void msbfs_finish(const int64_t *off, int64_t n, int Kp, unsigned long long *seen,
                             unsigned long long *front, unsigned long long *next, int32_t *depth_vm,
                             const MsCounters *c, unsigned long long *p_verts, unsigned long long *p_edges) {
  if (reinterpret_cast<const volatile MsCounters *>(c)->done) return;
  const int32_t level = reinterpret_cast<const volatile MsCounters *>(c)->level;
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  const int lane = lane_id();
  const int G = kWave / Kp;  // depth rows one store instruction covers
  unsigned long long verts = 0, edges = 0;
  // the loop bound is wave-uniform (every wave's 64 vertices start at a
  // multiple of 64), so the shuffles below run on whole waves
  const int64_t nr = (n + kWave - 1) / kWave * kWave;
  for (int64_t v = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; v < nr; v += stride) {
    unsigned long long add = 0;
    if (v < n) {
      const unsigned long long nx = next[v];
      const unsigned long long sv = seen[v];
      add = nx & ~sv;
      if (front[v] != add) front[v] = add;
      if (nx) next[v] = 0;
      if (add) {
        seen[v] = sv | add;
        verts++;
        edges += static_cast<unsigned long long>(off[v + 1] - off[v]);
      }
    }
    // depth rows of the wave's new vertices: lane group g of Kp lanes
    // writes row v0 + j0 + g (lane k: source k), one coalesced masked
    // store per G rows instead of one scattered store per new source
    const uint64_t any = __ballot(add != 0);
    if (any == 0) continue;
    const int64_t v0 = v - lane;
    const int k = lane % Kp;
    for (int j0 = 0; j0 < kWave; j0 += G) {
      const uint64_t grp = (G == kWave ? ~0ull : ((1ull << G) - 1)) << j0;
      if ((any & grp) == 0) continue;  // wave-uniform
      const int j = j0 + lane / Kp;
      const unsigned long long aj = __shfl(add, j);
      if ((aj >> k) & 1ull) depth_vm[(v0 + j) * Kp + k] = level;
    }
  }
  block_partial(verts, p_verts);
  block_partial(edges, p_edges);
}

// Ends a level: the new frontier's size and out-edges, then the next
// level's direction (bfs.cc:231-262: bottom-up while the frontier's
// out-edges exceed m / alpha) or the end of the traversal.
__global__ void msbfs_store(const unsigned long long *partial, int count, MsCounters *c, long long bu_edges) {
  if (reinterpret_cast<volatile MsCounters *>(c)->done) return;
  __shared__ unsigned long long red[2][1024 / kWave];
  unsigned long long a = 0, b = 0;
  for (int i = threadIdx.x; i < count; i += blockDim.x) {
    a += partial[i];
    b += partial[kGridCap + i];
  }
  a = wave_sum(a);
  b = wave_sum(b);
  if (lane_id() == 0) {
    red[0][threadIdx.x / kWave] = a;
    red[1][threadIdx.x / kWave] = b;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long x = 0, y = 0;
    for (int i = 0; i < static_cast<int>(blockDim.x) / kWave; i++) {
      x += red[0][i];
      y += red[1][i];
    }
    c->fverts = x;
    c->fedges = y;
    c->tail = 0;
    c->levels += 1;
    if (x == 0) {
      c->done = 1;
    } else {
      c->level += 1;
      c->dir = static_cast<long long>(y) > bu_edges ? 1 : 0;
    }
  }
}

static void launch_msbfs_bu(gbk_ctx *ctx, int G, const gbk_graph *g, const unsigned long long *seen,
                            const unsigned long long *front, unsigned long long *next,
                            unsigned long long full, unsigned grid, const MsCounters *c) {
  const int swz = ctx->tune.xcd_swizzle;
  hipStream_t s = ctx->stream;
  switch (G) {
    case 1: msbfs_bu<1><<<grid, kBlock, 0, s>>>(g->in_off, g->in_nbr, g->n, seen, front, next, full, swz, c); break;
    case 2: msbfs_bu<2><<<grid, kBlock, 0, s>>>(g->in_off, g->in_nbr, g->n, seen, front, next, full, swz, c); break;
    case 4: msbfs_bu<4><<<grid, kBlock, 0, s>>>(g->in_off, g->in_nbr, g->n, seen, front, next, full, swz, c); break;
    case 16: msbfs_bu<16><<<grid, kBlock, 0, s>>>(g->in_off, g->in_nbr, g->n, seen, front, next, full, swz, c); break;
    case 32: msbfs_bu<32><<<grid, kBlock, 0, s>>>(g->in_off, g->in_nbr, g->n, seen, front, next, full, swz, c); break;
    case 64: msbfs_bu<64><<<grid, kBlock, 0, s>>>(g->in_off, g->in_nbr, g->n, seen, front, next, full, swz, c); break;
    default: msbfs_bu<8><<<grid, kBlock, 0, s>>>(g->in_off, g->in_nbr, g->n, seen, front, next, full, swz, c); break;
  }
}

// ------------------------------------------------------------------ SSSP
struct BatchState {
  unsigned long long near_cnt[2];
  long long bucket;
  long long next_bucket;
  long long relax_rounds;
  long long buckets;
  unsigned long long rows_scanned;   // queue entries relaxed, over all rounds
  unsigned long long edges_scanned;  // their out-edges walked (each for every source)
  long long steps;                   // steps that found the search still open
  int cur, phase, done, overflow;
};

// The state's scratch: the state, then (256-byte aligned) the spread counter
// of edges walked in the current round.
constexpr size_t kBatchStateBytes = (sizeof(BatchState) + 64 + 255) / 256 * 256;

__global__ void sb_init_state(BatchState *st, long long first, unsigned long long *scan_slots) {
  if (threadIdx.x == 0) {
    for (int i = 0; i < kSpread; i++) scan_slots[i * kSpreadStride] = 0;
    st->near_cnt[0] = static_cast<unsigned long long>(first);
    st->near_cnt[1] = 0;
    st->bucket = 0;
    st->next_bucket = 0;
    st->relax_rounds = 0;
    st->buckets = 0;
    st->rows_scanned = 0;
    st->edges_scanned = 0;
    st->steps = 0;
    st->cur = 0;
    st->phase = 0;
    st->done = 0;
    st->overflow = 0;
  }
}

__global__ void sb_sources(const int32_t *src, int K, int Kp, int32_t *dist) {
  const int j = threadIdx.x;
  if (j < K) dist[static_cast<int64_t>(src[j]) * Kp + j] = 0;
}

// Per-lane relaxation context of a group of Kp lanes (lane k = source k).
template <int Kp>
struct GroupRelax {
  const int32_t *nbr;
  const int32_t *wgt;
  int32_t *dist;
  int32_t *pend;
  uint32_t *vtag;
  WaveQueue nearq;
  int64_t delta, lower, upper;
  uint32_t tag;
  int k;
  bool leader;

  __device__ void setup(int32_t *seg, const int32_t *nb, const int32_t *wg, int32_t *d, int32_t *p,
                        uint32_t *vt, int32_t *near0, int32_t *near1, int64_t cap, int64_t dl,
                        BatchState *st) {
    const volatile BatchState *vs = st;
    const int cur = vs->cur;
    nbr = nb;
    wgt = wg;
    dist = d;
    pend = p;
    vtag = vt;
    nearq.init(seg + (threadIdx.x / kWave) * kSeg, &st->near_cnt[1 - cur], cur ? near0 : near1, cap,
               &st->overflow);
    delta = dl;
    lower = dl * vs->bucket;
    upper = lower + dl;
    tag = static_cast<uint32_t>(vs->relax_rounds + 1);
    k = lane_id() % Kp;
    leader = k == 0;
  }

  // Lane k's distance of a row and whether it relaxes from it; a later
  // bucket of this lane is registered in pend[row] (by the group leader).
  __device__ void row_lane(int64_t row, bool valid, bool reg, int64_t &du, bool &act) {
    du = valid ? dist[row * Kp + k] : kBatchInf;
    act = valid && du >= lower && du < upper;
    const int32_t fut = (valid && reg && du >= upper && du < kBatchInf) ? static_cast<int32_t>(du / delta) : INT32_MAX;
    const int32_t f = group_min<Kp>(fut);
    if (leader && valid && f != INT32_MAX && f < pend[row]) atomicMin(&pend[row], f);
  }

  // Relaxes edges [s, s + len) of a row for every source of the group;
  // len is group-uniform, the trip count wave-uniform.
  __device__ void walk(int64_t s, int64_t len, int64_t du, bool act) {
    if constexpr (Kp == kWave) {
      walk_wave(s, len, du, act);
      return;
    }
    const int64_t trips = wave_max(len);
    for (int64_t j0 = 0; j0 < trips; j0 += kRelaxUnroll) {
      int32_t x[kRelaxUnroll], w[kRelaxUnroll], dv[kRelaxUnroll];
      bool valid[kRelaxUnroll];
#pragma unroll
      for (int j = 0; j < kRelaxUnroll; j++) {
        valid[j] = j0 + j < len;
        const int64_t e = valid[j] ? s + j0 + j : 0;
        x[j] = nbr[e];
        w[j] = wgt[e];
      }
#pragma unroll
      for (int j = 0; j < kRelaxUnroll; j++) dv[j] = dist[static_cast<int64_t>(x[j]) * Kp + k];
#pragma unroll
      for (int j = 0; j < kRelaxUnroll; j++) {
        bool here = false;
        int32_t fut = INT32_MAX;
        const int64_t cand = du + w[j];
        // Judged against the loaded distance (never below the current one),
        // so every real improvement queues x; a lane that lost a race to a
        // lower value queues x needlessly, which costs a visit, not
        // correctness.  The atomic's result is not waited for.
        const bool imp = valid[j] && act && cand < dv[j];
        if (__ballot(imp) == 0) continue;  // nothing improved in the wave
        if (imp) {
          atomicMin(&dist[static_cast<int64_t>(x[j]) * Kp + k], static_cast<int32_t>(cand));
          if (cand < upper) here = true;
          else fut = static_cast<int32_t>(cand / delta);
        }
        const bool anyh = group_any<Kp>(here);
        const int32_t f = group_min<Kp>(fut);
        bool push = false;
        if (leader && valid[j]) {
          if (anyh) push = atomicExch(&vtag[x[j]], tag) != tag;
          if (f != INT32_MAX && f < pend[x[j]]) atomicMin(&pend[x[j]], f);
        }
        nearq.push(push, x[j]);
      }
    }
  }

  // Kp == 64: the wave is one group, so the row and its edge range are
  // wave-uniform.  Each 64-edge slice of neighbour ids and weights comes in
  // with one coalesced load and is broadcast edge by edge (readlane); the
  // distance rows of kRelaxUnroll neighbours are in flight together; the
  // cross-lane reductions and the leader's atomics run only for edges on
  // which some lane improved.
  __device__ void walk_wave(int64_t s, int64_t len, int64_t du, bool act) {
    s = uniform64(s);
    len = uniform64(len);
    for (int64_t c0 = 0; c0 < len; c0 += kWave) {
      const int cnt = static_cast<int>(len - c0 < kWave ? len - c0 : kWave);
      const bool mine = k < cnt;
      const int32_t vx = mine ? nbr[s + c0 + k] : 0;
      const int32_t vw = mine ? wgt[s + c0 + k] : 0;
      for (int j0 = 0; j0 < cnt; j0 += kRelaxUnroll) {
        int32_t x[kRelaxUnroll], w[kRelaxUnroll], dv[kRelaxUnroll];
#pragma unroll
        for (int j = 0; j < kRelaxUnroll; j++) {
          const int at = j0 + j < kWave ? j0 + j : kWave - 1;  // lanes past cnt hold vertex 0
          x[j] = __builtin_amdgcn_readlane(vx, at);
          w[j] = __builtin_amdgcn_readlane(vw, at);
        }
#pragma unroll
        for (int j = 0; j < kRelaxUnroll; j++) dv[j] = dist[static_cast<int64_t>(x[j]) * Kp + k];
#pragma unroll
        for (int j = 0; j < kRelaxUnroll; j++) {
          const int64_t cand = du + w[j];
          // judged against the loaded distance, as in walk()
          // (non-short-circuit: the loads stay unconditional, not sunk under act)
          const bool imp = (j0 + j < cnt) & act & (cand < dv[j]);
          const uint64_t im = __ballot(imp);
          if (im == 0) continue;
          if (imp) atomicMin(&dist[static_cast<int64_t>(x[j]) * Kp + k], static_cast<int32_t>(cand));
          const bool here = imp && cand < upper;
          const uint64_t hm = __ballot(here);
          if (hm != im) {  // some improvement lies in a later bucket
            const int32_t f = group_min<Kp>(imp && !here ? static_cast<int32_t>(cand / delta) : INT32_MAX);
            if (leader && f < pend[x[j]]) atomicMin(&pend[x[j]], f);
          }
          if (hm) {
            bool push = false;
            if (leader) push = atomicExch(&vtag[x[j]], tag) != tag;
            nearq.push(push, x[j]);
          }
        }
      }
    }
  }
};

template <int Kp>
__global__ void __launch_bounds__(kBlock, 8)
sb_relax(const int64_t *off, const int32_t *nbr, const int32_t *wgt, int32_t *dist, int32_t *pend,
         uint32_t *vtag, int32_t *near0, int32_t *near1, int64_t cap, int64_t delta, BatchState *st,
         unsigned long long *scan_slots, HeavyQueue hq) {
  const volatile BatchState *vs = st;
  if (vs->done || vs->phase != 0) return;
  __shared__ int32_t seg[kWavesPerBlock * kSeg];
  constexpr int G = kWave / Kp;
  const int grp = lane_id() / Kp;
  const int cur = vs->cur;
  // the tail counts past cap only when a push overflowed (reported by the host)
  const int64_t count = min(static_cast<int64_t>(vs->near_cnt[cur]), cap);
  const int32_t *list = cur ? near1 : near0;
  GroupRelax<Kp> f;
  f.setup(seg, nbr, wgt, dist, pend, vtag, near0, near1, cap, delta, st);
  const int64_t waves = static_cast<int64_t>(gridDim.x) * kWavesPerBlock;
  const int64_t gw = static_cast<int64_t>(blockIdx.x) * kWavesPerBlock + threadIdx.x / kWave;
  // edges walked, counted in LDS (a register held across the loop spills)
  __shared__ unsigned long long walked;
  if (threadIdx.x == 0) walked = 0;
  __syncthreads();
  for (int64_t base = gw * G; base < count; base += waves * G) {
    const int64_t i = base + grp;
    const bool valid = i < count;
    int64_t row = 0, s = 0, len = 0;
    if (valid) {
      row = list[i];
      s = off[row];
      len = off[row + 1] - s;
    }
    const bool heavy = len > kBatchHeavy;
    if (heavy && f.leader) {
      heavy_push(hq, static_cast<int32_t>(row), s, len);
    }
    int64_t du;
    bool act;
    f.row_lane(row, valid, !heavy, du, act);
    f.walk(s, heavy ? 0 : len, du, act);
    if (f.leader && !heavy) atomicAdd(&walked, static_cast<unsigned long long>(len));
  }
  f.nearq.flush_block();
  __syncthreads();
  if (threadIdx.x == 0 && walked) atomicAdd(&scan_slots[(blockIdx.x % kSpread) * kSpreadStride], walked);
}

template <int Kp>
__global__ void __launch_bounds__(kBlock, 8)
sb_relax_heavy(const int32_t *nbr, const int32_t *wgt, int32_t *dist, int32_t *pend, uint32_t *vtag,
               int32_t *near0, int32_t *near1, int64_t cap, int64_t delta, BatchState *st,
               unsigned long long *scan_slots, HeavyQueue hq) {
  const volatile BatchState *vs = st;
  if (vs->done || vs->phase != 0) return;
  int64_t rows, total;
  heavy_extent(hq, rows, total);
  if (rows == 0) return;
  __shared__ int32_t seg[kWavesPerBlock * kSeg];
  constexpr int G = kWave / Kp;
  const int grp = lane_id() / Kp;
  GroupRelax<Kp> f;
  f.setup(seg, nbr, wgt, dist, pend, vtag, near0, near1, cap, delta, st);
  const int64_t stride = static_cast<int64_t>(gridDim.x) * kWavesPerBlock * G;
  const int64_t gw = static_cast<int64_t>(blockIdx.x) * kWavesPerBlock + threadIdx.x / kWave;
  // edges walked, counted in LDS (a register held across the loop spills)
  __shared__ unsigned long long walked;
  if (threadIdx.x == 0) walked = 0;
  __syncthreads();
  for (int64_t c0 = gw * G; c0 < total; c0 += stride) {
    const int64_t c = c0 + grp;
    const bool valid = c < total;
    int64_t row = 0, first = 0, last = 0, start = 0;
    if (valid) {
      int64_t lo = 0, hi = rows;  // last row with prefix[row] <= c
      while (hi - lo > 1) {
        const int64_t mid = (lo + hi) >> 1;
        if (hq.prefix[mid] <= c) lo = mid;
        else hi = mid;
      }
      row = hq.row[lo];
      start = hq.start[lo];
      first = (c - hq.prefix[lo]) * hq.chunk;
      last = min(hq.len[lo], first + hq.chunk);
    }
    int64_t du;
    bool act;
    f.row_lane(row, valid, first == 0, du, act);
    f.walk(start + first, valid ? last - first : 0, du, act);
    if (f.leader && valid) atomicAdd(&walked, static_cast<unsigned long long>(last - first));
  }
  f.nearq.flush_block();
  __syncthreads();
  if (threadIdx.x == 0 && walked) atomicAdd(&scan_slots[(blockIdx.x % kSpread) * kSpreadStride], walked);
}

// One wave: closes a round (and folds the round's walked edges into the
// state).
__global__ void sb_relax_done(BatchState *st, unsigned long long *heavy_count, unsigned long long *scan_slots) {
  if (st->done || st->phase != 0) return;
  const unsigned long long walked = wave_take_spread(scan_slots);
  if (threadIdx.x != 0) return;
  st->steps += 1;
  st->edges_scanned += walked;
  *heavy_count = 0;
  const int cur = st->cur;
  st->rows_scanned += st->near_cnt[cur];
  st->near_cnt[cur] = 0;
  st->cur = 1 - cur;
  st->relax_rounds += 1;
  if (st->near_cnt[1 - cur] == 0) st->phase = 1;
}

__global__ void sb_min_pend(const int32_t *pend, int64_t n, const BatchState *st, unsigned long long *partial) {
  const volatile BatchState *vs = st;
  if (vs->done || vs->phase != 1) return;
  int32_t best = INT32_MAX;
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t v = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; v < n; v += stride) {
    const int32_t p = pend[v];
    best = p < best ? p : best;
  }
  best = group_min<kWave>(best);
  __shared__ int32_t red[kWavesPerBlock];
  if (lane_id() == 0) red[threadIdx.x / kWave] = best;
  __syncthreads();
  if (threadIdx.x == 0) {
    int32_t m = red[0];
    for (int i = 1; i < kWavesPerBlock; i++) m = red[i] < m ? red[i] : m;
    partial[blockIdx.x] = static_cast<unsigned long long>(m);
  }
}

__global__ void sb_min_done(const unsigned long long *partial, int count, BatchState *st) {
  if (st->done || st->phase != 1) return;
  __shared__ unsigned long long red[1024 / kWave];
  unsigned long long m = INT32_MAX;
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
    st->next_bucket = static_cast<long long>(m);
    if (m >= static_cast<unsigned long long>(INT32_MAX)) st->done = 1;
    else st->phase = 2;
  }
}

__global__ void __launch_bounds__(kBlock)
sb_split(int32_t *pend, int64_t n, int32_t *near0, int32_t *near1, int64_t cap, BatchState *st) {
  const volatile BatchState *vs = st;
  if (vs->done || vs->phase != 2) return;
  __shared__ int32_t seg[kWavesPerBlock * kSeg];
  const int cur = vs->cur;
  const int32_t nb = static_cast<int32_t>(vs->next_bucket);
  WaveQueue q;
  q.init(seg + (threadIdx.x / kWave) * kSeg, &st->near_cnt[cur], cur ? near1 : near0, cap, &st->overflow);
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t base = static_cast<int64_t>(blockIdx.x) * blockDim.x; base < n; base += stride) {
    const int64_t v = base + threadIdx.x;
    bool take = false;
    if (v < n && pend[v] == nb) {
      pend[v] = INT32_MAX;
      take = true;
    }
    q.push(take, static_cast<int32_t>(v));
  }
  q.flush_block();
}

__global__ void sb_split_done(BatchState *st) {
  if (threadIdx.x != 0 || st->done || st->phase != 2) return;
  st->bucket = st->next_bucket;
  st->buckets += 1;
  st->phase = 0;
}

// ------------------------------------------------------------------ BC
// Brandes from up to 64 sources at once (bc.cc:71-231 per source).  Lane k
// of a group holds source k; depth, path count and the backward coefficient
// of a vertex are stored for all sources together (vertex-major), so one
// random access of a neighbour serves the whole group.  The union frontier
// of level d lists every vertex that some source reaches at depth d (once:
// a per-vertex level tag de-duplicates the pushes), and a row's lanes act
// only where the vertex is at depth d for their source.  A lane's successor
// test is depth[v] == depth[u] + 1, the relation the reference's successor
// bitmap records (bc.cc:104-112), so no per-source bitmap is kept.
// Forward state of a BC batch, on the device: the host queues levels blind
// and reads it once per block of levels.  Level d's frontier is
// queue[lb, le); lvl[d] holds the level starts (lvl[0] = 0).
struct BcbCounters {
  unsigned long long tail;
  unsigned long long edges_scanned;
  unsigned long long listed_edges;  // out-edges of the level lists after the first
  unsigned long long open_edges;    // in-edges of vertices some source has not reached
  long long lb, le;                 // the current level's list
  long long listed;                 // listed_edges when the current level began
  long long weighed;                // its frontier edges x bc_pull_ratio
  int overflow;
  int d;                            // current level
  int done;                         // the last level listed nothing: later launches return at once
  int pull;                         // this level runs bottom-up
  int need_count;                   // the direction waits for the open-edge count
  int levels;                       // levels executed
};

__device__ __forceinline__ bool bcb_skip(const BcbCounters *c, int pull) {
  const volatile BcbCounters *v = c;
  return v->done || v->need_count || v->pull != pull;
}

template <int Kp>
__device__ __forceinline__ double group_sum_d(double v) {
#pragma unroll
  for (int o = Kp / 2; o >= 1; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

// Calls f(x, on) on batches of kBcbUnroll<Kp> edges of [s, s + len): for
// Kp == 64 the row is wave-uniform and 64 ids come in with one coalesced
// load (broadcast by readlane), otherwise every group walks its own row.
template <int Kp, class F>
__device__ void bcb_edges(const int32_t *nbr, int64_t s, int64_t len, F &f) {
  constexpr int U = kBcbUnroll<Kp>;
  int32_t x[U];
  bool on[U];
  if constexpr (Kp == kWave) {
    s = uniform64(s);
    len = uniform64(len);
    const int lane = lane_id();
    for (int64_t c0 = 0; c0 < len; c0 += kWave) {
      const int cnt = static_cast<int>(len - c0 < kWave ? len - c0 : kWave);
      const int32_t vx = lane < cnt ? nbr[s + c0 + lane] : 0;
      for (int j0 = 0; j0 < cnt; j0 += U) {
#pragma unroll
        for (int j = 0; j < U; j++) {
          const int at = j0 + j < kWave ? j0 + j : kWave - 1;  // lanes past cnt hold vertex 0
          x[j] = __builtin_amdgcn_readlane(vx, at);
          on[j] = j0 + j < cnt;
        }
        f(x, on);
      }
    }
  } else {
    const int64_t trips = wave_max(len);
    for (int64_t j0 = 0; j0 < trips; j0 += U) {
#pragma unroll
      for (int j = 0; j < U; j++) {
        on[j] = j0 + j < len;
        x[j] = on[j] ? nbr[s + j0 + j] : 0;
      }
      f(x, on);
    }
  }
}

template <int Kp>
struct BcbForward {
  int32_t *depth;
  double *sigma;
  uint32_t *ltag;
  WaveQueue q;
  int32_t d;
  uint32_t tag;
  double su;
  bool act;
  int k;
  bool leader;
  static constexpr int U = kBcbUnroll<Kp>;
  __device__ void operator()(const int32_t (&x)[U], const bool (&on)[U]) {
    int32_t dv[U];
#pragma unroll
    for (int j = 0; j < U; j++) dv[j] = depth[static_cast<int64_t>(x[j]) * Kp + k];
#pragma unroll
    for (int j = 0; j < U; j++) {
      const bool a = on[j] && act && (dv[j] == -1 || dv[j] == d + 1);
      if (__ballot(a) == 0) continue;
      bool claimed = false;
      if (a) {
        int32_t dd = dv[j];
        if (dd == -1) {
          const int32_t old = atomicCAS(&depth[static_cast<int64_t>(x[j]) * Kp + k], -1, d + 1);
          claimed = old == -1;
          dd = claimed ? d + 1 : old;
        }
        // path counts are integers below 2^53: the double sum is exact in any order
        if (dd == d + 1) unsafeAtomicAdd(&sigma[static_cast<int64_t>(x[j]) * Kp + k], su);
      }
      const bool anyc = group_any<Kp>(claimed);
      bool push = false;
      if (leader && on[j] && anyc) push = atomicExch(&ltag[x[j]], tag) != tag;
      q.push(push, x[j]);
    }
  }
};

// The backward pass reads a successor's depth and coefficient as one 64-bit
// word dc[v * Kp + k] = depth << 32 | coefficient bits: one gathered line per
// edge instead of one from each of two arrays.  A level stores only the low
// halves of its own rows' words, which no reader of that level uses.
__device__ __forceinline__ float *dc_coef(uint64_t *dc, int64_t i) { return reinterpret_cast<float *>(dc + i); }

template <int Kp>
struct BcbBackward {
  const uint64_t *dc;
  int32_t d;
  bool act;
  int k;
  double acc;
  static constexpr int U = kBcbUnroll<Kp>;
  __device__ void operator()(const int32_t (&x)[U], const bool (&on)[U]) {
    uint64_t w[U];
#pragma unroll
    for (int j = 0; j < U; j++) w[j] = dc[static_cast<int64_t>(x[j]) * Kp + k];
#pragma unroll
    for (int j = 0; j < U; j++)
      if (on[j] && act && static_cast<int32_t>(w[j] >> 32) == d + 1)
        acc += static_cast<double>(__uint_as_float(static_cast<uint32_t>(w[j])));
  }
};

// dc = depth << 32 (coefficients 0) after the forward pass.
__global__ void bcb_pack(const int32_t *depth, int64_t total, uint64_t *dc) {
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < total; i += stride)
    dc[i] = static_cast<uint64_t>(static_cast<uint32_t>(depth[i])) << 32;
}

__global__ void bcb_init_counters(BcbCounters *c, long long first, int pull, long long *lvl) {
  if (threadIdx.x == 0) {
    c->tail = static_cast<unsigned long long>(first);
    c->edges_scanned = 0;
    c->listed_edges = 0;
    c->open_edges = 0;
    c->lb = 0;
    c->le = first;
    c->listed = 0;
    c->weighed = 0;
    c->overflow = 0;
    c->d = 0;
    c->done = 0;
    c->pull = pull;
    c->need_count = 0;
    c->levels = 0;
    lvl[0] = 0;
    lvl[1] = first;
  }
}

// Ends forward level d: the next level's list is queue[le, tail); none ends
// the traversal.  Otherwise the next level's direction (bfs.cc:231-262
// weighs frontier edges against unexplored ones the same way): mode 2 always
// bottom-up; mode 1 bottom-up while the frontier's out-edges (x ratio)
// exceed the edges of the vertices some source has not reached -- bounded
// below by m - listed, counted exactly (bcb_open_edges) past that bound
// unless the rows are wide (count_wide 0), where a top-down frontier edge
// costs more than the bound's error.
__global__ void bcb_level_end(BcbCounters *c, long long *lvl, long long cap, long long m, int mode, long long ratio,
                              int count_wide) {
  if (threadIdx.x != 0 || c->done) return;
  const long long tail = static_cast<long long>(c->tail);
  if (c->overflow || tail > cap) {
    c->overflow = 1;
    c->done = 1;
    return;
  }
  c->levels = c->d + 1;
  if (tail == c->le) {
    c->done = 1;
    return;
  }
  lvl[c->d + 2] = tail;
  const long long frontier = static_cast<long long>(c->listed_edges) - c->listed;
  c->listed = static_cast<long long>(c->listed_edges);
  c->lb = c->le;
  c->le = tail;
  c->d += 1;
  c->pull = mode == 2;
  c->weighed = frontier * ratio;
  if (mode == 1 && c->weighed > m - c->listed) {
    if (count_wide) {
      c->open_edges = 0;
      c->need_count = 1;
    } else {
      c->pull = 1;
    }
  }
}

// The direction of a level that waited for the open-edge count.
__global__ void bcb_pull_decide(BcbCounters *c) {
  if (threadIdx.x != 0 || c->done || !c->need_count) return;
  c->pull = c->weighed > static_cast<long long>(c->open_edges);
  c->need_count = 0;
}

__global__ void bcb_sources(const int32_t *src, int K, int Kp, int32_t *depth, double *sigma) {
  const int j = threadIdx.x;
  if (j < K) {
    depth[static_cast<int64_t>(src[j]) * Kp + j] = 0;
    sigma[static_cast<int64_t>(src[j]) * Kp + j] = 1.0;
  }
}

// Heavy chunk c of the queue: its row and edge range (binary search in the
// chunk prefix).
__device__ __forceinline__ void bcb_chunk(const HeavyQueue &hq, int64_t rows, int64_t c, int64_t &h,
                                          int64_t &first, int64_t &len) {
  int64_t lo = 0, hi = rows;
  while (hi - lo > 1) {
    const int64_t mid = (lo + hi) >> 1;
    if (hq.prefix[mid] <= c) lo = mid;
    else hi = mid;
  }
  h = lo;
  first = (c - hq.prefix[lo]) * hq.chunk;
  const int64_t rest = hq.len[lo] - first;
  len = rest < hq.chunk ? rest : hq.chunk;
}

// Forward level d over the union frontier list[0, count); rows above
// kBatchHeavy go to the chunked kernel.
template <int Kp>
__global__ void __launch_bounds__(kBlock, kBcbWaves<Kp>)
bcb_forward(const int64_t *off, const int32_t *nbr, int32_t *depth, double *sigma, uint32_t *ltag,
            int32_t *queue, int64_t cap, BcbCounters *c, HeavyQueue hq) {
  if (bcb_skip(c, 0)) return;
  const int32_t d = c->d;
  const int32_t *list = queue + c->lb;
  const int64_t count = c->le - c->lb;
  __shared__ int32_t seg[kWavesPerBlock * kSeg];
  constexpr int G = kWave / Kp;
  const int grp = lane_id() / Kp;
  BcbForward<Kp> f;
  f.depth = depth;
  f.sigma = sigma;
  f.ltag = ltag;
  f.q.init(seg + (threadIdx.x / kWave) * kSeg, &c->tail, queue, cap, &c->overflow);
  f.d = d;
  f.tag = static_cast<uint32_t>(d) + 2;
  f.k = lane_id() % Kp;
  f.leader = f.k == 0;
  const int64_t waves = static_cast<int64_t>(gridDim.x) * kWavesPerBlock;
  const int64_t gw = static_cast<int64_t>(blockIdx.x) * kWavesPerBlock + threadIdx.x / kWave;
  unsigned long long walked = 0;
  for (int64_t base = gw * G; base < count; base += waves * G) {
    const int64_t i = base + grp;
    const bool valid = i < count;
    int64_t row = 0, s = 0, len = 0;
    if (valid) {
      row = list[i];
      s = off[row];
      len = off[row + 1] - s;
    }
    f.act = valid && depth[row * Kp + f.k] == d;
    f.su = f.act ? sigma[row * Kp + f.k] : 0.0;
    const bool heavy = len > kBatchHeavy;
    if (heavy && f.leader) heavy_push(hq, static_cast<int32_t>(row), s, len);
    bcb_edges<Kp>(nbr, s, heavy ? 0 : len, f);
    if (f.leader && !heavy) walked += static_cast<unsigned long long>(len);
  }
  f.q.flush_block();
  block_add(walked, &c->edges_scanned);
}

template <int Kp>
__global__ void __launch_bounds__(kBlock, kBcbWaves<Kp>)
bcb_forward_heavy(const int32_t *nbr, int32_t *depth, double *sigma, uint32_t *ltag, int32_t *queue,
                  int64_t cap, BcbCounters *c, HeavyQueue hq) {
  if (bcb_skip(c, 0)) return;
  const int32_t d = c->d;
  int64_t rows, total;
  heavy_extent(hq, rows, total);
  if (rows == 0) return;
  __shared__ int32_t seg[kWavesPerBlock * kSeg];
  constexpr int G = kWave / Kp;
  const int grp = lane_id() / Kp;
  BcbForward<Kp> f;
  f.depth = depth;
  f.sigma = sigma;
  f.ltag = ltag;
  f.q.init(seg + (threadIdx.x / kWave) * kSeg, &c->tail, queue, cap, &c->overflow);
  f.d = d;
  f.tag = static_cast<uint32_t>(d) + 2;
  f.k = lane_id() % Kp;
  f.leader = f.k == 0;
  const int64_t stride = static_cast<int64_t>(gridDim.x) * kWavesPerBlock * G;
  const int64_t gw = static_cast<int64_t>(blockIdx.x) * kWavesPerBlock + threadIdx.x / kWave;
  unsigned long long walked = 0;
  for (int64_t c0 = gw * G; c0 < total; c0 += stride) {
    const int64_t cc = c0 + grp;
    const bool valid = cc < total;
    int64_t h = 0, first = 0, len = 0, row = 0;
    if (valid) {
      bcb_chunk(hq, rows, cc, h, first, len);
      row = hq.row[h];
    }
    f.act = valid && depth[row * Kp + f.k] == d;
    f.su = f.act ? sigma[row * Kp + f.k] : 0.0;
    bcb_edges<Kp>(nbr, valid ? hq.start[h] + first : 0, len, f);
    if (f.leader) walked += static_cast<unsigned long long>(len);
  }
  f.q.flush_block();
  block_add(walked, &c->edges_scanned);
}

// Bottom-up forward level d, for levels whose frontier has more out-edges
// than the vertices not yet reached: every vertex that some lane has not
// reached sums, lane by lane, sigma over its in-neighbours at depth d and
// takes depth d + 1 where the sum is positive (sigma >= 1 on every reached
// vertex).  One group owns a vertex, so there are no atomics; a vertex set
// here is never at depth d, so concurrent readers decide the same either
// way.  The sum of integer-valued doubles is exact in any order, as the
// top-down atomics' is.
template <int Kp>
struct BcbPull {
  const int32_t *depth;
  const double *sigma;
  int32_t d;
  bool act;
  int k;
  double acc;
  static constexpr int U = kBcbUnroll<Kp>;
  __device__ void operator()(const int32_t (&x)[U], const bool (&on)[U]) {
    int32_t du[U];
#pragma unroll
    for (int j = 0; j < U; j++) du[j] = depth[static_cast<int64_t>(x[j]) * Kp + k];
#pragma unroll
    for (int j = 0; j < U; j++)
      if (on[j] && act && du[j] == d) acc += sigma[static_cast<int64_t>(x[j]) * Kp + k];
  }
};

// Light rows of a bottom-up level; rows above kBatchHeavy go to 'hq'.
template <int Kp>
__global__ void __launch_bounds__(kBlock, kBcbWaves<Kp>)
bcb_forward_pull(const int64_t *ioff, const int32_t *inbr, int64_t n, int K, int32_t *depth, double *sigma,
                 int32_t *queue, int64_t cap, BcbCounters *c, HeavyQueue hq) {
  if (bcb_skip(c, 1)) return;
  const int32_t d = c->d;
  __shared__ int32_t seg[kWavesPerBlock * kSeg];
  constexpr int G = kWave / Kp;
  const int grp = lane_id() / Kp;
  WaveQueue q;
  q.init(seg + (threadIdx.x / kWave) * kSeg, &c->tail, queue, cap, &c->overflow);
  BcbPull<Kp> f;
  f.depth = depth;
  f.sigma = sigma;
  f.d = d;
  f.k = lane_id() % Kp;
  const bool leader = f.k == 0;
  const int64_t waves = static_cast<int64_t>(gridDim.x) * kWavesPerBlock;
  const int64_t gw = static_cast<int64_t>(blockIdx.x) * kWavesPerBlock + threadIdx.x / kWave;
  unsigned long long walked = 0;
  for (int64_t base = gw * G; base < n; base += waves * G) {
    const int64_t v = base + grp;
    const bool valid = v < n;
    f.act = valid && f.k < K && depth[v * Kp + f.k] == -1;
    const bool any = group_any<Kp>(f.act);
    int64_t s = 0, len = 0;
    if (any) {
      s = ioff[v];
      len = ioff[v + 1] - s;
    }
    const bool heavy = len > kBatchHeavy;
    if (heavy && leader) heavy_push(hq, static_cast<int32_t>(v), s, len);
    f.acc = 0.0;
    bcb_edges<Kp>(inbr, s, heavy ? 0 : len, f);
    const bool found = f.act && !heavy && f.acc > 0.0;
    if (found) {
      depth[v * Kp + f.k] = d + 1;
      sigma[v * Kp + f.k] = f.acc;
    }
    const bool anyf = group_any<Kp>(found);
    q.push(leader && anyf, static_cast<int32_t>(v));
    if (leader && !heavy) walked += static_cast<unsigned long long>(len);
  }
  q.flush_block();
  block_add(walked, &c->edges_scanned);
}

// Heavy rows of a bottom-up level: hpart[chunk][k] = lane k's sigma sum over
// the chunk's in-neighbours at depth d.
template <int Kp>
__global__ void __launch_bounds__(kBlock, kBcbWaves<Kp>)
bcb_pull_heavy(const int32_t *inbr, int K, const int32_t *depth, const double *sigma, BcbCounters *c,
               HeavyQueue hq, double *hpart) {
  if (bcb_skip(c, 1)) return;
  const int32_t d = c->d;
  int64_t rows, total;
  heavy_extent(hq, rows, total);
  if (rows == 0) return;
  constexpr int G = kWave / Kp;
  const int grp = lane_id() / Kp;
  BcbPull<Kp> f;
  f.depth = depth;
  f.sigma = sigma;
  f.d = d;
  f.k = lane_id() % Kp;
  const int64_t stride = static_cast<int64_t>(gridDim.x) * kWavesPerBlock * G;
  const int64_t gw = static_cast<int64_t>(blockIdx.x) * kWavesPerBlock + threadIdx.x / kWave;
  unsigned long long walked = 0;
  for (int64_t c0 = gw * G; c0 < total; c0 += stride) {
    const int64_t cc = c0 + grp;
    const bool valid = cc < total;
    int64_t h = 0, first = 0, len = 0, row = 0;
    if (valid) {
      bcb_chunk(hq, rows, cc, h, first, len);
      row = hq.row[h];
    }
    f.act = valid && f.k < K && depth[row * Kp + f.k] == -1;
    f.acc = 0.0;
    bcb_edges<Kp>(inbr, valid ? hq.start[h] + first : 0, len, f);
    if (valid) hpart[cc * Kp + f.k] = f.acc;
    if (f.k == 0) walked += static_cast<unsigned long long>(len);
  }
  block_add(walked, &c->edges_scanned);
}

// One wave per heavy row of a bottom-up level: its chunks' sums meet as in
// bcb_backward_finish; lanes with a positive sum take depth d + 1 and the
// row joins the next level's list.
template <int Kp>
__global__ void __launch_bounds__(kBlock)
bcb_pull_finish(int K, int32_t *depth, double *sigma, HeavyQueue hq, const double *hpart,
                int32_t *queue, int64_t cap, BcbCounters *c) {
  if (bcb_skip(c, 1)) return;
  const int32_t d = c->d;
  __shared__ int32_t seg[kWavesPerBlock * kSeg];
  WaveQueue q;
  q.init(seg + (threadIdx.x / kWave) * kSeg, &c->tail, queue, cap, &c->overflow);
  int64_t rows, total;
  heavy_extent(hq, rows, total);
  constexpr int G = kWave / Kp;
  const int lane = lane_id();
  const int k = lane % Kp;
  const int slot = lane / Kp;
  const int64_t waves = static_cast<int64_t>(gridDim.x) * kWavesPerBlock;
  for (int64_t h = static_cast<int64_t>(blockIdx.x) * kWavesPerBlock + threadIdx.x / kWave; h < rows; h += waves) {
    const int32_t row = hq.row[h];
    const bool need = k < K && depth[static_cast<int64_t>(row) * Kp + k] == -1;
    const int64_t end = h + 1 < rows ? hq.prefix[h + 1] : total;
    double tot = 0.0;
    if (need)
      for (int64_t cc = hq.prefix[h] + slot; cc < end; cc += G) tot += hpart[cc * Kp + k];
#pragma unroll
    for (int o = Kp; o < kWave; o <<= 1) tot += __shfl_xor(tot, o);
    const bool found = slot == 0 && need && tot > 0.0;
    if (found) {
      depth[static_cast<int64_t>(row) * Kp + k] = d + 1;
      sigma[static_cast<int64_t>(row) * Kp + k] = tot;
    }
    const bool anyf = __ballot(found) != 0;
    q.push(lane == 0 && anyf, row);
  }
  q.flush_block();
}

// In-edges of the vertices that some source (lane k < K) has not reached,
// added into c->open_edges: what a bottom-up level would walk.  16-byte
// loads of four lanes; a vertex's Kp / 4 items sit in one aligned lane
// group of the wave and are OR-ed over shuffles.
__global__ void __launch_bounds__(kBlock)
bcb_open_edges(const int32_t *depth, int64_t n, int log_kp, int K, const int64_t *ioff, BcbCounters *c) {
  {
    const volatile BcbCounters *v = c;
    if (v->done || !v->need_count) return;
  }
  const int4 *d4 = reinterpret_cast<const int4 *>(depth);
  const int log_q = log_kp - 2;
  const int q = 1 << log_q;
  const int64_t total = (n << log_kp) >> 2;
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  unsigned long long e = 0;
  for (int64_t i0 = static_cast<int64_t>(blockIdx.x) * blockDim.x; i0 < total; i0 += stride) {
    const int64_t i = i0 + threadIdx.x;
    int open = 0;
    if (i < total) {
      const int4 x = d4[i];
      const int k0 = static_cast<int>((i & (q - 1)) << 2);
      open = (k0 < K && x.x == -1) || (k0 + 1 < K && x.y == -1) || (k0 + 2 < K && x.z == -1) ||
             (k0 + 3 < K && x.w == -1);
    }
    for (int o = 1; o < q; o <<= 1) open |= __shfl_xor(open, o);
    if (i < total && (i & (q - 1)) == 0 && open) {
      const int64_t v = i >> log_q;
      e += static_cast<unsigned long long>(ioff[v + 1] - ioff[v]);
    }
  }
  block_add(e, &c->open_edges);
}

// Out-edges of the level list [le, tail) added into c->listed_edges: the
// next level's direction compares the frontier's edges with those not yet
// listed.
__global__ void __launch_bounds__(kBlock)
bcb_list_edges(const int64_t *off, const int32_t *queue, int64_t cap, BcbCounters *c) {
  if (reinterpret_cast<const volatile BcbCounters *>(c)->done) return;
  const int64_t le = c->le;
  const int64_t tail = min(static_cast<int64_t>(c->tail), cap);
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  unsigned long long e = 0;
  for (int64_t i = le + static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < tail; i += stride) {
    const int32_t v = queue[i];
    e += static_cast<unsigned long long>(off[v + 1] - off[v]);
  }
  block_add(e, &c->listed_edges);
}

// Backward level d: delta of a lane = sigma[u] * sum of coef[v] over its
// successors; coef[u] = (1 + delta) / sigma[u] for the level above, and the
// row's score gains the deltas of all its lanes at this depth.
template <int Kp>
__global__ void __launch_bounds__(kBlock, kBcbWaves<Kp>)
bcb_backward(const int64_t *off, const int32_t *nbr, const int32_t *depth, const double *sigma, uint64_t *dc,
             float *scores, const int32_t *list, int64_t count, int32_t d, BcbCounters *c, HeavyQueue hq) {
  constexpr int G = kWave / Kp;
  const int grp = lane_id() / Kp;
  BcbBackward<Kp> f;
  f.dc = dc;
  f.d = d;
  f.k = lane_id() % Kp;
  const bool leader = f.k == 0;
  const int64_t waves = static_cast<int64_t>(gridDim.x) * kWavesPerBlock;
  const int64_t gw = static_cast<int64_t>(blockIdx.x) * kWavesPerBlock + threadIdx.x / kWave;
  unsigned long long walked = 0;
  for (int64_t base = gw * G; base < count; base += waves * G) {
    const int64_t i = base + grp;
    const bool valid = i < count;
    int64_t row = 0, s = 0, len = 0;
    if (valid) {
      row = list[i];
      s = off[row];
      len = off[row + 1] - s;
    }
    f.act = valid && depth[row * Kp + f.k] == d;
    const bool heavy = len > kBatchHeavy;
    if (heavy && leader) heavy_push(hq, static_cast<int32_t>(row), s, len);
    f.acc = 0.0;
    bcb_edges<Kp>(nbr, s, heavy ? 0 : len, f);
    double delta = 0.0;
    if (f.act && !heavy) {
      const double su = sigma[row * Kp + f.k];
      delta = su * f.acc;
      *dc_coef(dc, row * Kp + f.k) = static_cast<float>((1.0 + delta) / su);
    }
    const double tot = group_sum_d<Kp>(delta);
    if (leader && valid && !heavy && tot != 0.0) scores[row] += static_cast<float>(tot);
    if (leader && !heavy) walked += static_cast<unsigned long long>(len);
  }
  block_add(walked, &c->edges_scanned);
}

// Heavy rows of a backward level: hpart[chunk][k] = lane k's coefficient
// sum over the chunk's successors.
template <int Kp>
__global__ void __launch_bounds__(kBlock, kBcbWaves<Kp>)
bcb_backward_heavy(const int32_t *nbr, const int32_t *depth, const uint64_t *dc, int32_t d, BcbCounters *c,
                   HeavyQueue hq, double *hpart) {
  int64_t rows, total;
  heavy_extent(hq, rows, total);
  if (rows == 0) return;
  constexpr int G = kWave / Kp;
  const int grp = lane_id() / Kp;
  BcbBackward<Kp> f;
  f.dc = dc;
  f.d = d;
  f.k = lane_id() % Kp;
  const int64_t stride = static_cast<int64_t>(gridDim.x) * kWavesPerBlock * G;
  const int64_t gw = static_cast<int64_t>(blockIdx.x) * kWavesPerBlock + threadIdx.x / kWave;
  unsigned long long walked = 0;
  for (int64_t c0 = gw * G; c0 < total; c0 += stride) {
    const int64_t cc = c0 + grp;
    const bool valid = cc < total;
    int64_t h = 0, first = 0, len = 0, row = 0;
    if (valid) {
      bcb_chunk(hq, rows, cc, h, first, len);
      row = hq.row[h];
    }
    f.act = valid && depth[row * Kp + f.k] == d;
    f.acc = 0.0;
    bcb_edges<Kp>(nbr, valid ? hq.start[h] + first : 0, len, f);
    if (valid) hpart[cc * Kp + f.k] = f.acc;
    if (f.k == 0) walked += static_cast<unsigned long long>(len);
  }
  block_add(walked, &c->edges_scanned);
}

// One wave per heavy row: lane (slot, k) adds source k's partials of chunks
// slot, slot + G, ... (G = 64 / Kp slots), and the slots' sums meet over
// shuffles -- a fixed order, so the result does not depend on the schedule.
// (One lane group walking a 162K-edge row's 635 chunks one by one took
// 0.4 ms a level.)
template <int Kp>
__global__ void __launch_bounds__(kBlock)
bcb_backward_finish(const int32_t *depth, const double *sigma, uint64_t *dc, float *scores, int32_t d,
                    HeavyQueue hq, const double *hpart) {
  int64_t rows, total;
  heavy_extent(hq, rows, total);
  constexpr int G = kWave / Kp;
  const int lane = lane_id();
  const int k = lane % Kp;
  const int slot = lane / Kp;
  const int64_t waves = static_cast<int64_t>(gridDim.x) * kWavesPerBlock;
  for (int64_t h = static_cast<int64_t>(blockIdx.x) * kWavesPerBlock + threadIdx.x / kWave; h < rows; h += waves) {
    const int32_t row = hq.row[h];
    const bool here = depth[static_cast<int64_t>(row) * Kp + k] == d;
    const int64_t end = h + 1 < rows ? hq.prefix[h + 1] : total;
    double tot = 0.0;
    if (here)
      for (int64_t cc = hq.prefix[h] + slot; cc < end; cc += G) tot += hpart[cc * Kp + k];
#pragma unroll
    for (int o = Kp; o < kWave; o <<= 1) tot += __shfl_xor(tot, o);
    double delta = 0.0;
    if (slot == 0 && here) {
      const double su = sigma[static_cast<int64_t>(row) * Kp + k];
      delta = su * tot;
      *dc_coef(dc, static_cast<int64_t>(row) * Kp + k) = static_cast<float>((1.0 + delta) / su);
    }
    const double sum = group_sum_d<Kp>(delta);
    if (lane == 0 && sum != 0.0) scores[row] += static_cast<float>(sum);
  }
}

// The relaxation launches fill the GPU with exactly the workgroups it keeps
// resident (grid-stride loops; a second round of workgroups would leave a
// tail).
template <int Kp>
static unsigned sb_grid(gbk_ctx *ctx) {
  const unsigned a = resident_grid(ctx, reinterpret_cast<const void *>(&sb_relax<Kp>));
  const unsigned b = resident_grid(ctx, reinterpret_cast<const void *>(&sb_relax_heavy<Kp>));
  return a < b ? a : b;
}

template <int Kp>
static void launch_sb_relax(gbk_ctx *ctx, const gbk_graph *g, int32_t *dist, int32_t *pend, uint32_t *vtag,
                            int32_t *near0, int32_t *near1, int64_t cap, int64_t delta, BatchState *st,
                            const HeavyQueue &hq, unsigned grid) {
  hipStream_t s = ctx->stream;
  auto *slots = reinterpret_cast<unsigned long long *>(reinterpret_cast<char *>(st) + kBatchStateBytes);
  sb_relax<Kp><<<grid, kBlock, 0, s>>>(g->off, g->nbr, g->wgt, dist, pend, vtag, near0, near1, cap, delta,
                                       st, slots, hq);
  sb_relax_heavy<Kp><<<grid, kBlock, 0, s>>>(g->nbr, g->wgt, dist, pend, vtag, near0, near1, cap, delta,
                                             st, slots, hq);
}

}  // namespace gbk

using namespace gbk;

namespace {

// Per-source reached / edge counts (and optional source-major output) from
// a vertex-major array; adds into the caller's arrays at offset 'first'.
int collect(gbk_ctx *ctx, const gbk_graph *g, const int32_t *vm, int Kp, int K, int32_t unreached,
            int32_t *out_dev, int64_t *reached, int64_t *edges, int64_t first, int64_t *edge_total) {
  hipStream_t s = ctx->stream;
  const unsigned grid = 1024;
  auto *part = static_cast<unsigned long long *>(scratch(ctx, kS14, sizeof(unsigned long long) * 2 * grid * kBatchMax + 1024));
  auto *sums = static_cast<unsigned long long *>(pinned(ctx, sizeof(unsigned long long) * 2 * kBatchMax));
  if (!part || !sums) return set_error(ctx, GBK_ENOMEM, "batch collect scratch");
  if (out_dev) {
    batch_collect<<<grid, kBlock, 0, s>>>(vm, g->n, Kp, K, unreached, g->off, out_dev, part);
  } else {
    // 16-byte loads of four sources: Kp >= 4 and an aligned array
    if (Kp < 4 || (reinterpret_cast<uintptr_t>(vm) & 15))
      return set_error(ctx, GBK_EINVAL, "batch count: misaligned vertex-major array");
    int log_kp = 0;
    while ((1 << log_kp) < Kp) log_kp++;
    batch_count<<<grid, kBlock, 0, s>>>(vm, g->n, log_kp, unreached, g->off, part);
  }
  batch_collect_sum<<<1, 1024, 0, s>>>(part, static_cast<int>(grid), part + 2 * grid * kBatchMax);
  GBK_LAUNCH_CHECK(ctx, "batch_collect");
  GBK_HIP(ctx, hipMemcpyAsync(sums, part + 2 * grid * kBatchMax, sizeof(unsigned long long) * 2 * kBatchMax,
                              hipMemcpyDeviceToHost, s));
  GBK_HIP(ctx, hipStreamSynchronize(s));
  for (int k = 0; k < K; k++) {
    if (reached) reached[first + k] = static_cast<int64_t>(sums[k]);
    if (edges) edges[first + k] = static_cast<int64_t>(sums[kBatchMax + k]);
    *edge_total += static_cast<int64_t>(sums[kBatchMax + k]);
  }
  return GBK_OK;
}

int check_sources(gbk_ctx *ctx, const gbk_graph *g, const int32_t *sources, int num) {
  if (!ctx || !g || !sources || num <= 0) return set_error(ctx, GBK_EINVAL, "batch: bad arguments");
  for (int i = 0; i < num; i++)
    if (sources[i] < 0 || sources[i] >= g->n) return set_error(ctx, GBK_EINVAL, "batch: source out of range");
  return GBK_OK;
}

int upload_sources(gbk_ctx *ctx, const int32_t *src, int K, int32_t **dev) {
  *dev = static_cast<int32_t *>(scratch(ctx, kS13, sizeof(int32_t) * kBatchMax));
  if (!*dev) return set_error(ctx, GBK_ENOMEM, "batch sources");
  GBK_HIP(ctx, hipMemcpyAsync(*dev, src, sizeof(int32_t) * K, hipMemcpyHostToDevice, ctx->stream));
  return GBK_OK;
}

int msbfs_chunk(gbk_ctx *ctx, const gbk_graph *g, const int32_t *src, int K, int32_t *depth_host,
                int64_t *reached, int64_t *edges, int64_t first, gbk_batch_stats *acc) {
  const int64_t n = g->n;
  hipStream_t s = ctx->stream;
  const int Kp = batch_width(K);
  const int64_t cap = n + 64;
  auto *seen = static_cast<unsigned long long *>(scratch(ctx, kS0, sizeof(unsigned long long) * (n + 1)));
  auto *front = static_cast<unsigned long long *>(scratch(ctx, kS1, sizeof(unsigned long long) * (n + 1)));
  auto *next = static_cast<unsigned long long *>(scratch(ctx, kS2, sizeof(unsigned long long) * (n + 1)));
  auto *dvm = static_cast<int32_t *>(scratch(ctx, kS3, sizeof(int32_t) * n * Kp + 64));
  auto *queue = static_cast<int32_t *>(scratch(ctx, kS4, sizeof(int32_t) * cap));
  auto *c = static_cast<MsCounters *>(scratch(ctx, kS5, sizeof(MsCounters) + 64));
  auto *partial = static_cast<unsigned long long *>(scratch(ctx, kS6, sizeof(unsigned long long) * 2 * kGridCap));
  auto *out = depth_host ? static_cast<int32_t *>(scratch(ctx, kS8, sizeof(int32_t) * n * K + 64)) : nullptr;
  auto *hc = static_cast<MsCounters *>(pinned(ctx, sizeof(MsCounters)));
  HeavyQueue hq;
  if (!seen || !front || !next || !dvm || !queue || !c || !partial || !hc || (depth_host && !out) ||
      !make_heavy_queue(ctx, g, kS7, &hq))
    return set_error(ctx, GBK_ENOMEM, "bfs batch scratch");
  hq.overflow = &c->overflow;
  int32_t *dsrc = nullptr;
  if (int rc = upload_sources(ctx, src, K, &dsrc)) return rc;
  const unsigned long long full = K == 64 ? ~0ull : ((1ull << K) - 1);
  const unsigned n_grid = expand_grid(n);
  // the bottom-up step gives each vertex msbfs_group lanes (grid-stride beyond the cap)
  const unsigned bu_grid = expand_grid(n * static_cast<int64_t>(ctx->tune.msbfs_group));
  const long long bu_edges = g->m / ctx->tune.bfs_alpha;
  const int G = ctx->tune.msbfs_group;
  // one block of levels: every level queues both directions; the one the
  // device state did not choose, and every level after the last, returns at
  // once, so blocks are queued blind (and replayed from a captured graph)
  // with one host check of the state per block
  auto levels_block = [&]() {
    for (int i = 0; i < kMsLevelsPerBlock; i++) {
      const int kt = kt_begin(ctx, GBK_KT_BFS_STEP);
      launch_msbfs_bu(ctx, G, g, seen, front, next, full, bu_grid, c);
      fill<unsigned long long>(ctx, hq.count, 0ull, 1);
      msbfs_front_queue<<<n_grid, kBlock, 0, s>>>(front, n, queue, cap, c);
      msbfs_td<<<kGridCap, kBlock, 0, s>>>(g->off, g->nbr, queue, c, front, seen, next, ctx->tune.heavy_degree, hq);
      msbfs_td_heavy<<<kHeavyGrid, kBlock, 0, s>>>(g->nbr, front, seen, next, hq, c);
      kt_end(ctx, kt);
      msbfs_finish<<<n_grid, kBlock, 0, s>>>(g->off, n, Kp, seen, front, next, dvm, c, partial, partial + kGridCap);
      msbfs_store<<<1, 1024, 0, s>>>(partial, static_cast<int>(n_grid), c, bu_edges);
    }
  };
  LoopGraph loop(ctx);
  if (int rc = loop.capture(levels_block)) return rc;
  EventTimer timer(ctx);
  fill<unsigned long long>(ctx, seen, 0ull, n);
  fill<unsigned long long>(ctx, front, 0ull, n);
  fill<unsigned long long>(ctx, next, 0ull, n);
  fill<int32_t>(ctx, dvm, -1, n * Kp);
  msbfs_sources<<<1, kBatchMax, 0, s>>>(dsrc, K, Kp, seen, front, dvm, c);
  GBK_LAUNCH_CHECK(ctx, "bfs batch sources");
  int64_t levels = 0;
  for (int64_t queued = 0;; queued += kMsLevelsPerBlock) {
    if (queued > n + 1) return set_error(ctx, GBK_EHIP, "bfs batch: more levels than vertices");
    if (int rc = loop.run(levels_block)) return rc;
    GBK_LAUNCH_CHECK(ctx, "bfs batch level");
    GBK_HIP(ctx, hipMemcpyAsync(hc, c, sizeof(MsCounters), hipMemcpyDeviceToHost, s));
    GBK_HIP(ctx, hipStreamSynchronize(s));
    if (hc->overflow) return set_error(ctx, GBK_EHIP, "bfs batch: queue overflow");
    levels = hc->levels;
    if (hc->done) break;
  }
  int64_t etot = 0;
  if (int rc = collect(ctx, g, dvm, Kp, K, -1, out, reached, edges, first, &etot)) return rc;
  const double ms = timer.stop_ms();
  acc->device_ms += ms;
  acc->rounds += levels;
  acc->edges_reached += etot;
  if (depth_host) {
    GBK_HIP(ctx, hipMemcpyAsync(depth_host + first * n, out, sizeof(int32_t) * n * K, hipMemcpyDeviceToHost, s));
    GBK_HIP(ctx, hipStreamSynchronize(s));
  }
  return GBK_OK;
}

int sssp_batch_chunk(gbk_ctx *ctx, const gbk_graph *g, const int32_t *src, int K, int32_t delta,
                     int32_t *dist_host, int64_t *reached, int64_t *edges, int64_t first,
                     gbk_batch_stats *acc) {
  const int64_t n = g->n;
  hipStream_t s = ctx->stream;
  // 17..32 sources take 64-lane rows too when the wave-uniform walk wins
  const int Kp = K > 16 && ctx->tune.sssp_wide ? kWave : batch_width(K);
  const int64_t cap = n + 64;
  auto *dist = static_cast<int32_t *>(scratch(ctx, kS0, sizeof(int32_t) * n * Kp + 64));
  auto *pend = static_cast<int32_t *>(scratch(ctx, kS1, sizeof(int32_t) * (n + 1)));
  auto *vtag = static_cast<uint32_t *>(scratch(ctx, kS2, sizeof(uint32_t) * (n + 1)));
  auto *q = static_cast<int32_t *>(scratch(ctx, kS4, sizeof(int32_t) * cap * 2));
  auto *st = static_cast<BatchState *>(scratch(ctx, kS5, kBatchStateBytes + kSpreadBytes));
  auto *scan_slots = reinterpret_cast<unsigned long long *>(reinterpret_cast<char *>(st) + kBatchStateBytes);
  auto *partial = static_cast<unsigned long long *>(scratch(ctx, kS6, sizeof(unsigned long long) * kGridCap));
  auto *out = dist_host ? static_cast<int32_t *>(scratch(ctx, kS8, sizeof(int32_t) * n * K + 64)) : nullptr;
  auto *hs = static_cast<BatchState *>(pinned(ctx, sizeof(BatchState)));
  HeavyQueue hq;
  if (!dist || !pend || !vtag || !q || !st || !partial || !hs || (dist_host && !out) ||
      !make_heavy_queue(ctx, g, kS7, &hq))
    return set_error(ctx, GBK_ENOMEM, "sssp batch scratch");
  // rows above kBatchHeavy are queued: at most m / kBatchHeavy of them
  if (hq.cap < std::min<int64_t>(n, g->m / kBatchHeavy + 1) + 1) {
    const int64_t H = ctx->tune.heavy_degree;
    ctx->tune.heavy_degree = kBatchHeavy;
    const bool ok = make_heavy_queue(ctx, g, kS7, &hq);
    ctx->tune.heavy_degree = H;
    if (!ok) return set_error(ctx, GBK_ENOMEM, "sssp batch heavy queue");
  }
  hq.overflow = &st->overflow;
  hq.chunk = kBatchChunk;
  int32_t *near0 = q, *near1 = q + cap;
  // distinct sources form the first near queue
  std::vector<int32_t> uniq(src, src + K);
  std::sort(uniq.begin(), uniq.end());
  uniq.erase(std::unique(uniq.begin(), uniq.end()), uniq.end());
  int32_t *dsrc = nullptr;
  if (int rc = upload_sources(ctx, src, K, &dsrc)) return rc;
  unsigned grid = 0;
  switch (Kp) {
    case 8: grid = sb_grid<8>(ctx); break;
    case 16: grid = sb_grid<16>(ctx); break;
    case 32: grid = sb_grid<32>(ctx); break;
    default: grid = sb_grid<64>(ctx); break;
  }
  const unsigned scan_grid = expand_grid(n);
  // one step: relaxation round, then the bucket bookkeeping (each kernel
  // returns at once unless the state's phase is its own)
  auto step = [&]() {
      const int kt = kt_begin(ctx, GBK_KT_SSSP_RELAX);
      switch (Kp) {
        case 8: launch_sb_relax<8>(ctx, g, dist, pend, vtag, near0, near1, cap, delta, st, hq, grid); break;
        case 16: launch_sb_relax<16>(ctx, g, dist, pend, vtag, near0, near1, cap, delta, st, hq, grid); break;
        case 32: launch_sb_relax<32>(ctx, g, dist, pend, vtag, near0, near1, cap, delta, st, hq, grid); break;
        default: launch_sb_relax<64>(ctx, g, dist, pend, vtag, near0, near1, cap, delta, st, hq, grid); break;
      }
      kt_end(ctx, kt);
      sb_relax_done<<<1, kWave, 0, s>>>(st, hq.count, scan_slots);
      sb_min_pend<<<scan_grid, kBlock, 0, s>>>(pend, n, st, partial);
      sb_min_done<<<1, 1024, 0, s>>>(partial, static_cast<int>(scan_grid), st);
      sb_split<<<scan_grid, kBlock, 0, s>>>(pend, n, near0, near1, cap, st);
      sb_split_done<<<1, kWave, 0, s>>>(st);
  };
  // steps are queued blindly in blocks of kStepsPerGraph (a captured graph)
  // and the host reads the state after 4, 8, ... 64 of them -- or first
  // after as many as the last batch of this shape (vertices, edges,
  // sources, delta) took: the bench relabels into a new graph every step
  constexpr int kStepsPerGraph = 4;
  auto block = [&]() {
    for (int j = 0; j < kStepsPerGraph; j++) step();
  };
  LoopGraph loop(ctx);
  if (int rc = loop.capture(block)) return rc;
  const gbk_ctx::BatchHint shape{n, g->m, K, delta, 0};
  const bool hinted = ctx->sssp_batch_hint.n == shape.n && ctx->sssp_batch_hint.m == shape.m &&
                      ctx->sssp_batch_hint.k == shape.k && ctx->sssp_batch_hint.delta == shape.delta &&
                      ctx->sssp_batch_hint.steps > 0;
  EventTimer timer(ctx);
  GBK_HIP(ctx, hipMemcpyAsync(near0, uniq.data(), sizeof(int32_t) * uniq.size(), hipMemcpyHostToDevice, s));
  fill<int32_t>(ctx, dist, kBatchInf, n * Kp);
  fill<int32_t>(ctx, pend, INT32_MAX, n);
  fill<uint32_t>(ctx, vtag, 0u, n);
  sb_sources<<<1, kBatchMax, 0, s>>>(dsrc, K, Kp, dist);
  sb_init_state<<<1, 1, 0, s>>>(st, static_cast<long long>(uniq.size()), scan_slots);
  int64_t batch = kStepsPerGraph;
  if (hinted)
    batch = std::min<int64_t>((ctx->sssp_batch_hint.steps + kStepsPerGraph - 1) / kStepsPerGraph * kStepsPerGraph,
                              256);
  int64_t launched = 0;
  for (;;) {
    for (int64_t j = 0; j < batch; j += kStepsPerGraph)
      if (int rc = loop.run(block)) return rc;
    launched += batch;
    GBK_LAUNCH_CHECK(ctx, "sssp batch step");
    GBK_HIP(ctx, hipMemcpyAsync(hs, st, sizeof(BatchState), hipMemcpyDeviceToHost, s));
    GBK_HIP(ctx, hipStreamSynchronize(s));
    kt_collect(ctx);
    if (hs->overflow) return set_error(ctx, GBK_EHIP, "sssp batch: queue overflow");
    if (hs->done) break;
    if (hinted && launched == batch) batch = kStepsPerGraph;
    else if (batch < 64) batch *= 2;
    if (launched > 4 * (n + 16) + 1024) return set_error(ctx, GBK_EHIP, "sssp batch: no convergence");
  }
  const BatchState fin = *hs;  // collect() may regrow the pinned staging buffer
  ctx->sssp_batch_hint = shape;
  ctx->sssp_batch_hint.steps = fin.steps;
  int64_t etot = 0;
  if (int rc = collect(ctx, g, dist, Kp, K, kBatchInf, out, reached, edges, first, &etot)) return rc;
  const double ms = timer.stop_ms();
  acc->device_ms += ms;
  acc->rounds += fin.relax_rounds;
  acc->buckets += fin.buckets + 1;
  acc->rows_scanned += static_cast<int64_t>(fin.rows_scanned);
  acc->edges_scanned += static_cast<int64_t>(fin.edges_scanned);
  acc->edges_reached += etot;
  if (dist_host) {
    GBK_HIP(ctx, hipMemcpyAsync(dist_host + first * n, out, sizeof(int32_t) * n * K, hipMemcpyDeviceToHost, s));
    GBK_HIP(ctx, hipStreamSynchronize(s));
  }
  return GBK_OK;
}

// Brandes for one chunk of K <= 64 sources; deltas are added into 'scores'
// (device, not normalised).
template <int Kp>
int bc_batch_run(gbk_ctx *ctx, const gbk_graph *g, const int32_t *src, int K, float *scores,
                 gbk_batch_stats *acc) {
  const int64_t n = g->n;
  hipStream_t s = ctx->stream;
  // every vertex enters at most one level per distinct depth of its lanes
  const int64_t cap = n * Kp + 64;
  auto *depth = static_cast<int32_t *>(scratch(ctx, kS0, sizeof(int32_t) * n * Kp + 64));
  auto *sigma = static_cast<double *>(scratch(ctx, kS1, sizeof(double) * n * Kp + 64));
  auto *dc = static_cast<uint64_t *>(scratch(ctx, kS2, sizeof(uint64_t) * n * Kp + 64));
  auto *queue = static_cast<int32_t *>(scratch(ctx, kS3, sizeof(int32_t) * cap));
  auto *c = static_cast<BcbCounters *>(scratch(ctx, kS5, sizeof(BcbCounters) + 64));
  auto *ltag = static_cast<uint32_t *>(scratch(ctx, kS8, sizeof(uint32_t) * (n + 1)));
  auto *hc = static_cast<BcbCounters *>(pinned(ctx, sizeof(BcbCounters)));
  if (!depth || !sigma || !dc || !queue || !c || !ltag || !hc) return set_error(ctx, GBK_ENOMEM, "bc batch scratch");
  HeavyQueue hq;
  {
    const int64_t H = ctx->tune.heavy_degree;
    ctx->tune.heavy_degree = kBatchHeavy;  // rows above kBatchHeavy are chunked
    const bool ok = make_heavy_queue(ctx, g, kS7, &hq);
    ctx->tune.heavy_degree = H;
    if (!ok) return set_error(ctx, GBK_ENOMEM, "bc batch heavy queue");
  }
  hq.overflow = &c->overflow;
  hq.chunk = kBatchChunk;
  auto *hpart = static_cast<double *>(scratch(ctx, kS9, sizeof(double) * (g->m / kBatchChunk + hq.cap + 1) * Kp));
  if (!hpart) return set_error(ctx, GBK_ENOMEM, "bc batch heavy partials");
  std::vector<int32_t> uniq(src, src + K);
  std::sort(uniq.begin(), uniq.end());
  uniq.erase(std::unique(uniq.begin(), uniq.end()), uniq.end());
  int32_t *dsrc = nullptr;
  if (int rc = upload_sources(ctx, src, K, &dsrc)) return rc;
  // forward and backward launches fill the GPU with resident workgroups
  const unsigned grid_f = std::min(resident_grid(ctx, reinterpret_cast<const void *>(&bcb_forward<Kp>)),
                                   resident_grid(ctx, reinterpret_cast<const void *>(&bcb_forward_heavy<Kp>)));
  const unsigned grid_b = std::min(resident_grid(ctx, reinterpret_cast<const void *>(&bcb_backward<Kp>)),
                                   resident_grid(ctx, reinterpret_cast<const void *>(&bcb_backward_heavy<Kp>)));
  const unsigned grid_p = std::min(resident_grid(ctx, reinterpret_cast<const void *>(&bcb_forward_pull<Kp>)),
                                   resident_grid(ctx, reinterpret_cast<const void *>(&bcb_pull_heavy<Kp>)));
  auto *lvl = static_cast<long long *>(scratch(ctx, kS6, sizeof(long long) * (n + 3)));
  if (!lvl) return set_error(ctx, GBK_ENOMEM, "bc batch levels");
  int log_kp = 0;
  while ((1 << log_kp) < Kp) log_kp++;
  const int mode = ctx->tune.bc_pull;
  // one block of forward levels, queued blind: the direction of each level,
  // its list and the end of the traversal are device state (bcb_level_end),
  // so a block is captured once per call and replayed with one host check of
  // the state per block
  auto levels_block = [&]() {
    for (int i = 0; i < kBcbLevelsPerBlock; i++) {
      if (mode == 1 && Kp < 32) {
        bcb_open_edges<<<1024, kBlock, 0, s>>>(depth, n, log_kp, K, g->in_off, c);
        bcb_pull_decide<<<1, 1, 0, s>>>(c);
      }
      fill<unsigned long long>(ctx, hq.count, 0ull, 1);
      const int kt = kt_begin(ctx, GBK_KT_BC_LEVEL);
      bcb_forward_pull<Kp><<<grid_p, kBlock, 0, s>>>(g->in_off, g->in_nbr, n, K, depth, sigma, queue, cap, c, hq);
      bcb_pull_heavy<Kp><<<grid_p, kBlock, 0, s>>>(g->in_nbr, K, depth, sigma, c, hq, hpart);
      bcb_pull_finish<Kp><<<1024, kBlock, 0, s>>>(K, depth, sigma, hq, hpart, queue, cap, c);
      bcb_forward<Kp><<<grid_f, kBlock, 0, s>>>(g->off, g->nbr, depth, sigma, ltag, queue, cap, c, hq);
      bcb_forward_heavy<Kp><<<grid_f, kBlock, 0, s>>>(g->nbr, depth, sigma, ltag, queue, cap, c, hq);
      kt_end(ctx, kt);
      if (mode == 1) bcb_list_edges<<<256, kBlock, 0, s>>>(g->off, queue, cap, c);
      bcb_level_end<<<1, 1, 0, s>>>(c, lvl, cap, g->m, mode, ctx->tune.bc_pull_ratio, Kp < 32);
    }
  };
  LoopGraph loop(ctx);
  if (int rc = loop.capture(levels_block)) return rc;
  EventTimer timer(ctx);
  GBK_HIP(ctx, hipMemcpyAsync(queue, uniq.data(), sizeof(int32_t) * uniq.size(), hipMemcpyHostToDevice, s));
  fill<int32_t>(ctx, depth, -1, n * Kp);
  fill<double>(ctx, sigma, 0.0, n * Kp);
  fill<uint32_t>(ctx, ltag, 0u, n);
  bcb_init_counters<<<1, 1, 0, s>>>(c, static_cast<long long>(uniq.size()), mode == 2, lvl);
  bcb_sources<<<1, kBatchMax, 0, s>>>(dsrc, K, Kp, depth, sigma);
  GBK_LAUNCH_CHECK(ctx, "bc batch sources");
  for (int64_t queued = 0;; queued += kBcbLevelsPerBlock) {
    if (queued > n + 1) return set_error(ctx, GBK_EHIP, "bc batch: more levels than vertices");
    if (int rc = loop.run(levels_block)) return rc;
    GBK_LAUNCH_CHECK(ctx, "bc batch forward");
    GBK_HIP(ctx, hipMemcpyAsync(hc, c, sizeof(BcbCounters), hipMemcpyDeviceToHost, s));
    GBK_HIP(ctx, hipStreamSynchronize(s));
    if (hc->overflow) return set_error(ctx, GBK_EHIP, "bc batch: queue overflow");
    if (hc->done) break;
  }
  std::vector<int64_t> level(static_cast<size_t>(hc->levels) + 1);
  GBK_HIP(ctx, hipMemcpyAsync(level.data(), lvl, sizeof(long long) * level.size(), hipMemcpyDeviceToHost, s));
  GBK_HIP(ctx, hipStreamSynchronize(s));
  const int levels = static_cast<int>(level.size()) - 1;
  bcb_pack<<<grid_for(n * Kp), kBlock, 0, s>>>(depth, n * Kp, dc);
  for (int32_t d = levels - 1; d >= 0; d--) {
    const int64_t lb = level[d], cnt = level[d + 1] - lb;
    fill<unsigned long long>(ctx, hq.count, 0ull, 1);
    const int kt = kt_begin(ctx, GBK_KT_BC_LEVEL);
    bcb_backward<Kp><<<grid_b, kBlock, 0, s>>>(g->off, g->nbr, depth, sigma, dc, scores, queue + lb, cnt, d, c, hq);
    bcb_backward_heavy<Kp><<<grid_b, kBlock, 0, s>>>(g->nbr, depth, dc, d, c, hq, hpart);
    bcb_backward_finish<Kp><<<1024, kBlock, 0, s>>>(depth, sigma, dc, scores, d, hq, hpart);
    kt_end(ctx, kt);
  }
  GBK_LAUNCH_CHECK(ctx, "bc batch backward");
  int64_t etot = 0;
  if (int rc = collect(ctx, g, depth, Kp, K, -1, nullptr, nullptr, nullptr, 0, &etot)) return rc;
  acc->device_ms += timer.stop_ms();
  GBK_HIP(ctx, hipMemcpyAsync(hc, c, sizeof(BcbCounters), hipMemcpyDeviceToHost, s));
  GBK_HIP(ctx, hipStreamSynchronize(s));
  if (hc->overflow) return set_error(ctx, GBK_EHIP, "bc batch: heavy queue overflow");
  acc->rounds += levels;
  acc->edges_reached += etot;
  acc->rows_scanned += level.back();
  acc->edges_scanned += static_cast<int64_t>(hc->edges_scanned);
  return GBK_OK;
}

int bc_batch_chunk(gbk_ctx *ctx, const gbk_graph *g, const int32_t *src, int K, float *scores,
                   gbk_batch_stats *acc) {
  // up to 4 sources take 16 B depth rows: half the footprint of width 8, and
  // 16 edges in flight per wave; 17..32 sources take 64-lane rows, whose
  // wave-uniform walk beats two 32-lane groups per wave
  switch (K <= 4 ? 4 : K > 16 ? 64 : batch_width(K)) {
    case 4: return bc_batch_run<4>(ctx, g, src, K, scores, acc);
    case 8: return bc_batch_run<8>(ctx, g, src, K, scores, acc);
    case 16: return bc_batch_run<16>(ctx, g, src, K, scores, acc);
    case 32: return bc_batch_run<32>(ctx, g, src, K, scores, acc);
    default: return bc_batch_run<64>(ctx, g, src, K, scores, acc);
  }
}

}  // namespace

extern "C" int gbk_bfs_batch(gbk_ctx *ctx, const gbk_graph *g, const int32_t *sources, int num_sources,
                             int32_t *depth, int64_t *reached, int64_t *edges_reached,
                             gbk_batch_stats *stats) {
  if (int rc = check_sources(ctx, g, sources, num_sources)) return rc;
  gbk_batch_stats acc{};
  for (int first = 0; first < num_sources; first += kBatchMax) {
    const int K = std::min(kBatchMax, num_sources - first);
    if (int rc = msbfs_chunk(ctx, g, sources + first, K, depth, reached, edges_reached, first, &acc)) return rc;
  }
  if (stats) *stats = acc;
  return GBK_OK;
}

extern "C" int gbk_sssp_batch(gbk_ctx *ctx, const gbk_graph *g, const int32_t *sources, int num_sources,
                              int32_t delta, int32_t *dist, int64_t *reached, int64_t *edges_reached,
                              gbk_batch_stats *stats) {
  if (int rc = check_sources(ctx, g, sources, num_sources)) return rc;
  if (!g->weighted) return set_error(ctx, GBK_ESTATE, "sssp batch: graph has no weights");
  if (delta <= 0) return set_error(ctx, GBK_EINVAL, "sssp batch: delta must be positive");
  gbk_batch_stats acc{};
  for (int first = 0; first < num_sources; first += kBatchMax) {
    const int K = std::min(kBatchMax, num_sources - first);
    if (int rc = sssp_batch_chunk(ctx, g, sources + first, K, delta, dist, reached, edges_reached, first, &acc))
      return rc;
  }
  if (stats) *stats = acc;
  return GBK_OK;
}

extern "C" int gbk_bc_batch(gbk_ctx *ctx, const gbk_graph *g, const int32_t *sources, int num_sources,
                            float *scores_out, gbk_batch_stats *stats) {
  if (int rc = check_sources(ctx, g, sources, num_sources)) return rc;
  const int64_t n = g->n;
  hipStream_t s = ctx->stream;
  auto *scores = static_cast<float *>(scratch(ctx, kS10, sizeof(float) * (n + 1)));
  auto *pmax = static_cast<unsigned int *>(scratch(ctx, kS11, sizeof(unsigned int) * kGridCap));
  if (!scores || !pmax) return set_error(ctx, GBK_ENOMEM, "bc batch scores");
  gbk_batch_stats acc{};
  // device time: each chunk times its own region (after capturing its level
  // loop), plus the clearing and the normalisation
  {
    EventTimer t0(ctx);
    fill<float>(ctx, scores, 0.0f, n);
    acc.device_ms += t0.stop_ms();
  }
  for (int first = 0; first < num_sources; first += kBatchMax) {
    const int K = std::min(kBatchMax, num_sources - first);
    if (int rc = bc_batch_chunk(ctx, g, sources + first, K, scores, &acc)) return rc;
  }
  EventTimer timer(ctx);
  const unsigned ng = expand_grid(n);
  bc_max<<<ng, kBlock, 0, s>>>(scores, n, pmax);
  bc_normalize<<<ng, kBlock, 0, s>>>(scores, n, pmax, static_cast<int>(ng));
  GBK_LAUNCH_CHECK(ctx, "bc batch normalize");
  acc.device_ms += timer.stop_ms();
  if (scores_out) {
    GBK_HIP(ctx, hipMemcpyAsync(scores_out, scores, sizeof(float) * n, hipMemcpyDeviceToHost, s));
    GBK_HIP(ctx, hipStreamSynchronize(s));
  }
  if (stats) *stats = acc;
  return GBK_OK;
}
