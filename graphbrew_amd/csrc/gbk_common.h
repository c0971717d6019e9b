// gbk_common.h -- shared device/host definitions for the gfx950 graph kernels.
//
// Data layout in HBM (one gbk_graph):
//   off[n+1]  int64   CSR row offsets (SGOffset, gapbs/graph.h:215)
//   nbr[m]    int32   neighbour ids, rows sorted ascending
//   wgt[m]    int32   optional edge weights, parallel to nbr (SoA)
//   in_*              the in-CSR of a directed graph; aliases the out arrays
//                     for an undirected graph (gapbs/graph.h:293-299)
//   heavy_out/heavy_in  ids of rows whose degree exceeds kHeavyDegree, sorted,
//                     built once at upload; those rows get block-wide work.
// Everything a kernel needs stays resident; no host round trip per launch
// except where a frontier decision needs a count on the host.
#pragma once

#include <unordered_map>
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/gbk.h"

namespace gbk {

constexpr int kWave = 64;
constexpr int kBlock = 256;                 // 4 waves per workgroup
constexpr int kWavesPerBlock = kBlock / kWave;

// ----------------------------------------------------------------- tunables
struct Tuning {
  int64_t heavy_degree = 2048;  // rows above this get a whole workgroup
  int bu_group = 1;             // lanes per vertex in the bottom-up BFS step
  int bfs_alpha = 15;           // bfs.cc:165 defaults
  int bfs_beta = 18;
  int xcd_swizzle = 1;          // contiguous vertex ranges per XCD
  int relax_grid = 1024;        // workgroups of a frontier (queue) launch
  int graphs = 1;               // iteration loops replayed as captured hipGraphs
  int batch_grid = 0;           // workgroups of a source batch's launches (0: as many as stay resident)
  int pr_blocks = 0;            // row blocks of a Gauss-Seidel PageRank sweep (0: by size)
  int pr_inplace = 0;           // Gauss-Seidel blocks update contributions in place (1) or publish them
  int pr_unroll = 0;            // 64-edge chunks a PageRank wave has in flight (1, 2, 4, 8; 0: 8 for
                                // Gauss-Seidel blocks, 1 for Jacobi's one launch)
  int bc_pull = 1;              // BC batch forward levels: 0 top-down, 1 bottom-up when the
                                // frontier has more out-edges than the unlisted vertices, 2 bottom-up
  int bc_pull_ratio = 1;        // a frontier edge weighs this many open edges in that choice
  int sssp_wide = 1;            // SSSP batches of 17..32 sources on 64-lane rows
  int64_t sssp_heavy = 256;     // one-source SSSP: rows with more edges to walk go to the
                                // workgroup-wide heavy kernel
  int sssp_tail = 1;            // one-source SSSP: after a pulled heavy pass, the vertices left are
                                // relaxed Bellman-Ford style to the end (0: buckets to the end)
  int sssp_pull = 1;            // one-source SSSP heavy pass: 0 pushed from the bucket's vertices,
                                // 1 pulled into the unsettled vertices when that reads fewer
                                // edges, 2 always pulled
  int msbfs_group = 32;         // lanes per vertex in the BFS batch's bottom-up step (1..64)
  int pr_segment = 0;           // PageRank segmented sweep: 0 by size, -1 never, k > 0 smallest
                                // segment width 2^k tried (6..26)
};
constexpr int kGridCapRuntime = 2048;  // upper bound of relax_grid
constexpr int64_t kHeavyChunkEdges = 1024;  // heavy-row chunk of the all-rows kernels

// ------------------------------------------------------------------ helpers
__device__ __forceinline__ int lane_id() { return __lane_id(); }

__device__ __forceinline__ int lanemask_lt_count(uint64_t mask) {
  // popcount of mask bits below this lane (mbcnt).
  return __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(mask >> 32),
                                   __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(mask), 0));
}

// Block index remapped so each of the 8 XCDs (each with its own L2) gets a
// contiguous range of logical blocks instead of every 8th one.
__device__ __forceinline__ int64_t swizzled_block(int enable) {
  const int64_t g = gridDim.x, b = blockIdx.x;
  if (!enable || g < 16) return b;
  const int64_t xcd = b % 8, local = b / 8, per = g / 8, rem = g % 8;
  return xcd * per + (xcd < rem ? xcd : rem) + local;
}

// Wave-aggregated append: every lane with pred=true gets a unique slot in
// [*tail, *tail + count).  One atomic per wave.
__device__ __forceinline__ int64_t wave_append(unsigned long long *tail, bool pred) {
  const uint64_t mask = __ballot(pred);
  if (mask == 0) return -1;
  const int leader = __ffsll(static_cast<long long>(mask)) - 1;
  unsigned long long base = 0;
  if (lane_id() == leader) base = atomicAdd(tail, static_cast<unsigned long long>(__popcll(mask)));
  base = __shfl(base, leader);
  return pred ? static_cast<int64_t>(base) + lanemask_lt_count(mask) : -1;
}

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
  for (int o = kWave / 2; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

template <typename T>
__device__ __forceinline__ T wave_max(T v) {
  for (int o = kWave / 2; o > 0; o >>= 1) {
    T w = __shfl_xor(v, o);
    v = w > v ? w : v;
  }
  return v;
}

template <typename T>
__device__ __forceinline__ T wave_min(T v) {
  for (int o = kWave / 2; o > 0; o >>= 1) {
    T w = __shfl_xor(v, o);
    v = w < v ? w : v;
  }
  return v;
}

// The value of x held by lane (lane ^ J), J in {1, 2, 4, 8, 16, 32}, moved
// by the VALU instead of the LDS crossbar (ds_bpermute): DPP quad permutes
// for 1 and 2, DPP row shifts for 4 and 8 (a lane takes the shift toward
// its partner), gfx950's v_permlane16_swap / v_permlane32_swap for 16 and
// 32 (with both operands x, the swapped halves hold every lane's partner).
// Every lane of the wave must be active.
template <int J>
__device__ __forceinline__ uint32_t lane_xor(uint32_t x) {
  static_assert(J == 1 || J == 2 || J == 4 || J == 8 || J == 16 || J == 32, "lane_xor stride");
  if constexpr (J == 1) {
    return __builtin_amdgcn_update_dpp(0u, x, 0xB1, 0xF, 0xF, false);  // quad_perm [1,0,3,2]
  } else if constexpr (J == 2) {
    return __builtin_amdgcn_update_dpp(0u, x, 0x4E, 0xF, 0xF, false);  // quad_perm [2,3,0,1]
  } else if constexpr (J == 4 || J == 8) {
    const uint32_t from_up = __builtin_amdgcn_update_dpp(0u, x, 0x100 + J, 0xF, 0xF, false);  // row_shl:J
    const uint32_t from_dn = __builtin_amdgcn_update_dpp(0u, x, 0x110 + J, 0xF, 0xF, false);  // row_shr:J
    return (__lane_id() & J) ? from_dn : from_up;
  } else if constexpr (J == 16) {
    const auto r = __builtin_amdgcn_permlane16_swap(x, x, false, false);
    return (__lane_id() & 16) ? r[0] : r[1];
  } else {
    const auto r = __builtin_amdgcn_permlane32_swap(x, x, false, false);
    return (__lane_id() & 32) ? r[0] : r[1];
  }
}

template <int J, typename T>
__device__ __forceinline__ T lane_xor_t(T x) {
  static_assert(sizeof(T) == 4 || sizeof(T) == 8, "lane_xor_t width");
  if constexpr (sizeof(T) == 8) {
    const uint64_t u = static_cast<uint64_t>(x);
    const uint32_t lo = lane_xor<J>(static_cast<uint32_t>(u));
    const uint32_t hi = lane_xor<J>(static_cast<uint32_t>(u >> 32));
    return static_cast<T>((static_cast<uint64_t>(hi) << 32) | lo);
  } else {
    return static_cast<T>(lane_xor<J>(static_cast<uint32_t>(x)));
  }
}

// lane_xor_t for a stride known only after unrolling (j a power of two
// below 64).
template <typename T>
__device__ __forceinline__ T lane_xor_rt(T x, int j) {
  switch (j) {
    case 1: return lane_xor_t<1>(x);
    case 2: return lane_xor_t<2>(x);
    case 4: return lane_xor_t<4>(x);
    case 8: return lane_xor_t<8>(x);
    case 16: return lane_xor_t<16>(x);
    default: return lane_xor_t<32>(x);
  }
}

// Inclusive scan across the wave.
template <typename T>
__device__ __forceinline__ T wave_incl_scan(T v) {
  const int l = lane_id();
  for (int o = 1; o < kWave; o <<= 1) {
    T w = __shfl_up(v, o);
    if (l >= o) v += w;
  }
  return v;
}

// Block-wide sum; every thread gets the result.  'smem' holds kWavesPerBlock.
template <typename T>
__device__ __forceinline__ T block_sum(T v, T *smem) {
  v = wave_sum(v);
  const int w = threadIdx.x / kWave;
  __syncthreads();
  if (lane_id() == 0) smem[w] = v;
  __syncthreads();
  T t = 0;
  for (int i = 0; i < kWavesPerBlock; i++) t += smem[i];
  return t;
}

__device__ __forceinline__ bool bit_get(const uint32_t *bm, int64_t i) {
  return (bm[i >> 5] >> (i & 31)) & 1u;
}
__device__ __forceinline__ void bit_set_atomic(uint32_t *bm, int64_t i) {
  atomicOr(&bm[i >> 5], 1u << (i & 31));
}

// Fill kernel used instead of hipMemsetAsync so that every clear is an
// ordinary kernel node (safe under stream capture).
template <typename T>
__global__ void fill_kernel(T *p, T v, int64_t n) {
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride)
    p[i] = v;
}

inline unsigned grid_for(int64_t items, int per_block = kBlock, int64_t cap = 1 << 20) {
  int64_t g = (items + per_block - 1) / per_block;
  if (g < 1) g = 1;
  if (g > cap) g = cap;
  return static_cast<unsigned>(g);
}

}  // namespace gbk

// ------------------------------------------------------------ host objects
struct gbk_graph {
  int64_t n = 0;
  int64_t m = 0;          // directed edges in the out-CSR
  int directed = 0;
  int weighted = 0;
  int64_t *off = nullptr;
  int32_t *nbr = nullptr;
  int32_t *wgt = nullptr;
  int64_t *in_off = nullptr;
  int32_t *in_nbr = nullptr;
  int32_t *in_wgt = nullptr;
  int32_t *org_ids = nullptr;  // device copy of original ids (new -> original)
  // rows with degree > heavy_degree, ascending ids
  int32_t *heavy_out = nullptr;
  int64_t n_heavy_out = 0;
  int32_t *heavy_in = nullptr;
  int64_t n_heavy_in = 0;
  int64_t max_out_degree = 0;
  int64_t max_in_degree = 0;
  int64_t heavy_deg = 0;       // threshold the heavy lists were built with
  // Edge-balanced work units of the all-rows kernels, per CSR: tile starts
  // (ntiles + 1 entries, the last = n; a tile has at most 64 rows and about
  // 2048 light edges) and the chunk prefix of the heavy rows (nheavy + 1).
  int32_t *tiles_out = nullptr, *tiles_in = nullptr;
  int64_t ntiles_out = 0, ntiles_in = 0;
  int64_t *hchunk_out = nullptr, *hchunk_in = nullptr;
  // Smallest and largest out-edge weight, found by the first call that needs
  // them (-1 until then); the graph is immutable, so they stay valid.
  int64_t wmin = -1, wmax = -1;
  // One-source SSSP's light/heavy layout of the out-CSR for delta
  // 'part_delta' (-1: none yet): each row's edges lighter than delta first,
  // then the rest, each class in CSR order, as (neighbour, weight) pairs (one
  // 8-byte load per edge); plc[u] = light edges of row u.
  int64_t part_delta = -1;
  int2 *pnw = nullptr;
  int32_t *plc = nullptr;
  // PageRank's segmented layout of the in-CSR for segment width 2^seg_lw
  // (0: none yet), one allocation of seg_bytes, built by the first call.
  int seg_lw = 0;
  void *seg_mem = nullptr;
  size_t seg_bytes = 0;
  std::vector<void *> owned;   // device allocations freed with the graph
};

struct gbk_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  std::string err;
  gbk::Tuning tune;
  // grow-only scratch slots
  struct Slot {
    void *p = nullptr;
    size_t bytes = 0;
  };
  Slot slot[18];
  void *pinned = nullptr;      // small pinned staging for counters
  size_t pinned_bytes = 0;
  // Per-launch timing of one selected kernel (GBK_KT_*): a start/stop event
  // pair is recorded on the stream around each of its launches and read back
  // after the entry point's closing synchronisation.  kt_dev holds device
  // counters [rows, edges] plus per-workgroup partials for kernels whose work
  // per launch is data dependent.
  int ktime = 0;
  std::vector<hipEvent_t> kt_ev;
  size_t kt_used = 0;
  double kt_ms = 0;
  int64_t kt_launches = 0;
  unsigned long long *kt_dev = nullptr;
  // resident workgroups per kernel (occupancy x compute units), cached
  int cu_count = 0;
  std::unordered_map<const void *, unsigned> resident;
  // Graph arrays (gbk_graph_alloc) of freed graphs, kept for the next graphs
  // of about the same size instead of hipFree / hipMalloc (0.5-1 GB arrays
  // took 1-3 ms of host time per relabel call); released with the context.
  struct PoolBlock {
    void *p;
    size_t bytes;
  };
  std::vector<PoolBlock> gpool;
  std::unordered_map<void *, size_t> gsize;  // live and cached graph arrays
  // Steps the last one-source SSSP call took, with its graph and delta: the
  // next call on them queues that many before its first look at the state.
  const gbk_graph *sssp_hint_graph = nullptr;
  int64_t sssp_hint_delta = 0;
  int64_t sssp_hint_steps = 0;
  // The same for SSSP source batches, keyed by the batch's shape.
  struct BatchHint {
    int64_t n, m, k, delta, steps;
  };
  BatchHint sssp_batch_hint{0, 0, 0, 0, 0};
};

namespace gbk {

// Scratch slot ids (each kernel family documents which it uses).
enum SlotId {
  kS0 = 0, kS1, kS2, kS3, kS4, kS5, kS6, kS7, kS8, kS9, kS10, kS11, kS12, kS13, kS14, kS15, kS16, kS17
};

int set_error(gbk_ctx *ctx, int code, const std::string &msg);
// Workgroups of kBlock lanes of 'kernel' that the device keeps resident at
// once (its occupancy times the compute units), or the "batch_grid"
// parameter when set: the grid of a grid-stride launch that fills the GPU
// exactly, with no second round of workgroups.
unsigned resident_grid(gbk_ctx *ctx, const void *kernel);
void *scratch(gbk_ctx *ctx, int slot, size_t bytes);   // nullptr on failure
// Pinned host staging for counters read back by entry points: one fixed
// area of kPinnedBytes per context, so pointers stay valid for its lifetime.
constexpr size_t kPinnedBytes = 64 * 1024;
void *pinned(gbk_ctx *ctx, size_t bytes);
int check_hip(gbk_ctx *ctx, hipError_t e, const char *what);

template <typename T>
inline void fill(gbk_ctx *ctx, T *p, T v, int64_t n) {
  if (n <= 0) return;
  int64_t g = (n + kBlock - 1) / kBlock;
  if (g > 4096) g = 4096;
  fill_kernel<T><<<static_cast<unsigned>(g), kBlock, 0, ctx->stream>>>(p, v, n);
}

// Device-wide exclusive scan of int64 counts (out[i] = sum in[0..i), out[n] = total).
int exclusive_scan_i64(gbk_ctx *ctx, const int64_t *in, int64_t *out, int64_t n);
int exclusive_scan_i32(gbk_ctx *ctx, const int32_t *in, int64_t *out, int64_t n);
// Sorts every CSR row by (neighbour, weight) into new arrays (identity relabel).
int sort_rows(gbk_ctx *ctx, int64_t n, const int64_t *off, const int32_t *nbr, const int32_t *wgt,
              const int32_t *iota, int64_t *deg, int64_t *off_out, int32_t *nbr_out,
              int32_t *wgt_out);
// Stable compaction of ids [0,n) where flag(i) != 0 (flags given as int32).
int compact_flags(gbk_ctx *ctx, const int32_t *flags, int64_t n, int32_t *out, int64_t *count);
// The reference's random mapping (SplitMix64 Fisher-Yates, reorder_basic.h:
// 104-135) of n vertices into new_ids (device), gbk_shuffle.hip.
int random_mapping(gbk_ctx *ctx, int64_t n, int32_t *new_ids);
// Graph arrays from the context's cache of freed graph arrays (best fit
// within 1/8 above the request) or hipMalloc; graph_mem_free returns them to
// the cache (pointers not from graph_mem_alloc are hipFree'd).  Work on the
// context's stream is ordered, so a cached array is reused only after the
// work of its former graph.
void *graph_mem_alloc(gbk_ctx *ctx, size_t bytes);
void graph_mem_free(gbk_ctx *ctx, void *p);
// Smallest and largest edge weight of a weighted graph, found once and kept
// with the graph (g->wmin, g->wmax; 0 when unweighted or empty); 'partial'
// holds 2048 entries of device scratch.
int graph_weight_range(gbk_ctx *ctx, const gbk_graph *g, long long *partial);
// Heavy-row list for a CSR.
int build_heavy_list(gbk_ctx *ctx, const int64_t *off, int64_t n, int64_t threshold,
                     int32_t **list, int64_t *count, int64_t *max_degree,
                     std::vector<void *> &owned);
// Records ev0/ev1 around a region on the context stream.
struct EventTimer {
  gbk_ctx *ctx;
  explicit EventTimer(gbk_ctx *c);
  double stop_ms();  // synchronises, then collects kernel timer pairs
};

// Kernel timer: kt_begin returns a pair index when 'id' is the selected
// kernel (else -1); kt_end closes it; kt_collect reads finished pairs after
// the stream has been synchronised.
int kt_begin(gbk_ctx *ctx, int id);
void kt_end(gbk_ctx *ctx, int pair);
void kt_collect(gbk_ctx *ctx);
constexpr int kKtPartials = 2048;  // partial slots per counter in kt_dev

}  // namespace gbk

#define GBK_HIP(ctx, expr)                                              \
  do {                                                                  \
    hipError_t _e = (expr);                                             \
    if (_e != hipSuccess) return gbk::check_hip((ctx), _e, #expr);      \
  } while (0)

#define GBK_LAUNCH_CHECK(ctx, what)                                     \
  do {                                                                  \
    hipError_t _e = hipGetLastError();                                  \
    if (_e != hipSuccess) return gbk::check_hip((ctx), _e, what);       \
  } while (0)

namespace gbk {

// A block of loop iterations recorded once per call as a hipGraph (stream
// capture of the kernels 'body' launches on ctx->stream) and replayed.
// Callers capture before their timer starts, so the device time covers the
// replays only.  With capture off (parameter "graphs" = 0) or per-launch
// timing on (its events stay outside graphs), run() launches the body
// eagerly.  A body launches kernels only: no memset nodes, no host
// synchronisation.
class LoopGraph {
 public:
  explicit LoopGraph(gbk_ctx *ctx) : ctx_(ctx) {}
  LoopGraph(const LoopGraph &) = delete;
  LoopGraph &operator=(const LoopGraph &) = delete;
  ~LoopGraph() {
    if (exec_) (void)hipGraphExecDestroy(exec_);
  }
  template <class Body>
  int capture(Body &&body) {
    if (!ctx_->tune.graphs || ctx_->ktime != 0) return 0;
    hipGraph_t graph = nullptr;
    GBK_HIP(ctx_, hipStreamBeginCapture(ctx_->stream, hipStreamCaptureModeThreadLocal));
    body();
    const hipError_t e = hipStreamEndCapture(ctx_->stream, &graph);
    if (e != hipSuccess) return check_hip(ctx_, e, "loop graph capture");
    const hipError_t e2 = hipGraphInstantiate(&exec_, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e2 != hipSuccess) {
      exec_ = nullptr;
      return check_hip(ctx_, e2, "loop graph instantiate");
    }
    return 0;
  }
  template <class Body>
  int run(Body &&body) {
    if (!exec_) {
      body();
      return 0;
    }
    GBK_HIP(ctx_, hipGraphLaunch(exec_, ctx_->stream));
    return 0;
  }

 private:
  gbk_ctx *ctx_;
  hipGraphExec_t exec_ = nullptr;
};

}  // namespace gbk
