// gbk_pr.hip -- PageRank pull on gfx950.
//
// PageRankPullGS (bench/src/pr.cc:40-94): Gauss-Seidel.  The reference
// updates contributions in place, so a row may read neighbours' values of
// this sweep or the last one (its relaxed atomics allow either).  On the
// GPU, thousands of waves updating in place in a different order every
// sweep keep only the contraction bound (the L1 change shrinks ~0.82x a
// sweep at Kronecker scale 20, so 20 sweeps do not reach 1e-4 where the
// reference needs 7).  Here a sweep is a block Gauss-Seidel in a fixed
// order instead: the rows' tiles are cut into equal runs (pr_blocks, or by
// default one per ~4096 tiles, at most 2), one launch each,
// and a row of block b reads the new contributions of the
// rows before block b and the previous sweep's for the rest -- no races:
// every run reads the same values (heavy rows' chunk sums are added by
// float atomics, so their last bits can round differently).  The gathers read one buffer: a block's new
// contributions go to a side buffer and a small close launch copies them
// in before the next block (selecting between two buffers per edge costs
// 20%).  A graph too small to fill the GPU from a block's launch runs one
// launch that updates contributions in place, as the reference does: on
// such graphs it converges like the reference's sweep.
// PageRankPull (bench/src/pr_spmv.cc:40-83): Jacobi, the same kernel with
// one block that reads only the previous sweep; contrib_next = score /
// out_degree is the float the reference recomputes.
//
// A block's launch works on the graph's fixed, edge-balanced units: heavy
// in-rows (more than the graph's heavy threshold) in 1024-edge chunks, each
// summed by a workgroup and added into the row's accumulator, then tiles of
// at most 64 rows and about 2048 edges, one per wave.  A tile's in-edges are
// contiguous in the CSR, so the wave streams them in 64-edge chunks with
// coalesced neighbour reads, maps each edge to its owner lane by a 6-step
// shuffle search over the tile's edge prefix, gathers the contribution and
// accumulates into the owner's LDS slot.  A small close launch then
// finishes the block's heavy rows (and, for Gauss-Seidel, publishes the
// block's contributions); finishing a heavy row in the workgroup that adds
// its last chunk would need an agent-scope acquire/release per chunk, which
// writes back and invalidates the L2 of a multi-XCD GPU (2.4x slower).  A
// one-workgroup finish launch adds the per-workgroup L1 changes, decides
// convergence and raises a device flag that makes the remaining queued
// launches return at once -- no host synchronisation inside the iteration
// loop.
#include <algorithm>
#include <mutex>
#include <optional>
#include <string>

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

// Units of one block of rows: its tiles and its heavy rows' chunks.
struct PrBlock {
  int64_t row_lo;             // first row (rows below it are earlier blocks)
  int64_t tile_lo, tile_hi;
  int64_t chunk_lo, chunk_hi;
  int64_t heavy_lo, heavy_hi;  // the block's entries of the heavy list
};

__device__ __forceinline__ int64_t lower_index(const int32_t *a, int64_t len, int64_t key) {
  int64_t lo = 0, hi = len;
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if (a[mid] < key) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

// Block b holds tiles [b*T/B, (b+1)*T/B) of the T tiles (edge-balanced, so
// the blocks are too, whatever the order of the rows -- hubs packed at the
// front by a degree ordering included); its heavy rows are the heavy-list
// entries in its row range (the list is ascending).  Also clears the heavy
// rows' accumulators.
__global__ void pr_blocks_init(RowTiles rt, int B, PrBlock *blk, float *hacc) {
  const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t < rt.nheavy) hacc[t] = 0.0f;
  if (t < B) {
    auto tile_at = [&](int64_t k) -> int64_t { return rt.ntiles * k / B; };
    PrBlock pb;
    pb.tile_lo = tile_at(t);
    pb.tile_hi = tile_at(t + 1);
    pb.row_lo = rt.tiles[pb.tile_lo];
    const int64_t h0 = lower_index(rt.heavy, rt.nheavy, pb.row_lo);
    const int64_t h1 = lower_index(rt.heavy, rt.nheavy, rt.tiles[pb.tile_hi]);
    pb.chunk_lo = rt.nheavy ? rt.hprefix[h0] : 0;
    pb.chunk_hi = rt.nheavy ? rt.hprefix[h1] : 0;
    pb.heavy_lo = h0;
    pb.heavy_hi = h1;
    blk[t] = pb;
  }
}

// One workgroup: adds the per-workgroup L1 changes of iteration 'iter' and
// decides convergence (pr.cc:77-88).
__global__ void pr_finish(const double *partial, int count, PrState *st, double *errs, int iter, double tol,
                          int fixed) {
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

// New score and contribution of row u from its in-sum; returns |change|.
__device__ __forceinline__ double pr_store(int64_t u, float sum, const int64_t *out_off, float *scores,
                                           float *contrib_next, float base, float damp) {
  const float old = scores[u];
  const float nw = base + damp * sum;
  scores[u] = nw;
  const int64_t d = out_off[u + 1] - out_off[u];
  contrib_next[u] = d ? nw / static_cast<float>(d) : 0.0f;
  return fabs(static_cast<double>(nw - old));
}

// One block of rows: heavy chunks first (summed into hacc), then light rows
// tile by tile (one wave per tile of at most 64 rows, the owner of each edge
// found by a shuffle search over the tile's edge prefix).  Contributions are
// read from contrib_cur and the light rows' new ones written to
// contrib_next.
template <int U>
__global__ void __launch_bounds__(kBlock)
pr_pull_block(RowTiles rt, const PrBlock *blocks, int b, const int32_t *in_nbr,
              const int64_t *out_off, float *scores, const float *contrib_cur, float *contrib_next,
              const PrState *st, double *partial, float *hacc, float base, float damp) {
  if (*reinterpret_cast<const volatile int *>(&st->done)) return;
  __shared__ float acc[kBlock];
  __shared__ float hred[kWavesPerBlock];
  const PrBlock blk = blocks[b];
  const int lane = lane_id();
  const int wv = threadIdx.x / kWave;
  double err = 0;
  for (int64_t gc = blk.chunk_lo + blockIdx.x; gc < blk.chunk_hi; gc += gridDim.x) {
    int64_t h, first, last;
    heavy_chunk_of(rt, gc, h, first, last);
    const int32_t *row = in_nbr + rt.off[rt.heavy[h]];
    float part = 0;
    // U neighbour loads, then U gathers, in flight per thread
    for (int64_t k = first + threadIdx.x; k < last; k += U * blockDim.x) {
      int32_t v[U];
#pragma unroll
      for (int j = 0; j < U; j++) v[j] = k + j * blockDim.x < last ? row[k + j * blockDim.x] : -1;
#pragma unroll
      for (int j = 0; j < U; j++) part += v[j] >= 0 ? contrib_cur[v[j]] : 0.0f;
    }
    const float sum = block_sum(part, hred);
    if (threadIdx.x == 0) atomicAdd(&hacc[h], sum);
  }
  const int64_t waves = static_cast<int64_t>(gridDim.x) * kWavesPerBlock;
  for (int64_t t = blk.tile_lo + static_cast<int64_t>(blockIdx.x) * kWavesPerBlock + wv; t < blk.tile_hi;
       t += waves) {
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
    for (int64_t c = 0; c < total; c += U * kWave) {
      int owner[U];
      int32_t v[U];
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
        const int64_t ostart = __shfl(start, lo);
        const int64_t oexcl = __shfl(excl, lo);
        const bool act = e < total;
        owner[j] = act ? lo : kWave;
        v[j] = act ? in_nbr[ostart + (e - oexcl)] : -1;
      }
      float val[U];
#pragma unroll
      for (int j = 0; j < U; j++) val[j] = v[j] >= 0 ? contrib_cur[v[j]] : 0.0f;
      // one writer per owner per chunk; chunks apply in order in this wave
#pragma unroll
      for (int j = 0; j < U; j++) {
        bool tail;
        const float run = seg_run_sum(val[j], owner[j], tail);
        if (owner[j] < kWave && tail) acc[wv * kWave + owner[j]] += run;
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "workgroup");
    __builtin_amdgcn_wave_barrier();
    if (light) err += pr_store(u, acc[threadIdx.x], out_off, scores, contrib_next, base, damp);
    __builtin_amdgcn_wave_barrier();
  }
  block_partial(err, partial + static_cast<int64_t>(b) * gridDim.x);
}

// Closes block b: finishes its heavy rows from their accumulators (cleared
// for the next sweep), writing their contributions to 'contrib'; with a
// side buffer (Gauss-Seidel) also copies the block's light rows' new
// contributions from it into 'contrib'.  L1 changes into the partials.
__global__ void __launch_bounds__(kBlock)
pr_block_close(RowTiles rt, const PrBlock *blocks, int b, int64_t row_hi, const int64_t *out_off, float *scores,
               float *contrib, const float *side, const PrState *st, double *partial, float *hacc, float base,
               float damp) {
  if (*reinterpret_cast<const volatile int *>(&st->done)) return;
  const PrBlock blk = blocks[b];
  const int64_t tid = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const int64_t nth = static_cast<int64_t>(gridDim.x) * blockDim.x;
  double err = 0;
  for (int64_t h = blk.heavy_lo + tid; h < blk.heavy_hi; h += nth) {
    const float sum = hacc[h];
    hacc[h] = 0.0f;
    err += pr_store(rt.heavy[h], sum, out_off, scores, contrib, base, damp);
  }
  if (side) {
    const int64_t hi = row_hi < 0 ? blocks[b + 1].row_lo : row_hi;
    for (int64_t u = blk.row_lo + tid; u < hi; u += nth)
      if (rt.off[u + 1] - rt.off[u] <= rt.heavy_deg) contrib[u] = side[u];
  }
  block_partial(err, partial + static_cast<int64_t>(b) * gridDim.x);
}

// ------------------------------------------------------- segmented sweep
// Once the contributions outgrow an XCD's 4 MB L2 (more than 2^19 vertices),
// each 4-byte gather of the row-tile sweep above pulls a whole line from the
// Infinity Cache.  The segmented sweep cuts the source range into S <= 32
// segments of 2^lw vertices (2 MiB of contributions at the smallest width)
// and has every wave walk its rows' in-edges segment by segment, so the
// waves of an XCD gather from one L2-sized slice at a time (CSR segmenting:
// Zhang, Kiriansky, Mendis, Amarasinghe, Zaharia and Shun, "Making caches
// work for graph analytics", IEEE BigData 2017).
//
// Layout, rebuilt by each call from the in-CSR and its tiles:
//  * records: tile t's light in-edges ordered by (segment, row, position in
//    the row), stored from in_off[tiles[t]] on -- inside the tile's own
//    edge range; a record is (row - tiles[t]) << 26 | (source - segment
//    base);
//  * tseg: per tile, the S + 1 offsets of its segments' records;
//  * heavy rows: each row's ascending in-edges cut at the segment bounds
//    into pieces, pieces into 256-edge chunks; hpc[s] is segment s's chunk
//    prefix over the heavy list, hps[s][h] the first edge of row h's piece.
// A sweep launch gives each wave G consecutive tiles of the block, their
// rows' sums in LDS; for s = 0 .. S-1 the wave adds the tiles' segment-s
// records, then takes its share of the block's segment-s heavy chunks,
// whose sums go to the heavy rows' accumulators (closed as above).  The
// reads and writes of contributions are the row-tile sweep's, so both
// variants keep their semantics; only the order of the float sums differs.
constexpr int kSegMax = 32;          // segments of the source range
constexpr int kSegMinLog2 = 19;      // 2^19 sources = 2 MiB of contributions
constexpr int kSegMaxLog2 = 26;      // records keep 26 bits of local source
constexpr int kSegMaxG = 16;         // tiles of one wave (64 rows each)
constexpr int64_t kSegChunk = 256;   // heavy-row edges per chunk
constexpr int kSegLaneShift = 26;
constexpr int kSegBuildU = 4;        // chunks in flight per wave in the layout build

struct SegLayout {
  int S = 0;
  int lw = 0;
  bool build = false;       // the call builds it (else it is the graph's, built before)
  int32_t *rec = nullptr;
  int32_t *tseg = nullptr;
  int64_t *hps = nullptr;
  int64_t *hpc = nullptr;
  int64_t *hcnt = nullptr;  // chunk counts per (segment, heavy row)
  int64_t *cbase = nullptr; // S + 1: first entry of segment s in crow
  int32_t *crow = nullptr;  // heavy row of each chunk, segment by segment
};

// Bytes of the layout's arrays, carved from one allocation (256-byte
// aligned) when 'base' is given.
size_t seg_carve(char *base, int64_t m, int64_t ntiles, int64_t nheavy, int S, SegLayout *L) {
  size_t off = 0;
  auto take = [&](size_t bytes) {
    char *q = base ? base + off : nullptr;
    off += (bytes + 255) / 256 * 256;
    return q;
  };
  const int64_t nh = std::max<int64_t>(nheavy, 1);
  L->rec = reinterpret_cast<int32_t *>(take(sizeof(int32_t) * std::max<int64_t>(m, 1)));
  L->tseg = reinterpret_cast<int32_t *>(take(sizeof(int32_t) * ntiles * (S + 1)));
  L->hps = reinterpret_cast<int64_t *>(take(sizeof(int64_t) * S * nh));
  L->hpc = reinterpret_cast<int64_t *>(take(sizeof(int64_t) * S * (nheavy + 1)));
  L->hcnt = reinterpret_cast<int64_t *>(take(sizeof(int64_t) * S * nh));
  L->cbase = reinterpret_cast<int64_t *>(take(sizeof(int64_t) * (S + 1)));
  // every piece ends at most one chunk early: m / 256 + S (nheavy + 1) bounds the chunks
  L->crow = reinterpret_cast<int32_t *>(take(sizeof(int32_t) * ((m + kSegChunk - 1) / kSegChunk + S * (nheavy + 1))));
  return off;
}

// A record's bits: the row's lane in its tile (6) over the local source (26).
__device__ __forceinline__ int32_t seg_record(int lane, int32_t local) {
  return static_cast<int32_t>((static_cast<uint32_t>(lane) << kSegLaneShift) | static_cast<uint32_t>(local));
}

__device__ __forceinline__ int64_t lower_edge(const int32_t *nbr, int64_t lo, int64_t hi, int64_t key) {
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if (nbr[mid] < key) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

// 6-step owner search over 32-bit exclusive prefixes (a tile's edges).
__device__ __forceinline__ int tile_owner32(int32_t excl, int32_t e) {
  int lo = 0;
#pragma unroll
  for (int step = 32; step >= 1; step >>= 1) {
    const int probe = lo + step;
    const int32_t x = __shfl(excl, probe & (kWave - 1));
    if (probe < kWave && x <= e) lo = probe;
  }
  return lo;
}

// One wave per tile: the tile's light in-edges (rows of at most heavy_deg
// entries), in row order, are counting-sorted by source segment, stably, so
// the records come out ordered by (segment, row, position) whatever the
// order of the sources inside a row.  Pass 1 counts each segment's edges (a
// ballot per segment and 64 edges; lane s keeps segment s's count), their
// exclusive scan gives each segment's first record, pass 2 ranks every edge
// among its batch's edges of the same segment (ballot + mbcnt) and writes
// its record.  kSegBuildU batches of 64 edges are in flight per wave.
// Records stay inside the tile's edge range.
__device__ __forceinline__ void seg_batch_load(const int32_t *in_nbr, int64_t tbase, int32_t excl, int32_t rel,
                                               int32_t total, int32_t e, int lw, int &own, int32_t &v, int &sg) {
  const int lo = tile_owner32(excl, e);
  // shuffles in uniform control flow: the owner lane may be past the end
  const int32_t k = e - __shfl(excl, lo);
  const int32_t orel = __shfl(rel, lo);
  own = lo;
  v = e < total ? in_nbr[tbase + orel + k] : 0;
  sg = e < total ? (v >> lw) : -1;
}

__global__ void __launch_bounds__(kBlock)
seg_records(RowTiles rt, const int32_t *in_nbr, int S, int lw, int32_t *rec, int32_t *tseg) {
  const int lane = lane_id();
  const int wv = threadIdx.x / kWave;
  const int32_t mask = (1 << lw) - 1;
  for (int64_t t = static_cast<int64_t>(blockIdx.x) * kWavesPerBlock + wv; t < rt.ntiles;
       t += static_cast<int64_t>(gridDim.x) * kWavesPerBlock) {
    const int64_t row0 = rt.tiles[t];
    const int64_t u = row0 + lane;
    const int64_t tbase = rt.off[row0];
    int64_t start = 0, len = 0;
    if (u < rt.tiles[t + 1]) {
      start = rt.off[u];
      len = rt.off[u + 1] - start;
    }
    if (len > rt.heavy_deg) len = 0;
    const int32_t len32 = static_cast<int32_t>(len);
    const int32_t rel = static_cast<int32_t>(start - tbase);
    const int32_t incl = wave_incl_scan(len32);
    const int32_t excl = incl - len32;
    const int32_t total = __shfl(incl, kWave - 1);
    // pass 1: lane s counts segment s
    int32_t mine = 0;
    for (int32_t c = 0; c < total; c += kSegBuildU * kWave) {
      int own[kSegBuildU], sg[kSegBuildU];
      int32_t v[kSegBuildU];
#pragma unroll
      for (int j = 0; j < kSegBuildU; j++)
        seg_batch_load(in_nbr, tbase, excl, rel, total, c + j * kWave + lane, lw, own[j], v[j], sg[j]);
#pragma unroll
      for (int j = 0; j < kSegBuildU; j++)
        for (int q = 0; q < S; q++) {
          const uint64_t m = __ballot(sg[j] == q);
          if (lane == q) mine += __popcll(m);
        }
    }
    const int32_t first = wave_incl_scan(mine) - mine;
    if (lane < S) tseg[t * (S + 1) + lane] = first;
    if (lane == 0) tseg[t * (S + 1) + S] = total;
    // pass 2: lane s holds segment s's next record
    int32_t run = first;
    int32_t *out = rec + tbase;
    for (int32_t c = 0; c < total; c += kSegBuildU * kWave) {
      int own[kSegBuildU], sg[kSegBuildU];
      int32_t v[kSegBuildU];
#pragma unroll
      for (int j = 0; j < kSegBuildU; j++)
        seg_batch_load(in_nbr, tbase, excl, rel, total, c + j * kWave + lane, lw, own[j], v[j], sg[j]);
#pragma unroll
      for (int j = 0; j < kSegBuildU; j++) {
        int32_t pos = 0;
        for (int q = 0; q < S; q++) {
          const uint64_t m = __ballot(sg[j] == q);
          const int32_t at = __shfl(run, q);
          if (sg[j] == q) pos = at + lanemask_lt_count(m);
          if (lane == q) run += __popcll(m);
        }
        if (sg[j] >= 0 && pos >= 0 && pos < total) out[pos] = seg_record(own[j], v[j] & mask);
      }
    }
  }
}


// One workgroup per heavy row: its pieces' first edges and chunk counts per
// segment.  A row whose sources are not ascending is one piece in segment
// 0 (every edge is still summed once; it only loses the locality).
__global__ void __launch_bounds__(kBlock)
seg_heavy_pieces(RowTiles rt, const int32_t *in_nbr, int S, int lw, int64_t *hps, int64_t *hcnt) {
  __shared__ int bad;
  for (int64_t h = blockIdx.x; h < rt.nheavy; h += gridDim.x) {
    const int64_t a = rt.off[rt.heavy[h]], b = rt.off[rt.heavy[h] + 1];
    if (threadIdx.x == 0) bad = 0;
    __syncthreads();
    bool mine = false;
#pragma unroll 8
    for (int64_t k = a + 1 + threadIdx.x; k < b; k += blockDim.x) mine |= in_nbr[k] < in_nbr[k - 1];
    if (mine) atomicOr(&bad, 1);
    __syncthreads();
    if (static_cast<int>(threadIdx.x) < S) {
      const int s = threadIdx.x;
      int64_t lo, hi;
      if (bad == 0) {
        lo = s == 0 ? a : lower_edge(in_nbr, a, b, static_cast<int64_t>(s) << lw);
        hi = s + 1 == S ? b : lower_edge(in_nbr, a, b, static_cast<int64_t>(s + 1) << lw);
      } else {
        lo = s == 0 ? a : b;
        hi = b;
      }
      hps[s * rt.nheavy + h] = lo;
      hcnt[s * rt.nheavy + h] = (hi - lo + kSegChunk - 1) / kSegChunk;
    }
    __syncthreads();
  }
}

// One workgroup per segment: hpc[s][h] = chunks of segment s before heavy
// row h (nheavy + 1 entries).
__global__ void __launch_bounds__(kBlock)
seg_heavy_prefix(const int64_t *hcnt, int64_t nheavy, int64_t *hpc) {
  __shared__ int64_t red[kWavesPerBlock];
  const int64_t s = blockIdx.x;
  const int64_t *c = hcnt + s * nheavy;
  int64_t *o = hpc + s * (nheavy + 1);
  const int wv = threadIdx.x / kWave;
  int64_t carry = 0;
  for (int64_t base = 0; base < nheavy; base += blockDim.x) {
    const int64_t i = base + threadIdx.x;
    const int64_t x = i < nheavy ? c[i] : 0;
    const int64_t wi = wave_incl_scan(x);
    if (lane_id() == kWave - 1) red[wv] = wi;
    __syncthreads();
    int64_t before = 0, tot = 0;
    for (int w = 0; w < kWavesPerBlock; w++) {
      if (w < wv) before += red[w];
      tot += red[w];
    }
    if (i < nheavy) o[i] = carry + before + wi - x;
    carry += tot;
    __syncthreads();
  }
  if (threadIdx.x == 0) o[nheavy] = carry;
}

// cbase[s] = chunks of the segments before s (one thread).
__global__ void seg_chunk_bases(const int64_t *hpc, int64_t nheavy, int S, int64_t *cbase) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  int64_t at = 0;
  for (int s = 0; s < S; s++) {
    cbase[s] = at;
    at += hpc[static_cast<int64_t>(s) * (nheavy + 1) + nheavy];
  }
  cbase[S] = at;
}

// crow[cbase[s] + c] = the heavy row of chunk c of segment s.
__global__ void seg_chunk_rows(const int64_t *hpc, const int64_t *cbase, int64_t nheavy, int S, int32_t *crow) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= static_cast<int64_t>(S) * nheavy) return;
  const int64_t s = i / nheavy, h = i % nheavy;
  const int64_t *pc = hpc + s * (nheavy + 1);
  for (int64_t c = pc[h]; c < pc[h + 1]; c++) crow[cbase[s] + c] = static_cast<int32_t>(h);
}

// One row block of a sweep, segment by segment (see above).  Each wave takes
// parts of G tiles, one per round (the rounds are the same for every wave);
// the heavy chunks are shared out in round 0.
template <int U>
__global__ void __launch_bounds__(kBlock)
pr_seg_block(SegLayout L, RowTiles rt, const PrBlock *blocks, int b, int G, const int32_t *in_nbr,
             const int64_t *out_off, float *scores, const float *contrib_cur, float *contrib_next,
             const PrState *st, double *partial, float *hacc, float base, float damp) {
  if (*reinterpret_cast<const volatile int *>(&st->done)) return;
  __shared__ float acc_all[kWavesPerBlock][kWave * kSegMaxG];
  const PrBlock blk = blocks[b];
  const int lane = lane_id();
  const int wv = threadIdx.x / kWave;
  float *acc = acc_all[wv];
  const int S = L.S;
  const int32_t mask = (1 << L.lw) - 1;
  const int64_t nw = static_cast<int64_t>(gridDim.x) * kWavesPerBlock;
  const int64_t gw = static_cast<int64_t>(blockIdx.x) * kWavesPerBlock + wv;
  const int64_t nparts = (blk.tile_hi - blk.tile_lo + G - 1) / G;
  double err = 0;
  for (int64_t round = 0; round == 0 || round * nw < nparts; round++) {
    const int64_t p = round * nw + gw;
    const bool has = p < nparts;
    const int64_t t0 = blk.tile_lo + p * G;
    const int64_t t1 = has ? min(t0 + G, blk.tile_hi) : t0;
    const int64_t row0 = has ? rt.tiles[t0] : 0;
    const int rows = has ? static_cast<int>(rt.tiles[t1] - row0) : 0;
    for (int k = lane; k < rows; k += kWave) acc[k] = 0.0f;
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "workgroup");
    __builtin_amdgcn_wave_barrier();
    // lane j < t1 - t0 holds tile t0 + j: its records' base and first row
    const bool tl = lane < t1 - t0;
    const int64_t tj = t0 + lane;
    const int64_t tbase = tl ? rt.off[rt.tiles[tj]] : 0;
    const int trow = tl ? static_cast<int>(rt.tiles[tj] - row0) : 0;
    const int32_t *ts = L.tseg + (tl ? tj : 0) * (S + 1);
    int32_t lo = tl ? ts[0] : 0;
    int32_t hi = tl ? ts[1] : 0;
    for (int s = 0; s < S; s++) {
      const int32_t nhi = tl && s + 2 <= S ? ts[s + 2] : 0;
      if (has) {
        const int32_t len = hi - lo;
        const int64_t pstart = tbase + lo;
        const int32_t incl = wave_incl_scan(len);
        const int32_t excl = incl - len;
        const int32_t total = __shfl(incl, kWave - 1);
        const int64_t sbase = static_cast<int64_t>(s) << L.lw;
        for (int32_t c = 0; c < total; c += U * kWave) {
          int owner[U];
          int32_t r[U];
#pragma unroll
          for (int j = 0; j < U; j++) {
            const int32_t e = c + j * kWave + lane;
            const int pc = tile_owner32(excl, e);
            const bool act = e < total;
            const int64_t ostart = __shfl(pstart, pc);
            const int32_t oexcl = __shfl(excl, pc);
            const int orow = __shfl(trow, pc);
            // records stream through once: non-temporal, so they do not
            // push the contributions' slice out of the L2
            r[j] = act ? __builtin_nontemporal_load(L.rec + ostart + (e - oexcl)) : 0;
            owner[j] = act ? orow + static_cast<int>(static_cast<uint32_t>(r[j]) >> kSegLaneShift) : kWave * kSegMaxG;
          }
          float val[U];
#pragma unroll
          for (int j = 0; j < U; j++)
            val[j] = owner[j] < kWave * kSegMaxG ? contrib_cur[sbase + (r[j] & mask)] : 0.0f;
#pragma unroll
          for (int j = 0; j < U; j++) {
            bool tail;
            const float run = seg_run_sum(val[j], owner[j], tail);
            if (owner[j] < kWave * kSegMaxG && tail) acc[owner[j]] += run;
          }
        }
      }
      if (round == 0 && rt.nheavy) {
        const int64_t *pc = L.hpc + static_cast<int64_t>(s) * (rt.nheavy + 1);
        const int64_t ce = pc[blk.heavy_hi];
        const int32_t *cr = L.crow + L.cbase[s];
        for (int64_t ci = pc[blk.heavy_lo] + gw; ci < ce; ci += nw) {
          const int64_t hh = cr[ci];
          const int64_t first = L.hps[s * rt.nheavy + hh] + (ci - pc[hh]) * kSegChunk;
          const int64_t pend = s + 1 < S ? L.hps[(s + 1) * rt.nheavy + hh] : rt.off[rt.heavy[hh] + 1];
          const int64_t last = min(first + kSegChunk, pend);
          int32_t v[kSegChunk / kWave];
#pragma unroll
          for (int j = 0; j < kSegChunk / kWave; j++) {
            const int64_t k = first + j * kWave + lane;
            v[j] = k < last ? in_nbr[k] : -1;
          }
          float part = 0.0f;
#pragma unroll
          for (int j = 0; j < kSegChunk / kWave; j++) part += v[j] >= 0 ? contrib_cur[v[j]] : 0.0f;
          part = wave_sum(part);
          if (lane == 0) atomicAdd(&hacc[hh], part);
        }
      }
      lo = hi;
      hi = nhi;
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "workgroup");
    __builtin_amdgcn_wave_barrier();
    for (int k = lane; k < rows; k += kWave) {
      const int64_t u = row0 + k;
      if (rt.off[u + 1] - rt.off[u] <= rt.heavy_deg)
        err += pr_store(u, acc[k], out_off, scores, contrib_next, base, damp);
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "workgroup");
    __builtin_amdgcn_wave_barrier();
  }
  block_partial(err, partial + static_cast<int64_t>(b) * gridDim.x);
}

// ------------------------------------------------------------ host side
// Launch shape of one PageRank call, the same for both entry points.
struct PrPlan {
  int B;             // Gauss-Seidel row blocks (1 for Jacobi)
  unsigned grid;     // workgroups of a block's launch
  unsigned cgrid;    // workgroups of a close launch
  bool publish;      // Gauss-Seidel blocks publish through a side buffer
  bool close;        // a close launch follows each block
  int unroll;        // 64-edge chunks in flight per wave
  size_t nparts;     // per-workgroup L1 partials per iteration
  bool seg = false;  // segmented sweep (pr_seg_block)
  int G = 1;         // tiles per wave part of the segmented sweep
};

// Buffers of one call (scores may be the caller's).
struct PrBufs {
  float *scores, *c0, *c1;
  PrState *st;
  double *errs, *partial;
  float *hacc;
  PrBlock *blocks;
};

// Workgroups of kBlock lanes of pr_pull_block that stay resident, read from
// the context's cache when present, else queried without caching (the
// no-allocation entry point inserts nothing into the cache).
unsigned pr_resident(gbk_ctx *ctx, bool cache) {
  const void *k = reinterpret_cast<const void *>(pr_pull_block<4>);
  if (cache || ctx->tune.batch_grid > 0) return resident_grid(ctx, k);
  auto it = ctx->resident.find(k);
  if (it != ctx->resident.end()) return it->second;
  int cu = ctx->cu_count, per_cu = 0;
  if (cu == 0 && (hipDeviceGetAttribute(&cu, hipDeviceAttributeMultiprocessorCount, ctx->device) != hipSuccess || cu < 1))
    cu = 256;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k, kBlock, 0) != hipSuccess || per_cu < 1) per_cu = 4;
  return static_cast<unsigned>(per_cu) * static_cast<unsigned>(cu);
}

// log2 of the segment width of the segmented sweep for n vertices, 0 when
// it does not apply (one segment, or the "pr_segment" parameter is -1); a
// positive parameter is the smallest width tried.
int seg_log2(const gbk_ctx *ctx, int64_t n) {
  const int force = ctx->tune.pr_segment;
  if (force < 0) return 0;
  int lw = force > 0 ? force : kSegMinLog2;
  auto segs = [&](int w) { return (n + (int64_t(1) << w) - 1) >> w; };
  while (lw < kSegMaxLog2 && segs(lw) > kSegMax) lw++;
  return segs(lw) >= 2 && segs(lw) <= kSegMax ? lw : 0;
}

const void *seg_kernel(int unroll) {
  switch (unroll) {
    case 1: return reinterpret_cast<const void *>(pr_seg_block<1>);
    case 2: return reinterpret_cast<const void *>(pr_seg_block<2>);
    case 8: return reinterpret_cast<const void *>(pr_seg_block<8>);
    default: return reinterpret_cast<const void *>(pr_seg_block<4>);
  }
}

PrPlan pr_plan(gbk_ctx *ctx, const RowTiles &rt, int64_t n, bool gs, bool cache, bool seg = false) {
  PrPlan p;
  // Gauss-Seidel blocks: one per ~4096 tiles (a block's launch then keeps
  // the GPU's waves busy), at most 2 -- at Kronecker scale 22 two blocks
  // converge in the sweeps eight do (6) and save six launch tails a sweep
  // (8.9 against 10.0 ms); a small graph's one block is Jacobi
  int64_t want = ctx->tune.pr_blocks ? ctx->tune.pr_blocks : std::min<int64_t>(2, rt.ntiles / 4096);
  p.B = gs ? static_cast<int>(std::max<int64_t>(1, std::min(want, std::max<int64_t>(rt.ntiles, 1)))) : 1;
  // a block's launch covers about 1/B of the tiles, about one per wave;
  // Jacobi's one launch all of them
  const int64_t block_units = (rt.ntiles / p.B) / kWavesPerBlock + 1;
  p.grid = gs ? static_cast<unsigned>(std::min<int64_t>(pr_resident(ctx, cache), std::max<int64_t>(block_units, 64)))
              : rows_grid(rt);
  // the close launch copies about n/B contributions (Gauss-Seidel) or
  // finishes the heavy rows
  p.publish = gs && p.B > 1 && !ctx->tune.pr_inplace;
  p.cgrid = p.publish ? static_cast<unsigned>(std::min<int64_t>(grid_for(n / p.B + 1), 1024)) : 16;
  p.close = p.publish || rt.nheavy;
  // block launches are short and latency-bound: more loads in flight pay
  // (8 chunks 1% over 4); Jacobi's one launch is bound by the random-line
  // rate and loses 7%
  p.unroll = ctx->tune.pr_unroll ? ctx->tune.pr_unroll : (gs && p.B > 1 ? 8 : 1);
  if (seg) {
    // resident waves; each takes parts of G tiles so that a block's parts
    // are about one per wave
    p.seg = true;
    p.unroll = ctx->tune.pr_unroll ? ctx->tune.pr_unroll : 4;
    p.grid = resident_grid(ctx, seg_kernel(p.unroll));
    const int64_t waves = static_cast<int64_t>(p.grid) * kWavesPerBlock;
    const int64_t per_block = (rt.ntiles + p.B - 1) / p.B;
    p.G = static_cast<int>(std::min<int64_t>(kSegMaxG, std::max<int64_t>(1, (per_block + waves - 1) / waves)));
  }
  p.nparts = static_cast<size_t>(p.B) * (p.grid + (p.close ? p.cgrid : 0));
  return p;
}

// Carves the buffers (scores excluded) from 'base' at 256-byte alignment
// and returns the bytes they take; base may be null to measure.
size_t pr_carve(char *base, int64_t n, int max_iters, const PrPlan &p, int64_t nheavy, PrBufs *b) {
  size_t off = 0;
  auto take = [&](size_t bytes) {
    char *q = base ? base + off : nullptr;
    off += (bytes + 255) / 256 * 256;
    return q;
  };
  b->c0 = reinterpret_cast<float *>(take(sizeof(float) * n));
  b->c1 = reinterpret_cast<float *>(take(sizeof(float) * n));
  b->st = reinterpret_cast<PrState *>(take(sizeof(PrState)));
  b->errs = reinterpret_cast<double *>(take(sizeof(double) * max_iters));
  b->partial = reinterpret_cast<double *>(take(sizeof(double) * p.nparts));
  b->hacc = reinterpret_cast<float *>(take(sizeof(float) * (nheavy + 1)));
  b->blocks = reinterpret_cast<PrBlock *>(take(sizeof(PrBlock) * p.B));
  return off;
}

// Queues the whole call on the context stream: block table, initial
// scores, then max_iters iterations (captured once and replayed when
// 'loop' is given, else launched one by one); kernels only.
int pr_enqueue(gbk_ctx *ctx, const gbk_graph *g, const RowTiles &rt, const PrPlan &p, const PrBufs &b, bool gs,
               int max_iters, double tol, int fixed_work, LoopGraph *loop, std::optional<EventTimer> *timer,
               const SegLayout *L = nullptr) {
  const int64_t n = g->n;
  hipStream_t s = ctx->stream;
  const float kDamp = 0.85f;
  const float init = 1.0f / static_cast<float>(n);
  const float base = (1.0f - kDamp) / static_cast<float>(n);
  // all iterations are queued; after convergence the kernels return at once
  // on the state's done flag
  auto iterations = [&]() {
    for (int it = 0; it < max_iters; it++) {
      const int kt = loop ? kt_begin(ctx, GBK_KT_PR_ITER) : -1;
      // Jacobi: ping-pong buffers; Gauss-Seidel in blocks: c0 read and
      // published into, c1 the side buffer; in one launch: c0 in place
      float *cur = gs ? b.c0 : ((it & 1) ? b.c1 : b.c0);
      float *nxt = gs ? (p.publish ? b.c1 : b.c0) : ((it & 1) ? b.c0 : b.c1);
      for (int blk = 0; blk < p.B; blk++) {
        if (p.seg) {
          switch (p.unroll) {
            case 1:
              pr_seg_block<1><<<p.grid, kBlock, 0, s>>>(*L, rt, b.blocks, blk, p.G, g->in_nbr, g->off, b.scores, cur,
                                                        nxt, b.st, b.partial, b.hacc, base, kDamp);
              break;
            case 2:
              pr_seg_block<2><<<p.grid, kBlock, 0, s>>>(*L, rt, b.blocks, blk, p.G, g->in_nbr, g->off, b.scores, cur,
                                                        nxt, b.st, b.partial, b.hacc, base, kDamp);
              break;
            case 8:
              pr_seg_block<8><<<p.grid, kBlock, 0, s>>>(*L, rt, b.blocks, blk, p.G, g->in_nbr, g->off, b.scores, cur,
                                                        nxt, b.st, b.partial, b.hacc, base, kDamp);
              break;
            default:
              pr_seg_block<4><<<p.grid, kBlock, 0, s>>>(*L, rt, b.blocks, blk, p.G, g->in_nbr, g->off, b.scores, cur,
                                                        nxt, b.st, b.partial, b.hacc, base, kDamp);
          }
        } else switch (p.unroll) {
          case 1:
            pr_pull_block<1><<<p.grid, kBlock, 0, s>>>(rt, b.blocks, blk, g->in_nbr, g->off, b.scores, cur, nxt,
                                                       b.st, b.partial, b.hacc, base, kDamp);
            break;
          case 2:
            pr_pull_block<2><<<p.grid, kBlock, 0, s>>>(rt, b.blocks, blk, g->in_nbr, g->off, b.scores, cur, nxt,
                                                       b.st, b.partial, b.hacc, base, kDamp);
            break;
          case 8:
            pr_pull_block<8><<<p.grid, kBlock, 0, s>>>(rt, b.blocks, blk, g->in_nbr, g->off, b.scores, cur, nxt,
                                                       b.st, b.partial, b.hacc, base, kDamp);
            break;
          default:
            pr_pull_block<4><<<p.grid, kBlock, 0, s>>>(rt, b.blocks, blk, g->in_nbr, g->off, b.scores, cur, nxt,
                                                       b.st, b.partial, b.hacc, base, kDamp);
        }
        if (p.close)
          pr_block_close<<<p.cgrid, kBlock, 0, s>>>(rt, b.blocks, blk, blk + 1 == p.B ? n : -1, g->off, b.scores,
                                                    p.publish ? cur : nxt, p.publish ? nxt : nullptr, b.st,
                                                    b.partial + static_cast<size_t>(p.B) * p.grid, b.hacc, base,
                                                    kDamp);
      }
      kt_end(ctx, kt);
      pr_finish<<<1, 1024, 0, s>>>(b.partial, static_cast<int>(p.nparts), b.st, b.errs, it, tol, fixed_work);
    }
  };
  if (loop)
    if (int rc = loop->capture(iterations)) return rc;
  if (timer) timer->emplace(ctx);
  if (p.seg && L->build) {
    // the graph's segmented layout, built by its first call inside the
    // timed region and kept with the graph
    seg_records<<<grid_for(rt.ntiles, kWavesPerBlock), kBlock, 0, s>>>(rt, g->in_nbr, L->S, L->lw, L->rec, L->tseg);
    if (rt.nheavy)
      seg_heavy_pieces<<<static_cast<unsigned>(std::min<int64_t>(rt.nheavy, 4096)), kBlock, 0, s>>>(
          rt, g->in_nbr, L->S, L->lw, L->hps, L->hcnt);
    seg_heavy_prefix<<<L->S, kBlock, 0, s>>>(L->hcnt, rt.nheavy, L->hpc);
    seg_chunk_bases<<<1, 1, 0, s>>>(L->hpc, rt.nheavy, L->S, L->cbase);
    if (rt.nheavy)
      seg_chunk_rows<<<grid_for(static_cast<int64_t>(L->S) * rt.nheavy), kBlock, 0, s>>>(L->hpc, L->cbase, rt.nheavy,
                                                                                     L->S, L->crow);
    GBK_LAUNCH_CHECK(ctx, "pr segment layout");
  }
  pr_blocks_init<<<grid_for(std::max<int64_t>(rt.nheavy, p.B)), kBlock, 0, s>>>(rt, p.B, b.blocks, b.hacc);
  pr_init<<<grid_for(n), kBlock, 0, s>>>(g->off, n, init, b.scores, b.c0, b.st);
  if (loop) {
    if (int rc = loop->run(iterations)) return rc;
  } else {
    iterations();
  }
  GBK_LAUNCH_CHECK(ctx, "pr_pull_block");
  return GBK_OK;
}

// Iterations executed, last L1 change and first converged iteration (-1
// if none) as doubles, for callers that keep the results on the device.
__global__ void pr_report(const PrState *st, double *out) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    out[0] = st->iters;
    out[1] = st->last_err;
    out[2] = st->converged;
  }
}

// The no-allocation variant's workspace: one array per GPU, held until the
// process ends, shared by every call on that GPU.  Host code uses its
// device address (hipGetSymbolAddress, once per device), never the name.
constexpr size_t kNoAllocBytes = size_t(64) << 20;
static __device__ __align__(256) char noalloc_scratch[kNoAllocBytes];
constexpr int kMaxDevices = 64;
std::once_flag noalloc_once[kMaxDevices];
void *noalloc_ptr[kMaxDevices];
hipError_t noalloc_err[kMaxDevices];

int noalloc_workspace(gbk_ctx *ctx, char **out) {
  const int dev = ctx->device;
  if (dev < 0 || dev >= kMaxDevices) return set_error(ctx, GBK_EINVAL, "pr noalloc: device index");
  std::call_once(noalloc_once[dev], [dev]() {
    int prev = 0;
    const bool restore = hipGetDevice(&prev) == hipSuccess && prev != dev;
    noalloc_err[dev] = hipSetDevice(dev);
    if (noalloc_err[dev] == hipSuccess)
      noalloc_err[dev] = hipGetSymbolAddress(&noalloc_ptr[dev], HIP_SYMBOL(noalloc_scratch));
    if (restore) (void)hipSetDevice(prev);
  });
  if (noalloc_err[dev] != hipSuccess || !noalloc_ptr[dev])
    return check_hip(ctx, noalloc_err[dev] != hipSuccess ? noalloc_err[dev] : hipErrorInvalidSymbol,
                     "pr noalloc: workspace address");
  *out = static_cast<char *>(noalloc_ptr[dev]);
  return GBK_OK;
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
  const bool gs = variant == GBK_PR_GS;
  SegLayout L;
  // 32-bit record counts and in-tile offsets: light rows up to 2^24 edges,
  // rows below 2^30
  L.lw = rt.ntiles > 0 && rt.heavy_deg <= (int64_t(1) << 24) && g->max_in_degree < (int64_t(1) << 30)
             ? seg_log2(ctx, n) : 0;
  if (L.lw) {
    // the layout lives with the graph (immutable), built once per segment
    // width; a call with another width rebuilds it in place
    L.S = static_cast<int>((n + (int64_t(1) << L.lw) - 1) >> L.lw);
    const size_t need = seg_carve(nullptr, g->m, rt.ntiles, rt.nheavy, L.S, &L);
    auto *mg = const_cast<gbk_graph *>(g);
    if (mg->seg_bytes < need) {
      if (mg->seg_mem) {
        GBK_HIP(ctx, hipStreamSynchronize(s));
        graph_mem_free(ctx, mg->seg_mem);
        mg->owned.erase(std::find(mg->owned.begin(), mg->owned.end(), mg->seg_mem));
        mg->seg_mem = nullptr;
        mg->seg_bytes = 0;
        mg->seg_lw = 0;
      }
      void *mem = graph_mem_alloc(ctx, need);
      if (!mem) return set_error(ctx, GBK_ENOMEM, "pr segment layout");
      mg->owned.push_back(mem);
      mg->seg_mem = mem;
      mg->seg_bytes = need;
    }
    seg_carve(static_cast<char *>(mg->seg_mem), g->m, rt.ntiles, rt.nheavy, L.S, &L);
    L.build = mg->seg_lw != L.lw;
    mg->seg_lw = 0;  // set again once the build is queued
  }
  const PrPlan p = pr_plan(ctx, rt, n, gs, true, L.lw != 0);
  PrBufs b;
  b.scores = static_cast<float *>(scratch(ctx, kS0, sizeof(float) * n));
  b.c0 = static_cast<float *>(scratch(ctx, kS1, sizeof(float) * n));
  b.c1 = static_cast<float *>(scratch(ctx, kS2, sizeof(float) * n));
  b.st = static_cast<PrState *>(scratch(ctx, kS3, sizeof(PrState)));
  b.errs = static_cast<double *>(scratch(ctx, kS4, sizeof(double) * max_iters));
  b.partial = static_cast<double *>(scratch(ctx, kS5, sizeof(double) * p.nparts));
  b.hacc = static_cast<float *>(scratch(ctx, kS6, sizeof(float) * (rt.nheavy + 1)));
  b.blocks = static_cast<PrBlock *>(scratch(ctx, kS8, sizeof(PrBlock) * p.B));
  if (!b.scores || !b.c0 || !b.c1 || !b.st || !b.errs || !b.partial || !b.hacc || !b.blocks)
    return set_error(ctx, GBK_ENOMEM, "pr scratch");
  LoopGraph loop(ctx);
  std::optional<EventTimer> timer;
  if (int rc = pr_enqueue(ctx, g, rt, p, b, gs, max_iters, tol, fixed_work, &loop, &timer, &L)) return rc;
  if (L.lw) const_cast<gbk_graph *>(g)->seg_lw = L.lw;
  const double ms = timer->stop_ms();
  PrState h{};
  GBK_HIP(ctx, hipMemcpyAsync(&h, b.st, sizeof(h), hipMemcpyDeviceToHost, s));
  if (scores_out) GBK_HIP(ctx, hipMemcpyAsync(scores_out, b.scores, sizeof(float) * n, hipMemcpyDeviceToHost, s));
  GBK_HIP(ctx, hipStreamSynchronize(s));
  if (stats) {
    stats->device_ms = ms;
    stats->iterations = h.iters;
    stats->converged_iter = h.converged;
    stats->final_error = h.last_err;
  }
  return GBK_OK;
}

extern "C" int64_t gbk_pr_noalloc_max_bytes(void) { return static_cast<int64_t>(kNoAllocBytes); }

extern "C" int gbk_pr_noalloc(gbk_ctx *ctx, const gbk_graph *g, int variant, int max_iters, double tol,
                              int fixed_work, float *scores_dev, double *report_dev) {
  if (!ctx || !g || !scores_dev || max_iters <= 0 || (variant != GBK_PR_GS && variant != GBK_PR_SPMV))
    return set_error(ctx, GBK_EINVAL, "pr noalloc: bad arguments");
  const int64_t n = g->n;
  if (n == 0) return set_error(ctx, GBK_EINVAL, "pr noalloc: empty graph");
  const RowTiles rt = row_tiles(g, true);
  const bool gs = variant == GBK_PR_GS;
  const PrPlan p = pr_plan(ctx, rt, n, gs, false);
  PrBufs b;
  const size_t need = pr_carve(nullptr, n, max_iters, p, rt.nheavy, &b);
  if (need > kNoAllocBytes)
    return set_error(ctx, GBK_EINVAL,
                     "pr noalloc: needs " + std::to_string(need) + " bytes of workspace, more than its " +
                         std::to_string(kNoAllocBytes) + " (about 8 bytes per vertex: n up to ~8M)");
  char *ws = nullptr;
  if (int rc = noalloc_workspace(ctx, &ws)) return rc;
  pr_carve(ws, n, max_iters, p, rt.nheavy, &b);
  b.scores = scores_dev;
  if (int rc = pr_enqueue(ctx, g, rt, p, b, gs, max_iters, tol, fixed_work, nullptr, nullptr)) return rc;
  if (report_dev) {
    pr_report<<<1, 1, 0, ctx->stream>>>(b.st, report_dev);
    GBK_LAUNCH_CHECK(ctx, "pr noalloc report");
  }
  return GBK_OK;
}
