// gbk_runtime.hip -- context, scratch memory, device scan/compaction and
// graph upload/download for libgbk.so.
#include <algorithm>
#include <cstring>

#include "gbk_common.h"

namespace gbk {

int set_error(gbk_ctx *ctx, int code, const std::string &msg) {
  if (ctx) ctx->err = msg;
  return code;
}

int check_hip(gbk_ctx *ctx, hipError_t e, const char *what) {
  if (e == hipSuccess) return GBK_OK;
  const int code = (e == hipErrorOutOfMemory) ? GBK_ENOMEM : GBK_EHIP;
  return set_error(ctx, code, std::string(what) + ": " + hipGetErrorString(e));
}

// Frees the cached arrays of freed graphs (graph_mem_alloc's cache).
static void drop_graph_cache(gbk_ctx *ctx) {
  hipStreamSynchronize(ctx->stream);
  for (auto &b : ctx->gpool) {
    ctx->gsize.erase(b.p);
    hipFree(b.p);
  }
  ctx->gpool.clear();
}

void *scratch(gbk_ctx *ctx, int slot, size_t bytes) {
  auto &s = ctx->slot[slot];
  if (bytes == 0) bytes = 1;
  if (s.bytes >= bytes) return s.p;
  if (s.p) {
    hipStreamSynchronize(ctx->stream);
    hipFree(s.p);
    s.p = nullptr;
    s.bytes = 0;
  }
  // Round up so small growth steps do not reallocate every call.
  size_t want = bytes + bytes / 8 + 4096;
  if (hipMalloc(&s.p, want) != hipSuccess) {
    // the cache of freed graph arrays may hold the memory: release it once
    (void)hipGetLastError();
    drop_graph_cache(ctx);
    if (hipMalloc(&s.p, want) != hipSuccess) {
      (void)hipGetLastError();
      s.p = nullptr;
      return nullptr;
    }
  }
  s.bytes = want;
  return s.p;
}

constexpr size_t kPoolMaxBlocks = 24;
constexpr size_t kPoolMaxBytes = size_t{64} << 30;

void *graph_mem_alloc(gbk_ctx *ctx, size_t bytes) {
  if (bytes == 0) bytes = 4;
  size_t best = ctx->gpool.size();
  for (size_t i = 0; i < ctx->gpool.size(); i++) {
    const size_t b = ctx->gpool[i].bytes;
    if (b >= bytes && b <= bytes + bytes / 8 + (size_t{1} << 20) &&
        (best == ctx->gpool.size() || b < ctx->gpool[best].bytes))
      best = i;
  }
  if (best < ctx->gpool.size()) {
    void *p = ctx->gpool[best].p;
    ctx->gpool.erase(ctx->gpool.begin() + static_cast<std::ptrdiff_t>(best));
    return p;
  }
  void *p = nullptr;
  if (hipMalloc(&p, bytes) != hipSuccess) {
    // the cache may hold the memory: give it back and try once more
    (void)hipGetLastError();
    drop_graph_cache(ctx);
    if (hipMalloc(&p, bytes) != hipSuccess) {
      (void)hipGetLastError();
      return nullptr;
    }
  }
  ctx->gsize[p] = bytes;
  return p;
}

void graph_mem_free(gbk_ctx *ctx, void *p) {
  const auto it = ctx->gsize.find(p);
  if (it == ctx->gsize.end()) {
    hipFree(p);
    return;
  }
  ctx->gpool.push_back({p, it->second});
  size_t total = 0;
  for (const auto &b : ctx->gpool) total += b.bytes;
  while (!ctx->gpool.empty() && (ctx->gpool.size() > kPoolMaxBlocks || total > kPoolMaxBytes)) {
    total -= ctx->gpool.front().bytes;
    ctx->gsize.erase(ctx->gpool.front().p);
    hipFree(ctx->gpool.front().p);
    ctx->gpool.erase(ctx->gpool.begin());
  }
}

unsigned resident_grid(gbk_ctx *ctx, const void *kernel) {
  if (ctx->tune.batch_grid > 0) return static_cast<unsigned>(ctx->tune.batch_grid);
  auto it = ctx->resident.find(kernel);
  if (it != ctx->resident.end()) return it->second;
  if (ctx->cu_count == 0) {
    int cu = 0;
    if (hipDeviceGetAttribute(&cu, hipDeviceAttributeMultiprocessorCount, ctx->device) != hipSuccess || cu < 1)
      cu = 256;
    ctx->cu_count = cu;
  }
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, kBlock, 0) != hipSuccess || per_cu < 1)
    per_cu = 4;
  const unsigned grid = static_cast<unsigned>(per_cu) * static_cast<unsigned>(ctx->cu_count);
  ctx->resident[kernel] = grid;
  return grid;
}

// One fixed staging area per context, allocated on first use and never
// reallocated: a pointer handed out earlier in an entry point stays valid
// after later, larger requests (callers hold counters across helpers that
// also stage through it).  Requests above kPinnedBytes fail.
void *pinned(gbk_ctx *ctx, size_t bytes) {
  if (bytes > kPinnedBytes) return nullptr;
  if (!ctx->pinned) {
    if (hipHostMalloc(&ctx->pinned, kPinnedBytes, hipHostMallocDefault) != hipSuccess) {
      ctx->pinned = nullptr;
      return nullptr;
    }
    ctx->pinned_bytes = kPinnedBytes;
  }
  return ctx->pinned;
}

EventTimer::EventTimer(gbk_ctx *c) : ctx(c) { hipEventRecord(ctx->ev0, ctx->stream); }

double EventTimer::stop_ms() {
  hipEventRecord(ctx->ev1, ctx->stream);
  hipEventSynchronize(ctx->ev1);
  float ms = 0;
  hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1);
  kt_collect(ctx);
  return ms;
}

int kt_begin(gbk_ctx *ctx, int id) {
  if (ctx->ktime == 0 || ctx->ktime != id) return -1;
  if (ctx->kt_used * 2 + 2 > ctx->kt_ev.size()) {
    hipEvent_t a = nullptr, b = nullptr;
    if (hipEventCreate(&a) != hipSuccess) return -1;
    if (hipEventCreate(&b) != hipSuccess) {
      hipEventDestroy(a);
      return -1;
    }
    ctx->kt_ev.push_back(a);
    ctx->kt_ev.push_back(b);
  }
  const int pair = static_cast<int>(ctx->kt_used++);
  hipEventRecord(ctx->kt_ev[2 * pair], ctx->stream);
  return pair;
}

void kt_end(gbk_ctx *ctx, int pair) {
  if (pair >= 0) hipEventRecord(ctx->kt_ev[2 * pair + 1], ctx->stream);
}

void kt_collect(gbk_ctx *ctx) {
  for (size_t i = 0; i < ctx->kt_used; i++) {
    float ms = 0;
    if (hipEventElapsedTime(&ms, ctx->kt_ev[2 * i], ctx->kt_ev[2 * i + 1]) == hipSuccess) {
      ctx->kt_ms += ms;
      ctx->kt_launches += 1;
    }
  }
  ctx->kt_used = 0;
}

// ------------------------------------------------------------------- scan
constexpr int kScanItems = 8;
constexpr int kScanTile = kBlock * kScanItems;

template <typename In>
__global__ void scan_reduce_tiles(const In *in, int64_t n, int64_t *tile_sums) {
  __shared__ int64_t red[kWavesPerBlock];
  const int64_t base = static_cast<int64_t>(blockIdx.x) * kScanTile;
  int64_t s = 0;
  for (int i = 0; i < kScanItems; i++) {
    const int64_t idx = base + static_cast<int64_t>(i) * kBlock + threadIdx.x;
    if (idx < n) s += static_cast<int64_t>(in[idx]);
  }
  s = block_sum(s, red);
  if (threadIdx.x == 0) tile_sums[blockIdx.x] = s;
}

// One workgroup scans the tile sums in place (exclusive) and writes the total.
__global__ void scan_tile_sums(int64_t *tile_sums, int64_t tiles, int64_t *total_out) {
  __shared__ int64_t part[1024];
  const int t = threadIdx.x;
  const int64_t per = (tiles + blockDim.x - 1) / blockDim.x;
  const int64_t b = t * per, e = (b + per < tiles) ? b + per : tiles;
  int64_t s = 0;
  for (int64_t i = b; i < e; i++) s += tile_sums[i];
  part[t] = s;
  __syncthreads();
  if (t == 0) {
    int64_t run = 0;
    for (int i = 0; i < static_cast<int>(blockDim.x); i++) {
      const int64_t x = part[i];
      part[i] = run;
      run += x;
    }
    *total_out = run;
  }
  __syncthreads();
  int64_t run = part[t];
  for (int64_t i = b; i < e; i++) {
    const int64_t x = tile_sums[i];
    tile_sums[i] = run;
    run += x;
  }
}

template <typename In>
__global__ void scan_tiles(const In *in, int64_t n, const int64_t *tile_prefix, int64_t *out) {
  __shared__ int64_t wsum[kWavesPerBlock];
  const int64_t base = static_cast<int64_t>(blockIdx.x) * kScanTile;
  // each thread owns kScanItems consecutive elements
  const int64_t my = base + static_cast<int64_t>(threadIdx.x) * kScanItems;
  int64_t v[kScanItems];
  int64_t s = 0;
  for (int i = 0; i < kScanItems; i++) {
    v[i] = (my + i < n) ? static_cast<int64_t>(in[my + i]) : 0;
    s += v[i];
  }
  int64_t incl = wave_incl_scan(s);
  const int w = threadIdx.x / kWave;
  if (lane_id() == kWave - 1) wsum[w] = incl;
  __syncthreads();
  int64_t wprefix = 0;
  for (int i = 0; i < w; i++) wprefix += wsum[i];
  int64_t run = tile_prefix[blockIdx.x] + wprefix + incl - s;
  for (int i = 0; i < kScanItems; i++) {
    if (my + i < n) out[my + i] = run;
    run += v[i];
  }
}

template <typename In>
static int exclusive_scan_impl(gbk_ctx *ctx, const In *in, int64_t *out, int64_t n) {
  const int64_t tiles = (n + kScanTile - 1) / kScanTile;
  auto *tile_sums = static_cast<int64_t *>(scratch(ctx, kS15, sizeof(int64_t) * (tiles + 1)));
  if (!tile_sums) return set_error(ctx, GBK_ENOMEM, "scan scratch");
  if (tiles > 0) {
    scan_reduce_tiles<In><<<tiles, kBlock, 0, ctx->stream>>>(in, n, tile_sums);
  }
  scan_tile_sums<<<1, 1024, 0, ctx->stream>>>(tile_sums, tiles, out + n);
  if (tiles > 0) {
    scan_tiles<In><<<tiles, kBlock, 0, ctx->stream>>>(in, n, tile_sums, out);
  }
  GBK_LAUNCH_CHECK(ctx, "exclusive_scan");
  return GBK_OK;
}

int exclusive_scan_i64(gbk_ctx *ctx, const int64_t *in, int64_t *out, int64_t n) {
  return exclusive_scan_impl<int64_t>(ctx, in, out, n);
}

int exclusive_scan_i32(gbk_ctx *ctx, const int32_t *in, int64_t *out, int64_t n) {
  return exclusive_scan_impl<int32_t>(ctx, in, out, n);
}

__global__ void scatter_flagged(const int32_t *flags, const int64_t *pos, int64_t n, int32_t *out) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n && flags[i]) out[pos[i]] = static_cast<int32_t>(i);
}

int compact_flags(gbk_ctx *ctx, const int32_t *flags, int64_t n, int32_t *out, int64_t *count) {
  auto *pos = static_cast<int64_t *>(scratch(ctx, kS14, sizeof(int64_t) * (n + 1)));
  if (!pos) return set_error(ctx, GBK_ENOMEM, "compact scratch");
  int rc = exclusive_scan_impl<int32_t>(ctx, flags, pos, n);
  if (rc) return rc;
  if (n > 0) scatter_flagged<<<grid_for(n), kBlock, 0, ctx->stream>>>(flags, pos, n, out);
  GBK_LAUNCH_CHECK(ctx, "scatter_flagged");
  GBK_HIP(ctx, hipMemcpyAsync(count, pos + n, sizeof(int64_t), hipMemcpyDeviceToHost, ctx->stream));
  GBK_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return GBK_OK;
}

// ------------------------------------------------------------ heavy lists
__global__ void flag_heavy(const int64_t *off, int64_t n, int64_t threshold, int32_t *flags,
                           unsigned long long *partial_max) {
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  unsigned long long mx = 0;
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
    const int64_t d = off[i + 1] - off[i];
    flags[i] = d > threshold ? 1 : 0;
    mx = max(mx, static_cast<unsigned long long>(d));
  }
  mx = wave_max(mx);
  __shared__ unsigned long long red[kWavesPerBlock];
  if (lane_id() == 0) red[threadIdx.x / kWave] = mx;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int i = 0; i < kWavesPerBlock; i++) mx = max(mx, red[i]);
    partial_max[blockIdx.x] = mx;
  }
}

int build_heavy_list(gbk_ctx *ctx, const int64_t *off, int64_t n, int64_t threshold,
                     int32_t **list, int64_t *count, int64_t *max_degree,
                     std::vector<void *> &owned) {
  const unsigned grid = grid_for(n, kBlock, 2048);
  auto *flags = static_cast<int32_t *>(scratch(ctx, kS13, sizeof(int32_t) * (n + 1)));
  auto *mx = static_cast<unsigned long long *>(scratch(ctx, kS12, sizeof(unsigned long long) * grid));
  if (!flags || !mx) return set_error(ctx, GBK_ENOMEM, "heavy scratch");
  fill<unsigned long long>(ctx, mx, 0ull, grid);
  if (n > 0) flag_heavy<<<grid, kBlock, 0, ctx->stream>>>(off, n, threshold, flags, mx);
  GBK_LAUNCH_CHECK(ctx, "flag_heavy");
  int32_t *out = nullptr;
  // worst case n entries; most graphs have few heavy rows, so compact into
  // scratch first and copy the exact size.
  auto *tmp = static_cast<int32_t *>(scratch(ctx, kS11, sizeof(int32_t) * (n + 1)));
  if (!tmp) return set_error(ctx, GBK_ENOMEM, "heavy scratch");
  int64_t cnt = 0;
  int rc = compact_flags(ctx, flags, n, tmp, &cnt);
  if (rc) return rc;
  GBK_HIP(ctx, hipMalloc(&out, sizeof(int32_t) * (cnt + 1)));
  owned.push_back(out);
  if (cnt > 0)
    GBK_HIP(ctx, hipMemcpyAsync(out, tmp, sizeof(int32_t) * cnt, hipMemcpyDeviceToDevice, ctx->stream));
  std::vector<unsigned long long> hm(grid);
  GBK_HIP(ctx, hipMemcpyAsync(hm.data(), mx, sizeof(unsigned long long) * grid, hipMemcpyDeviceToHost,
                              ctx->stream));
  GBK_HIP(ctx, hipStreamSynchronize(ctx->stream));
  unsigned long long h = 0;
  for (auto x : hm) h = std::max(h, x);
  *list = out;
  *count = cnt;
  *max_degree = static_cast<int64_t>(h);
  return GBK_OK;
}

// ------------------------------------------------------------------ tiles
constexpr int64_t kTileRows = 64;
constexpr int64_t kTileEdges = 2048;

__global__ void light_degrees(const int64_t *off, int64_t n, int64_t threshold, int64_t *deg) {
  const int64_t v = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (v < n) {
    const int64_t d = off[v + 1] - off[v];
    deg[v] = d > threshold ? 0 : d;
  }
}

// A tile starts every kTileRows rows and wherever the light-edge prefix
// crosses a multiple of kTileEdges.
__global__ void tile_flags(const int64_t *lp, int64_t n, int32_t *flags) {
  const int64_t v = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (v < n) flags[v] = (v % kTileRows == 0) || (lp[v] / kTileEdges != lp[v - 1] / kTileEdges);
}

__global__ void append_sentinel(int32_t *tiles, int64_t at, int64_t n) {
  if (threadIdx.x == 0 && blockIdx.x == 0) tiles[at] = static_cast<int32_t>(n);
}

__global__ void heavy_chunk_counts(const int64_t *off, const int32_t *heavy, int64_t nheavy, int64_t *cnt) {
  const int64_t h = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (h < nheavy) {
    const int64_t d = off[heavy[h] + 1] - off[heavy[h]];
    cnt[h] = (d + kHeavyChunkEdges - 1) / kHeavyChunkEdges;
  }
}

int build_tiles(gbk_ctx *ctx, const int64_t *off, int64_t n, int64_t threshold, const int32_t *heavy,
                int64_t nheavy, int32_t **tiles, int64_t *ntiles, int64_t **hchunk,
                std::vector<void *> &owned) {
  hipStream_t s = ctx->stream;
  auto *deg = static_cast<int64_t *>(scratch(ctx, kS12, sizeof(int64_t) * (n + 1)));
  auto *lp = static_cast<int64_t *>(scratch(ctx, kS11, sizeof(int64_t) * (n + 1)));
  auto *flags = static_cast<int32_t *>(scratch(ctx, kS13, sizeof(int32_t) * (n + 1)));
  auto *tmp = static_cast<int32_t *>(scratch(ctx, kS10, sizeof(int32_t) * (n + 2)));
  if (!deg || !lp || !flags || !tmp) return set_error(ctx, GBK_ENOMEM, "tile scratch");
  int64_t cnt = 0;
  if (n > 0) {
    light_degrees<<<grid_for(n), kBlock, 0, s>>>(off, n, threshold, deg);
    if (int rc = exclusive_scan_i64(ctx, deg, lp, n)) return rc;
    tile_flags<<<grid_for(n), kBlock, 0, s>>>(lp, n, flags);
    GBK_LAUNCH_CHECK(ctx, "tile_flags");
    if (int rc = compact_flags(ctx, flags, n, tmp, &cnt)) return rc;
  }
  int32_t *out = nullptr;
  GBK_HIP(ctx, hipMalloc(&out, sizeof(int32_t) * (cnt + 1)));
  owned.push_back(out);
  if (cnt > 0) GBK_HIP(ctx, hipMemcpyAsync(out, tmp, sizeof(int32_t) * cnt, hipMemcpyDeviceToDevice, s));
  append_sentinel<<<1, 1, 0, s>>>(out, cnt, n);
  int64_t *pre = nullptr;
  GBK_HIP(ctx, hipMalloc(&pre, sizeof(int64_t) * (nheavy + 1)));
  owned.push_back(pre);
  auto *hc = static_cast<int64_t *>(scratch(ctx, kS12, sizeof(int64_t) * (nheavy + 1)));
  if (!hc) return set_error(ctx, GBK_ENOMEM, "heavy chunk scratch");
  if (nheavy > 0) heavy_chunk_counts<<<grid_for(nheavy), kBlock, 0, s>>>(off, heavy, nheavy, hc);
  if (int rc = exclusive_scan_i64(ctx, hc, pre, nheavy)) return rc;
  GBK_LAUNCH_CHECK(ctx, "build_tiles");
  *tiles = out;
  *ntiles = cnt;
  *hchunk = pre;
  return GBK_OK;
}

}  // namespace gbk

using namespace gbk;

// ----------------------------------------------------------------- C ABI
extern "C" {

int gbk_ctx_create(int device, gbk_ctx **out) {
  if (!out) return GBK_EINVAL;
  *out = nullptr;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || device < 0 || device >= count) return GBK_EINVAL;
  if (hipSetDevice(device) != hipSuccess) return GBK_EHIP;
  auto *ctx = new gbk_ctx();
  ctx->device = device;
  if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&ctx->ev0) != hipSuccess || hipEventCreate(&ctx->ev1) != hipSuccess) {
    delete ctx;
    return GBK_EHIP;
  }
  *out = ctx;
  return GBK_OK;
}

int gbk_ctx_destroy(gbk_ctx *ctx) {
  if (!ctx) return GBK_EINVAL;
  hipSetDevice(ctx->device);
  hipStreamSynchronize(ctx->stream);
  for (auto &s : ctx->slot)
    if (s.p) hipFree(s.p);
  for (auto &b : ctx->gpool) hipFree(b.p);
  if (ctx->pinned) hipHostFree(ctx->pinned);
  if (ctx->kt_dev) hipFree(ctx->kt_dev);
  for (hipEvent_t e : ctx->kt_ev) hipEventDestroy(e);
  hipEventDestroy(ctx->ev0);
  hipEventDestroy(ctx->ev1);
  hipStreamDestroy(ctx->stream);
  delete ctx;
  return GBK_OK;
}

const char *gbk_last_error(const gbk_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

void *gbk_ctx_stream(gbk_ctx *ctx) { return ctx ? static_cast<void *>(ctx->stream) : nullptr; }

int gbk_ctx_sync(gbk_ctx *ctx) {
  if (!ctx) return GBK_EINVAL;
  GBK_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return GBK_OK;
}

int gbk_ctx_set_param(gbk_ctx *ctx, const char *name, int64_t value) {
  if (!ctx || !name) return GBK_EINVAL;
  const std::string k(name);
  auto &t = ctx->tune;
  if (k == "heavy_degree" && value >= 64) t.heavy_degree = value;
  else if (k == "bu_group" && (value == 1 || value == 2 || value == 4 || value == 8 ||
                               value == 16 || value == 32 || value == 64))
    t.bu_group = static_cast<int>(value);
  else if (k == "bfs_alpha" && value > 0) t.bfs_alpha = static_cast<int>(value);
  else if (k == "bfs_beta" && value > 0) t.bfs_beta = static_cast<int>(value);
  else if (k == "xcd_swizzle" && (value == 0 || value == 1)) t.xcd_swizzle = static_cast<int>(value);
  else if (k == "relax_grid" && value >= 64 && value <= kGridCapRuntime && value % 8 == 0)
    t.relax_grid = static_cast<int>(value);
  else if (k == "graphs" && (value == 0 || value == 1)) t.graphs = static_cast<int>(value);
  else if (k == "batch_grid" && value >= 0 && value <= 16384 && value % 8 == 0) t.batch_grid = static_cast<int>(value);
  else if (k == "pr_blocks" && value >= 0 && value <= 1024) t.pr_blocks = static_cast<int>(value);
  else if (k == "pr_inplace" && (value == 0 || value == 1)) t.pr_inplace = static_cast<int>(value);
  else if (k == "pr_unroll" && (value == 0 || value == 1 || value == 2 || value == 4 || value == 8)) t.pr_unroll = static_cast<int>(value);
  else if (k == "bc_pull" && value >= 0 && value <= 2) t.bc_pull = static_cast<int>(value);
  else if (k == "bc_pull_ratio" && value >= 1 && value <= 1024) t.bc_pull_ratio = static_cast<int>(value);
  else if (k == "sssp_wide" && (value == 0 || value == 1)) t.sssp_wide = static_cast<int>(value);
  else if (k == "sssp_pull" && value >= 0 && value <= 2) t.sssp_pull = static_cast<int>(value);
  else if (k == "sssp_heavy" && value >= 64) t.sssp_heavy = value;
  else if (k == "sssp_tail" && (value == 0 || value == 1)) t.sssp_tail = static_cast<int>(value);
  else if (k == "msbfs_group" && value >= 1 && value <= 64 && (value & (value - 1)) == 0)
    t.msbfs_group = static_cast<int>(value);
  else if (k == "pr_segment" && (value == -1 || value == 0 || (value >= 6 && value <= 26)))
    t.pr_segment = static_cast<int>(value);
  else if (k == "ktime" && value >= GBK_KT_NONE && value <= GBK_KT_LAST) {
    if (value != GBK_KT_NONE && !ctx->kt_dev) {
      const size_t bytes = sizeof(unsigned long long) * (2 + 2 * kKtPartials);
      hipSetDevice(ctx->device);
      if (hipMalloc(&ctx->kt_dev, bytes) != hipSuccess) {
        ctx->kt_dev = nullptr;
        return set_error(ctx, GBK_ENOMEM, "kernel timer counters");
      }
      if (hipMemset(ctx->kt_dev, 0, bytes) != hipSuccess) return set_error(ctx, GBK_EHIP, "kernel timer counters");
    }
    ctx->ktime = static_cast<int>(value);
  } else return set_error(ctx, GBK_EINVAL, "unknown parameter or bad value: " + k);
  return GBK_OK;
}

int gbk_ctx_get_param(gbk_ctx *ctx, const char *name, int64_t *value) {
  if (!ctx || !name || !value) return GBK_EINVAL;
  const std::string k(name);
  const auto &t = ctx->tune;
  if (k == "heavy_degree") *value = t.heavy_degree;
  else if (k == "bu_group") *value = t.bu_group;
  else if (k == "bfs_alpha") *value = t.bfs_alpha;
  else if (k == "bfs_beta") *value = t.bfs_beta;
  else if (k == "xcd_swizzle") *value = t.xcd_swizzle;
  else if (k == "relax_grid") *value = t.relax_grid;
  else if (k == "graphs") *value = t.graphs;
  else if (k == "batch_grid") *value = t.batch_grid;
  else if (k == "pr_blocks") *value = t.pr_blocks;
  else if (k == "pr_inplace") *value = t.pr_inplace;
  else if (k == "pr_unroll") *value = t.pr_unroll;
  else if (k == "bc_pull") *value = t.bc_pull;
  else if (k == "bc_pull_ratio") *value = t.bc_pull_ratio;
  else if (k == "sssp_wide") *value = t.sssp_wide;
  else if (k == "sssp_pull") *value = t.sssp_pull;
  else if (k == "sssp_heavy") *value = t.sssp_heavy;
  else if (k == "sssp_tail") *value = t.sssp_tail;
  else if (k == "msbfs_group") *value = t.msbfs_group;
  else if (k == "pr_segment") *value = t.pr_segment;
  else if (k == "ktime") *value = ctx->ktime;
  else return set_error(ctx, GBK_EINVAL, "unknown parameter: " + k);
  return GBK_OK;
}

int gbk_ktime_read(gbk_ctx *ctx, double *total_ms, int64_t *launches, int64_t *rows, int64_t *edges) {
  if (!ctx) return GBK_EINVAL;
  hipSetDevice(ctx->device);
  GBK_HIP(ctx, hipStreamSynchronize(ctx->stream));
  kt_collect(ctx);
  unsigned long long u[2] = {0, 0};
  if (ctx->kt_dev) {
    GBK_HIP(ctx, hipMemcpy(u, ctx->kt_dev, sizeof(u), hipMemcpyDeviceToHost));
    GBK_HIP(ctx, hipMemset(ctx->kt_dev, 0, sizeof(u)));
  }
  if (total_ms) *total_ms = ctx->kt_ms;
  if (launches) *launches = ctx->kt_launches;
  if (rows) *rows = static_cast<int64_t>(u[0]);
  if (edges) *edges = static_cast<int64_t>(u[1]);
  ctx->kt_ms = 0;
  ctx->kt_launches = 0;
  return GBK_OK;
}

int gbk_graph_info(const gbk_graph *g, int64_t *n, int64_t *m, int *directed, int *weighted) {
  if (!g) return GBK_EINVAL;
  if (n) *n = g->n;
  if (m) *m = g->m;
  if (directed) *directed = g->directed;
  if (weighted) *weighted = g->weighted;
  return GBK_OK;
}

int gbk_graph_free(gbk_ctx *ctx, gbk_graph *g) {
  if (!ctx || !g) return GBK_EINVAL;
  hipStreamSynchronize(ctx->stream);
  for (void *p : g->owned) graph_mem_free(ctx, p);
  if (ctx->sssp_hint_graph == g) ctx->sssp_hint_graph = nullptr;
  delete g;
  return GBK_OK;
}

}  // extern "C"


// Builds heavy lists for both CSRs of a freshly filled device graph.
int gbk_graph_finalize(gbk_ctx *ctx, gbk_graph *g) {
  g->heavy_deg = ctx->tune.heavy_degree;
  int rc = build_heavy_list(ctx, g->off, g->n, g->heavy_deg, &g->heavy_out, &g->n_heavy_out,
                            &g->max_out_degree, g->owned);
  if (rc) return rc;
  rc = build_tiles(ctx, g->off, g->n, g->heavy_deg, g->heavy_out, g->n_heavy_out, &g->tiles_out,
                   &g->ntiles_out, &g->hchunk_out, g->owned);
  if (rc) return rc;
  if (g->directed) {
    rc = build_heavy_list(ctx, g->in_off, g->n, g->heavy_deg, &g->heavy_in, &g->n_heavy_in,
                          &g->max_in_degree, g->owned);
    if (rc) return rc;
    rc = build_tiles(ctx, g->in_off, g->n, g->heavy_deg, g->heavy_in, g->n_heavy_in, &g->tiles_in,
                     &g->ntiles_in, &g->hchunk_in, g->owned);
    if (rc) return rc;
  } else {
    g->heavy_in = g->heavy_out;
    g->n_heavy_in = g->n_heavy_out;
    g->max_in_degree = g->max_out_degree;
    g->tiles_in = g->tiles_out;
    g->ntiles_in = g->ntiles_out;
    g->hchunk_in = g->hchunk_out;
  }
  return GBK_OK;
}

// Allocates the device arrays of a graph with the given shape.
int gbk_graph_alloc(gbk_ctx *ctx, int64_t n, int64_t m, int64_t m_in, int directed, int weighted,
                    gbk_graph **out) {
  auto *g = new gbk_graph();
  g->n = n;
  g->m = m;
  g->directed = directed;
  g->weighted = weighted;
  auto alloc = [&](void **p, size_t bytes) {
    *p = graph_mem_alloc(ctx, bytes);
    if (!*p) return false;
    g->owned.push_back(*p);
    return true;
  };
  bool ok = alloc(reinterpret_cast<void **>(&g->off), sizeof(int64_t) * (n + 1)) &&
            alloc(reinterpret_cast<void **>(&g->nbr), sizeof(int32_t) * m) &&
            alloc(reinterpret_cast<void **>(&g->org_ids), sizeof(int32_t) * n);
  if (ok && weighted) ok = alloc(reinterpret_cast<void **>(&g->wgt), sizeof(int32_t) * m);
  if (ok && directed) {
    ok = alloc(reinterpret_cast<void **>(&g->in_off), sizeof(int64_t) * (n + 1)) &&
         alloc(reinterpret_cast<void **>(&g->in_nbr), sizeof(int32_t) * m_in);
    if (ok && weighted) ok = alloc(reinterpret_cast<void **>(&g->in_wgt), sizeof(int32_t) * m_in);
  } else if (ok) {
    g->in_off = g->off;
    g->in_nbr = g->nbr;
    g->in_wgt = g->wgt;
  }
  if (!ok) {
    for (void *p : g->owned) graph_mem_free(ctx, p);
    delete g;
    return set_error(ctx, GBK_ENOMEM, "graph allocation");
  }
  *out = g;
  return GBK_OK;
}

__global__ void iota_i32(int32_t *a, int64_t n) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) a[i] = static_cast<int32_t>(i);
}

// A caller's CSR is checked on the host before anything reaches the device:
// offsets start at 0 and never decrease, and every neighbour id lies in
// [0, n).  Every kernel indexes per-vertex arrays by neighbour id, so one
// bad entry would otherwise be an out-of-bounds device access.
static const char *check_csr(int64_t n, const int64_t *off, const int32_t *nbr) {
  if (off[0] != 0) return "offsets must start at 0";
  for (int64_t i = 0; i < n; i++)
    if (off[i + 1] < off[i]) return "offsets decrease";
  const int64_t m = off[n];
  for (int64_t k = 0; k < m; k++)
    if (nbr[k] < 0 || nbr[k] >= n) return "neighbour id out of range";
  return nullptr;
}

extern "C" int gbk_graph_upload(gbk_ctx *ctx, int64_t n, int directed, const int64_t *off,
                                const int32_t *nbr, const int32_t *wgt, const int64_t *in_off,
                                const int32_t *in_nbr, const int32_t *in_wgt,
                                const int32_t *org_ids, gbk_graph **out) {
  if (!ctx || !out || !off || n < 0 || n > INT32_MAX) return set_error(ctx, GBK_EINVAL, "upload: bad arguments");
  *out = nullptr;
  if (off[0] != 0) return set_error(ctx, GBK_EINVAL, "upload: off[0] must be 0");
  const int64_t m = off[n];
  if (m < 0 || (m > 0 && !nbr)) return set_error(ctx, GBK_EINVAL, "upload: bad edge count");
  if (const char *bad = check_csr(n, off, nbr)) return set_error(ctx, GBK_EINVAL, std::string("upload: ") + bad);
  int64_t m_in = m;
  if (directed) {
    if (!in_off || (in_off[n] > 0 && !in_nbr) || in_off[0] != 0)
      return set_error(ctx, GBK_EINVAL, "upload: directed graph needs its in-CSR");
    if (const char *bad = check_csr(n, in_off, in_nbr))
      return set_error(ctx, GBK_EINVAL, std::string("upload: in-CSR ") + bad);
    m_in = in_off[n];
    if (wgt && !in_wgt) return set_error(ctx, GBK_EINVAL, "upload: in-CSR weights missing");
  }
  hipSetDevice(ctx->device);
  gbk_graph *g = nullptr;
  int rc = gbk_graph_alloc(ctx, n, m, m_in, directed, wgt != nullptr, &g);
  if (rc) return rc;
  auto cp = [&](void *d, const void *h, size_t b) {
    return b == 0 ? hipSuccess : hipMemcpyAsync(d, h, b, hipMemcpyHostToDevice, ctx->stream);
  };
  hipError_t e = cp(g->off, off, sizeof(int64_t) * (n + 1));
  if (e == hipSuccess) e = cp(g->nbr, nbr, sizeof(int32_t) * m);
  if (e == hipSuccess && wgt) e = cp(g->wgt, wgt, sizeof(int32_t) * m);
  if (e == hipSuccess && directed) {
    e = cp(g->in_off, in_off, sizeof(int64_t) * (n + 1));
    if (e == hipSuccess) e = cp(g->in_nbr, in_nbr, sizeof(int32_t) * m_in);
    if (e == hipSuccess && wgt) e = cp(g->in_wgt, in_wgt, sizeof(int32_t) * m_in);
  }
  if (e == hipSuccess) {
    if (org_ids) e = cp(g->org_ids, org_ids, sizeof(int32_t) * n);
    else if (n > 0) iota_i32<<<grid_for(n), kBlock, 0, ctx->stream>>>(g->org_ids, n);
  }
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) {
    gbk_graph_free(ctx, g);
    return check_hip(ctx, e, "upload copy");
  }
  rc = gbk_graph_finalize(ctx, g);
  if (rc) {
    gbk_graph_free(ctx, g);
    return rc;
  }
  *out = g;
  return GBK_OK;
}

extern "C" int gbk_graph_download(gbk_ctx *ctx, const gbk_graph *g, int64_t *off, int32_t *nbr,
                                  int32_t *wgt, int64_t *in_off, int32_t *in_nbr, int32_t *in_wgt,
                                  int32_t *org_ids) {
  if (!ctx || !g) return GBK_EINVAL;
  auto cp = [&](void *h, const void *d, size_t b) {
    return (!h || b == 0) ? hipSuccess : hipMemcpyAsync(h, d, b, hipMemcpyDeviceToHost, ctx->stream);
  };
  const int64_t m_in = g->directed ? -1 : g->m;
  GBK_HIP(ctx, cp(off, g->off, sizeof(int64_t) * (g->n + 1)));
  GBK_HIP(ctx, cp(nbr, g->nbr, sizeof(int32_t) * g->m));
  if (wgt && !g->weighted) return set_error(ctx, GBK_ESTATE, "download: graph is unweighted");
  if (g->weighted) GBK_HIP(ctx, cp(wgt, g->wgt, sizeof(int32_t) * g->m));
  GBK_HIP(ctx, cp(org_ids, g->org_ids, sizeof(int32_t) * g->n));
  if (g->directed) {
    int64_t mi = 0;
    GBK_HIP(ctx, hipMemcpyAsync(&mi, g->in_off + g->n, sizeof(int64_t), hipMemcpyDeviceToHost, ctx->stream));
    GBK_HIP(ctx, hipStreamSynchronize(ctx->stream));
    GBK_HIP(ctx, cp(in_off, g->in_off, sizeof(int64_t) * (g->n + 1)));
    GBK_HIP(ctx, cp(in_nbr, g->in_nbr, sizeof(int32_t) * mi));
    if (g->weighted) GBK_HIP(ctx, cp(in_wgt, g->in_wgt, sizeof(int32_t) * mi));
  } else {
    (void)m_in;
    GBK_HIP(ctx, cp(in_off, g->off, sizeof(int64_t) * (g->n + 1)));
    GBK_HIP(ctx, cp(in_nbr, g->nbr, sizeof(int32_t) * g->m));
    if (g->weighted) GBK_HIP(ctx, cp(in_wgt, g->wgt, sizeof(int32_t) * g->m));
  }
  GBK_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return GBK_OK;
}
