// gbk_tier0.hip -- the adaptive selector's Tier-0 graph features on the device:
// ComputeTier0SampledGraphFeatures (reorder/reorder_types.h:6066-6195).
//
// The reference samples S = max(1024, min(sqrt(n), 8192)) strided vertices
// (vertex ((2i+1) n / 2S + offset) mod n, offset = seed * 0x9E3779B97F4A7C15
// mod n), takes their out-degrees, ranks them by (degree desc, id asc), and
// walks the sampled rows once.  Here one workgroup of 1024 lanes does the
// sampling and the per-sample arithmetic with the sample's (degree, id) keys
// in LDS (64 KB for 8192 keys, sorted by a bitonic network), and one wave
// per sampled vertex walks its row for the edge span and window counts.
//
// Equal to the reference's doubles bit for bit where the reference's sums
// are integer-valued (average degree, hub concentration, packing factor,
// window overlap, internal density) or sequential: the degree variance is
// summed by one lane in the reference's sample order with non-contracted
// double operations.  The normalised edge span is the exact integer sum of
// |dst - src| divided once by (n - 1) instead of a sum of per-edge
// quotients: equal to the reference's within a few ulps per edge.
#include <algorithm>
#include <cmath>

#include "gbk_common.h"

namespace gbk {

constexpr int kT0Max = 8192;
constexpr int kT0Threads = 1024;

// res: avg_degree, degree_cv, hub_concentration, (span), (window),
// packing_factor, internal_density.  nodes_out: the sampled vertices.
// cnt: span sum, sampled edges, window edges (cleared here).
__global__ void __launch_bounds__(kT0Threads)
tier0_sample(const int64_t *off, int64_t n, int S, uint64_t offset, int32_t *nodes_out, double *res,
             unsigned long long *cnt) {
  __shared__ unsigned long long key[kT0Max];
  __shared__ unsigned int line[1024];
  __shared__ long long red[kT0Threads / kWave];
  __shared__ long long total_deg, hub_deg;
  __shared__ int distinct;
  const int tid = threadIdx.x;
  int P = 1;
  while (P < S) P <<= 1;
  for (int i = tid; i < P; i += kT0Threads) {
    if (i < S) {
      const uint64_t base = ((2 * static_cast<uint64_t>(i) + 1) * static_cast<uint64_t>(n)) / (2 * static_cast<uint64_t>(S));
      const int64_t node = static_cast<int64_t>((base + offset) % static_cast<uint64_t>(n));
      const int64_t deg = off[node + 1] - off[node];
      key[i] = (static_cast<unsigned long long>(0x7fffffffLL - deg) << 32) | static_cast<unsigned long long>(node);
      nodes_out[i] = static_cast<int32_t>(node);
    } else {
      key[i] = ~0ull;
    }
  }
  if (tid < 3) cnt[tid] = 0;
  __syncthreads();
  auto deg_of = [](unsigned long long k) { return 0x7fffffffLL - static_cast<long long>(k >> 32); };
  // exact degree sum
  long long ds = 0;
  for (int i = tid; i < S; i += kT0Threads) ds += deg_of(key[i]);
  ds = wave_sum(ds);
  if (lane_id() == 0) red[tid / kWave] = ds;
  __syncthreads();
  if (tid == 0) {
    long long t = 0;
    for (int w = 0; w < kT0Threads / kWave; w++) t += red[w];
    total_deg = t;
    // the reference's variance: one pass in sample order, no contraction
    const double avg = __ddiv_rn(static_cast<double>(t), static_cast<double>(S));
    double vs = 0.0;
    for (int i = 0; i < S; i++) {
      const double d = __dsub_rn(static_cast<double>(deg_of(key[i])), avg);
      vs = __dadd_rn(vs, __dmul_rn(d, d));
    }
    const double var = S > 1 ? __ddiv_rn(vs, static_cast<double>(S - 1)) : 0.0;
    res[0] = avg;
    res[1] = avg > 0.0 ? __ddiv_rn(__dsqrt_rn(var), avg) : 0.0;
    res[6] = n > 1 ? __ddiv_rn(avg, static_cast<double>(n - 1)) : 0.0;
  }
  __syncthreads();
  // bitonic sort of the keys: (degree desc, id asc)
  for (int k = 2; k <= P; k <<= 1) {
    for (int j = k >> 1; j > 0; j >>= 1) {
      for (int i = tid; i < P; i += kT0Threads) {
        const int l = i ^ j;
        if (l > i) {
          const unsigned long long a = key[i], b = key[l];
          const bool up = (i & k) == 0;
          if ((a > b) == up) {
            key[i] = b;
            key[l] = a;
          }
        }
      }
      __syncthreads();
    }
  }
  // hubs: the first max(1, S/10) entries; their degree sum and the distinct
  // 64-byte lines of 8-byte properties they occupy
  const int hubs = S / 10 > 1 ? S / 10 : 1;
  long long hs = 0;
  for (int i = tid; i < hubs; i += kT0Threads) hs += deg_of(key[i]);
  hs = wave_sum(hs);
  __syncthreads();
  if (lane_id() == 0) red[tid / kWave] = hs;
  int HP = 1;
  while (HP < hubs) HP <<= 1;
  for (int i = tid; i < HP; i += kT0Threads)
    line[i] = i < hubs ? static_cast<unsigned int>((key[i] & 0xffffffffull) / 8) : 0xffffffffu;
  __syncthreads();
  if (tid == 0) {
    long long t = 0;
    for (int w = 0; w < kT0Threads / kWave; w++) t += red[w];
    hub_deg = t;
  }
  for (int k = 2; k <= HP; k <<= 1) {
    for (int j = k >> 1; j > 0; j >>= 1) {
      for (int i = tid; i < HP; i += kT0Threads) {
        const int l = i ^ j;
        if (l > i) {
          const unsigned int a = line[i], b = line[l];
          const bool up = (i & k) == 0;
          if ((a > b) == up) {
            line[i] = b;
            line[l] = a;
          }
        }
      }
      __syncthreads();
    }
  }
  if (tid == 0) distinct = 0;
  __syncthreads();
  int mine = 0;
  for (int i = tid; i < hubs; i += kT0Threads) mine += (i == 0 || line[i] != line[i - 1]) ? 1 : 0;
  mine = wave_sum(mine);
  if (lane_id() == 0 && mine) atomicAdd(&distinct, mine);
  __syncthreads();
  if (tid == 0) {
    res[2] = total_deg > 0 ? __ddiv_rn(static_cast<double>(hub_deg), static_cast<double>(total_deg)) : 0.0;
    res[5] = distinct ? __ddiv_rn(static_cast<double>(hubs), static_cast<double>(distinct)) : 0.0;
  }
}

// One wave per sampled vertex: sum of |dst - src|, edges, and edges whose
// 64-byte property line is within one line of the source's.
__global__ void __launch_bounds__(kBlock)
tier0_rows(const int64_t *off, const int32_t *nbr, const int32_t *nodes, int S, unsigned long long *cnt) {
  const int lane = lane_id();
  const int64_t waves = static_cast<int64_t>(gridDim.x) * kWavesPerBlock;
  unsigned long long span = 0, edges = 0, window = 0;
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * kWavesPerBlock + threadIdx.x / kWave; i < S; i += waves) {
    const int64_t u = nodes[i];
    const int64_t ul = u / 8;
    for (int64_t k = off[u] + lane; k < off[u + 1]; k += kWave) {
      const int64_t v = nbr[k];
      span += static_cast<unsigned long long>(v > u ? v - u : u - v);
      edges += 1;
      const int64_t dl = v / 8 - ul;
      window += (dl >= -1 && dl <= 1) ? 1 : 0;
    }
  }
  span = wave_sum(span);
  edges = wave_sum(edges);
  window = wave_sum(window);
  if (lane == 0 && edges) {
    atomicAdd(&cnt[0], span);
    atomicAdd(&cnt[1], edges);
    atomicAdd(&cnt[2], window);
  }
}

}  // namespace gbk

using namespace gbk;

extern "C" int gbk_tier0_features(gbk_ctx *ctx, const gbk_graph *g, int64_t sample_size, uint64_t seed,
                                  double *feat) {
  if (!ctx || !g || !feat || sample_size < 0) return set_error(ctx, GBK_EINVAL, "tier0: bad arguments");
  const int64_t n = g->n;
  for (int i = 0; i < 7; i++) feat[i] = 0.0;
  if (n <= 0) return GBK_OK;
  int64_t S = sample_size;
  if (S == 0) S = std::max<int64_t>(1024, std::min<int64_t>(static_cast<int64_t>(std::sqrt(static_cast<double>(n))), 8192));
  S = std::min(S, n);
  if (S > kT0Max) return set_error(ctx, GBK_EINVAL, "tier0: at most 8192 samples");
  hipStream_t s = ctx->stream;
  auto *nodes = static_cast<int32_t *>(scratch(ctx, kS0, sizeof(int32_t) * S));
  auto *res = static_cast<double *>(scratch(ctx, kS1, sizeof(double) * 8 + sizeof(unsigned long long) * 4));
  if (!nodes || !res) return set_error(ctx, GBK_ENOMEM, "tier0 scratch");
  auto *cnt = reinterpret_cast<unsigned long long *>(res + 8);
  const uint64_t offset = (seed * 0x9E3779B97F4A7C15ULL) % static_cast<uint64_t>(n);
  tier0_sample<<<1, kT0Threads, 0, s>>>(g->off, n, static_cast<int>(S), offset, nodes, res, cnt);
  tier0_rows<<<static_cast<unsigned>((S + kWavesPerBlock - 1) / kWavesPerBlock), kBlock, 0, s>>>(
      g->off, g->nbr, nodes, static_cast<int>(S), cnt);
  GBK_LAUNCH_CHECK(ctx, "tier0");
  double h[8];
  unsigned long long c[3];
  GBK_HIP(ctx, hipMemcpyAsync(h, res, sizeof(h), hipMemcpyDeviceToHost, s));
  GBK_HIP(ctx, hipMemcpyAsync(c, cnt, sizeof(c), hipMemcpyDeviceToHost, s));
  GBK_HIP(ctx, hipStreamSynchronize(s));
  feat[0] = h[0];
  feat[1] = h[1];
  feat[2] = h[2];
  const double denom = static_cast<double>(std::max<int64_t>(1, n - 1));
  feat[3] = c[1] ? static_cast<double>(c[0]) / denom / static_cast<double>(c[1]) : 0.0;
  feat[4] = c[1] ? static_cast<double>(c[2]) / static_cast<double>(c[1]) : 0.0;
  feat[5] = h[5];
  feat[6] = h[6];
  return GBK_OK;
}
