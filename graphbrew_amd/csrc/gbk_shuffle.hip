// gbk_shuffle.hip -- the random ordering (-o 1) on the device, bit-exact with
// the reference's GenerateRandomMapping (reorder/reorder_basic.h:104-135).
//
// The reference shuffles permutation = 0..n-1 with Fisher-Yates: for
// remaining = n .. 2, swap(permutation[remaining - 1], permutation[j]) with
// j = UniformBounded(state, remaining) from one SplitMix64 stream (state
// starts at 0; a draw is rejected while below (2^64 - bound) mod bound,
// :20-37).  SplitMix64 is counter-based -- draw k (from 1) is the mix of
// k * 0x9e3779b97f4a7c15 -- so step t (remaining = n - t) takes draw t + 1
// unless an earlier draw was rejected.  All draws are computed in parallel;
// a rejection (probability below 2^-33 per step) shifts every later draw, so
// if any occurs the draws of that call are made on the host instead.
//
// The swaps run as deterministic reservations (Shun, Gu, Blelloch, Fineman
// and Gibbons, "Sequential random permutation, list contraction and tree
// contraction are highly parallel", SODA 2015): each round, every pending
// step t writes t into the reservation words of both slots it swaps by an
// atomic minimum, then commits its swap only if it holds both; the others
// stay pending.  The earliest pending step always commits, and a step
// commits only once every earlier step touching its slots has, so the
// result is the sequential loop's.  The pending list shrinks geometrically;
// the number of rounds grows with log n (32 at n = 20000).
#include <limits>
#include <vector>

#include "gbk_common.h"

namespace gbk {

constexpr uint64_t kGolden = 0x9e3779b97f4a7c15ull;

__device__ __host__ inline uint64_t splitmix_mix(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

// H[t] = the swap partner of step t assuming no rejection; counts rejections.
__global__ void shuffle_draws(int64_t n, int32_t *H, unsigned long long *rejected) {
  const int64_t steps = n - 1;
  int rej = 0;
  for (int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; t < steps;
       t += static_cast<int64_t>(gridDim.x) * blockDim.x) {
    const uint64_t bound = static_cast<uint64_t>(n - t);
    const uint64_t v = splitmix_mix(static_cast<uint64_t>(t + 1) * kGolden);
    const uint64_t threshold = (0ull - bound) % bound;
    if (v < threshold) rej = 1;
    H[t] = static_cast<int32_t>(v % bound);
  }
  if (__ballot(rej) && lane_id() == 0) atomicAdd(rejected, 1ull);
}

__global__ void shuffle_init(int64_t n, int32_t *A, int32_t *R, int32_t *pend) {
  const int64_t v = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (v < n) {
    A[v] = static_cast<int32_t>(v);
    R[v] = std::numeric_limits<int32_t>::max();
    if (v < n - 1) pend[v] = static_cast<int32_t>(v);
  }
}

__global__ void shuffle_reserve(int64_t n, const int32_t *H, const int32_t *pend, int64_t count, int32_t *R) {
  const int64_t x = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (x >= count) return;
  const int32_t t = pend[x];
  atomicMin(&R[n - 1 - t], t);
  atomicMin(&R[H[t]], t);
}

// Steps holding both slots swap; the rest go to the next pending list.  A
// step clears the reservation words it owns (every reserved word has one
// owner, so all are cleared for the next round).
__global__ void shuffle_commit(int64_t n, const int32_t *H, const int32_t *pend, int64_t count, int32_t *R,
                               int32_t *A, int32_t *next, unsigned long long *next_count) {
  const int64_t x = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const bool active = x < count;
  int32_t t = 0;
  bool wait = false;
  if (active) {
    t = pend[x];
    const int64_t i = n - 1 - t, j = H[t];
    const int32_t ri = R[i], rj = R[j];
    if (ri == t && rj == t) {
      const int32_t a = A[i];
      A[i] = A[j];
      A[j] = a;
    } else {
      wait = true;
    }
    if (ri == t) R[i] = std::numeric_limits<int32_t>::max();
    if (rj == t && j != i) R[j] = std::numeric_limits<int32_t>::max();
  }
  const int64_t slot = wave_append(next_count, wait);
  if (wait) next[slot] = t;
}

}  // namespace gbk

namespace gbk {

// new_ids (device, n) = the reference's random mapping of n vertices.
int random_mapping(gbk_ctx *ctx, int64_t n, int32_t *new_ids) {
  hipStream_t s = ctx->stream;
  if (n <= 0) return GBK_OK;
  if (n > std::numeric_limits<int32_t>::max()) return set_error(ctx, GBK_EINVAL, "random mapping: n too large");
  auto *H = static_cast<int32_t *>(scratch(ctx, kS10, sizeof(int32_t) * (2 * n + 2)));
  auto *pend = static_cast<int32_t *>(scratch(ctx, kS11, sizeof(int32_t) * (2 * n + 2)));
  auto *cnt = static_cast<unsigned long long *>(scratch(ctx, kS9, 64));
  auto *hc = static_cast<unsigned long long *>(pinned(ctx, sizeof(unsigned long long) * 2));
  if (!H || !pend || !cnt || !hc) return set_error(ctx, GBK_ENOMEM, "random mapping scratch");
  int32_t *R = H + n + 1;
  int32_t *pend2 = pend + n + 1;
  const int64_t steps = n - 1;
  fill<unsigned long long>(ctx, cnt, 0ull, 2);
  shuffle_init<<<grid_for(n), kBlock, 0, s>>>(n, new_ids, R, pend);
  if (steps > 0) shuffle_draws<<<grid_for(steps, kBlock, 4096), kBlock, 0, s>>>(n, H, cnt);
  GBK_LAUNCH_CHECK(ctx, "shuffle draws");
  GBK_HIP(ctx, hipMemcpyAsync(hc, cnt, sizeof(unsigned long long), hipMemcpyDeviceToHost, s));
  GBK_HIP(ctx, hipStreamSynchronize(s));
  if (hc[0] != 0) {
    // a rejected draw shifts the stream: make the draws on the host
    std::vector<int32_t> h(steps);
    uint64_t state = 0;
    for (int64_t t = 0; t < steps; t++) {
      const uint64_t bound = static_cast<uint64_t>(n - t);
      const uint64_t threshold = (0ull - bound) % bound;
      uint64_t v;
      do {
        state += kGolden;
        v = splitmix_mix(state);
      } while (v < threshold);
      h[t] = static_cast<int32_t>(v % bound);
    }
    GBK_HIP(ctx, hipMemcpyAsync(H, h.data(), sizeof(int32_t) * steps, hipMemcpyHostToDevice, s));
    GBK_HIP(ctx, hipStreamSynchronize(s));
  }
  int64_t count = steps;
  int32_t *cur = pend, *nxt = pend2;
  while (count > 0) {
    fill<unsigned long long>(ctx, cnt + 1, 0ull, 1);
    shuffle_reserve<<<grid_for(count), kBlock, 0, s>>>(n, H, cur, count, R);
    shuffle_commit<<<grid_for(count), kBlock, 0, s>>>(n, H, cur, count, R, new_ids, nxt, cnt + 1);
    GBK_LAUNCH_CHECK(ctx, "shuffle round");
    GBK_HIP(ctx, hipMemcpyAsync(hc + 1, cnt + 1, sizeof(unsigned long long), hipMemcpyDeviceToHost, s));
    GBK_HIP(ctx, hipStreamSynchronize(s));
    const int64_t left = static_cast<int64_t>(hc[1]);
    if (left >= count) return set_error(ctx, GBK_EHIP, "random mapping: a round committed nothing");
    count = left;
    std::swap(cur, nxt);
  }
  return GBK_OK;
}

}  // namespace gbk
