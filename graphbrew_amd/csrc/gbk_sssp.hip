// gbk_sssp.hip -- delta-stepping SSSP (bench/src/sssp.cc:129-283).
//
// Buckets of width delta as in the reference: a vertex is relaxed from the
// current bucket b only while dist[v] >= delta*b (stale copies are skipped,
// sssp.cc:191-192), relaxations take the minimum atomically, and vertices
// that land in bucket b are processed again in b before the next bucket
// opens (the reference's bucket fusion does the same for small bins).
//
// GPU organisation:
//   word[v]   one 64-bit word per vertex: distance in the high half, in the
//             low half the relaxation round that last queued v (0 = none).
//             A relaxation is a single 64-bit atomic minimum of
//             (candidate << 32 | tag): it lowers the distance and records
//             the round together, and the value it returns says both
//             whether the distance improved and whether v is already queued
//             for the next round -- one atomic instead of a minimum plus a
//             separate test-and-set of a queued flag.  The tag is the round
//             number only when the candidate falls in the current bucket,
//             so a later improvement into the bucket still queues v.
//   near[2]   ping-pong queues of the current bucket, staged per wave in LDS.
//   later buckets are not queued at all: when the current bucket drains, one
//             dense pass over the words finds the smallest bucket still
//             holding an unsettled vertex and a second pass queues that
//             bucket.  Two streaming passes per bucket cost far less than a
//             pending list would in random atomics (one per improvement).
// Light and heavy edges (Meyer and Sanders, "Delta-stepping: a
// parallelizable shortest path algorithm", J. Algorithms 2003): the rounds
// of a bucket relax only edges lighter than delta -- the only ones whose
// candidates can land in the bucket itself -- and once the bucket is
// settled one pass relaxes the heavy edges of each of its vertices, from
// final distances.  A vertex re-queued by an improvement therefore re-reads
// the distances behind its light edges only (on Kronecker scale 22 with
// weights U[1,255] and delta 64 the reference's loop scans each edge about
// four times per source).  The settled vertices are listed as the rounds
// process them, once per bucket (a per-vertex stamp of the last bucket).
// The distances are the reference's: both are exact shortest paths.
// Each lane relaxes four edges per trip (independent loads and atomics in
// flight together): the path is bound by the latency of dependent random
// accesses, not by bandwidth.
// Kernels of one step: relax and relax-heavy (a light round in phase 0, the
// heavy pass pushed in phase 3 or pulled in phase 4), then min-bucket and
// split (each acting only in its phase; split also opens the tail, phase 5);
// the last workgroup of each kernel advances the device state block.  After
// a pulled heavy pass the vertices left have few edges, so the buckets end:
// those vertices are queued once and relaxed Bellman-Ford style (every edge
// of each vertex improved in the round before) until a round improves
// nothing.  The host queues steps without waiting -- first as many as the
// previous call on the graph and delta took, then 4, 8, ... 64 -- and checks
// the 'done' flag after each batch.  Distances are integers, so the result is
// exact whatever the order of relaxations.
#include <algorithm>
#include <climits>
#include <cstddef>
#include <vector>

#include "gbk_expand.h"

namespace gbk {

constexpr int32_t kDistInf = INT32_MAX / 2;  // sssp.cc:67
constexpr int kRelaxBatch = 4;               // edges per lane per trip (2 and 8: the same or slower)
// Workgroups of the relax-heavy launch: its chunks of hub rows are most of a
// first bucket's light rounds (Kronecker scale 22, delta 40: 256 -> 1.81 ms
// per source, 512 -> 1.62, 1024 -> 1.54, 1536 -> 1.52, 2048 -> 1.78).
constexpr unsigned kSsspHeavyGrid = 1024;

__device__ __forceinline__ int64_t word_dist(unsigned long long w) { return static_cast<int64_t>(w >> 32); }
__device__ __forceinline__ unsigned long long make_word(int64_t d, uint32_t tag) {
  return (static_cast<unsigned long long>(d) << 32) | tag;
}

struct SsspState {
  unsigned long long near_cnt[2];
  unsigned long long settled_cnt;  // vertices of the bucket listed for its heavy pass
  unsigned long long next_bucket;
  long long bucket;
  long long relax_rounds;
  long long buckets;
  long long steps;  // steps that found the search still open
  // 1: the buckets are over and the rest runs Bellman-Ford style (rounds
  // over every edge of the vertices improved in the round before)
  int tail;
  int cur, phase, done, overflow;
};

// Device scratch beside the state that every workgroup of a launch updates:
// its counters are spread (kSpread slots 256 bytes apart, slot = workgroup id
// mod kSpread) and combined by one workgroup, and per-bucket edge counts are
// per-workgroup partials.
constexpr int kSlotWords = 2 * kSpreadStride;  // 32-bit words per 256-byte slot
struct SsspAux {
  // arrivals at the end of the relax-heavy, min and split kernels: a counter
  // per slot group, then one for the groups; the last arrival closes the phase
  unsigned int arrive[3][kSpread * kSlotWords];
  unsigned int arrive_top[3][kSlotWords];
  // smallest distance above the current bucket found by the pass that looks
  // for the next bucket (ULLONG_MAX: none), slot s at next_min[s * 32]
  unsigned long long next_min[kSpread * kSlotWords / 2];
  // per relax workgroup: heavy edges of the vertices the current bucket
  // listed (a push reads these) and in-edges of every vertex listed so far
  // (a pull reads the in-edges of the others)
  unsigned long long push_part[kGridCapRuntime];
  unsigned long long in_part[kGridCapRuntime];
};

__global__ void sssp_init(int64_t n, int32_t source, unsigned long long *word, int32_t *stamp, int32_t *near0,
                          SsspState *st, SsspAux *aux) {
  const int64_t v = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (v < n) {
    word[v] = make_word(v == source ? 0 : kDistInf, 0);
    stamp[v] = -1;
  }
  {
    // counters and partials 0, the spread minimum all ones (ULLONG_MAX)
    unsigned int *w = reinterpret_cast<unsigned int *>(aux);
    const int64_t words = sizeof(SsspAux) / sizeof(unsigned int);
    const int64_t nm0 = offsetof(SsspAux, next_min) / sizeof(unsigned int);
    const int64_t nm1 = nm0 + sizeof(aux->next_min) / sizeof(unsigned int);
    const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
    for (int64_t i = v; i < words; i += stride) w[i] = i >= nm0 && i < nm1 ? 0xffffffffu : 0u;
  }
  if (v == 0) {
    near0[0] = source;
    st->near_cnt[0] = 1;
    st->near_cnt[1] = 0;
    st->settled_cnt = 0;
    st->next_bucket = ULLONG_MAX;
    st->bucket = 0;
    st->relax_rounds = 0;
    st->buckets = 0;
    st->steps = 0;
    st->tail = 0;
    st->cur = 0;
    st->phase = 0;
    st->done = 0;
    st->overflow = 0;
  }
}

struct RelaxEdge {
  const int2 *nw;  // light/heavy layout: (neighbour, weight) pairs
  unsigned long long *word;
  WaveQueue nearq;
  int64_t upper;  // end of the current bucket: candidates below it are queued
  uint32_t tag;   // this round's tag, never 0
  bool heavy;     // the heavy pass: nothing queued

  __device__ __forceinline__ int64_t row_value(int64_t u) const { return word_dist(word[u]); }

  // du: the row's distance when it was picked (a later, lower one re-queues
  // the row, so relaxing from the value read then is safe)
  template <int U>
  __device__ __forceinline__ void batch(const int64_t (&u)[U], const int64_t (&k)[U], const int64_t (&du)[U],
                                        const bool (&act)[U]) {
    int32_t v[U];
    int64_t cand[U];
    unsigned long long wv[U];
#pragma unroll
    for (int j = 0; j < U; j++) {
      const int2 e = act[j] ? nw[k[j]] : make_int2(0, 0);
      v[j] = e.x;
      cand[j] = du[j] + e.y;
    }
#pragma unroll
    for (int j = 0; j < U; j++) wv[j] = act[j] ? word[v[j]] : 0ull;
    bool push[U];
#pragma unroll
    for (int j = 0; j < U; j++) {
      push[j] = false;
      // test first: most relaxations do not improve, and a plain load can
      // only overstate dist[v] (values only decrease), never hide a win
      if (act[j] && cand[j] < word_dist(wv[j])) {
        const bool here = !heavy && cand[j] < upper;
        const unsigned long long old = atomicMin(&word[v[j]], make_word(cand[j], here ? tag : 0u));
        push[j] = here && cand[j] < word_dist(old) && static_cast<uint32_t>(old) != tag;
      }
    }
#pragma unroll
    for (int j = 0; j < U; j++) nearq.push(push[j], v[j]);
  }
};

// The in-CSR (the out-CSR itself for undirected graphs) and its row count.
struct SsspIn {
  const int64_t *off;
  const int32_t *nbr;
  const int32_t *wgt;
  int64_t n;
};

// The heavy pass as a pull: each unsettled vertex v takes the minimum of
// dist[u] + w over all its in-edges (u, v).  Every distance is the length of
// a path, so relaxing from any vertex is safe; the edges out of the settled
// bucket are among those read, so the pass does what the push does.
// Candidates are at least the bucket's upper end, so nothing is queued.
struct PullEdge {
  const int32_t *nbr;  // in-CSR
  const int32_t *wgt;
  unsigned long long *word;
  int64_t best;        // smallest finite candidate or unsettled distance seen

  __device__ __forceinline__ int64_t row_value(int64_t v) const { return word_dist(word[v]); }

  // dv: the row's distance when it was picked (only this pass lowers it, so
  // a candidate below it that is not below the current value is just a
  // spare atomic)
  template <int U>
  __device__ __forceinline__ void batch(const int64_t (&v)[U], const int64_t (&k)[U], const int64_t (&dv)[U],
                                        const bool (&act)[U]) {
    int32_t u[U];
    int64_t cand[U];
#pragma unroll
    for (int j = 0; j < U; j++) {
      u[j] = act[j] ? nbr[k[j]] : 0;
      cand[j] = act[j] ? wgt[k[j]] : 0;
    }
#pragma unroll
    for (int j = 0; j < U; j++) cand[j] += act[j] ? word_dist(word[u[j]]) : kDistInf;
#pragma unroll
    for (int j = 0; j < U; j++) {
      if (act[j] && cand[j] < dv[j]) atomicMin(&word[v[j]], make_word(cand[j], 0u));
      if (act[j] && cand[j] < best) best = cand[j];
    }
  }
};

template <typename T>
__device__ __forceinline__ T agent(T *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// True in every thread of the last workgroup of the launch to arrive (the
// same answer in all threads of a workgroup).  Each thread first waits for its
// own memory operations (the barrier alone does not wait for atomics that
// return nothing), so every workgroup's atomics are performed before its
// arrival counts; the closing workgroup reads such values with agent-scope
// atomic loads.  What the workgroups stored plainly is read only by later
// kernels, so no release fence (a write-back of the L2 on this GPU) is
// needed.  The counters are left at 0.
__device__ __forceinline__ bool last_to_arrive(unsigned int *groups, unsigned int *top) {
  __shared__ bool last;
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  if (threadIdx.x == 0) {
    const unsigned g = blockIdx.x % kSpread;
    const unsigned members = (gridDim.x - 1 - g) / kSpread + 1;
    const unsigned ngroups = gridDim.x < kSpread ? gridDim.x : kSpread;
    bool l = false;
    if (atomicAdd(&groups[g * kSlotWords], 1u) == members - 1) {
      atomicExch(&groups[g * kSlotWords], 0u);
      if (atomicAdd(top, 1u) == ngroups - 1) {
        atomicExch(top, 0u);
        l = true;
      }
    }
    last = l;
  }
  __syncthreads();
  return last;
}

// Workgroup minimum of 'x' into this workgroup's slot of a spread minimum
// (one atomic per workgroup, slots 256 bytes apart); every thread calls it.
__device__ __forceinline__ void spread_min(int64_t x, unsigned long long *slots) {
  __shared__ int64_t red[kWavesPerBlock];
  x = wave_min(x);
  __syncthreads();
  if (lane_id() == 0) red[threadIdx.x / kWave] = x;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int i = 1; i < kWavesPerBlock; i++) x = min(x, red[i]);
    if (x < kDistInf)
      atomicMin(&slots[(blockIdx.x % kSpread) * (kSlotWords / 2)], static_cast<unsigned long long>(x));
  }
}

// Minimum over the slots, which are reset to ULLONG_MAX; called by every
// thread of one workgroup, returns the same value in all of them.
__device__ __forceinline__ unsigned long long spread_take_min(unsigned long long *slots) {
  __shared__ unsigned long long res;
  if (threadIdx.x < kWave) {
    unsigned long long x = ULLONG_MAX;
    if (threadIdx.x < kSpread) {
      x = agent(&slots[threadIdx.x * (kSlotWords / 2)]);
      atomicExch(&slots[threadIdx.x * (kSlotWords / 2)], ULLONG_MAX);
    }
    x = wave_min(x);
    if (threadIdx.x == 0) res = x;
  }
  __syncthreads();
  return res;
}

__device__ __forceinline__ void relax_setup(RelaxEdge &f, int32_t *seg, const int2 *nw, unsigned long long *word,
                                            int32_t *near0, int32_t *near1, int64_t cap,
                                            int64_t delta, SsspState *st) {
  const volatile SsspState *vs = st;
  const int cur = vs->cur;
  f.heavy = vs->phase == 3;
  f.nw = nw;
  f.word = word;
  f.nearq.init(seg + (threadIdx.x / kWave) * kSeg, &st->near_cnt[1 - cur], cur ? near0 : near1, cap,
               &st->overflow);
  // the tail has no bucket end: every improvement is queued
  f.upper = vs->tail ? INT64_MAX : delta * (vs->bucket + 1);
  f.tag = static_cast<uint32_t>(vs->relax_rounds + 1);
}

// A light round (phase 0) over the near queue -- each row not already
// listed in this bucket goes on the settled list -- or the heavy pass
// (phase 3) over the settled list.
// Rows are read in the light/heavy layout (pairs and lc): a light round
// walks a row's light edges, the heavy pass its heavy ones.  Without heavy
// edges in the graph no vertex is listed.
__global__ void __launch_bounds__(kBlock)
sssp_relax(const int64_t *off, const int2 *nw, const int32_t *lc, SsspIn in,
           unsigned long long *word, int32_t *near0, int32_t *near1, int32_t *settled, int32_t *stamp,
           int64_t cap, int64_t delta, int64_t heavy_deg, int has_heavy, int pull_mode, SsspState *st,
           SsspAux *aux, HeavyQueue hq) {
  const volatile SsspState *vs = st;
  const int phase = vs->phase;
  if (vs->done || (phase != 0 && phase != 3 && phase != 4)) return;
  if (phase == 4) {
    // the pull: rows are the vertices above the bucket, in id order
    const int64_t upper = delta * (vs->bucket + 1);
    PullEdge pf{in.nbr, in.wgt, word, kDistInf};
    for (int64_t tile = blockIdx.x; tile * kBlock < in.n; tile += gridDim.x) {
      const int64_t v = tile * kBlock + threadIdx.x;
      int64_t start = 0, len = 0;
      const int64_t dv = v < in.n ? word_dist(word[v]) : 0;
      if (dv >= upper) {
        pf.best = min(pf.best, dv);
        start = in.off[v];
        len = in.off[v + 1] - start;
        if (len > heavy_deg) {
          heavy_push(hq, static_cast<int32_t>(v), start, len);
          len = 0;
        }
      }
      wave_walk_batch_v<kRelaxBatch>(v, start, len, dv, pf);
    }
    spread_min(pf.best, aux->next_min);
    return;
  }
  __shared__ int32_t seg[kWavesPerBlock * kSeg];
  __shared__ int32_t sseg[kWavesPerBlock * kSeg];
  const int cur = vs->cur;
  const bool heavy = vs->phase == 3;
  // the tails count past cap only when a push overflowed (reported by the host)
  const int64_t count = min(static_cast<int64_t>(heavy ? vs->settled_cnt : vs->near_cnt[cur]), cap);
  const int32_t *list = heavy ? settled : (cur ? near1 : near0);
  RelaxEdge f;
  relax_setup(f, seg, nw, word, near0, near1, cap, delta, st);
  WaveQueue sq;
  sq.init(sseg + (threadIdx.x / kWave) * kSeg, &st->settled_cnt, settled, cap, &st->overflow);
  const bool tail = vs->tail;
  const int64_t lower = tail ? 0 : f.upper - delta;
  const int32_t bucket = static_cast<int32_t>(vs->bucket);
  unsigned long long push_edges = 0, settled_in = 0;
  for (int64_t tile = blockIdx.x; tile * kBlock < count; tile += gridDim.x) {
    const int64_t i = tile * kBlock + threadIdx.x;
    int64_t row = 0, start = 0, len = 0, drow = 0;
    bool first = false;
    if (i < count) {
      row = list[i];
      const int64_t deg = off[row + 1] - off[row];
      const int64_t nl = lc[row];
      drow = word_dist(word[row]);
      if (heavy) {
        start = off[row] + nl;
        len = deg - nl;
      } else if (drow >= lower) {
        start = off[row];
        len = tail ? deg : nl;
        // a round's list holds a vertex at most once, so no other lane
        // touches its stamp in this launch
        if (!tail && has_heavy && deg > nl && stamp[row] != bucket) {
          stamp[row] = bucket;
          first = true;
        }
        if (first && pull_mode == 1) {
          push_edges += static_cast<unsigned long long>(deg - nl);
          settled_in += static_cast<unsigned long long>(in.off[row + 1] - in.off[row]);
        }
      }
      if (len > heavy_deg) {
        heavy_push(hq, static_cast<int32_t>(row), start, len);
        len = 0;
      }
    }
    sq.push(first, static_cast<int32_t>(row));
    wave_walk_batch_v<kRelaxBatch>(row, start, len, drow, f);
  }
  f.nearq.flush_block();
  sq.flush_block();
  if (pull_mode == 1) {
    __shared__ unsigned long long red[2][kWavesPerBlock];
    push_edges = wave_sum(push_edges);
    settled_in = wave_sum(settled_in);
    if (lane_id() == 0) {
      red[0][threadIdx.x / kWave] = push_edges;
      red[1][threadIdx.x / kWave] = settled_in;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      for (int i = 1; i < kWavesPerBlock; i++) {
        push_edges += red[0][i];
        settled_in += red[1][i];
      }
      if (push_edges) {
        aux->push_part[blockIdx.x] += push_edges;
        aux->in_part[blockIdx.x] += settled_in;
      }
    }
  }
}

// Heavy rows of the round's list, spread over all workgroups.
__device__ void relax_close(SsspState *st, SsspAux *aux, int phase, int cur, int tail,
                            unsigned long long *heavy_count, int has_heavy, int pull_mode, int tail_mode,
                            int relax_blocks, int64_t n, int64_t m_in, int64_t delta);

// The last workgroup to finish closes the round (relax_close).
__global__ void __launch_bounds__(kBlock)
sssp_relax_heavy(const int2 *nw, SsspIn in, unsigned long long *word, int32_t *near0,
                 int32_t *near1, int64_t cap, int64_t delta, int has_heavy, int pull_mode, int tail_mode,
                 int relax_blocks, int64_t m_in, SsspState *st, SsspAux *aux, HeavyQueue hq) {
  const volatile SsspState *vs = st;
  const int phase = vs->phase;
  const int cur = vs->cur;
  const int tail = vs->tail;
  if (vs->done || (phase != 0 && phase != 3 && phase != 4)) return;
  if (phase == 4) {
    PullEdge pf{in.nbr, in.wgt, word, kDistInf};
    expand_heavy_batch_v<kRelaxBatch>(hq, pf);
    spread_min(pf.best, aux->next_min);
  } else {
    __shared__ int32_t seg[kWavesPerBlock * kSeg];
    RelaxEdge f;
    relax_setup(f, seg, nw, word, near0, near1, cap, delta, st);
    expand_heavy_batch_v<kRelaxBatch>(hq, f);
    f.nearq.flush_block();
  }
  if (last_to_arrive(aux->arrive[0], aux->arrive_top[0]))
    relax_close(st, aux, phase, cur, tail, hq.count, has_heavy, pull_mode, tail_mode, relax_blocks, in.n, m_in,
                delta);
}

// Work of the coming relaxation round, for the kernel timer only: rows of
// the near queue that are not stale and their edges, as per-workgroup
// partials (rows in partial[0, grid), edges in partial[grid, 2 grid)).
__global__ void sssp_count_round(const int64_t *off, const int32_t *lc, SsspIn in, const unsigned long long *word,
                                 const int32_t *near0, const int32_t *near1, const int32_t *settled,
                                 int64_t delta, const SsspState *st, unsigned long long *partial) {
  const volatile SsspState *vs = st;
  unsigned long long rows = 0, edges = 0;
  if (!vs->done && vs->phase == 4) {
    const int64_t upper = delta * (vs->bucket + 1);
    const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
    for (int64_t v = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; v < in.n; v += stride) {
      const int64_t deg = in.off[v + 1] - in.off[v];
      if (deg > 0 && word_dist(word[v]) >= upper) {
        rows++;
        edges += static_cast<unsigned long long>(deg);
      }
    }
  } else if (!vs->done && (vs->phase == 0 || vs->phase == 3)) {
    const int cur = vs->cur;
    const bool heavy = vs->phase == 3;
    const int32_t *q = heavy ? settled : (cur ? near1 : near0);
    const int64_t count = static_cast<int64_t>(heavy ? vs->settled_cnt : vs->near_cnt[cur]);
    const bool tail = vs->tail;
    const int64_t lower = heavy || tail ? 0 : delta * vs->bucket;
    const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < count; i += stride) {
      const int32_t u = q[i];
      if (word_dist(word[u]) >= lower) {
        rows++;
        const int64_t deg = off[u + 1] - off[u];
        edges += static_cast<unsigned long long>(heavy ? deg - lc[u] : tail ? deg : lc[u]);
      }
    }
  }
  block_partial(rows, partial);
  block_partial(edges, partial + gridDim.x);
}

// Closes a relaxation round (one thread, after every workgroup of the
// round's kernels): swap the near queues; when the next one is empty the
// bucket is done and its heavy pass opens.  The heavy pass is pulled
// (phase 4) rather than pushed (phase 3) when the in-edges of the vertices
// not yet listed -- an over-count of the unsettled ones' -- plus a pass over
// all vertices (n / 8 edges' worth) are fewer than the bucket's heavy edges.
// After a push the search for the next bucket opens (phase 1); a pull has
// found the next bucket itself (phase 2, or done when nothing is left).
__device__ void relax_close(SsspState *st, SsspAux *aux, int phase, int cur, int tail,
                            unsigned long long *heavy_count, int has_heavy, int pull_mode, int tail_mode,
                            int relax_blocks, int64_t n, int64_t m_in, int64_t delta) {
  // Every thread of the closing workgroup runs this.  phase, cur and tail
  // are the launch's own (read by every workgroup at its start); values other
  // workgroups updated by atomics are read with agent loads, counters are
  // bumped by atomics, and the state is written by thread 0.
  if (phase == 4) {
    // after a pull the vertices left have few edges: with tail_mode they are
    // all queued (phase 5) and relaxed Bellman-Ford style to the end
    const unsigned long long nm = spread_take_min(aux->next_min);
    if (threadIdx.x == 0) {
      if (nm == ULLONG_MAX) {
        st->done = 1;
      } else {
        st->next_bucket = nm / static_cast<unsigned long long>(delta);
        st->phase = tail_mode ? 5 : 2;
      }
    }
  }
  if (threadIdx.x != 0 && phase != 0) return;
  if (phase == 3 || phase == 4) {
    __hip_atomic_fetch_add(&st->steps, 1ll, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    *heavy_count = 0;
    st->settled_cnt = 0;
    if (phase == 3) st->phase = 1;
    return;
  }
  // the end of a light round
  __shared__ int open;
  if (threadIdx.x == 0) {
    __hip_atomic_fetch_add(&st->steps, 1ll, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    *heavy_count = 0;
    st->near_cnt[cur] = 0;
    st->cur = 1 - cur;
    __hip_atomic_fetch_add(&st->relax_rounds, 1ll, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    open = agent(&st->near_cnt[1 - cur]) != 0;
  }
  __syncthreads();
  if (open) return;
  if (tail) {
    if (threadIdx.x == 0) st->done = 1;
    return;
  }
  // the bucket's light rounds are over: push or pull its heavy edges
  if (has_heavy && pull_mode == 1) {
    __shared__ unsigned long long red[2][kWavesPerBlock];
    unsigned long long push = 0, listed = 0;
    for (int i = threadIdx.x; i < relax_blocks; i += blockDim.x) {
      push += aux->push_part[i];
      listed += aux->in_part[i];
      aux->push_part[i] = 0;
    }
    push = wave_sum(push);
    listed = wave_sum(listed);
    if (lane_id() == 0) {
      red[0][threadIdx.x / kWave] = push;
      red[1][threadIdx.x / kWave] = listed;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      for (int i = 1; i < kWavesPerBlock; i++) {
        push += red[0][i];
        listed += red[1][i];
      }
      const long long l = static_cast<long long>(listed);
      const long long pull = (m_in > l ? m_in - l : 0) + n / 8;
      st->phase = static_cast<unsigned long long>(pull) < push ? 4 : 3;
    }
    return;
  }
  if (threadIdx.x == 0) st->phase = !has_heavy ? 1 : pull_mode == 2 ? 4 : 3;
}

// ------------------------------------------------ light/heavy row layout
// Rows of the out-CSR's tiles are walked by one wave per tile in 64-edge
// chunks; a lane's rank within its row's run of the chunk comes from
// ballots of the light flag, the run's earlier chunks from per-row carries
// in LDS.  Rows above the heavy degree take one workgroup each.  Positions
// are fixed by the CSR order, so the layout is the same every run.

__device__ __forceinline__ uint64_t run_before(int run_lo) {
  // lanes run_lo .. lane-1
  const uint64_t lt = (uint64_t(1) << lane_id()) - 1;
  return lt & ~((uint64_t(1) << run_lo) - 1);
}

// Pass 1: light edges per row.
__global__ void __launch_bounds__(kBlock)
part_count(RowTiles rt, const int32_t *wgt, int64_t delta, int32_t *lc) {
  __shared__ int32_t cnt_all[kWavesPerBlock][kWave];
  __shared__ int32_t red[kWavesPerBlock];
  const int lane = lane_id();
  const int wv = threadIdx.x / kWave;
  int32_t *cnt = cnt_all[wv];
  for (int64_t t = static_cast<int64_t>(blockIdx.x) * kWavesPerBlock + wv; t < rt.ntiles;
       t += static_cast<int64_t>(gridDim.x) * kWavesPerBlock) {
    const int64_t u = rt.tiles[t] + lane;
    const bool valid = u < rt.tiles[t + 1];
    int64_t start = 0, len = 0;
    if (valid) {
      start = rt.off[u];
      len = rt.off[u + 1] - start;
    }
    const bool tiled = valid && len <= rt.heavy_deg;
    if (!tiled) len = 0;
    cnt[lane] = 0;
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "workgroup");
    __builtin_amdgcn_wave_barrier();
    const int64_t incl = wave_incl_scan(len);
    const int64_t excl = incl - len;
    const int64_t total = __shfl(incl, kWave - 1);
    for (int64_t c = 0; c < total; c += kWave) {
      const int64_t e = c + lane;
      const int lo = walk_owner(excl, e);
      const bool act = e < total;
      const int64_t oexcl = __shfl(excl, lo);
      const int64_t ostart = __shfl(start, lo);
      const bool light = act && wgt[ostart + (e - oexcl)] < delta;
      const uint64_t lm = __ballot(light);
      const int nxt = __shfl_down(lo, 1);
      const bool tail = act && (lane == kWave - 1 || e + 1 >= total || nxt != lo);
      const int run_lo = static_cast<int>(oexcl > c ? oexcl - c : 0);
      if (tail) cnt[lo] += __popcll(lm & (run_before(run_lo) | (uint64_t(1) << lane)));
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "workgroup");
    __builtin_amdgcn_wave_barrier();
    if (tiled) lc[u] = cnt[lane];
  }
  for (int64_t h = blockIdx.x; h < rt.nheavy; h += gridDim.x) {
    const int64_t row = rt.heavy[h];
    int32_t c = 0;
    for (int64_t k = rt.off[row] + threadIdx.x; k < rt.off[row + 1]; k += blockDim.x) c += wgt[k] < delta;
    c = block_sum(c, red);
    if (threadIdx.x == 0) lc[row] = c;
  }
}

// Pass 2: each edge to its place.
__global__ void __launch_bounds__(kBlock)
part_place(RowTiles rt, const int32_t *nbr, const int32_t *wgt, int64_t delta, const int32_t *lc,
           int2 *pnw) {
  __shared__ int32_t carry_all[kWavesPerBlock][2 * kWave];
  __shared__ int32_t wl[kWavesPerBlock], wh[kWavesPerBlock];
  const int lane = lane_id();
  const int wv = threadIdx.x / kWave;
  int32_t *carry = carry_all[wv];  // [0, 64): light placed so far, [64, 128): heavy
  for (int64_t t = static_cast<int64_t>(blockIdx.x) * kWavesPerBlock + wv; t < rt.ntiles;
       t += static_cast<int64_t>(gridDim.x) * kWavesPerBlock) {
    const int64_t u = rt.tiles[t] + lane;
    const bool valid = u < rt.tiles[t + 1];
    int64_t start = 0, len = 0;
    if (valid) {
      start = rt.off[u];
      len = rt.off[u + 1] - start;
    }
    const bool tiled = valid && len <= rt.heavy_deg;
    if (!tiled) len = 0;
    const int32_t my_light = tiled ? lc[u] : 0;
    carry[lane] = 0;
    carry[kWave + lane] = 0;
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "workgroup");
    __builtin_amdgcn_wave_barrier();
    const int64_t incl = wave_incl_scan(len);
    const int64_t excl = incl - len;
    const int64_t total = __shfl(incl, kWave - 1);
    for (int64_t c = 0; c < total; c += kWave) {
      const int64_t e = c + lane;
      const int lo = walk_owner(excl, e);
      const bool act = e < total;
      const int64_t oexcl = __shfl(excl, lo);
      const int64_t ostart = __shfl(start, lo);
      const int32_t olight = __shfl(my_light, lo);
      const int nxt = __shfl_down(lo, 1);
      int32_t v = 0, w = 0;
      if (act) {
        v = nbr[ostart + (e - oexcl)];
        w = wgt[ostart + (e - oexcl)];
      }
      const bool light = act && w < delta;
      const uint64_t lm = __ballot(light);
      const uint64_t hm = __ballot(act && !light);
      const int run_lo = static_cast<int>(oexcl > c ? oexcl - c : 0);
      const uint64_t before = run_before(run_lo);
      const int32_t cl = carry[lo], ch = carry[kWave + lo];
      if (act) {
        const int64_t pos = light ? ostart + cl + __popcll(lm & before)
                                  : ostart + olight + ch + __popcll(hm & before);
        pnw[pos] = make_int2(v, w);
      }
      const bool tail = act && (lane == kWave - 1 || e + 1 >= total || nxt != lo);
      if (tail) {
        const uint64_t run = before | (uint64_t(1) << lane);
        carry[lo] = cl + __popcll(lm & run);
        carry[kWave + lo] = ch + __popcll(hm & run);
      }
      __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "workgroup");
      __builtin_amdgcn_wave_barrier();
    }
  }
  for (int64_t h = blockIdx.x; h < rt.nheavy; h += gridDim.x) {
    const int64_t row = rt.heavy[h];
    const int64_t a = rt.off[row], b = rt.off[row + 1];
    const int32_t L = lc[row];
    int64_t cl = 0, ch = 0;
    for (int64_t base = a; base < b; base += blockDim.x) {
      const int64_t k = base + threadIdx.x;
      const bool act = k < b;
      int32_t v = 0, w = 0;
      if (act) {
        v = nbr[k];
        w = wgt[k];
      }
      const bool light = act && w < delta;
      const uint64_t lm = __ballot(light), hm = __ballot(act && !light);
      __syncthreads();
      if (lane == 0) {
        wl[wv] = __popcll(lm);
        wh[wv] = __popcll(hm);
      }
      __syncthreads();
      int64_t bl = 0, bh = 0, tl = 0, th = 0;
      for (int i = 0; i < kWavesPerBlock; i++) {
        if (i < wv) {
          bl += wl[i];
          bh += wh[i];
        }
        tl += wl[i];
        th += wh[i];
      }
      const uint64_t lt = (uint64_t(1) << lane) - 1;
      if (act) {
        const int64_t pos = light ? a + cl + bl + __popcll(lm & lt) : a + L + ch + bh + __popcll(hm & lt);
        pnw[pos] = make_int2(v, w);
      }
      cl += tl;
      ch += th;
    }
  }
}

// Smallest and largest weight (per-workgroup partials, then one workgroup).
__global__ void weight_range(const int32_t *w, int64_t m, long long *partial) {
  long long lo = LLONG_MAX, hi = LLONG_MIN;
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < m; i += stride) {
    lo = min(lo, static_cast<long long>(w[i]));
    hi = max(hi, static_cast<long long>(w[i]));
  }
  for (int o = kWave / 2; o > 0; o >>= 1) {
    lo = min(lo, __shfl_xor(lo, o));
    hi = max(hi, __shfl_xor(hi, o));
  }
  __shared__ long long red[2][kWavesPerBlock];
  if (lane_id() == 0) {
    red[0][threadIdx.x / kWave] = lo;
    red[1][threadIdx.x / kWave] = hi;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int i = 1; i < kWavesPerBlock; i++) {
      lo = min(lo, red[0][i]);
      hi = max(hi, red[1][i]);
    }
    partial[2 * blockIdx.x] = lo;
    partial[2 * blockIdx.x + 1] = hi;
  }
}

int graph_weight_range(gbk_ctx *ctx, const gbk_graph *g, long long *partial) {
  if (g->wmin >= 0) return GBK_OK;
  constexpr unsigned kRangeGrid = 1024;
  hipStream_t s = ctx->stream;
  std::vector<long long> h(2 * kRangeGrid);
  int64_t lo = 0, hi = 0;
  if (g->m > 0 && g->wgt) {
    weight_range<<<kRangeGrid, kBlock, 0, s>>>(g->wgt, g->m, partial);
    GBK_LAUNCH_CHECK(ctx, "weight range");
    GBK_HIP(ctx, hipMemcpyAsync(h.data(), partial, sizeof(long long) * h.size(), hipMemcpyDeviceToHost, s));
    GBK_HIP(ctx, hipStreamSynchronize(s));
    lo = LLONG_MAX;
    hi = LLONG_MIN;
    for (unsigned i = 0; i < kRangeGrid; i++) {
      lo = std::min<int64_t>(lo, h[2 * i]);
      hi = std::max<int64_t>(hi, h[2 * i + 1]);
    }
  }
  auto *mg = const_cast<gbk_graph *>(g);
  mg->wmin = std::max<int64_t>(lo, 0);
  mg->wmax = std::max<int64_t>(hi, 0);
  return GBK_OK;
}

// Smallest distance above the current bucket (after a pushed heavy pass);
// the last workgroup opens that bucket's split, or ends the search.
__global__ void sssp_min_bucket(const unsigned long long *word, int64_t n, int64_t delta, SsspState *st,
                                SsspAux *aux) {
  const volatile SsspState *vs = st;
  if (vs->done || vs->phase != 1) return;
  const int64_t upper = delta * (vs->bucket + 1);
  int64_t best = kDistInf;
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t v = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; v < n; v += stride) {
    const int64_t d = word_dist(word[v]);
    if (d >= upper && d < best) best = d;
  }
  spread_min(best, aux->next_min);
  if (last_to_arrive(aux->arrive[1], aux->arrive_top[1])) {
    const unsigned long long nm = spread_take_min(aux->next_min);
    if (threadIdx.x == 0) {
      if (nm == ULLONG_MAX) {
        st->done = 1;
      } else {
        st->next_bucket = nm / static_cast<unsigned long long>(delta);
        st->phase = 2;
      }
    }
  }
}

// Queues every vertex of the chosen bucket (each vertex once); the last
// workgroup makes it the current bucket.  Opening the tail (phase 5), it
// queues every reached vertex above the last bucket instead.
__global__ void __launch_bounds__(kBlock)
sssp_split(const unsigned long long *word, int64_t n, int32_t *near0, int32_t *near1, int64_t cap,
           int64_t delta, SsspState *st, SsspAux *aux) {
  const volatile SsspState *vs = st;
  const int phase = vs->phase;
  if (vs->done || (phase != 2 && phase != 5)) return;
  __shared__ int32_t seg[kWavesPerBlock * kSeg];
  const int cur = vs->cur;
  const int64_t nb = static_cast<int64_t>(vs->next_bucket);
  // phase 5 opens the tail: every reached vertex above the last bucket
  const int64_t above = delta * (vs->bucket + 1);
  WaveQueue q;
  q.init(seg + (threadIdx.x / kWave) * kSeg, &st->near_cnt[cur], cur ? near1 : near0, cap,
         &st->overflow);
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t base = static_cast<int64_t>(blockIdx.x) * blockDim.x; base < n; base += stride) {
    const int64_t v = base + threadIdx.x;
    bool take = false;
    if (v < n) {
      const int64_t d = word_dist(word[v]);
      take = d < kDistInf && (phase == 5 ? d >= above : d / delta == nb);
    }
    q.push(take, static_cast<int32_t>(v));
  }
  q.flush_block();
  if (last_to_arrive(aux->arrive[2], aux->arrive_top[2]) && threadIdx.x == 0) {
    st->bucket = nb;
    __hip_atomic_fetch_add(&st->buckets, 1ll, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (phase == 5) st->tail = 1;
    st->phase = 0;
  }
}

// Distances out of the words, plus reached vertices and their out-edges.
__global__ void sssp_finish(const int64_t *off, const unsigned long long *word, int64_t n,
                            int32_t *dist, unsigned long long *p_reached, unsigned long long *p_edges) {
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  unsigned long long r = 0, e = 0;
  for (int64_t v = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; v < n; v += stride) {
    const int32_t d = static_cast<int32_t>(word_dist(word[v]));
    dist[v] = d;
    if (d != kDistInf) {
      r++;
      e += static_cast<unsigned long long>(off[v + 1] - off[v]);
    }
  }
  block_partial(r, p_reached);
  block_partial(e, p_edges);
}

}  // namespace gbk

using namespace gbk;

int gbk_sssp_device(gbk_ctx *ctx, const gbk_graph *g, int32_t source, int32_t delta, int32_t *dist,
                    gbk_sssp_stats *stats) {
  const int64_t n = g->n;
  hipStream_t s = ctx->stream;
  const int64_t cap = n + 64;
  auto *word = static_cast<unsigned long long *>(scratch(ctx, kS5, sizeof(unsigned long long) * cap));
  auto *q = static_cast<int32_t *>(scratch(ctx, kS7, sizeof(int32_t) * cap * 4));
  auto *st = static_cast<SsspState *>(scratch(ctx, kS8, sizeof(SsspState) + 64));
  auto *aux = static_cast<SsspAux *>(scratch(ctx, kS12, sizeof(SsspAux)));
  auto *fin = static_cast<unsigned long long *>(scratch(ctx, kS9, 64));
  auto *partial = static_cast<unsigned long long *>(scratch(ctx, kS10, sizeof(unsigned long long) * 2 * kGridCap));
  auto *hs = static_cast<SsspState *>(pinned(ctx, sizeof(SsspState)));
  HeavyQueue hq;
  // rows with more edges to walk than this go to the heavy kernel
  const int64_t H = ctx->tune.sssp_heavy;
  if (!word || !q || !st || !aux || !fin || !partial || !hs || !make_heavy_queue(ctx, g, kS11, &hq, H))
    return set_error(ctx, GBK_ENOMEM, "sssp scratch");
  hq.overflow = &st->overflow;
  int32_t *near0 = q, *near1 = q + cap, *settled = q + 2 * cap, *stamp = q + 3 * cap;
  // the graph's weight range (found once): which edge classes a delta leaves
  if (int rc = graph_weight_range(ctx, g, reinterpret_cast<long long *>(partial))) return rc;
  // heavy edges (weight >= delta) exist
  const int has_heavy = g->m > 0 && g->wmax >= delta;
  // The light/heavy layout for this delta, kept with the graph; its first
  // call with a delta pays for it (timed on its own and added to the call's
  // device time).  It exists before the step loop is captured: a captured
  // launch keeps the pointers it was recorded with.
  double layout_ms = 0;
  if (g->part_delta != delta) {
    EventTimer layout_timer(ctx);
    auto *mg = const_cast<gbk_graph *>(g);
    if (!mg->pnw) {
      void *a = graph_mem_alloc(ctx, sizeof(int2) * std::max<int64_t>(g->m, 1));
      void *c = a ? graph_mem_alloc(ctx, sizeof(int32_t) * std::max<int64_t>(n, 1)) : nullptr;
      if (!c) {
        if (a) graph_mem_free(ctx, a);
        return set_error(ctx, GBK_ENOMEM, "sssp light/heavy layout");
      }
      mg->owned.push_back(a);
      mg->owned.push_back(c);
      mg->pnw = static_cast<int2 *>(a);
      mg->plc = static_cast<int32_t *>(c);
    }
    const RowTiles rt = row_tiles(g, false);
    const unsigned pgrid = grid_for(std::max<int64_t>(rt.ntiles, 1), kWavesPerBlock);
    part_count<<<pgrid, kBlock, 0, s>>>(rt, g->wgt, delta, mg->plc);
    part_place<<<pgrid, kBlock, 0, s>>>(rt, g->nbr, g->wgt, delta, mg->plc, mg->pnw);
    GBK_LAUNCH_CHECK(ctx, "sssp light/heavy layout");
    layout_ms = layout_timer.stop_ms();
    mg->part_delta = delta;
  }
  const unsigned relax_grid = static_cast<unsigned>(ctx->tune.relax_grid);
  // the pull reads the in-CSR's weights (the out-CSR's on undirected graphs)
  const SsspIn in{g->in_off, g->in_nbr, g->in_wgt, n};
  const int pull_mode = in.off && in.nbr && in.wgt ? ctx->tune.sssp_pull : 0;
  const int tail_mode = ctx->tune.sssp_tail;
  const bool count = ctx->ktime == GBK_KT_SSSP_RELAX && ctx->kt_dev;
  const unsigned count_grid = static_cast<unsigned>(kKtPartials < 256 ? kKtPartials : 256);
  const unsigned scan_grid = expand_grid(n);
  // one bucket step: relaxation round, then the bucket bookkeeping (each
  // kernel returns at once unless the state's phase is its own)
  auto step = [&]() {
      if (count) {
        unsigned long long *part = ctx->kt_dev + 2;
        sssp_count_round<<<count_grid, kBlock, 0, s>>>(g->off, g->plc, in, word, near0, near1, settled, delta,
                                                       st, part);
        sum_partials<unsigned long long><<<1, 1024, 0, s>>>(part, static_cast<int>(count_grid), ctx->kt_dev, 1);
        sum_partials<unsigned long long><<<1, 1024, 0, s>>>(part + count_grid, static_cast<int>(count_grid),
                                                          ctx->kt_dev + 1, 1);
      }
      const int kt = kt_begin(ctx, GBK_KT_SSSP_RELAX);
      sssp_relax<<<relax_grid, kBlock, 0, s>>>(g->off, g->pnw, g->plc, in, word, near0, near1,
                                               settled, stamp, cap, delta, H, has_heavy, pull_mode, st, aux, hq);
      sssp_relax_heavy<<<kSsspHeavyGrid, kBlock, 0, s>>>(g->pnw, in, word, near0, near1, cap, delta,
                                                     has_heavy, pull_mode, tail_mode,
                                                     static_cast<int>(relax_grid), g->m, st, aux, hq);
      kt_end(ctx, kt);
      sssp_min_bucket<<<scan_grid, kBlock, 0, s>>>(word, n, delta, st, aux);
      sssp_split<<<scan_grid, kBlock, 0, s>>>(word, n, near0, near1, cap, delta, st, aux);
  };
  // steps are queued blindly in blocks of kStepsPerGraph (a captured graph)
  // and the host reads the state after 4, 8, ... 64 of them -- or first
  // after as many as the last call on this graph and delta took
  constexpr int kStepsPerGraph = 4;
  auto block = [&]() {
    for (int j = 0; j < kStepsPerGraph; j++) step();
  };
  LoopGraph loop(ctx);
  if (int rc = loop.capture(block)) return rc;
  EventTimer timer(ctx);
  sssp_init<<<grid_for(n), kBlock, 0, s>>>(n, source, word, stamp, near0, st, aux);
  const bool hinted = ctx->sssp_hint_graph == g && ctx->sssp_hint_delta == delta && ctx->sssp_hint_steps > 0;
  int64_t batch = kStepsPerGraph;
  if (hinted)
    batch = std::min<int64_t>((ctx->sssp_hint_steps + kStepsPerGraph - 1) / kStepsPerGraph * kStepsPerGraph, 256);
  int64_t launched = 0;
  for (;;) {
    for (int64_t j = 0; j < batch; j += kStepsPerGraph)
      if (int rc = loop.run(block)) return rc;
    launched += batch;
    GBK_LAUNCH_CHECK(ctx, "sssp step");
    GBK_HIP(ctx, hipMemcpyAsync(hs, st, sizeof(SsspState), hipMemcpyDeviceToHost, s));
    GBK_HIP(ctx, hipStreamSynchronize(s));
    kt_collect(ctx);
    if (hs->overflow) return set_error(ctx, GBK_EHIP, "sssp: queue overflow");
    if (hs->done) break;
    if (hinted && launched == batch) batch = kStepsPerGraph;
    else if (batch < 64) batch *= 2;
    // every step either settles a bucket or shrinks the work of the current
    // one; more steps than buckets times rounds would mean a broken state
    if (launched > 4 * (n + 16) + 1024) return set_error(ctx, GBK_EHIP, "sssp: no convergence");
  }
  sssp_finish<<<scan_grid, kBlock, 0, s>>>(g->off, word, n, dist, partial, partial + kGridCap);
  sum_partials<unsigned long long><<<1, 1024, 0, s>>>(partial, static_cast<int>(scan_grid), fin, 0);
  sum_partials<unsigned long long><<<1, 1024, 0, s>>>(partial + kGridCap, static_cast<int>(scan_grid), fin + 1, 0);
  GBK_LAUNCH_CHECK(ctx, "sssp_finish");
  const double ms = timer.stop_ms() + layout_ms;
  ctx->sssp_hint_graph = g;
  ctx->sssp_hint_delta = delta;
  ctx->sssp_hint_steps = hs->steps;
  unsigned long long hf[2] = {0, 0};
  GBK_HIP(ctx, hipMemcpyAsync(hf, fin, sizeof(hf), hipMemcpyDeviceToHost, s));
  GBK_HIP(ctx, hipStreamSynchronize(s));
  if (stats) {
    stats->device_ms = ms;
    stats->buckets = hs->buckets + 1;
    stats->relax_rounds = hs->relax_rounds;
    stats->reached = static_cast<int64_t>(hf[0]);
    stats->edges_reached = static_cast<int64_t>(hf[1]);
  }
  return GBK_OK;
}

extern "C" int gbk_sssp(gbk_ctx *ctx, const gbk_graph *g, int32_t source, int32_t delta,
                        int32_t *dist_out, gbk_sssp_stats *stats) {
  if (!ctx || !g) return set_error(ctx, GBK_EINVAL, "sssp: null argument");
  if (!g->weighted) return set_error(ctx, GBK_ESTATE, "sssp: graph has no weights");
  if (source < 0 || source >= g->n) return set_error(ctx, GBK_EINVAL, "sssp: source out of range");
  if (delta <= 0) return set_error(ctx, GBK_EINVAL, "sssp: delta must be positive");
  auto *dist = static_cast<int32_t *>(scratch(ctx, kS4, sizeof(int32_t) * (g->n + 1)));
  if (!dist) return set_error(ctx, GBK_ENOMEM, "sssp dist");
  int rc = gbk_sssp_device(ctx, g, source, delta, dist, stats);
  if (rc) return rc;
  if (dist_out) {
    GBK_HIP(ctx, hipMemcpyAsync(dist_out, dist, sizeof(int32_t) * g->n, hipMemcpyDeviceToHost, ctx->stream));
    GBK_HIP(ctx, hipStreamSynchronize(ctx->stream));
  }
  return GBK_OK;
}
