// gbk_learn.hip -- GPU side of the learned ordering selector: cheap graph
// statistics, and a small fp32 MLP (in -> 64 -> 64 -> 1, ReLU) with its own
// forward, backward and Adam kernels.
//
// The reference's selector (reorder/reorder_adaptive.h, the frozen rule of
// :302-333) computes sampled features on the CPU (ComputeTier0Sampled-
// GraphFeatures, reorder_types.h:6066-6195) and applies thresholds; here the
// features are computed on the device and a network trained on this build's
// own timings predicts each candidate ordering's end-to-end time.
//
// MLP layout: one wave per sample, lane j owning hidden unit j of both
// hidden layers, so a sample's forward pass is two dot products per lane and
// one wave reduction, and its backward pass reuses the activations held in
// LDS.  A block keeps the weights and its gradient accumulators in LDS and
// adds them to the global gradient once per launch.  Parameters are one flat
// fp32 array: W1[H][D], b1[H], W2[H][H], b2[H], w3[H], b3.
#include <cmath>
#include <random>
#include <vector>

#include "gbk_expand.h"

namespace gbk {

constexpr int kHidden = 64;
constexpr int kMaxIn = 48;  // keeps weights + gradients of a block within 64 KB of LDS
constexpr int kFeatures = 12;

// ---------------------------------------------------------------- features
struct EdgeLocality {
  double logspan = 0;
  unsigned long long near64 = 0, near4k = 0;
  int64_t n;
  const int32_t *nbr;
  __device__ void operator()(int64_t u, int64_t k, bool active) {
    if (!active) return;
    const int64_t d = llabs(static_cast<int64_t>(nbr[k]) - u);
    logspan += log2(1.0 + static_cast<double>(d));
    near64 += d < 64;
    near4k += d < 4096;
  }
};

// per-workgroup partials, 8 doubles per group:
// [sum deg, sum deg^2, max deg, sum log2 span, near64, near4k, hub edges, -]
__global__ void __launch_bounds__(kBlock)
feat_degrees(const int64_t *off, const int32_t *nbr, int64_t n, int64_t avg, double *partial) {
  EdgeLocality f;
  f.n = n;
  f.nbr = nbr;
  double s1 = 0, s2 = 0, mx = 0, hub = 0;
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t v = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; v < n; v += stride) {
    const double d = static_cast<double>(off[v + 1] - off[v]);
    s1 += d;
    s2 += d * d;
    mx = d > mx ? d : mx;
    if (d > avg) hub += d;
  }
  auto rowlen = [&](int64_t u, int64_t &s, int64_t &l) {
    s = off[u];
    l = off[u + 1] - s;
  };
  expand_all(n, 2048, rowlen, f);
  __shared__ double red[kWavesPerBlock];
  const double v[8] = {s1, s2, mx, f.logspan, static_cast<double>(f.near64), static_cast<double>(f.near4k),
                       hub, 0.0};
  for (int i = 0; i < 8; i++) {
    double x = v[i];
    if (i == 2) {
      x = wave_max(x);
      if (lane_id() == 0) red[threadIdx.x / kWave] = x;
      __syncthreads();
      if (threadIdx.x == 0) {
        double m = red[0];
        for (int w = 1; w < kWavesPerBlock; w++) m = red[w] > m ? red[w] : m;
        partial[blockIdx.x * 8 + i] = m;
      }
      __syncthreads();
    } else {
      x = block_sum(x, red);
      if (threadIdx.x == 0) partial[blockIdx.x * 8 + i] = x;
      __syncthreads();
    }
  }
}

// Local clustering coefficient of sampled vertices: one wave per sample,
// the first 'cap' neighbours of u (sorted), pairs (a, b) with a < b closed
// when b is in N(a) (binary search in a's sorted list).
__global__ void __launch_bounds__(kBlock)
feat_clustering(const int64_t *off, const int32_t *nbr, int64_t n, int samples, int cap, double *out_cc,
                int *out_cnt) {
  const int wave = blockIdx.x * kWavesPerBlock + threadIdx.x / kWave;
  if (wave >= samples) return;
  // deterministic spread of the sample over the id range
  const int64_t u = static_cast<int64_t>((static_cast<unsigned long long>(wave) * 0x9E3779B97F4A7C15ull) % n);
  const int64_t s = off[u];
  const int64_t deg = off[u + 1] - s;
  const int64_t d = deg < cap ? deg : cap;
  if (d < 2) {
    if (lane_id() == 0) {
      out_cc[wave] = 0;
      out_cnt[wave] = 0;
    }
    return;
  }
  const int64_t pairs = d * (d - 1) / 2;
  unsigned long long closed = 0;
  for (int64_t p = lane_id(); p < pairs; p += kWave) {
    // p -> (i, j), i < j, row-major over the upper triangle
    int64_t i = 0, rem = p;
    while (rem >= d - 1 - i) {
      rem -= d - 1 - i;
      i++;
    }
    const int64_t j = i + 1 + rem;
    const int32_t a = nbr[s + i], b = nbr[s + j];
    int64_t lo = off[a], hi = off[a + 1];
    while (lo < hi) {
      const int64_t mid = (lo + hi) >> 1;
      if (nbr[mid] < b) lo = mid + 1;
      else hi = mid;
    }
    closed += (lo < off[a + 1] && nbr[lo] == b) ? 1 : 0;
  }
  closed = wave_sum(closed);
  if (lane_id() == 0) {
    out_cc[wave] = static_cast<double>(closed) / static_cast<double>(pairs);
    out_cnt[wave] = 1;
  }
}

// --------------------------------------------------------------------- MLP
__host__ __device__ inline int64_t mlp_params(int d) {
  return static_cast<int64_t>(kHidden) * d + kHidden + kHidden * kHidden + kHidden + kHidden + 1;
}

struct MlpView {
  const float *W1, *b1, *W2, *b2, *w3, *b3;
};

__device__ inline MlpView mlp_view(const float *p, int d) {
  MlpView v;
  v.W1 = p;
  v.b1 = v.W1 + kHidden * d;
  v.W2 = v.b1 + kHidden;
  v.b2 = v.W2 + kHidden * kHidden;
  v.w3 = v.b2 + kHidden;
  v.b3 = v.w3 + kHidden;
  return v;
}

// Forward (and, with 'grad', backward) over rows [0, rows): wave w of the
// grid takes samples w, w + W, ...  Loss = mean over rows of (y_hat - y)^2.
template <bool kGrad>
__global__ void __launch_bounds__(kBlock)
mlp_pass(const float *params, int d, const float *X, const float *y, int64_t rows, float *pred,
         float *grad, double *loss) {
  extern __shared__ float lds[];
  const int64_t P = mlp_params(d);
  float *w = lds;                       // weights
  float *g = w + P;                     // gradient accumulators (kGrad)
  float *act = g + (kGrad ? P : 0);     // per wave: x[kMaxIn], a1[64], dz2[64]
  for (int64_t i = threadIdx.x; i < P; i += blockDim.x) {
    w[i] = params[i];
    if (kGrad) g[i] = 0.f;
  }
  __syncthreads();
  const MlpView m = mlp_view(w, d);
  const int lane = lane_id(), wv = threadIdx.x / kWave;
  float *xs = act + wv * (kMaxIn + 2 * kHidden);
  float *a1s = xs + kMaxIn;
  float *dz2s = a1s + kHidden;
  const int64_t waves = static_cast<int64_t>(gridDim.x) * kWavesPerBlock;
  double lsum = 0;
  for (int64_t r = static_cast<int64_t>(blockIdx.x) * kWavesPerBlock + wv; r < rows; r += waves) {
    if (lane < d) xs[lane] = X[r * d + lane];
    __builtin_amdgcn_wave_barrier();
    float z1 = m.b1[lane];
    for (int i = 0; i < d; i++) z1 += m.W1[lane * d + i] * xs[i];
    const float a1 = z1 > 0.f ? z1 : 0.f;
    a1s[lane] = a1;
    __builtin_amdgcn_wave_barrier();
    float z2 = m.b2[lane];
    for (int k = 0; k < kHidden; k++) z2 += m.W2[lane * kHidden + k] * a1s[k];
    const float a2 = z2 > 0.f ? z2 : 0.f;
    const float yh = wave_sum(m.w3[lane] * a2) + m.b3[0];
    if (!kGrad) {
      if (lane == 0) pred[r] = yh;
    } else {
      const float e = yh - y[r];
      lsum += static_cast<double>(e) * e;
      const float gy = 2.f * e / static_cast<float>(rows);
      const float dz2 = z2 > 0.f ? gy * m.w3[lane] : 0.f;
      atomicAdd(&g[P - 1 - kHidden + lane], gy * a2);              // w3
      if (lane == 0) atomicAdd(&g[P - 1], gy);                     // b3
      dz2s[lane] = dz2;
      __builtin_amdgcn_wave_barrier();
      float *gW2 = g + kHidden * d + kHidden;
      for (int k = 0; k < kHidden; k++) atomicAdd(&gW2[lane * kHidden + k], dz2 * a1s[k]);
      atomicAdd(&gW2[kHidden * kHidden + lane], dz2);               // b2
      float da1 = 0.f;                                              // lane = input unit k of W2
      for (int j = 0; j < kHidden; j++) da1 += m.W2[j * kHidden + lane] * dz2s[j];
      const float dz1 = z1 > 0.f ? da1 : 0.f;
      for (int i = 0; i < d; i++) atomicAdd(&g[lane * d + i], dz1 * xs[i]);
      atomicAdd(&g[kHidden * d + lane], dz1);                      // b1
    }
    __builtin_amdgcn_wave_barrier();
  }
  if (kGrad) {
    __syncthreads();
    for (int64_t i = threadIdx.x; i < P; i += blockDim.x)
      if (g[i] != 0.f) atomicAdd(&grad[i], g[i]);
    __shared__ double red[kWavesPerBlock];
    const double t = block_sum(lane == 0 ? lsum : 0.0, red);
    if (threadIdx.x == 0) atomicAdd(loss, t / static_cast<double>(rows));
  }
}

__global__ void adam_step(float *p, const float *g, float *m1, float *m2, int64_t P, float lr, float b1,
                          float b2, float eps, float c1, float c2) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= P) return;
  const float gi = g[i];
  const float a = b1 * m1[i] + (1.f - b1) * gi;
  const float b = b2 * m2[i] + (1.f - b2) * gi * gi;
  m1[i] = a;
  m2[i] = b;
  p[i] -= lr * (a / c1) / (sqrtf(b / c2) + eps);
}

// Gathers the rows of a minibatch (indices 'idx') into contiguous buffers.
__global__ void gather_rows(const float *X, const float *y, const int32_t *idx, int64_t rows, int d,
                            float *Xb, float *yb) {
  const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (t >= rows * d) return;
  const int64_t r = t / d, c = t % d;
  Xb[t] = X[static_cast<int64_t>(idx[r]) * d + c];
  if (c == 0) yb[r] = y[idx[r]];
}

}  // namespace gbk

using namespace gbk;

struct gbk_mlp {
  int d = 0;
  int64_t P = 0;
  float *params = nullptr, *grad = nullptr, *m1 = nullptr, *m2 = nullptr;
  double *loss = nullptr;
  int64_t step = 0;
};

namespace {

size_t mlp_lds(int d, bool grad) {
  return sizeof(float) * (mlp_params(d) * (grad ? 2 : 1) + kWavesPerBlock * (kMaxIn + 2 * kHidden));
}

int mlp_check(gbk_ctx *ctx, const gbk_mlp *m) {
  if (!ctx || !m) return set_error(ctx, GBK_EINVAL, "mlp: null argument");
  return GBK_OK;
}

template <typename T>
int to_device(gbk_ctx *ctx, int slot, const T *host, int64_t count, T **dev) {
  *dev = static_cast<T *>(scratch(ctx, slot, sizeof(T) * (count > 0 ? count : 1)));
  if (!*dev) return set_error(ctx, GBK_ENOMEM, "mlp: device buffer");
  if (count > 0)
    GBK_HIP(ctx, hipMemcpyAsync(*dev, host, sizeof(T) * count, hipMemcpyHostToDevice, ctx->stream));
  return GBK_OK;
}

unsigned mlp_grid(int64_t rows) {
  const int64_t g = (rows + kWavesPerBlock - 1) / kWavesPerBlock;
  return static_cast<unsigned>(g < 1 ? 1 : (g > 256 ? 256 : g));
}

// grad = d loss / d params over rows (device X, y); adds loss into m->loss.
int mlp_grad_device(gbk_ctx *ctx, gbk_mlp *m, const float *X, const float *y, int64_t rows) {
  hipStream_t s = ctx->stream;
  fill<float>(ctx, m->grad, 0.f, m->P);
  fill<double>(ctx, m->loss, 0.0, 1);
  mlp_pass<true><<<mlp_grid(rows), kBlock, mlp_lds(m->d, true), s>>>(m->params, m->d, X, y, rows, nullptr,
                                                                    m->grad, m->loss);
  GBK_LAUNCH_CHECK(ctx, "mlp backward");
  return GBK_OK;
}

}  // namespace

extern "C" {

int gbk_graph_features(gbk_ctx *ctx, const gbk_graph *g, float *features, int count) {
  if (!ctx || !g || !features || count < kFeatures) return set_error(ctx, GBK_EINVAL, "features: bad arguments");
  const int64_t n = g->n, m = g->m;
  if (n == 0) return set_error(ctx, GBK_EINVAL, "features: empty graph");
  hipStream_t s = ctx->stream;
  const unsigned grid = expand_grid(n);
  constexpr int kSamples = 1024, kCap = 64;
  auto *part = static_cast<double *>(scratch(ctx, kS0, sizeof(double) * (8 * grid + kSamples)));
  auto *cnt = static_cast<int *>(scratch(ctx, kS1, sizeof(int) * kSamples));
  if (!part || !cnt) return set_error(ctx, GBK_ENOMEM, "features scratch");
  const int64_t avg = m / n;
  feat_degrees<<<grid, kBlock, 0, s>>>(g->off, g->nbr, n, avg, part);
  feat_clustering<<<kSamples / kWavesPerBlock, kBlock, 0, s>>>(g->off, g->nbr, n, kSamples, kCap,
                                                               part + 8 * grid, cnt);
  GBK_LAUNCH_CHECK(ctx, "features");
  std::vector<double> hp(8 * grid + kSamples);
  std::vector<int> hc(kSamples);
  GBK_HIP(ctx, hipMemcpyAsync(hp.data(), part, sizeof(double) * hp.size(), hipMemcpyDeviceToHost, s));
  GBK_HIP(ctx, hipMemcpyAsync(hc.data(), cnt, sizeof(int) * kSamples, hipMemcpyDeviceToHost, s));
  GBK_HIP(ctx, hipStreamSynchronize(s));
  double t[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (unsigned b = 0; b < grid; b++)
    for (int i = 0; i < 8; i++) t[i] = i == 2 ? std::max(t[i], hp[b * 8 + i]) : t[i] + hp[b * 8 + i];
  double cc = 0;
  int cc_n = 0;
  for (int i = 0; i < kSamples; i++)
    if (hc[i]) {
      cc += hp[8 * grid + i];
      cc_n++;
    }
  const double mean = t[0] / n;
  const double var = std::max(0.0, t[1] / n - mean * mean);
  const double me = std::max<double>(1.0, static_cast<double>(m));
  const double ws = 8.0 * (n + 1) + 4.0 * m + 8.0 * n;  // offsets + neighbours + two per-vertex arrays
  const float f[kFeatures] = {
      static_cast<float>(std::log2(static_cast<double>(n))),
      static_cast<float>(std::log2(me)),
      static_cast<float>(mean),
      static_cast<float>(std::log2(1.0 + t[2])),
      static_cast<float>(mean > 0 ? std::sqrt(var) / mean : 0.0),
      static_cast<float>(t[6] / me),                       // share of edges owned by above-average rows
      static_cast<float>(cc_n ? cc / cc_n : 0.0),          // sampled local clustering coefficient
      static_cast<float>(t[3] / me / std::max(1.0, std::log2(static_cast<double>(n)))),  // normalised log span
      static_cast<float>(t[4] / me),                       // edges within 64 ids
      static_cast<float>(t[5] / me),                       // edges within 4096 ids
      static_cast<float>(std::log2(ws / (4.0 * 1024 * 1024))),    // working set vs one XCD's L2
      static_cast<float>(std::log2(ws / (256.0 * 1024 * 1024))),  // working set vs Infinity Cache
  };
  for (int i = 0; i < kFeatures; i++) features[i] = f[i];
  return kFeatures;
}

int gbk_mlp_create(gbk_ctx *ctx, int in_dim, uint64_t seed, gbk_mlp **out) {
  if (!ctx || !out || in_dim <= 0 || in_dim > kMaxIn) return set_error(ctx, GBK_EINVAL, "mlp: bad arguments");
  auto *m = new gbk_mlp();
  m->d = in_dim;
  m->P = mlp_params(in_dim);
  const size_t bytes = sizeof(float) * m->P;
  if (hipMalloc(&m->params, bytes) != hipSuccess || hipMalloc(&m->grad, bytes) != hipSuccess ||
      hipMalloc(&m->m1, bytes) != hipSuccess || hipMalloc(&m->m2, bytes) != hipSuccess ||
      hipMalloc(&m->loss, sizeof(double)) != hipSuccess) {
    hipFree(m->params);
    hipFree(m->grad);
    hipFree(m->m1);
    hipFree(m->m2);
    hipFree(m->loss);
    delete m;
    return set_error(ctx, GBK_ENOMEM, "mlp parameters");
  }
  // Glorot-uniform weights, zero biases
  std::vector<float> p(m->P, 0.f);
  std::mt19937_64 rng(seed);
  auto init = [&](int64_t at, int fan_in, int fan_out, int64_t count) {
    const double a = std::sqrt(6.0 / (fan_in + fan_out));
    std::uniform_real_distribution<double> u(-a, a);
    for (int64_t i = 0; i < count; i++) p[at + i] = static_cast<float>(u(rng));
  };
  init(0, in_dim, kHidden, static_cast<int64_t>(kHidden) * in_dim);
  init(static_cast<int64_t>(kHidden) * in_dim + kHidden, kHidden, kHidden, kHidden * kHidden);
  init(m->P - 1 - kHidden, kHidden, 1, kHidden);
  hipMemcpy(m->params, p.data(), bytes, hipMemcpyHostToDevice);
  hipMemset(m->m1, 0, bytes);
  hipMemset(m->m2, 0, bytes);
  *out = m;
  return GBK_OK;
}

int gbk_mlp_destroy(gbk_ctx *ctx, gbk_mlp *m) {
  if (!m) return GBK_EINVAL;
  if (ctx) hipStreamSynchronize(ctx->stream);
  hipFree(m->params);
  hipFree(m->grad);
  hipFree(m->m1);
  hipFree(m->m2);
  hipFree(m->loss);
  delete m;
  return GBK_OK;
}

int64_t gbk_mlp_num_params(const gbk_mlp *m) { return m ? m->P : 0; }

int gbk_mlp_get_params(gbk_ctx *ctx, const gbk_mlp *m, float *params) {
  if (int rc = mlp_check(ctx, m)) return rc;
  if (!params) return set_error(ctx, GBK_EINVAL, "mlp: null buffer");
  GBK_HIP(ctx, hipMemcpyAsync(params, m->params, sizeof(float) * m->P, hipMemcpyDeviceToHost, ctx->stream));
  GBK_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return GBK_OK;
}

int gbk_mlp_set_params(gbk_ctx *ctx, gbk_mlp *m, const float *params) {
  if (int rc = mlp_check(ctx, m)) return rc;
  if (!params) return set_error(ctx, GBK_EINVAL, "mlp: null buffer");
  GBK_HIP(ctx, hipMemcpyAsync(m->params, params, sizeof(float) * m->P, hipMemcpyHostToDevice, ctx->stream));
  GBK_HIP(ctx, hipStreamSynchronize(ctx->stream));
  m->step = 0;
  fill<float>(ctx, m->m1, 0.f, m->P);
  fill<float>(ctx, m->m2, 0.f, m->P);
  return GBK_OK;
}

int gbk_mlp_forward(gbk_ctx *ctx, const gbk_mlp *m, const float *X, int64_t rows, float *pred) {
  if (int rc = mlp_check(ctx, m)) return rc;
  if (!X || !pred || rows <= 0) return set_error(ctx, GBK_EINVAL, "mlp: bad rows");
  float *dX = nullptr, *dp = nullptr;
  if (int rc = to_device(ctx, kS2, X, rows * m->d, &dX)) return rc;
  dp = static_cast<float *>(scratch(ctx, kS3, sizeof(float) * rows));
  if (!dp) return set_error(ctx, GBK_ENOMEM, "mlp: predictions");
  mlp_pass<false><<<mlp_grid(rows), kBlock, mlp_lds(m->d, false), ctx->stream>>>(m->params, m->d, dX, nullptr,
                                                                                rows, dp, nullptr, nullptr);
  GBK_LAUNCH_CHECK(ctx, "mlp forward");
  GBK_HIP(ctx, hipMemcpyAsync(pred, dp, sizeof(float) * rows, hipMemcpyDeviceToHost, ctx->stream));
  GBK_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return GBK_OK;
}

int gbk_mlp_grad(gbk_ctx *ctx, gbk_mlp *m, const float *X, const float *y, int64_t rows, float *grad,
                 double *loss) {
  if (int rc = mlp_check(ctx, m)) return rc;
  if (!X || !y || rows <= 0) return set_error(ctx, GBK_EINVAL, "mlp: bad rows");
  float *dX = nullptr, *dy = nullptr;
  if (int rc = to_device(ctx, kS2, X, rows * m->d, &dX)) return rc;
  if (int rc = to_device(ctx, kS4, y, rows, &dy)) return rc;
  if (int rc = mlp_grad_device(ctx, m, dX, dy, rows)) return rc;
  if (grad) GBK_HIP(ctx, hipMemcpyAsync(grad, m->grad, sizeof(float) * m->P, hipMemcpyDeviceToHost, ctx->stream));
  double l = 0;
  GBK_HIP(ctx, hipMemcpyAsync(&l, m->loss, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  GBK_HIP(ctx, hipStreamSynchronize(ctx->stream));
  if (loss) *loss = l;
  return GBK_OK;
}

int gbk_mlp_train(gbk_ctx *ctx, gbk_mlp *m, const float *X, const float *y, int64_t rows, int epochs,
                  int batch, float lr, uint64_t seed, double *loss_per_epoch) {
  if (int rc = mlp_check(ctx, m)) return rc;
  if (!X || !y || rows <= 0 || epochs <= 0 || batch <= 0 || !(lr > 0.f))
    return set_error(ctx, GBK_EINVAL, "mlp: bad training arguments");
  hipStream_t s = ctx->stream;
  float *dX = nullptr, *dy = nullptr;
  if (int rc = to_device(ctx, kS2, X, rows * m->d, &dX)) return rc;
  if (int rc = to_device(ctx, kS4, y, rows, &dy)) return rc;
  const int64_t B = std::min<int64_t>(batch, rows);
  auto *Xb = static_cast<float *>(scratch(ctx, kS5, sizeof(float) * B * m->d));
  auto *yb = static_cast<float *>(scratch(ctx, kS6, sizeof(float) * B));
  auto *idx = static_cast<int32_t *>(scratch(ctx, kS7, sizeof(int32_t) * rows));
  if (!Xb || !yb || !idx) return set_error(ctx, GBK_ENOMEM, "mlp: batch buffers");
  std::vector<int32_t> order(rows);
  for (int64_t i = 0; i < rows; i++) order[i] = static_cast<int32_t>(i);
  std::mt19937_64 rng(seed);
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  for (int e = 0; e < epochs; e++) {
    std::shuffle(order.begin(), order.end(), rng);
    GBK_HIP(ctx, hipMemcpyAsync(idx, order.data(), sizeof(int32_t) * rows, hipMemcpyHostToDevice, s));
    double epoch_loss = 0;
    int64_t nb = 0;
    for (int64_t at = 0; at < rows; at += B) {
      const int64_t r = std::min<int64_t>(B, rows - at);
      gather_rows<<<grid_for(r * m->d), kBlock, 0, s>>>(dX, dy, idx + at, r, m->d, Xb, yb);
      if (int rc = mlp_grad_device(ctx, m, Xb, yb, r)) return rc;
      m->step++;
      const float c1 = 1.f - std::pow(b1, static_cast<float>(m->step));
      const float c2 = 1.f - std::pow(b2, static_cast<float>(m->step));
      adam_step<<<grid_for(m->P), kBlock, 0, s>>>(m->params, m->grad, m->m1, m->m2, m->P, lr, b1, b2, eps, c1,
                                                 c2);
      if (loss_per_epoch) {
        double l = 0;
        GBK_HIP(ctx, hipMemcpyAsync(&l, m->loss, sizeof(double), hipMemcpyDeviceToHost, s));
        GBK_HIP(ctx, hipStreamSynchronize(s));
        epoch_loss += l * static_cast<double>(r);
      }
      nb += r;
    }
    GBK_LAUNCH_CHECK(ctx, "mlp train");
    if (loss_per_epoch) loss_per_epoch[e] = epoch_loss / static_cast<double>(nb);
  }
  GBK_HIP(ctx, hipStreamSynchronize(s));
  return GBK_OK;
}

}  // extern "C"
