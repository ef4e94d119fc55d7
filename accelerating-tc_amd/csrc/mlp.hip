// Learned scheduling for the AUTO engine.
//
//  * per-vertex features of the oriented CSR (k_features);
//  * measured per-vertex cost of the block and wave variants (k_time_*);
//  * an 8-32-32-2 MLP with BF16 weights and FP32 accumulation, trained by
//    this file's forward / backward / Adam kernels (deterministic gradient
//    reduction: per-tile partial sums, then a fixed-order tile sum);
//  * the learned schedule: each vertex goes to the variant with the lower
//    predicted cost, both task lists ordered by descending predicted cost.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "engine.h"
#include "internal.h"

namespace tcx {
namespace {

using namespace dev;

constexpr int F = TC_MODEL_FEATURES;
constexpr int H = 32;
constexpr int O = TC_MODEL_OUTPUTS;
constexpr int OFF_W1 = 0;
constexpr int OFF_B1 = OFF_W1 + F * H;
constexpr int OFF_W2 = OFF_B1 + H;
constexpr int OFF_B2 = OFF_W2 + H * H;
constexpr int OFF_W3 = OFF_B2 + H;
constexpr int OFF_B3 = OFF_W3 + H * O;
constexpr int P = OFF_B3 + O;  // 1410
constexpr int TILE = 128;      // samples per backward workgroup

struct Norm {
  float xm[F], xs[F], ym[O], ys[O];
};

// fp32 -> bf16 with round-to-nearest-even (finite inputs)
__host__ __device__ inline uint16_t to_bf16(float f) {
  uint32_t u;
  memcpy(&u, &f, 4);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__host__ __device__ inline float from_bf16(uint16_t b) {
  uint32_t u = (uint32_t)b << 16;
  float f;
  memcpy(&f, &u, 4);
  return f;
}

__device__ inline void load_weights(const uint16_t* __restrict__ wbf, float* sw) {
  for (int i = threadIdx.x; i < P; i += blockDim.x) sw[i] = from_bf16(wbf[i]);
  __syncthreads();
}

// forward of one sample; h1/h2 keep the post-ReLU activations
__device__ inline void mlp_forward(const float* sw, const float* x, float* h1, float* h2,
                                   float* y) {
  for (int j = 0; j < H; ++j) {
    float a = sw[OFF_B1 + j];
    for (int i = 0; i < F; ++i) a += x[i] * sw[OFF_W1 + i * H + j];
    h1[j] = a > 0.f ? a : 0.f;
  }
  for (int j = 0; j < H; ++j) {
    float a = sw[OFF_B2 + j];
    for (int i = 0; i < H; ++i) a += h1[i] * sw[OFF_W2 + i * H + j];
    h2[j] = a > 0.f ? a : 0.f;
  }
  for (int j = 0; j < O; ++j) {
    float a = sw[OFF_B3 + j];
    for (int i = 0; i < H; ++i) a += h2[i] * sw[OFF_W3 + i * O + j];
    y[j] = a;
  }
}

__global__ __launch_bounds__(256) void k_mlp_forward(const float* __restrict__ X, uint32_t N,
                                                     const uint16_t* __restrict__ wbf, Norm nm,
                                                     float* __restrict__ Y) {
  __shared__ float sw[P];
  load_weights(wbf, sw);
  for (uint64_t s = gtid(); s < N; s += gstride()) {
    float x[F], h1[H], h2[H], y[O];
    for (int i = 0; i < F; ++i) x[i] = (X[s * F + i] - nm.xm[i]) / nm.xs[i];
    mlp_forward(sw, x, h1, h2, y);
    for (int j = 0; j < O; ++j) Y[s * O + j] = y[j] * nm.ys[j] + nm.ym[j];
  }
}

// One tile of TILE samples (optionally through an index list): forward,
// backward to every parameter, and the tile's partial gradient sums written
// to partial[tile][P] in a fixed summation order.  T is in standardised
// units; the loss term is sum M (y - T)^2 / msum.
__global__ __launch_bounds__(TILE) void k_mlp_backward(
    const float* __restrict__ X, const float* __restrict__ T, const float* __restrict__ M,
    const uint32_t* __restrict__ idx, uint32_t N, float msum, const uint16_t* __restrict__ wbf,
    Norm nm, float* __restrict__ partial, float* __restrict__ loss_partial) {
  __shared__ float sw[P];
  __shared__ float sx[TILE][F];
  __shared__ float sh1[TILE][H];
  __shared__ float sh2[TILE][H];
  __shared__ float sd1[TILE][H];
  __shared__ float sd2[TILE][H];
  __shared__ float sdy[TILE][O];
  __shared__ float sl[TILE];
  load_weights(wbf, sw);
  const uint32_t s = threadIdx.x;
  const uint64_t g = (uint64_t)blockIdx.x * TILE + s;
  float loss = 0.f;
  if (g < N) {
    const uint64_t r = idx ? idx[g] : g;
    float x[F], h1[H], h2[H], y[O];
    for (int i = 0; i < F; ++i) x[i] = (X[r * F + i] - nm.xm[i]) / nm.xs[i];
    mlp_forward(sw, x, h1, h2, y);
    float dy[O];
    for (int j = 0; j < O; ++j) {
      const float m = M[r * O + j];
      const float e = y[j] - T[r * O + j];
      loss += m * e * e;
      dy[j] = m * 2.f * e / msum;
    }
    float d2[H], d1[H];
    for (int i = 0; i < H; ++i) {
      float a = 0.f;
      for (int j = 0; j < O; ++j) a += sw[OFF_W3 + i * O + j] * dy[j];
      d2[i] = h2[i] > 0.f ? a : 0.f;
    }
    for (int i = 0; i < H; ++i) {
      float a = 0.f;
      for (int j = 0; j < H; ++j) a += sw[OFF_W2 + i * H + j] * d2[j];
      d1[i] = h1[i] > 0.f ? a : 0.f;
    }
    for (int i = 0; i < F; ++i) sx[s][i] = x[i];
    for (int i = 0; i < H; ++i) {
      sh1[s][i] = h1[i];
      sh2[s][i] = h2[i];
      sd1[s][i] = d1[i];
      sd2[s][i] = d2[i];
    }
    for (int j = 0; j < O; ++j) sdy[s][j] = dy[j];
  } else {
    for (int i = 0; i < F; ++i) sx[s][i] = 0.f;
    for (int i = 0; i < H; ++i) sh1[s][i] = sh2[s][i] = sd1[s][i] = sd2[s][i] = 0.f;
    for (int j = 0; j < O; ++j) sdy[s][j] = 0.f;
  }
  sl[s] = loss;
  __syncthreads();
  const uint32_t rows = (uint32_t)min<uint64_t>(TILE, N - (uint64_t)blockIdx.x * TILE);
  float* out = partial + (uint64_t)blockIdx.x * P;
  for (int p = threadIdx.x; p < P; p += TILE) {
    float a = 0.f;
    if (p < OFF_B1) {
      const int i = (p - OFF_W1) / H, j = (p - OFF_W1) % H;
      for (uint32_t t = 0; t < rows; ++t) a += sx[t][i] * sd1[t][j];
    } else if (p < OFF_W2) {
      const int j = p - OFF_B1;
      for (uint32_t t = 0; t < rows; ++t) a += sd1[t][j];
    } else if (p < OFF_B2) {
      const int i = (p - OFF_W2) / H, j = (p - OFF_W2) % H;
      for (uint32_t t = 0; t < rows; ++t) a += sh1[t][i] * sd2[t][j];
    } else if (p < OFF_W3) {
      const int j = p - OFF_B2;
      for (uint32_t t = 0; t < rows; ++t) a += sd2[t][j];
    } else if (p < OFF_B3) {
      const int i = (p - OFF_W3) / O, j = (p - OFF_W3) % O;
      for (uint32_t t = 0; t < rows; ++t) a += sh2[t][i] * sdy[t][j];
    } else {
      const int j = p - OFF_B3;
      for (uint32_t t = 0; t < rows; ++t) a += sdy[t][j];
    }
    out[p] = a;
  }
  if (threadIdx.x == 0) {
    float l = 0.f;
    for (uint32_t t = 0; t < rows; ++t) l += sl[t];
    loss_partial[blockIdx.x] = l / msum;
  }
}

// grad[p] = sum over tiles in tile order; loss = sum of tile losses
__global__ void k_reduce_tiles(const float* __restrict__ partial, const float* __restrict__ lp,
                               uint32_t ntiles, float* __restrict__ grad, float* __restrict__ loss) {
  for (uint64_t p = gtid(); p <= (uint64_t)P; p += gstride()) {
    float a = 0.f;
    if (p < (uint64_t)P) {
      for (uint32_t t = 0; t < ntiles; ++t) a += partial[(uint64_t)t * P + p];
      grad[p] = a;
    } else {
      for (uint32_t t = 0; t < ntiles; ++t) a += lp[t];
      *loss = a;
    }
  }
}

__global__ void k_adam(float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m,
                       float* __restrict__ v, uint16_t* __restrict__ wbf, float lr, float b1,
                       float b2, float eps, float c1, float c2) {
  for (uint64_t p = gtid(); p < (uint64_t)P; p += gstride()) {
    const float gp = g[p];
    const float mp = b1 * m[p] + (1.f - b1) * gp;
    const float vp = b2 * v[p] + (1.f - b2) * gp * gp;
    m[p] = mp;
    v[p] = vp;
    const float wn = w[p] - lr * (mp / c1) / (sqrtf(vp / c2) + eps);
    w[p] = wn;
    wbf[p] = to_bf16(wn);
  }
}

__global__ void k_to_bf16(const float* __restrict__ w, uint16_t* __restrict__ wbf) {
  for (uint64_t p = gtid(); p < (uint64_t)P; p += gstride()) wbf[p] = to_bf16(w[p]);
}

// ---- features ------------------------------------------------------------
__global__ void k_features(const uint32_t* __restrict__ rp, const uint32_t* __restrict__ col,
                           const uint32_t* __restrict__ indeg, uint32_t n, float* __restrict__ X) {
  for (uint64_t u = gtid(); u < n; u += gstride()) {
    const uint32_t b = rp[u], e = rp[u + 1], du = e - b;
    double w2 = 0, smin = 0, chunks = 0;
    uint32_t mx = 0, shortc = 0;
    for (uint32_t k = b; k < e; ++k) {
      const uint32_t v = col[k];
      const uint32_t dv = rp[v + 1] - rp[v];
      w2 += dv;
      smin += dv < du ? dv : du;
      chunks += (dv + 63) / 64;
      mx = dv > mx ? dv : mx;
      shortc += dv < 64;
    }
    float* x = X + u * F;
    x[0] = log2f(1.f + du);
    x[1] = log2f(1.f + indeg[u]);
    x[2] = (float)log2(1.0 + w2);
    x[3] = log2f(1.f + mx);
    x[4] = (float)log2(1.0 + smin);
    x[5] = floorf(log2f(1.f + du));
    x[6] = du ? (float)shortc / du : 0.f;
    x[7] = (float)log2(1.0 + chunks);
  }
}

// ---- timed variants (cost measurement) -------------------------------------
template <int CAP>
__global__ __launch_bounds__(TB) void k_time_block(const uint32_t* __restrict__ rp,
                                                   const uint32_t* __restrict__ col,
                                                   const uint32_t* __restrict__ verts, uint32_t n,
                                                   uint32_t* cursor, float* __restrict__ cost,
                                                   unsigned long long* total) {
  __shared__ BlockLds<CAP> S;
  const uint32_t tid = threadIdx.x;
  clear_bitmap(S);
  unsigned long long cnt = 0;
  for (;;) {
    // the task's cost includes fetching it (cursor atomic + barrier): for
    // very small vertices that overhead is most of what the variant costs
    uint64_t t0 = 0;
    if (tid == 0) {
      t0 = wall_clock64();
      S.t = atomicAdd(cursor, 1u);
    }
    __syncthreads();
    const uint32_t t = S.t;
    if (t >= n) break;
    cnt += block_vertex<CAP>(rp, col, verts[t], S);  // ends with a barrier
    if (tid == 0) cost[verts[t]] = (float)(wall_clock64() - t0);
  }
  flush_count(cnt, total);
}

__global__ __launch_bounds__(TB) void k_time_wave(const uint32_t* __restrict__ rp,
                                                  const uint32_t* __restrict__ col,
                                                  const uint32_t* __restrict__ verts, uint32_t n,
                                                  uint32_t* cursor, float* __restrict__ cost,
                                                  unsigned long long* total) {
  __shared__ BlockLds<2048> S;
  const uint32_t lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  clear_bitmap(S);
  const WaveSlot W = wave_slot(S, wv);
  unsigned long long cnt = 0;
  // tasks are fetched kWaveChunk at a time, as the count kernel does; each
  // task's cost carries its share of the fetch
  for (;;) {
    uint64_t t0 = wall_clock64();
    uint32_t i0 = 0;
    if (lane == 0) i0 = atomicAdd(cursor, kWaveChunk);
    i0 = __shfl(i0, 0, 64);
    if (i0 >= n) break;
    const uint32_t i1 = min(n, i0 + kWaveChunk);
    for (uint32_t i = i0; i < i1; ++i) {
      cnt += wave_vertex(rp, col, verts[i], lane, W);
      const uint64_t t1 = wall_clock64();
      if (lane == 0) cost[verts[i]] = (float)(t1 - t0);
      t0 = t1;
    }
  }
  flush_count(cnt, total);
}

__global__ void k_deg_range_flags(const uint32_t* __restrict__ rp, uint32_t n, uint32_t lo,
                                  uint32_t hi, uint8_t* __restrict__ f) {
  for (uint64_t u = gtid(); u < n; u += gstride()) {
    const uint32_t d = rp[u + 1] - rp[u];
    f[u] = (d >= lo && d <= hi) ? 1 : 0;
  }
}

// variant per vertex + sort key (descending predicted cost of that variant)
// predicted total wave-time (us) of the cheaper variant per vertex
__global__ void k_predicted_work(const uint32_t* __restrict__ rp, uint32_t n,
                                 const float* __restrict__ Y, float* __restrict__ total) {
  float acc = 0.f;
  for (uint64_t u = gtid(); u < n; u += gstride()) {
    const uint32_t d = rp[u + 1] - rp[u];
    if (d < 2) continue;
    const float cb = Y[u * O + TC_VARIANT_BLOCK] + 2.f, cw = Y[u * O + TC_VARIANT_WAVE];
    acc += exp2f(fminf(d <= kWaveMax ? fminf(cb, cw) : cb, 40.f));
  }
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
  if ((threadIdx.x & 63) == 0) atomicAdd(total, acc);
}

// A degree bin leaves the fixed rule's variant only for a predicted saving
// above this fraction of its wave-time.
constexpr float kSwitchMargin = 0.25f;

__global__ void k_fill_f32(float* __restrict__ p, uint32_t n) {
  for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0.f;
}

// Per degree bin b = floor(log2 d+) (2 <= d+ <= kWaveMax): predicted total
// wave-time (us) of the bin's vertices under each variant.  A block task
// holds four waves for its duration, a wave task one.
__global__ void k_bin_costs(const uint32_t* __restrict__ rp, uint32_t n,
                            const float* __restrict__ Y, float* __restrict__ sums /*32 x 2*/) {
  __shared__ float acc[32][2];
  if (threadIdx.x < 64) acc[threadIdx.x >> 1][threadIdx.x & 1] = 0.f;
  __syncthreads();
  for (uint64_t u = gtid(); u < n; u += gstride()) {
    const uint32_t d = rp[u + 1] - rp[u];
    if (d < 2 || d > kWaveMax) continue;
    const uint32_t b = 31u - (uint32_t)__builtin_clz(d);
    atomicAdd(&acc[b][TC_VARIANT_BLOCK], exp2f(fminf(Y[u * O + TC_VARIANT_BLOCK] + 2.f, 40.f)));
    atomicAdd(&acc[b][TC_VARIANT_WAVE], exp2f(fminf(Y[u * O + TC_VARIANT_WAVE], 40.f)));
  }
  __syncthreads();
  if (threadIdx.x < 64) {
    const float v = acc[threadIdx.x >> 1][threadIdx.x & 1];
    if (v != 0.f) atomicAdd(sums + threadIdx.x, v);
  }
}

// Variant per vertex: block unless its degree bin is in `wave_bins` (bins
// whose vertices cost less wave-time as wave tasks) and its own predicted
// wave time stays under `guard` (log2 us): a longer wave task would outlast
// the other waves and leave the machine idle.  key: block tasks by
// descending predicted cost.  wkey: wave tasks predicted above `big` first
// (descending cost), then the rest in id order (neighbouring rows share
// cache lines); other vertices sort last.
__global__ void k_assign(const uint32_t* __restrict__ rp, uint32_t n, const float* __restrict__ Y,
                         uint32_t wave_bins, float guard, float big, uint8_t* __restrict__ var,
                         uint64_t* __restrict__ key, uint64_t* __restrict__ wkey, uint32_t* ids,
                         uint32_t* ids2) {
  for (uint64_t u = gtid(); u < n; u += gstride()) {
    const uint32_t d = rp[u + 1] - rp[u];
    uint8_t v = 255;
    float c = 0.f;
    if (d >= 2) {
      const float cb = Y[u * O + TC_VARIANT_BLOCK] + 2.f, cw = Y[u * O + TC_VARIANT_WAVE];
      const uint32_t b = 31u - (uint32_t)__builtin_clz(d);
      const bool wave = d <= kWaveMax && ((wave_bins >> b) & 1u) && cw <= guard;
      v = wave ? TC_VARIANT_WAVE : TC_VARIANT_BLOCK;
      c = wave ? cw : cb;
    }
    var[u] = v;
    // predicted log2 cost -> order-preserving integer, descending after sort
    const float clamped = fminf(fmaxf(c, -30.f), 60.f);
    const uint64_t desc = ~(uint64_t)(uint32_t)((clamped + 32.f) * 16777216.f);
    key[u] = desc;
    wkey[u] = v != TC_VARIANT_WAVE ? ~0ull
              : c > big ? (desc & ~(1ull << 63)) : ((1ull << 63) | (uint64_t)u);
    ids[u] = (uint32_t)u;
    ids2[u] = (uint32_t)u;
  }
}

// fixed rule as costs: the chosen variant gets the analytic work estimate,
// the other one a larger value
__global__ void k_fixed_costs(const uint32_t* __restrict__ rp, const float* __restrict__ X,
                              uint32_t n, float* __restrict__ Y) {
  for (uint64_t u = gtid(); u < n; u += gstride()) {
    const uint32_t d = rp[u + 1] - rp[u];
    const float w = X[u * F + 2];
    const bool wave = d <= 64;
    Y[u * O + TC_VARIANT_BLOCK] = wave ? w + 1.f : w - 2.f;
    Y[u * O + TC_VARIANT_WAVE] = wave ? w : w + 1.f;
  }
}

__global__ void k_pick(const uint8_t* __restrict__ var, const uint32_t* __restrict__ order,
                       const uint32_t* __restrict__ rp, uint32_t n, uint8_t want, uint32_t dmax,
                       uint8_t* __restrict__ flag) {
  for (uint64_t i = gtid(); i < n; i += gstride()) {
    const uint32_t u = order[i];
    flag[i] = (var[u] == want && rp[u + 1] - rp[u] <= dmax) ? 1 : 0;
  }
}

__global__ void k_pick_by_id(const uint8_t* __restrict__ var, const uint32_t* __restrict__ rp,
                             uint32_t n, uint8_t want, uint32_t dmax, uint8_t* __restrict__ flag) {
  for (uint64_t u = gtid(); u < n; u += gstride())
    flag[u] = (var[u] == want && rp[u + 1] - rp[u] <= dmax) ? 1 : 0;
}

__global__ void k_pick_huge(const uint8_t* __restrict__ var, const uint32_t* __restrict__ order,
                            const uint32_t* __restrict__ rp, uint32_t n, uint32_t dmin,
                            uint8_t* __restrict__ flag) {
  for (uint64_t i = gtid(); i < n; i += gstride()) {
    const uint32_t u = order[i];
    flag[i] = (var[u] != 255 && rp[u + 1] - rp[u] > dmin) ? 1 : 0;
  }
}

unsigned grid_for(uint64_t n, int tb = 256) {
  uint64_t b = (n + tb - 1) / tb;
  return (unsigned)std::max<uint64_t>(1, std::min<uint64_t>(b, 1u << 20));
}

uint64_t splitmix(uint64_t& s) {
  uint64_t z = (s += 0x9e3779b97f4a7c15ull);
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

}  // namespace
}  // namespace tcx

using namespace tcx;

struct tc_model {
  tc_ctx* ctx = nullptr;
  DevBuf w, wbf, m, v, grad, partial, lossp, loss;
  Norm nm;
  uint64_t step = 0;
  uint64_t id = 0;
};

namespace {

uint64_t g_model_ids = 1;

int model_upload_bf16(tc_model* md) {
  hipLaunchKernelGGL(k_to_bf16, dim3(grid_for(P)), dim3(256), 0, md->ctx->stream, md->w.as<float>(),
                     md->wbf.as<uint16_t>());
  TCX_HIP(hipGetLastError());
  return TC_OK;
}

int upload(tc_ctx* c, DevBuf& b, const void* h, size_t bytes) {
  TCX_TRY(b.reserve(std::max<size_t>(bytes, 64)));
  if (bytes) TCX_HIP(hipMemcpyAsync(b.p, h, bytes, hipMemcpyHostToDevice, c->stream));
  return TC_OK;
}

// loss + gradient over samples (optionally a subset through idx) into md->grad
int model_grad(tc_ctx* c, tc_model* md, const float* dX, const float* dT, const float* dM,
               const uint32_t* didx, uint32_t N, float msum) {
  const uint32_t ntiles = (N + TILE - 1) / TILE;
  TCX_TRY(md->partial.reserve(4ull * ntiles * P));
  TCX_TRY(md->lossp.reserve(4ull * ntiles + 64));
  hipLaunchKernelGGL(k_mlp_backward, dim3(ntiles), dim3(TILE), 0, c->stream, dX, dT, dM, didx, N,
                     msum, md->wbf.as<uint16_t>(), md->nm, md->partial.as<float>(),
                     md->lossp.as<float>());
  hipLaunchKernelGGL(k_reduce_tiles, dim3(grid_for(P + 1)), dim3(256), 0, c->stream,
                     md->partial.as<float>(), md->lossp.as<float>(), ntiles, md->grad.as<float>(),
                     md->loss.as<float>());
  TCX_HIP(hipGetLastError());
  return TC_OK;
}

// standardised targets on the host
void standardise_targets(const Norm& nm, const float* T, uint32_t N, std::vector<float>& out) {
  out.resize((size_t)N * O);
  for (uint64_t s = 0; s < N; ++s)
    for (int j = 0; j < O; ++j) out[s * O + j] = (T[s * O + j] - nm.ym[j]) / nm.ys[j];
}

int check_model(tc_ctx* c, tc_model* md) {
  if (!c || !md) return fail(TC_ERR_ARG, "NULL argument");
  if (md->ctx != c) return fail(TC_ERR_ARG, "model belongs to another context");
  TCX_HIP(hipSetDevice(c->device));
  return TC_OK;
}

}  // namespace

namespace {

int graph_features(tc_ctx* c, tc_graph* g, DevBuf& X) {
  const uint32_t n = g->n;
  DevBuf indeg;
  TCX_TRY(indeg.reserve(4 * ((size_t)n + 1)));
  TCX_TRY(X.reserve(4ull * F * ((size_t)n + 1)));
  TCX_HIP(hipMemsetAsync(indeg.p, 0, 4 * ((size_t)n + 1), c->stream));
  if (g->m_or)
    hipLaunchKernelGGL(k_count_ids<8192>, dim3(std::min<uint64_t>(grid_for(g->m_or), (uint64_t)c->num_cus * 4u)),
                       dim3(256), 0, c->stream, g->col.as<uint32_t>(), (const uint32_t*)nullptr,
                       g->m_or, n, indeg.as<uint32_t>());
  hipLaunchKernelGGL(k_features, dim3(grid_for(n)), dim3(256), 0, c->stream,
                     g->rowptr.as<uint32_t>(), g->col.as<uint32_t>(), indeg.as<uint32_t>(), n,
                     X.as<float>());
  TCX_HIP(hipGetLastError());
  TCX_HIP(hipStreamSynchronize(c->stream));
  return TC_OK;
}

// per-vertex predicted [block, wave] log2 cost, or the fixed rule
int assignment(tc_ctx* c, tc_graph* g, tc_model* md, DevBuf& var, DevBuf& key, DevBuf& wkey,
               DevBuf& ids, DevBuf& ids2) {
  const uint32_t n = g->n;
  TCX_TRY(var.reserve((size_t)n + 1));
  TCX_TRY(key.reserve(8 * ((size_t)n + 1)));
  TCX_TRY(wkey.reserve(8 * ((size_t)n + 1)));
  TCX_TRY(ids2.reserve(4 * ((size_t)n + 1)));
  TCX_TRY(ids.reserve(4 * ((size_t)n + 1)));
  DevBuf Y;
  TCX_TRY(Y.reserve(4ull * O * ((size_t)n + 1)));
  if (md) {
    DevBuf X;
    TCX_TRY(graph_features(c, g, X));
    hipLaunchKernelGGL(k_mlp_forward, dim3(std::min<unsigned>(grid_for(n), 8192)), dim3(256), 0,
                       c->stream, X.as<float>(), n, md->wbf.as<uint16_t>(), md->nm, Y.as<float>());
  } else {
    // fallback: the fixed degree-bin rule, ordered by the analytic work
    // log2(1 + sum_v d+(v)) (the AUTO schedule of tc_count)
    DevBuf X;
    TCX_TRY(graph_features(c, g, X));
    hipLaunchKernelGGL(k_fixed_costs, dim3(grid_for(n)), dim3(256), 0, c->stream,
                       g->rowptr.as<uint32_t>(), X.as<float>(), n, Y.as<float>());
  }
  // per-bin variant: the one with the smaller predicted total wave-time
  TCX_TRY(c->prim_out.reserve(64 * 4 + 64));
  float* dsum = c->prim_out.as<float>();
  hipLaunchKernelGGL(k_fill_f32, dim3(1), dim3(64), 0, c->stream, dsum, 65u);
  hipLaunchKernelGGL(k_bin_costs, dim3(std::min<unsigned>(grid_for(n), 2048)), dim3(256), 0,
                     c->stream, g->rowptr.as<uint32_t>(), n, Y.as<float>(), dsum);
  uint32_t wave_bins = 0;
  {
    float sums[64];
    TCX_HIP(hipMemcpyAsync(sums, dsum, sizeof(sums), hipMemcpyDeviceToHost, c->stream));
    TCX_HIP(hipStreamSynchronize(c->stream));
    // the fixed rule (wave iff d+ <= 64, bins 1..5 and d+ = 64 of bin 6) is
    // the prior: a bin changes variant only when the prediction says the
    // other one saves more than kSwitchMargin of the bin's wave-time
    for (int b = 0; b < 32; ++b) {
      const float w = sums[2 * b + TC_VARIANT_WAVE], bl = sums[2 * b + TC_VARIANT_BLOCK];
      if (!(w > 0.f)) continue;
      const bool prior_wave = b <= 5;
      const bool wave = prior_wave ? !(bl < (1.f - kSwitchMargin) * w)
                                   : (w < (1.f - kSwitchMargin) * bl);
      if (wave) wave_bins |= 1u << b;
    }
  }
  // per-wave share of the predicted wave-time (log2 us): a wave task may take
  // at most half of it; tasks above 1/16 of it (a fetch holds kWaveChunk
  // tasks) go first so that they cannot end up in the tail
  float guard = 1e30f, big = 1e30f;
  if (md) {
    float* dtot = c->prim_out.as<float>() + 64;
    hipLaunchKernelGGL(k_predicted_work, dim3(std::min<unsigned>(grid_for(n), 2048)), dim3(256), 0,
                       c->stream, g->rowptr.as<uint32_t>(), n, Y.as<float>(), dtot);
    float tot = 0;
    TCX_HIP(hipMemcpyAsync(&tot, dtot, 4, hipMemcpyDeviceToHost, c->stream));
    TCX_HIP(hipStreamSynchronize(c->stream));
    const double waves = (double)c->num_cus * 32.0;
    const float share = (float)std::log2(std::max(1e-3, tot / waves));
    guard = share - 1.f;
    big = share - std::log2((float)kWaveChunk);
  }
  hipLaunchKernelGGL(k_assign, dim3(grid_for(n)), dim3(256), 0, c->stream,
                     g->rowptr.as<uint32_t>(), n, Y.as<float>(), wave_bins, guard, big,
                     var.as<uint8_t>(), key.as<uint64_t>(), wkey.as<uint64_t>(), ids.as<uint32_t>(),
                     ids2.as<uint32_t>());
  TCX_HIP(hipGetLastError());
  TCX_HIP(hipStreamSynchronize(c->stream));
  return TC_OK;
}

int learned_schedule(tc_ctx* c, tc_graph* g, tc_model* md) {
  SchedCache& L = g->sc;
  const uint64_t mid = md ? md->id : 0;
  if (L.learned_version == g->sched_version && L.learned_model == mid) return TC_OK;
  const uint32_t n = g->n;
  DevBuf var, key, wkey, ids, ids2, ktmp, itmp, flag;
  TCX_TRY(assignment(c, g, md, var, key, wkey, ids, ids2));
  TCX_TRY(ktmp.reserve(8 * ((size_t)n + 1)));
  TCX_TRY(itmp.reserve(4 * ((size_t)n + 1)));
  TCX_TRY(flag.reserve((size_t)n + 1));
  uint64_t* sk;
  uint32_t* order;
  TCX_TRY(sort_pairs_u64_u32(c, key.as<uint64_t>(), ktmp.as<uint64_t>(), ids.as<uint32_t>(),
                             itmp.as<uint32_t>(), n, 64, &sk, &order));
  TCX_TRY(L.l_blk.reserve(4 * ((size_t)n + 1)));
  TCX_TRY(L.l_wav.reserve(4 * ((size_t)n + 1)));
  TCX_TRY(L.l_huge.reserve(4 * ((size_t)n + 1)));
  const uint32_t* rp = g->rowptr.as<uint32_t>();
  hipLaunchKernelGGL(k_pick, dim3(grid_for(n)), dim3(256), 0, c->stream, var.as<uint8_t>(), order,
                     rp, n, (uint8_t)TC_VARIANT_BLOCK, 2048u, flag.as<uint8_t>());
  TCX_TRY(select_flagged_u32(c, order, flag.as<uint8_t>(), L.l_blk.as<uint32_t>(), n, &L.nl_blk));
  // wave tasks: the predicted-long ones first, the rest in id order (many
  // short tasks gain more from the locality of neighbouring rows than from a
  // cost order); the count comes from the flags, the order from wkey
  hipLaunchKernelGGL(k_pick_by_id, dim3(grid_for(n)), dim3(256), 0, c->stream, var.as<uint8_t>(),
                     rp, n, (uint8_t)TC_VARIANT_WAVE, kWaveMax, flag.as<uint8_t>());
  TCX_TRY(select_indices(c, flag.as<uint8_t>(), n, L.l_wav.as<uint32_t>(), &L.nl_wav));
  hipLaunchKernelGGL(k_pick_huge, dim3(grid_for(n)), dim3(256), 0, c->stream, var.as<uint8_t>(),
                     order, rp, n, 2048u, flag.as<uint8_t>());
  TCX_TRY(select_flagged_u32(c, order, flag.as<uint8_t>(), L.l_huge.as<uint32_t>(), n, &L.nl_huge));
  // (after the last use of `order`: this sort reuses its temporaries)
  uint64_t* swk;
  uint32_t* worder;
  TCX_TRY(sort_pairs_u64_u32(c, wkey.as<uint64_t>(), ktmp.as<uint64_t>(), ids2.as<uint32_t>(),
                             itmp.as<uint32_t>(), n, 64, &swk, &worder));
  if (L.nl_wav)
    TCX_HIP(hipMemcpyAsync(L.l_wav.p, worder, 4 * L.nl_wav, hipMemcpyDeviceToDevice, c->stream));
  TCX_HIP(hipGetLastError());
  L.learned_version = g->sched_version;
  L.learned_model = mid;
  return TC_OK;
}

}  // namespace

namespace tcx {
uint64_t model_id(const tc_model* md) { return md ? md->id : 0; }

int model_predict(tc_ctx* c, tc_model* md, const float* dX, uint32_t n, float* dY) {
  TCX_TRY(check_model(c, md));
  if (n == 0) return TC_OK;
  hipLaunchKernelGGL(k_mlp_forward, dim3(std::min<unsigned>(grid_for(n), 8192)), dim3(256), 0,
                     c->stream, dX, n, md->wbf.as<uint16_t>(), md->nm, dY);
  TCX_HIP(hipGetLastError());
  return TC_OK;
}
}  // namespace tcx

extern "C" {

uint32_t tc_model_num_params(void) { return P; }

int tc_model_create(tc_ctx* c, uint64_t seed, tc_model** out) {
  if (!c || !out) return fail(TC_ERR_ARG, "NULL argument");
  *out = nullptr;
  TCX_HIP(hipSetDevice(c->device));
  tc_model* md = new tc_model();
  md->ctx = c;
  md->id = g_model_ids++;
  for (int i = 0; i < F; ++i) { md->nm.xm[i] = 0.f; md->nm.xs[i] = 1.f; }
  for (int j = 0; j < O; ++j) { md->nm.ym[j] = 0.f; md->nm.ys[j] = 1.f; }
  int st = md->w.reserve(4 * P);
  if (!st) st = md->wbf.reserve(2 * P);
  if (!st) st = md->m.reserve(4 * P);
  if (!st) st = md->v.reserve(4 * P);
  if (!st) st = md->grad.reserve(4 * P);
  if (!st) st = md->loss.reserve(64);
  if (st) { delete md; return st; }
  // He-uniform initialisation from a splitmix64 stream
  std::vector<float> w(P, 0.f);
  uint64_t s = seed * 0x9e3779b97f4a7c15ull + 1;
  auto uni = [&](float lim) { return (float)(((splitmix(s) >> 11) * (1.0 / 9007199254740992.0)) * 2 - 1) * lim; };
  for (int i = 0; i < F * H; ++i) w[OFF_W1 + i] = uni(std::sqrt(6.f / F));
  for (int i = 0; i < H * H; ++i) w[OFF_W2 + i] = uni(std::sqrt(6.f / H));
  for (int i = 0; i < H * O; ++i) w[OFF_W3 + i] = uni(std::sqrt(6.f / H));
  if (hipMemcpyAsync(md->w.p, w.data(), 4 * P, hipMemcpyHostToDevice, c->stream) != hipSuccess ||
      hipMemsetAsync(md->m.p, 0, 4 * P, c->stream) != hipSuccess ||
      hipMemsetAsync(md->v.p, 0, 4 * P, c->stream) != hipSuccess) {
    delete md;
    return fail(TC_ERR_HIP, "model upload failed");
  }
  st = model_upload_bf16(md);
  if (st) { delete md; return st; }
  (void)hipStreamSynchronize(c->stream);
  *out = md;
  return TC_OK;
}

int tc_model_destroy(tc_model* md) {
  if (!md) return TC_OK;
  if (md->ctx) (void)hipStreamSynchronize(md->ctx->stream);
  delete md;
  return TC_OK;
}

int tc_model_set_params(tc_model* md, const float* params, uint32_t n) {
  if (!md || !params || n != (uint32_t)P) return fail(TC_ERR_ARG, "need %d parameters", P);
  tc_ctx* c = md->ctx;
  TCX_HIP(hipSetDevice(c->device));
  TCX_HIP(hipMemcpyAsync(md->w.p, params, 4 * P, hipMemcpyHostToDevice, c->stream));
  TCX_HIP(hipMemsetAsync(md->m.p, 0, 4 * P, c->stream));
  TCX_HIP(hipMemsetAsync(md->v.p, 0, 4 * P, c->stream));
  TCX_TRY(model_upload_bf16(md));
  TCX_HIP(hipStreamSynchronize(c->stream));
  md->step = 0;
  md->id = g_model_ids++;
  return TC_OK;
}

int tc_model_get_params(const tc_model* md, float* params, uint32_t n) {
  if (!md || !params || n != (uint32_t)P) return fail(TC_ERR_ARG, "need %d parameters", P);
  TCX_HIP(hipSetDevice(md->ctx->device));
  TCX_HIP(hipMemcpyAsync(params, md->w.p, 4 * P, hipMemcpyDeviceToHost, md->ctx->stream));
  TCX_HIP(hipStreamSynchronize(md->ctx->stream));
  return TC_OK;
}

int tc_model_set_norm(tc_model* md, const float* xm, const float* xs, const float* ym,
                      const float* ys) {
  if (!md || !xm || !xs || !ym || !ys) return fail(TC_ERR_ARG, "NULL argument");
  for (int i = 0; i < F; ++i) {
    if (!(xs[i] > 0)) return fail(TC_ERR_ARG, "x_std must be positive");
    md->nm.xm[i] = xm[i];
    md->nm.xs[i] = xs[i];
  }
  for (int j = 0; j < O; ++j) {
    if (!(ys[j] > 0)) return fail(TC_ERR_ARG, "y_std must be positive");
    md->nm.ym[j] = ym[j];
    md->nm.ys[j] = ys[j];
  }
  md->id = g_model_ids++;
  return TC_OK;
}

int tc_model_get_norm(const tc_model* md, float* xm, float* xs, float* ym, float* ys) {
  if (!md || !xm || !xs || !ym || !ys) return fail(TC_ERR_ARG, "NULL argument");
  for (int i = 0; i < F; ++i) { xm[i] = md->nm.xm[i]; xs[i] = md->nm.xs[i]; }
  for (int j = 0; j < O; ++j) { ym[j] = md->nm.ym[j]; ys[j] = md->nm.ys[j]; }
  return TC_OK;
}

int tc_model_forward(tc_ctx* c, tc_model* md, const float* X, uint32_t N, float* Y) {
  TCX_TRY(check_model(c, md));
  if (N && (!X || !Y)) return fail(TC_ERR_ARG, "NULL arrays");
  if (N == 0) return TC_OK;
  DevBuf dX, dY;
  TCX_TRY(upload(c, dX, X, 4ull * F * N));
  TCX_TRY(dY.reserve(4ull * O * N));
  hipLaunchKernelGGL(k_mlp_forward, dim3(std::min<unsigned>(grid_for(N), 4096)), dim3(256), 0,
                     c->stream, dX.as<float>(), N, md->wbf.as<uint16_t>(), md->nm, dY.as<float>());
  TCX_HIP(hipGetLastError());
  TCX_HIP(hipMemcpyAsync(Y, dY.p, 4ull * O * N, hipMemcpyDeviceToHost, c->stream));
  TCX_HIP(hipStreamSynchronize(c->stream));
  return TC_OK;
}

int tc_model_loss_grad(tc_ctx* c, tc_model* md, const float* X, const float* T, const float* M,
                       uint32_t N, float* loss, float* grad) {
  TCX_TRY(check_model(c, md));
  if (!X || !T || !M || !loss || !grad || N == 0) return fail(TC_ERR_ARG, "bad arguments");
  double ms = 0;
  for (uint64_t i = 0; i < (uint64_t)N * O; ++i) ms += M[i];
  if (ms <= 0) return fail(TC_ERR_ARG, "mask selects no targets");
  std::vector<float> Tn;
  standardise_targets(md->nm, T, N, Tn);
  DevBuf dX, dT, dM;
  TCX_TRY(upload(c, dX, X, 4ull * F * N));
  TCX_TRY(upload(c, dT, Tn.data(), 4ull * O * N));
  TCX_TRY(upload(c, dM, M, 4ull * O * N));
  TCX_TRY(model_grad(c, md, dX.as<float>(), dT.as<float>(), dM.as<float>(), nullptr, N, (float)ms));
  TCX_HIP(hipMemcpyAsync(grad, md->grad.p, 4 * P, hipMemcpyDeviceToHost, c->stream));
  TCX_HIP(hipMemcpyAsync(loss, md->loss.p, 4, hipMemcpyDeviceToHost, c->stream));
  TCX_HIP(hipStreamSynchronize(c->stream));
  return TC_OK;
}

int tc_model_train(tc_ctx* c, tc_model* md, const float* X, const float* T, const float* M,
                   uint32_t N, const tc_train_opts* opts, float* final_loss) {
  TCX_TRY(check_model(c, md));
  if (!X || !T || !M || N == 0) return fail(TC_ERR_ARG, "bad arguments");
  tc_train_opts o;
  memset(&o, 0, sizeof(o));
  if (opts) o = *opts;
  const uint32_t steps = o.steps ? o.steps : 3000;
  const uint32_t batch = std::min<uint32_t>(o.batch ? o.batch : 4096, N);
  const float lr = o.lr > 0 ? o.lr : 3e-3f;
  // standardisation from the data (masked for the targets)
  Norm nm;
  for (int i = 0; i < F; ++i) {
    double s = 0, s2 = 0;
    for (uint64_t k = 0; k < N; ++k) { s += X[k * F + i]; s2 += (double)X[k * F + i] * X[k * F + i]; }
    double mu = s / N, var = s2 / N - mu * mu;
    nm.xm[i] = (float)mu;
    nm.xs[i] = (float)std::max(std::sqrt(std::max(var, 0.0)), 1e-3);
  }
  for (int j = 0; j < O; ++j) {
    double s = 0, s2 = 0, cnt = 0;
    for (uint64_t k = 0; k < N; ++k)
      if (M[k * O + j] > 0) { s += T[k * O + j]; s2 += (double)T[k * O + j] * T[k * O + j]; cnt++; }
    double mu = cnt ? s / cnt : 0, var = cnt ? s2 / cnt - mu * mu : 1;
    nm.ym[j] = (float)mu;
    nm.ys[j] = (float)std::max(std::sqrt(std::max(var, 0.0)), 1e-3);
  }
  md->nm = nm;
  std::vector<float> Tn;
  standardise_targets(nm, T, N, Tn);
  DevBuf dX, dT, dM, didx;
  TCX_TRY(upload(c, dX, X, 4ull * F * N));
  TCX_TRY(upload(c, dT, Tn.data(), 4ull * O * N));
  TCX_TRY(upload(c, dM, M, 4ull * O * N));
  // one shuffled order, walked in batch windows (reshuffled every epoch)
  std::vector<uint32_t> order(N);
  for (uint32_t i = 0; i < N; ++i) order[i] = i;
  uint64_t rs = 0x5eedull + o.seed;
  auto shuffle = [&]() {
    for (uint32_t i = N - 1; i > 0; --i) std::swap(order[i], order[splitmix(rs) % (i + 1)]);
  };
  shuffle();
  TCX_TRY(upload(c, didx, order.data(), 4ull * N));
  // the mask sum of one batch, estimated as the data's mean times the batch
  double msum_all = 0;
  for (uint64_t i = 0; i < (uint64_t)N * O; ++i) msum_all += M[i];
  const float msum = (float)std::max(1.0, msum_all / N * batch);
  uint32_t pos = 0;
  for (uint32_t it = 0; it < steps; ++it) {
    if (pos + batch > N) {
      shuffle();
      TCX_HIP(hipMemcpyAsync(didx.p, order.data(), 4ull * N, hipMemcpyHostToDevice, c->stream));
      pos = 0;
    }
    TCX_TRY(model_grad(c, md, dX.as<float>(), dT.as<float>(), dM.as<float>(),
                       didx.as<uint32_t>() + pos, batch, msum));
    pos += batch;
    md->step++;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.f - std::pow(b1, (float)md->step), c2 = 1.f - std::pow(b2, (float)md->step);
    hipLaunchKernelGGL(k_adam, dim3(grid_for(P)), dim3(256), 0, c->stream, md->w.as<float>(),
                       md->grad.as<float>(), md->m.as<float>(), md->v.as<float>(),
                       md->wbf.as<uint16_t>(), lr, b1, b2, 1e-8f, c1, c2);
    TCX_HIP(hipGetLastError());
  }
  // final loss over the whole set
  TCX_TRY(model_grad(c, md, dX.as<float>(), dT.as<float>(), dM.as<float>(), nullptr, N,
                     (float)std::max(1.0, msum_all)));
  float l = 0;
  TCX_HIP(hipMemcpyAsync(&l, md->loss.p, 4, hipMemcpyDeviceToHost, c->stream));
  TCX_HIP(hipStreamSynchronize(c->stream));
  if (final_loss) *final_loss = l;
  md->id = g_model_ids++;
  return TC_OK;
}

int tc_graph_features(tc_ctx* c, tc_graph* g, float* X) {
  if (!c || !g || !X) return fail(TC_ERR_ARG, "NULL argument");
  if (g->orient_mode < 0) return fail(TC_ERR_STATE, "graph must be oriented");
  TCX_HIP(hipSetDevice(c->device));
  if (g->n == 0) return TC_OK;
  DevBuf dX;
  TCX_TRY(graph_features(c, g, dX));
  TCX_HIP(hipMemcpyAsync(X, dX.p, 4ull * F * g->n, hipMemcpyDeviceToHost, c->stream));
  TCX_HIP(hipStreamSynchronize(c->stream));
  return TC_OK;
}

int tc_graph_measure_costs(tc_ctx* c, tc_graph* g, int variant, float* cost_us, uint64_t* count) {
  if (!c || !g || !cost_us) return fail(TC_ERR_ARG, "NULL argument");
  if (g->orient_mode < 0) return fail(TC_ERR_STATE, "graph must be oriented");
  if (variant != TC_VARIANT_BLOCK && variant != TC_VARIANT_WAVE)
    return fail(TC_ERR_ARG, "unknown variant %d", variant);
  TCX_HIP(hipSetDevice(c->device));
  const uint32_t n = g->n;
  if (n == 0) return TC_OK;
  DevBuf flags, verts, cost;
  TCX_TRY(flags.reserve((size_t)n + 1));
  TCX_TRY(verts.reserve(4 * ((size_t)n + 1)));
  TCX_TRY(cost.reserve(4 * ((size_t)n + 1)));
  TCX_TRY(c->scratch.reserve(64));
  TCX_TRY(c->accum.reserve(64));
  uint32_t* cursor = c->scratch.as<uint32_t>();
  unsigned long long* acc = c->accum.as<unsigned long long>();
  TCX_HIP(hipMemsetAsync(cost.p, 0, 4 * (size_t)n, c->stream));
  TCX_HIP(hipMemsetAsync(cursor, 0, 16, c->stream));
  TCX_HIP(hipMemsetAsync(acc, 0, 8, c->stream));
  const uint32_t* rp = g->rowptr.as<uint32_t>();
  const uint32_t* col = g->col.as<uint32_t>();
  uint64_t nv = 0;
  const uint32_t G = (uint32_t)c->num_cus * 8u;
  if (variant == TC_VARIANT_BLOCK) {
    hipLaunchKernelGGL(k_deg_range_flags, dim3(grid_for(n)), dim3(256), 0, c->stream, rp, n, 2u,
                       2048u, flags.as<uint8_t>());
    TCX_TRY(select_indices(c, flags.as<uint8_t>(), n, verts.as<uint32_t>(), &nv));
    if (nv)
      hipLaunchKernelGGL(k_time_block<2048>, dim3(G), dim3(TB), 0, c->stream, rp, col,
                         verts.as<uint32_t>(), (uint32_t)nv, cursor, cost.as<float>(), acc);
    hipLaunchKernelGGL(k_deg_range_flags, dim3(grid_for(n)), dim3(256), 0, c->stream, rp, n, 2049u,
                       0xFFFFFFFFu, flags.as<uint8_t>());
    TCX_TRY(select_indices(c, flags.as<uint8_t>(), n, verts.as<uint32_t>(), &nv));
    if (nv)
      hipLaunchKernelGGL(k_time_block<8192>, dim3(std::min<uint64_t>(nv, (uint64_t)c->num_cus * 4)),
                         dim3(TB), 0, c->stream, rp, col, verts.as<uint32_t>(), (uint32_t)nv,
                         cursor + 1, cost.as<float>(), acc);
  } else {
    hipLaunchKernelGGL(k_deg_range_flags, dim3(grid_for(n)), dim3(256), 0, c->stream, rp, n, 2u,
                       kWaveMax, flags.as<uint8_t>());
    TCX_TRY(select_indices(c, flags.as<uint8_t>(), n, verts.as<uint32_t>(), &nv));
    if (nv)
      hipLaunchKernelGGL(k_time_wave, dim3(G), dim3(TB), 0, c->stream, rp, col,
                         verts.as<uint32_t>(), (uint32_t)nv, cursor, cost.as<float>(), acc);
  }
  TCX_HIP(hipGetLastError());
  std::vector<float> h(n);
  TCX_HIP(hipMemcpyAsync(h.data(), cost.p, 4 * (size_t)n, hipMemcpyDeviceToHost, c->stream));
  unsigned long long total = 0;
  TCX_HIP(hipMemcpyAsync(&total, acc, 8, hipMemcpyDeviceToHost, c->stream));
  TCX_HIP(hipStreamSynchronize(c->stream));
  const float us_per_tick = 1000.f / (float)c->wall_clock_khz;
  for (uint32_t u = 0; u < n; ++u) cost_us[u] = h[u] * us_per_tick;
  if (count) *count = total;
  return TC_OK;
}

int tc_learned_assignment(tc_ctx* c, tc_graph* g, tc_model* md, uint8_t* variant) {
  if (!c || !g || !variant) return fail(TC_ERR_ARG, "NULL argument");
  if (md && md->ctx != c) return fail(TC_ERR_ARG, "model belongs to another context");
  if (g->orient_mode < 0) return fail(TC_ERR_STATE, "graph must be oriented");
  TCX_HIP(hipSetDevice(c->device));
  if (g->n == 0) return TC_OK;
  DevBuf var, key, wkey, ids, ids2;
  TCX_TRY(assignment(c, g, md, var, key, wkey, ids, ids2));
  TCX_HIP(hipMemcpyAsync(variant, var.p, g->n, hipMemcpyDeviceToHost, c->stream));
  TCX_HIP(hipStreamSynchronize(c->stream));
  return TC_OK;
}

int tc_count_learned(tc_ctx* c, tc_graph* g, tc_model* md, uint64_t* count, tc_timing* timing) {
  if (!c || !g || !count) return fail(TC_ERR_ARG, "NULL argument");
  if (g->ctx != c) return fail(TC_ERR_ARG, "graph belongs to another context");
  if (md && md->ctx != c) return fail(TC_ERR_ARG, "model belongs to another context");
  if (g->orient_mode < 0) return fail(TC_ERR_STATE, "graph must be oriented before counting");
  TCX_HIP(hipSetDevice(c->device));
  tc_timing tm;
  memset(&tm, 0, sizeof(tm));
  *count = 0;
  if (g->m_or == 0 || g->n == 0) {
    if (timing) *timing = tm;
    return TC_OK;
  }
  const SchedCache& L = g->sc;
  TCX_HIP(hipEventRecord(c->ev0, c->stream));
  TCX_TRY(learned_schedule(c, g, md));
  TCX_HIP(hipEventRecord(c->ev1, c->stream));
  TCX_HIP(hipEventSynchronize(c->ev1));
  TCX_HIP(hipEventElapsedTime(&tm.ms_prep, c->ev0, c->ev1));
  TCX_TRY(c->scratch.reserve(64));
  TCX_TRY(c->accum.reserve(64));
  uint32_t* cursor = c->scratch.as<uint32_t>();
  unsigned long long* acc = c->accum.as<unsigned long long>();
  TCX_HIP(hipMemsetAsync(cursor, 0, 16, c->stream));
  TCX_HIP(hipMemsetAsync(acc, 0, 8, c->stream));
  const uint32_t* rp = g->rowptr.as<uint32_t>();
  const uint32_t* col = g->col.as<uint32_t>();
  TCX_HIP(hipEventRecord(c->ev0, c->stream));
  if (L.nl_blk || L.nl_wav) {
    const uint32_t G = (uint32_t)c->num_cus * 8u;
    if (L.nl_wav <= (uint64_t)kStaticWaves * G * (TB / 64))
      hipLaunchKernelGGL((k_auto<2048, true>), dim3(G), dim3(TB), 0, c->stream, rp, col,
                         L.l_blk.as<uint32_t>(), (uint32_t)L.nl_blk, L.l_wav.as<uint32_t>(),
                         (uint32_t)L.nl_wav, cursor, acc);
    else
      hipLaunchKernelGGL(k_auto<2048>, dim3(G), dim3(TB), 0, c->stream, rp, col,
                         L.l_blk.as<uint32_t>(), (uint32_t)L.nl_blk, L.l_wav.as<uint32_t>(),
                         (uint32_t)L.nl_wav, cursor, acc);
    tm.launches++;
  }
  if (L.nl_huge) {
    hipLaunchKernelGGL(k_auto<8192>, dim3(std::min<uint64_t>((uint64_t)c->num_cus * 4, L.nl_huge)),
                       dim3(TB), 0, c->stream, rp, col, L.l_huge.as<uint32_t>(), (uint32_t)L.nl_huge,
                       (const uint32_t*)nullptr, 0u, cursor + 2, acc);
    tm.launches++;
  }
  TCX_HIP(hipGetLastError());
  TCX_HIP(hipEventRecord(c->ev1, c->stream));
  TCX_HIP(hipEventSynchronize(c->ev1));
  TCX_HIP(hipEventElapsedTime(&tm.ms_count, c->ev0, c->ev1));
  tm.ms_dominant = tm.ms_count;
  unsigned long long h = 0;
  TCX_HIP(hipMemcpyAsync(&h, acc, 8, hipMemcpyDeviceToHost, c->stream));
  TCX_HIP(hipStreamSynchronize(c->stream));
  *count = h;
  if (timing) *timing = tm;
  return TC_OK;
}

}  // extern "C"
