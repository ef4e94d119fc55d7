// Learned scheduling for the AUTO engine.
//
//  * an 8-32-32-2 MLP with BF16 weights and FP32 accumulation, trained by
//    this file's forward / backward / Adam kernels (deterministic gradient
//    reduction: per-tile partial sums, then a fixed-order tile sum);
//  * cell features of the work-list schedule (k_cell_accum) and timed count
//    launches of one cell at a time under either variant (the training data);
//  * the learned schedule: per cell, the variant the model predicts faster,
//    with the fixed rule as the prior;
//  * per-task features and per-task timed variants (k_wl_features,
//    k_time_wl_*), kept as diagnostics.
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

// ---- features of the work-list tasks ---------------------------------------
// The bandwidth model of tc_bw_model on the device: the rate (GB/s) at which
// a list of `len` entries is read.  Lists shorter than two entries take the
// two-entry rate (the model starts there).
struct BwTable {
  float table[10];
  double a, b;
};

__device__ inline float bw_rate(const BwTable& m, uint32_t len) {
  if (len < 10) return m.table[len < 2 ? 2 : len];
  return (float)(m.a * pow((double)len, m.b));
}

// Task x of the work-list schedule: N+(x) staged, N+(y) streamed for y in
// W(x).  Its eight features (0 for x without a task):
//   log2(1+d+(x)), log2(1+|W(x)|), log2(1+sum d+(y)), log2(1+max d+(y)),
//   log2(1+t) with t the ns its lists take to read under the bandwidth model
//   (sum over y of 4 d+(y) bytes / rate(d+(y))), floor(log2(1+d+(x))), the
//   fraction of y with d+(y) < 64, log2(1+sum ceil(d+(y)/64)).
__global__ void k_wl_features(const uint32_t* __restrict__ rp, const uint32_t* __restrict__ wrp,
                              const uint32_t* __restrict__ wcol, uint32_t n, BwTable bw,
                              float* __restrict__ X) {
  for (uint64_t x = gtid(); x < n; x += gstride()) {
    const uint32_t dx = rp[x + 1] - rp[x];
    const uint32_t wb = wrp[x], we = wrp[x + 1], nw = we - wb;
    double sum = 0, chunks = 0, tns = 0;
    uint32_t mx = 0, shortc = 0;
    for (uint32_t i = wb; i < we; ++i) {
      const uint32_t y = wcol[i];
      const uint32_t dy = rp[y + 1] - rp[y];
      sum += dy;
      chunks += (dy + 63) / 64;
      mx = dy > mx ? dy : mx;
      shortc += dy < 64;
      tns += 4.0 * dy / bw_rate(bw, dy);
    }
    float* f = X + x * F;
    const bool task = nw > 0 && dx >= 2;
    f[0] = task ? log2f(1.f + dx) : 0.f;
    f[1] = task ? log2f(1.f + nw) : 0.f;
    f[2] = task ? (float)log2(1.0 + sum) : 0.f;
    f[3] = task ? log2f(1.f + mx) : 0.f;
    f[4] = task ? (float)log2(1.0 + tns) : 0.f;
    f[5] = task ? floorf(log2f(1.f + dx)) : 0.f;
    f[6] = task ? (float)shortc / nw : 0.f;
    f[7] = task ? (float)log2(1.0 + chunks) : 0.f;
  }
}

// ---- timed variants (cost measurement) -------------------------------------
// Costs are wave-microseconds (in clock ticks here): the time a task holds
// one wave.  A block task holds four waves for its duration, a lane task one
// 64th of its wave's round.
template <int CAP>
__global__ __launch_bounds__(TB) void k_time_wl_block(const uint32_t* __restrict__ rp,
                                                      const uint32_t* __restrict__ col,
                                                      const uint32_t* __restrict__ wrp,
                                                      const uint2* __restrict__ wse,
                                                      const uint32_t* __restrict__ verts,
                                                      uint32_t n, uint32_t* cursor,
                                                      float* __restrict__ cost,
                                                      unsigned long long* total) {
  __shared__ RankLds<CAP> S;
  const uint32_t tid = threadIdx.x;
  for (uint32_t i = tid; i < RankLds<CAP>::kRegion; i += TB) S.region[i] = 0;
  __syncthreads();
  unsigned long long cnt = 0;
  for (;;) {
    // the task's cost includes fetching it (cursor atomic + barrier)
    uint64_t t0 = 0;
    if (tid == 0) {
      t0 = wall_clock64();
      S.t = atomicAdd(cursor, 1u);
    }
    __syncthreads();
    const uint32_t t = S.t;
    if (t >= n) break;
    const uint32_t x = verts[t], wb = wrp[x], wn = wrp[x + 1] - wb;
    if (uniform(rp[x + 1] - rp[x]) <= (uint32_t)CAP) {
      cnt += block_rank<CAP>(rp, col, x, S, wse + wb, wn);  // ends with a barrier
    } else {
      cnt += block_global(rp, col, x, wse + wb, wn);
      __syncthreads();
    }
    if (tid == 0) cost[x] = 4.f * (float)(wall_clock64() - t0);
  }
  flush_count(cnt, total);
}

__global__ __launch_bounds__(TB) void k_time_wl_wave(const uint32_t* __restrict__ rp,
                                                     const uint32_t* __restrict__ col,
                                                     const uint32_t* __restrict__ wrp,
                                                     const uint2* __restrict__ wse,
                                                     const uint32_t* __restrict__ verts, uint32_t n,
                                                     uint32_t* cursor, float* __restrict__ cost,
                                                     unsigned long long* total) {
  __shared__ RankLds<2048> S;
  const uint32_t lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  for (uint32_t i = threadIdx.x; i < RankLds<2048>::kRegion; i += TB) S.region[i] = 0;
  __syncthreads();
  const RankSlot W = rank_wave(S, wv);
  unsigned long long cnt = 0;
  // fetched kWaveChunk at a time, as the count kernel does; each task's cost
  // carries its share of the fetch
  for (;;) {
    uint64_t t0 = wall_clock64();
    uint32_t i0 = 0;
    if (lane == 0) i0 = atomicAdd(cursor, kWaveChunk);
    i0 = __shfl(i0, 0, 64);
    if (i0 >= n) break;
    const uint32_t i1 = min(n, i0 + kWaveChunk);
    for (uint32_t i = i0; i < i1; ++i) {
      const uint32_t x = verts[i], wb = wrp[x];
      cnt += wave_rank<2048>(rp, col, x, lane, W, wse + wb, wrp[x + 1] - wb);
      const uint64_t t1 = wall_clock64();
      if (lane == 0) cost[x] = (float)(t1 - t0);
      t0 = t1;
    }
  }
  flush_count(cnt, total);
}

// One lane task per lane; each task is charged 1/64 of its round (the wave
// is held until its longest lane finishes).
__global__ __launch_bounds__(TB) void k_time_wl_lane(const uint4* __restrict__ ell,
                                                     const uint32_t* __restrict__ wrp,
                                                     const uint32_t* __restrict__ wcol,
                                                     const uint32_t* __restrict__ verts, uint32_t n,
                                                     uint32_t* cursor, float* __restrict__ cost,
                                                     unsigned long long* total) {
  const uint32_t lane = threadIdx.x & 63;
  unsigned long long cnt = 0;
  for (;;) {
    const uint64_t t0 = wall_clock64();
    uint32_t i0 = 0;
    if (lane == 0) i0 = atomicAdd(cursor, 64u);
    i0 = __shfl(i0, 0, 64);
    if (i0 >= n) break;
    const uint32_t i = i0 + lane;
    if (i < n) cnt += lane_task(ell, wrp, wcol, verts[i]);
    const uint64_t t1 = wall_clock64();
    if (i < n) cost[verts[i]] = (float)(t1 - t0) / 64.f;
  }
  flush_count(cnt, total);
}

// Tasks of the work-list schedule (non-empty W(x), d+(x) >= 2) with d+ in
// [lo, hi].
__global__ void k_task_flags(const uint32_t* __restrict__ rp, const uint32_t* __restrict__ wrp,
                             uint32_t n, uint32_t lo, uint32_t hi, uint8_t* __restrict__ f) {
  for (uint64_t x = gtid(); x < n; x += gstride()) {
    const uint32_t d = rp[x + 1] - rp[x];
    f[x] = (wrp[x + 1] > wrp[x] && d >= 2 && d >= lo && d <= hi) ? 1 : 0;
  }
}

// ---- learned schedule -----------------------------------------------------
// Decision cells: d+ itself for 2..8 (cells 2..8), then (2^(b-1), 2^b] for
// b = 4..9, i.e. d+ 9..16, 17..32, ..., 257..512 (cells 9..14).  Each cell's
// tasks run their narrow variant (a lane for d+ <= 8, a wave above) or their
// wide one (a wave for d+ <= 8, a workgroup above); tasks above d+ 512 are
// block tasks.  Whatever its cell runs, a task keeps the fixed rule's limits:
// a lane holds at most kLaneMaxList streamed lists and a wave at most
// kWlWaveWork entries of work, else the task runs one size up.  Cell 15
// holds the tasks above d+ 512: it has no choice, but its features and its
// timed launch let the model predict a whole count (learned orientation
// ties compare candidate orientations by it).
constexpr int kCells = TC_CELLS;
constexpr int kChoiceCells = 15;  // cells 2..14 have a narrow variant
constexpr int kBlockCell = 15;
constexpr int kAcc = 8;  // per-cell accumulators of k_cell_accum
// The fixed rule (AUTO's work-list thresholds 64 / 8): narrow up to d+ 64.
constexpr uint32_t kFixedNarrow = 0xFFCu;  // cells 2..11
static_assert(kCells == 16, "cell layout");

__device__ __host__ inline int task_cell(uint32_t d) {
  if (d < 2 || d > kWaveMax) return -1;
  if (d <= kLaneMaxDeg) return (int)d;
  const int b = 32 - __builtin_clz(d - 1);  // ceil(log2 d): 4 .. 9
  return (int)kLaneMaxDeg + 1 + (b - 4);
}

// The cell a task's features and timings go to (cells 2..15).
__device__ __host__ inline int feature_cell(uint32_t d) {
  return d > kWaveMax ? kBlockCell : task_cell(d);
}

// TC_VARIANT_* of a task (d >= 2, nw >= 1 lists, work = d + sum d+(y)).
__device__ inline uint8_t task_variant(uint32_t d, uint32_t nw, uint64_t work,
                                       uint32_t narrow_cells) {
  const int cell = task_cell(d);
  const bool narrow = cell >= 0 && ((narrow_cells >> cell) & 1u);
  if (d <= kLaneMaxDeg) {
    if (narrow && nw <= kLaneMaxList) return TC_VARIANT_LANE;
    return work > kWlWaveWork ? TC_VARIANT_BLOCK : TC_VARIANT_WAVE;
  }
  return (narrow && work <= kWlWaveWork) ? TC_VARIANT_WAVE : TC_VARIANT_BLOCK;
}

// Tasks of sampling level s: all for s = 0, else those whose hashed id has
// s low zero bits (about 2^-s of each cell).
__device__ __host__ inline bool in_level(uint32_t x, uint32_t level) {
  if (level == 0) return true;
  uint32_t h = x * 0x9E3779B1u;
  h ^= h >> 16;
  return (h & ((1u << level) - 1u)) == 0;
}

// Per cell, over the tasks of one level: [0] tasks, [1] sum |W(x)|, [2] sum
// of streamed entries (d+(y) over W(x)), [3] sum d+(x), [4] the lists' read
// time under the bandwidth model in picoseconds (per task rounded, then
// summed), [5] streamed lists with d+(y) < 64, [6] the largest task work
// d+(x) + sum d+(y), [7] tasks within their narrow variant's limit.
__global__ void k_cell_accum(const uint32_t* __restrict__ rp, const uint32_t* __restrict__ wrp,
                             const uint32_t* __restrict__ wcol, uint32_t n, uint32_t level,
                             BwTable bw, unsigned long long* __restrict__ acc) {
  __shared__ unsigned long long sa[kCells * kAcc];
  for (int i = threadIdx.x; i < kCells * kAcc; i += blockDim.x) sa[i] = 0;
  __syncthreads();
  for (uint64_t x = gtid(); x < n; x += gstride()) {
    const uint32_t d = rp[x + 1] - rp[x];
    const uint32_t wb = wrp[x], we = wrp[x + 1];
    const int cell = feature_cell(d);
    if (we == wb || cell < 0 || !in_level((uint32_t)x, level)) continue;
    uint64_t ent = 0;
    uint32_t shortc = 0;
    double tns = 0;
    for (uint32_t i = wb; i < we; ++i) {
      const uint32_t y = wcol[i];
      const uint32_t dy = rp[y + 1] - rp[y];
      ent += dy;
      shortc += dy < 64;
      tns += 4.0 * dy / bw_rate(bw, dy);
    }
    const uint32_t nw = we - wb;
    const uint64_t work = d + ent;
    const bool ok = d <= kLaneMaxDeg ? nw <= kLaneMaxList
                                     : (d <= kWaveMax && work <= kWlWaveWork);
    unsigned long long* a = sa + cell * kAcc;
    atomicAdd(a + 0, 1ull);
    atomicAdd(a + 1, (unsigned long long)nw);
    atomicAdd(a + 2, (unsigned long long)ent);
    atomicAdd(a + 3, (unsigned long long)d);
    atomicAdd(a + 4, (unsigned long long)llround(tns * 1000.0));
    atomicAdd(a + 5, (unsigned long long)shortc);
    atomicMax(a + 6, (unsigned long long)work);
    atomicAdd(a + 7, ok ? 1ull : 0ull);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < kCells * kAcc; i += blockDim.x) {
    if (!sa[i]) continue;
    if (i % kAcc == 6) atomicMax(acc + i, sa[i]);
    else atomicAdd(acc + i, sa[i]);
  }
}

// Variant per task (TC_VARIANT_*, 255 none) under narrow_cells, restricted
// to cell `only` (all cells and the tasks above d+ 512 when only < 0) and to
// the tasks of `level`.  key: block tasks by descending work (the launch
// order, as the fixed rule's), other tasks sort last.
__global__ void k_task_var(const uint32_t* __restrict__ rp, const uint32_t* __restrict__ wrp,
                           const uint32_t* __restrict__ wcol, uint32_t n, uint32_t narrow_cells,
                           int only, uint32_t level, uint8_t* __restrict__ var,
                           uint64_t* __restrict__ key, uint32_t* __restrict__ ids) {
  for (uint64_t x = gtid(); x < n; x += gstride()) {
    const uint32_t d = rp[x + 1] - rp[x];
    const uint32_t wb = wrp[x], we = wrp[x + 1];
    uint8_t v = 255;
    uint64_t work = 0;
    if (we > wb && d >= 2 && (only < 0 || feature_cell(d) == only) &&
        in_level((uint32_t)x, level)) {
      work = d;
      for (uint32_t i = wb; i < we; ++i) {
        const uint32_t y = wcol[i];
        work += rp[y + 1] - rp[y];
      }
      v = task_variant(d, we - wb, work, narrow_cells);
    }
    var[x] = v;
    key[x] = v == TC_VARIANT_BLOCK ? ~work : ~0ull;
    ids[x] = (uint32_t)x;
  }
}

__global__ void k_var_flags(const uint8_t* __restrict__ var, const uint32_t* __restrict__ order,
                            const uint32_t* __restrict__ rp, uint64_t n, uint8_t want,
                            uint32_t dlo, uint32_t dhi, uint8_t* __restrict__ flag) {
  for (uint64_t i = gtid(); i < n; i += gstride()) {
    const uint32_t x = order ? order[i] : (uint32_t)i;
    const uint32_t d = rp[x + 1] - rp[x];
    flag[i] = (var[x] == want && d >= dlo && d <= dhi) ? 1 : 0;
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
  tc_bw_model bw;  // list-read bandwidth model behind feature 4
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

BwTable bw_table(const tc_bw_model& m) {
  BwTable t;
  for (int i = 0; i < 10; ++i) t.table[i] = m.table[i];
  t.a = m.a;
  t.b = m.b;
  return t;
}

// Work-list task features (k_wl_features) with the bandwidth model `bw`.
int graph_features(tc_ctx* c, tc_graph* g, const tc_bw_model& bw, DevBuf& X) {
  const uint32_t n = g->n;
  TCX_TRY(worklists(c, g));
  TCX_TRY(X.reserve(4ull * F * ((size_t)n + 1)));
  hipLaunchKernelGGL(k_wl_features, dim3(grid_for(n)), dim3(256), 0, c->stream,
                     g->rowptr.as<uint32_t>(), g->sc.wl_rp.as<uint32_t>(),
                     g->sc.wl_col.as<uint32_t>(), n, bw_table(bw), X.as<float>());
  TCX_HIP(hipGetLastError());
  TCX_HIP(hipStreamSynchronize(c->stream));
  return TC_OK;
}

// Per-cell accumulators (kCells x kAcc) of the tasks of `level`.
int cell_accum(tc_ctx* c, tc_graph* g, const tc_bw_model& bw, uint32_t level,
               unsigned long long* out) {
  const uint32_t n = g->n;
  TCX_TRY(worklists(c, g));
  TCX_TRY(c->prim_out.reserve(8 * kCells * kAcc));
  unsigned long long* d = c->prim_out.as<unsigned long long>();
  TCX_HIP(hipMemsetAsync(d, 0, 8 * kCells * kAcc, c->stream));
  hipLaunchKernelGGL(k_cell_accum, dim3(std::min<unsigned>(grid_for(n), 4096)), dim3(256), 0,
                     c->stream, g->rowptr.as<uint32_t>(), g->sc.wl_rp.as<uint32_t>(),
                     g->sc.wl_col.as<uint32_t>(), n, level, bw_table(bw), d);
  TCX_HIP(hipGetLastError());
  TCX_HIP(hipMemcpyAsync(out, d, 8 * kCells * kAcc, hipMemcpyDeviceToHost, c->stream));
  TCX_HIP(hipStreamSynchronize(c->stream));
  return TC_OK;
}

// The eight cell features from the accumulators (0 for an empty cell):
// log2(1+tasks), log2(1+lists), log2(1+streamed entries), log2(1+held
// entries), log2(1+list-read ns under the bandwidth model), the fraction of
// lists shorter than 64, log2(1+largest task work), the fraction of tasks
// within the narrow variant's limit.
void cell_features(const unsigned long long* acc, float* X) {
  for (int c = 0; c < kCells; ++c) {
    const unsigned long long* a = acc + c * kAcc;
    float* f = X + c * F;
    if (a[0] == 0) {
      for (int j = 0; j < F; ++j) f[j] = 0.f;
      continue;
    }
    f[0] = (float)std::log2(1.0 + (double)a[0]);
    f[1] = (float)std::log2(1.0 + (double)a[1]);
    f[2] = (float)std::log2(1.0 + (double)a[2]);
    f[3] = (float)std::log2(1.0 + (double)a[3]);
    f[4] = (float)std::log2(1.0 + (double)a[4] / 1000.0);
    f[5] = a[1] ? (float)((double)a[5] / (double)a[1]) : 0.f;
    f[6] = (float)std::log2(1.0 + (double)a[6]);
    f[7] = (float)((double)a[7] / (double)a[0]);
  }
}

// Task lists of one schedule: block tasks by descending work (split at
// d+ 2048 between the two block kernels), wave and lane tasks in id order.
struct TaskLists {
  DevBuf blk, wav, lan, huge;
  uint64_t nb = 0, nw = 0, nl = 0, nh = 0;
};

int lists_from_var(tc_ctx* c, tc_graph* g, DevBuf& var, DevBuf& key, DevBuf& ids, TaskLists& L) {
  const uint32_t n = g->n;
  const uint32_t* rp = g->rowptr.as<uint32_t>();
  DevBuf ktmp, itmp, flag;
  TCX_TRY(ktmp.reserve(8 * ((size_t)n + 1)));
  TCX_TRY(itmp.reserve(4 * ((size_t)n + 1)));
  TCX_TRY(flag.reserve((size_t)n + 1));
  TCX_TRY(L.wav.reserve(4 * ((size_t)n + 1)));
  TCX_TRY(L.lan.reserve(4 * ((size_t)n + 1)));
  TCX_TRY(L.blk.reserve(4 * ((size_t)n + 1)));
  TCX_TRY(L.huge.reserve(4 * ((size_t)n + 1)));
  hipLaunchKernelGGL(k_var_flags, dim3(grid_for(n)), dim3(256), 0, c->stream, var.as<uint8_t>(),
                     (const uint32_t*)nullptr, rp, (uint64_t)n, (uint8_t)TC_VARIANT_WAVE, 0u,
                     0xFFFFFFFFu, flag.as<uint8_t>());
  TCX_TRY(select_indices(c, flag.as<uint8_t>(), n, L.wav.as<uint32_t>(), &L.nw));
  hipLaunchKernelGGL(k_var_flags, dim3(grid_for(n)), dim3(256), 0, c->stream, var.as<uint8_t>(),
                     (const uint32_t*)nullptr, rp, (uint64_t)n, (uint8_t)TC_VARIANT_LANE, 0u,
                     0xFFFFFFFFu, flag.as<uint8_t>());
  TCX_TRY(select_indices(c, flag.as<uint8_t>(), n, L.lan.as<uint32_t>(), &L.nl));
  uint64_t* sk;
  uint32_t* order;
  TCX_TRY(sort_pairs_u64_u32(c, key.as<uint64_t>(), ktmp.as<uint64_t>(), ids.as<uint32_t>(),
                             itmp.as<uint32_t>(), n, 64, &sk, &order));
  hipLaunchKernelGGL(k_var_flags, dim3(grid_for(n)), dim3(256), 0, c->stream, var.as<uint8_t>(),
                     order, rp, (uint64_t)n, (uint8_t)TC_VARIANT_BLOCK, 0u, 2048u,
                     flag.as<uint8_t>());
  TCX_TRY(select_flagged_u32(c, order, flag.as<uint8_t>(), L.blk.as<uint32_t>(), n, &L.nb));
  hipLaunchKernelGGL(k_var_flags, dim3(grid_for(n)), dim3(256), 0, c->stream, var.as<uint8_t>(),
                     order, rp, (uint64_t)n, (uint8_t)TC_VARIANT_BLOCK, 2049u, 0xFFFFFFFFu,
                     flag.as<uint8_t>());
  TCX_TRY(select_flagged_u32(c, order, flag.as<uint8_t>(), L.huge.as<uint32_t>(), n, &L.nh));
  TCX_HIP(hipGetLastError());
  TCX_HIP(hipStreamSynchronize(c->stream));
  return TC_OK;
}

// The count launches of a schedule (k_auto_wl as tc_count's AUTO launches
// it, plus the 8192-entry block kernel for d+ > 2048); cursors and the
// accumulator must be cleared first.
int launch_lists(tc_ctx* c, tc_graph* g, const uint32_t* blk, uint64_t nb, const uint32_t* wav,
                 uint64_t nw, const uint32_t* lan, uint64_t nl, const uint32_t* huge, uint64_t nh,
                 uint32_t* cursor, unsigned long long* acc, uint32_t* launches) {
  const uint32_t* rp = g->rowptr.as<uint32_t>();
  const uint32_t* col = g->col.as<uint32_t>();
  const uint32_t* wrp = g->sc.wl_rp.as<uint32_t>();
  const uint32_t* wcol = g->sc.wl_col.as<uint32_t>();
  const uint2* wse = g->sc.wl_se.as<uint2>();
  const uint4* ell = g->sc.wl_ell.as<uint4>();
  if (nb || nw || nl) {
    const uint32_t G = (uint32_t)c->num_cus * 8u;
    if (nw <= (uint64_t)kStaticWaves * G * (TB / 64))
      hipLaunchKernelGGL((k_auto_wl<2048, true>), dim3(G), dim3(TB), 0, c->stream, rp, col, wrp,
                         wcol, wse, ell, blk, (uint32_t)nb, wav, (uint32_t)nw, lan, (uint32_t)nl, cursor,
                         acc);
    else
      hipLaunchKernelGGL((k_auto_wl<2048>), dim3(G), dim3(TB), 0, c->stream, rp, col, wrp, wcol,
                         wse, ell, blk, (uint32_t)nb, wav, (uint32_t)nw, lan, (uint32_t)nl, cursor, acc);
    if (launches) ++*launches;
  }
  if (nh) {
    hipLaunchKernelGGL((k_auto_wl<8192>), dim3(std::min<uint64_t>((uint64_t)c->num_cus * 4, nh)),
                       dim3(TB), 0, c->stream, rp, col, wrp, wcol, wse, ell, huge, (uint32_t)nh,
                       (const uint32_t*)nullptr, 0u, (const uint32_t*)nullptr, 0u, cursor + 4,
                       acc);
    if (launches) ++*launches;
  }
  TCX_HIP(hipGetLastError());
  return TC_OK;
}

int launch_lists(tc_ctx* c, tc_graph* g, const TaskLists& L, uint32_t* cursor,
                 unsigned long long* acc, uint32_t* launches) {
  return launch_lists(c, g, L.blk.as<uint32_t>(), L.nb, L.wav.as<uint32_t>(), L.nw,
                      L.lan.as<uint32_t>(), L.nl, L.huge.as<uint32_t>(), L.nh, cursor, acc,
                      launches);
}

// Variant per task (TC_VARIANT_*, 255 none) for narrow_cells.
int task_var(tc_ctx* c, tc_graph* g, uint32_t narrow_cells, int only, uint32_t level,
             DevBuf& var, DevBuf& key, DevBuf& ids) {
  const uint32_t n = g->n;
  TCX_TRY(var.reserve((size_t)n + 1));
  TCX_TRY(key.reserve(8 * ((size_t)n + 1)));
  TCX_TRY(ids.reserve(4 * ((size_t)n + 1)));
  hipLaunchKernelGGL(k_task_var, dim3(grid_for(n)), dim3(256), 0, c->stream,
                     g->rowptr.as<uint32_t>(), g->sc.wl_rp.as<uint32_t>(),
                     g->sc.wl_col.as<uint32_t>(), n, narrow_cells, only, level, var.as<uint8_t>(),
                     key.as<uint64_t>(), ids.as<uint32_t>());
  TCX_HIP(hipGetLastError());
  return TC_OK;
}

// The model's cells: predictions Y (kCells x 2, log2 us of the cell's
// launch, wide / narrow) and the narrow-cell mask.  The fixed rule is the
// prior: a cell leaves it only for a predicted saving above kSwitchMargin.
// total_us (optional): the predicted launch time summed over the non-empty
// cells under the chosen variants, cell 15 included.
constexpr float kSwitchMargin = 0.10f;
// A tie rule other than the reference's is taken for a predicted saving
// above this fraction of the count.
constexpr double kTieMargin = 0.02;
int learned_cells(tc_ctx* c, tc_graph* g, tc_model* md, uint32_t* mask, float* Y,
                  double* total_us = nullptr);

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

namespace {

int learned_cells(tc_ctx* c, tc_graph* g, tc_model* md, uint32_t* mask, float* Y,
                  double* total_us) {
  unsigned long long acc[kCells * kAcc];
  TCX_TRY(cell_accum(c, g, md->bw, 0, acc));
  float X[kCells * F];
  cell_features(acc, X);
  DevBuf dX, dY;
  TCX_TRY(upload(c, dX, X, sizeof(X)));
  TCX_TRY(dY.reserve(4 * kCells * O));
  TCX_TRY(model_predict(c, md, dX.as<float>(), kCells, dY.as<float>()));
  TCX_HIP(hipMemcpyAsync(Y, dY.p, 4 * kCells * O, hipMemcpyDeviceToHost, c->stream));
  TCX_HIP(hipStreamSynchronize(c->stream));
  uint32_t m = kFixedNarrow;
  double total = 0;
  for (int cell = 2; cell < kChoiceCells; ++cell) {
    if (acc[cell * kAcc] == 0) continue;
    const bool prior = (kFixedNarrow >> cell) & 1u;
    const double wide = std::exp2((double)Y[cell * O + 0]);
    const double narrow = std::exp2((double)Y[cell * O + 1]);
    const double keep = prior ? narrow : wide, other = prior ? wide : narrow;
    const bool flip = other < (1.0 - kSwitchMargin) * keep;
    if (flip) m ^= 1u << cell;
    total += flip ? other : keep;
  }
  if (acc[kBlockCell * kAcc]) total += std::exp2((double)Y[kBlockCell * O + 0]);
  *mask = m;
  if (total_us) *total_us = total;
  return TC_OK;
}

// The learned task lists, kept with the graph per model parameters.
int learned_schedule(tc_ctx* c, tc_graph* g, tc_model* md) {
  SchedCache& S = g->sc;
  TCX_TRY(worklists(c, g));
  if (S.learned_version == g->sched_version && S.learned_model == md->id) return TC_OK;
  uint32_t mask = 0;
  float Y[kCells * O];
  TCX_TRY(learned_cells(c, g, md, &mask, Y));
  DevBuf var, key, ids;
  TCX_TRY(task_var(c, g, mask, -1, 0, var, key, ids));
  TaskLists L;
  TCX_TRY(lists_from_var(c, g, var, key, ids, L));
  swap(S.l_blk, L.blk);
  swap(S.l_wav, L.wav);
  swap(S.l_lan, L.lan);
  swap(S.l_huge, L.huge);
  S.nl_blk = L.nb;
  S.nl_wav = L.nw;
  S.nl_lan = L.nl;
  S.nl_huge = L.nh;
  S.learned_mask = mask;
  S.learned_version = g->sched_version;
  S.learned_model = md->id;
  return TC_OK;
}

}  // namespace

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
  bw_model_reference(&md->bw);
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

int tc_model_set_bw_model(tc_model* md, const tc_bw_model* bw) {
  if (!md || !bw) return fail(TC_ERR_ARG, "NULL argument");
  for (int l = 2; l < 10; ++l)
    if (!(std::isfinite(bw->table[l]) && bw->table[l] > 0))
      return fail(TC_ERR_ARG, "bandwidth model entry %d is not finite and positive", l);
  if (!(std::isfinite(bw->a) && bw->a > 0 && std::isfinite(bw->b)))
    return fail(TC_ERR_ARG, "bandwidth model power law is not finite and positive");
  md->bw = *bw;
  md->id = g_model_ids++;  // schedules built with the old features are stale
  return TC_OK;
}

int tc_graph_features(tc_ctx* c, tc_graph* g, const tc_model* md, float* X) {
  if (!c || !g || !X) return fail(TC_ERR_ARG, "NULL argument");
  if (g->ctx != c) return fail(TC_ERR_ARG, "graph belongs to another context");
  if (g->orient_mode < 0) return fail(TC_ERR_STATE, "graph must be oriented");
  TCX_HIP(hipSetDevice(c->device));
  if (g->n == 0) return TC_OK;
  tc_bw_model bw;
  if (md) bw = md->bw;
  else bw_model_reference(&bw);
  if (g->m_or == 0) {
    memset(X, 0, 4ull * F * g->n);
    return TC_OK;
  }
  DevBuf dX;
  TCX_TRY(graph_features(c, g, bw, dX));
  TCX_HIP(hipMemcpyAsync(X, dX.p, 4ull * F * g->n, hipMemcpyDeviceToHost, c->stream));
  TCX_HIP(hipStreamSynchronize(c->stream));
  return TC_OK;
}

int tc_graph_measure_costs(tc_ctx* c, tc_graph* g, int variant, float* cost_us, uint64_t* count) {
  if (!c || !g || !cost_us) return fail(TC_ERR_ARG, "NULL argument");
  if (g->ctx != c) return fail(TC_ERR_ARG, "graph belongs to another context");
  if (g->orient_mode < 0) return fail(TC_ERR_STATE, "graph must be oriented");
  if (variant != TC_VARIANT_BLOCK && variant != TC_VARIANT_WAVE && variant != TC_VARIANT_LANE)
    return fail(TC_ERR_ARG, "unknown variant %d", variant);
  TCX_HIP(hipSetDevice(c->device));
  const uint32_t n = g->n;
  if (n == 0) return TC_OK;
  if (count) *count = 0;
  if (g->m_or == 0) {
    memset(cost_us, 0, 4ull * n);
    return TC_OK;
  }
  TCX_TRY(worklists(c, g));
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
  const uint32_t* wrp = g->sc.wl_rp.as<uint32_t>();
  const uint32_t* wcol = g->sc.wl_col.as<uint32_t>();
  const uint2* wse = g->sc.wl_se.as<uint2>();
  uint64_t nv = 0;
  const uint32_t G = (uint32_t)c->num_cus * 8u;
  auto select = [&](uint32_t lo, uint32_t hi) -> int {
    hipLaunchKernelGGL(k_task_flags, dim3(grid_for(n)), dim3(256), 0, c->stream, rp, wrp, n, lo,
                       hi, flags.as<uint8_t>());
    TCX_HIP(hipGetLastError());
    return select_indices(c, flags.as<uint8_t>(), n, verts.as<uint32_t>(), &nv);
  };
  if (variant == TC_VARIANT_BLOCK) {
    TCX_TRY(select(2u, 2048u));
    if (nv)
      hipLaunchKernelGGL(k_time_wl_block<2048>, dim3(G), dim3(TB), 0, c->stream, rp, col, wrp, wse,
                         verts.as<uint32_t>(), (uint32_t)nv, cursor, cost.as<float>(), acc);
    TCX_TRY(select(2049u, 0xFFFFFFFFu));
    if (nv)
      hipLaunchKernelGGL(k_time_wl_block<8192>,
                         dim3(std::min<uint64_t>(nv, (uint64_t)c->num_cus * 4)), dim3(TB), 0,
                         c->stream, rp, col, wrp, wse, verts.as<uint32_t>(), (uint32_t)nv,
                         cursor + 1, cost.as<float>(), acc);
  } else if (variant == TC_VARIANT_WAVE) {
    TCX_TRY(select(2u, kWaveMax));
    if (nv)
      hipLaunchKernelGGL(k_time_wl_wave, dim3(G), dim3(TB), 0, c->stream, rp, col, wrp, wse,
                         verts.as<uint32_t>(), (uint32_t)nv, cursor, cost.as<float>(), acc);
  } else {
    TCX_TRY(select(2u, kLaneMaxDeg));
    if (nv)
      hipLaunchKernelGGL(k_time_wl_lane, dim3(G), dim3(TB), 0, c->stream,
                         g->sc.wl_ell.as<uint4>(), wrp, wcol,
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

namespace {
int check_graph(tc_ctx* c, tc_graph* g, tc_model* md) {
  if (!c || !g) return fail(TC_ERR_ARG, "NULL argument");
  if (g->ctx != c) return fail(TC_ERR_ARG, "graph belongs to another context");
  if (md && md->ctx != c) return fail(TC_ERR_ARG, "model belongs to another context");
  if (g->orient_mode < 0) return fail(TC_ERR_STATE, "graph must be oriented");
  TCX_HIP(hipSetDevice(c->device));
  return TC_OK;
}
}  // namespace

int tc_graph_cell_features(tc_ctx* c, tc_graph* g, const tc_model* md, uint32_t level,
                           float* X) {
  TCX_TRY(check_graph(c, g, nullptr));
  if (!X) return fail(TC_ERR_ARG, "NULL argument");
  if (md && md->ctx != c) return fail(TC_ERR_ARG, "model belongs to another context");
  if (level > TC_CELL_MAX_LEVEL) return fail(TC_ERR_ARG, "level %u > %d", level, TC_CELL_MAX_LEVEL);
  unsigned long long acc[kCells * kAcc] = {};
  if (g->n && g->m_or) {
    tc_bw_model bw;
    if (md) bw = md->bw;
    else bw_model_reference(&bw);
    TCX_TRY(cell_accum(c, g, bw, level, acc));
  }
  cell_features(acc, X);
  return TC_OK;
}

int tc_graph_measure_cells(tc_ctx* c, tc_graph* g, int narrow, uint32_t level, uint32_t reps,
                           float* us, uint64_t* counts) {
  TCX_TRY(check_graph(c, g, nullptr));
  if (!us) return fail(TC_ERR_ARG, "NULL argument");
  if (level > TC_CELL_MAX_LEVEL) return fail(TC_ERR_ARG, "level %u > %d", level, TC_CELL_MAX_LEVEL);
  if (reps == 0 || reps > 64) return fail(TC_ERR_ARG, "reps must be 1..64");
  for (int i = 0; i < kCells; ++i) {
    us[i] = 0.f;
    if (counts) counts[i] = 0;
  }
  if (g->n == 0 || g->m_or == 0) return TC_OK;
  unsigned long long acc[kCells * kAcc];
  tc_bw_model bw;
  bw_model_reference(&bw);
  TCX_TRY(cell_accum(c, g, bw, level, acc));
  TCX_TRY(c->scratch.reserve(64));
  TCX_TRY(c->accum.reserve(64));
  uint32_t* cursor = c->scratch.as<uint32_t>();
  unsigned long long* dacc = c->accum.as<unsigned long long>();
  DevBuf var, key, ids;
  for (int cell = 2; cell < kCells; ++cell) {
    if (acc[cell * kAcc] == 0 || (narrow && cell >= kChoiceCells)) continue;
    TCX_TRY(task_var(c, g, narrow ? (1u << cell) : 0u, cell, level, var, key, ids));
    TaskLists L;
    TCX_TRY(lists_from_var(c, g, var, key, ids, L));
    float best = 0.f;
    unsigned long long first = 0;
    for (uint32_t r = 0; r < reps; ++r) {
      TCX_HIP(hipMemsetAsync(cursor, 0, 8 * sizeof(uint32_t), c->stream));
      TCX_HIP(hipMemsetAsync(dacc, 0, 8, c->stream));
      TCX_HIP(hipEventRecord(c->ev0, c->stream));
      TCX_TRY(launch_lists(c, g, L, cursor, dacc, nullptr));
      TCX_HIP(hipEventRecord(c->ev1, c->stream));
      TCX_HIP(hipEventSynchronize(c->ev1));
      float ms = 0.f;
      TCX_HIP(hipEventElapsedTime(&ms, c->ev0, c->ev1));
      unsigned long long h = 0;
      TCX_HIP(hipMemcpyAsync(&h, dacc, 8, hipMemcpyDeviceToHost, c->stream));
      TCX_HIP(hipStreamSynchronize(c->stream));
      if (r == 0) first = h;
      else if (h != first) return fail(TC_ERR_STATE, "cell %d counted %llu then %llu", cell, first, h);
      if (r == 0 || ms < best) best = ms;
    }
    us[cell] = best * 1000.f;
    if (counts) counts[cell] = first;
  }
  return TC_OK;
}

int tc_learned_cells(tc_ctx* c, tc_graph* g, tc_model* md, uint32_t* narrow_cells, float* Y) {
  TCX_TRY(check_graph(c, g, md));
  if (!narrow_cells) return fail(TC_ERR_ARG, "NULL argument");
  float y[kCells * O] = {};
  *narrow_cells = kFixedNarrow;
  if (md && g->n && g->m_or) TCX_TRY(learned_cells(c, g, md, narrow_cells, y));
  if (Y) memcpy(Y, y, sizeof(y));
  return TC_OK;
}

int tc_graph_orient_learned(tc_ctx* c, tc_graph* g, int mode, tc_model* md, int* chosen,
                            double* predicted_us, double* adir_cost) {
  if (!c || !g) return fail(TC_ERR_ARG, "NULL argument");
  if (g->ctx != c) return fail(TC_ERR_ARG, "graph belongs to another context");
  if (md && md->ctx != c) return fail(TC_ERR_ARG, "model belongs to another context");
  if (mode != TC_ORIENT_ADIR && mode != TC_ORIENT_DEGREE)
    return fail(TC_ERR_ARG, "learned ties apply to TC_ORIENT_ADIR and TC_ORIENT_DEGREE");
  double pred[TC_TIE_CANDIDATES] = {}, cost[TC_TIE_CANDIDATES] = {};
  for (int t = 0; t < TC_TIE_CANDIDATES; ++t) {
    TCX_TRY(tc_graph_orient_ties(c, g, mode, t));
    TCX_TRY(tc_graph_adir_cost(c, g, &cost[t]));
    if (md && g->n && g->m_or) {
      uint32_t mask = 0;
      float Y[kCells * O];
      TCX_TRY(worklists(c, g));
      TCX_TRY(learned_cells(c, g, md, &mask, Y, &pred[t]));
    }
  }
  // the model's prediction picks, with the reference's rule as the prior;
  // without a model the A-direction cost function does
  int best = 0;
  for (int t = 1; t < TC_TIE_CANDIDATES; ++t) {
    if (md) {
      if (pred[t] < pred[best]) best = t;
    } else if (cost[t] < cost[best]) {
      best = t;
    }
  }
  if (md && !(pred[best] < (1.0 - kTieMargin) * pred[0])) best = 0;
  if (best != TC_TIE_CANDIDATES - 1) TCX_TRY(tc_graph_orient_ties(c, g, mode, best));
  if (chosen) *chosen = best;
  for (int t = 0; t < TC_TIE_CANDIDATES; ++t) {
    if (predicted_us) predicted_us[t] = pred[t];
    if (adir_cost) adir_cost[t] = cost[t];
  }
  return TC_OK;
}

int tc_learned_assignment(tc_ctx* c, tc_graph* g, tc_model* md, uint8_t* variant) {
  TCX_TRY(check_graph(c, g, md));
  if (!variant) return fail(TC_ERR_ARG, "NULL argument");
  const uint32_t n = g->n;
  if (n == 0) return TC_OK;
  memset(variant, 255, n);
  if (g->m_or == 0) return TC_OK;
  TCX_TRY(worklists(c, g));
  if (!md) {  // the fixed rule: the work-list schedule's own lists
    std::vector<uint32_t> ids;
    const SchedCache& L = g->sc;
    auto mark = [&](const DevBuf& list, uint64_t cnt, uint8_t v) -> int {
      ids.resize(std::max<uint64_t>(cnt, 1));
      if (cnt)
        TCX_HIP(hipMemcpyAsync(ids.data(), list.p, 4 * cnt, hipMemcpyDeviceToHost, c->stream));
      TCX_HIP(hipStreamSynchronize(c->stream));
      for (uint64_t i = 0; i < cnt; ++i) variant[ids[i]] = v;
      return TC_OK;
    };
    TCX_TRY(mark(L.wl_large, L.nw_large, TC_VARIANT_BLOCK));
    TCX_TRY(mark(L.wl_huge, L.nw_huge, TC_VARIANT_BLOCK));
    TCX_TRY(mark(L.wl_small, L.nw_small, TC_VARIANT_WAVE));
    TCX_TRY(mark(L.wl_lane, L.nw_lane, TC_VARIANT_LANE));
    return TC_OK;
  }
  uint32_t mask = 0;
  float Y[kCells * O];
  TCX_TRY(learned_cells(c, g, md, &mask, Y));
  DevBuf var, key, ids;
  TCX_TRY(task_var(c, g, mask, -1, 0, var, key, ids));
  TCX_HIP(hipMemcpyAsync(variant, var.p, n, hipMemcpyDeviceToHost, c->stream));
  TCX_HIP(hipStreamSynchronize(c->stream));
  return TC_OK;
}

int tc_count_learned(tc_ctx* c, tc_graph* g, tc_model* md, uint64_t* count, tc_timing* timing) {
  TCX_TRY(check_graph(c, g, md));
  if (!count) return fail(TC_ERR_ARG, "NULL argument");
  if (!md) {  // the fixed rule: AUTO's work-list schedule
    tc_sched s;
    memset(&s, 0, sizeof(s));
    s.kernel = TC_KERNEL_AUTO;
    return tc_count(c, g, &s, count, timing);
  }
  tc_timing tm;
  memset(&tm, 0, sizeof(tm));
  *count = 0;
  if (g->m_or == 0 || g->n == 0) {
    if (timing) *timing = tm;
    return TC_OK;
  }
  const SchedCache& S = g->sc;
  // the schedule is built once per (graph, model) and timed then; a count
  // that reuses it enqueues nothing before the kernel and waits once
  if (S.learned_version != g->sched_version || S.learned_model != md->id) {
    TCX_HIP(hipEventRecord(c->ev0, c->stream));
    TCX_TRY(learned_schedule(c, g, md));
    TCX_HIP(hipEventRecord(c->ev1, c->stream));
    TCX_HIP(hipEventSynchronize(c->ev1));
    TCX_HIP(hipEventElapsedTime(&tm.ms_prep, c->ev0, c->ev1));
  } else {
    TCX_TRY(learned_schedule(c, g, md));
  }
  // claim cursors (8 words) and the count (8 bytes) share one buffer and
  // one clear
  TCX_TRY(c->scratch.reserve(64));
  uint32_t* cursor = c->scratch.as<uint32_t>();
  unsigned long long* acc = reinterpret_cast<unsigned long long*>(cursor + 8);
  TCX_HIP(hipMemsetAsync(cursor, 0, 8 * sizeof(uint32_t) + 8, c->stream));
  TCX_HIP(hipEventRecord(c->ev0, c->stream));
  TCX_TRY(launch_lists(c, g, S.l_blk.as<uint32_t>(), S.nl_blk, S.l_wav.as<uint32_t>(), S.nl_wav,
                       S.l_lan.as<uint32_t>(), S.nl_lan, S.l_huge.as<uint32_t>(), S.nl_huge,
                       cursor, acc, &tm.launches));
  TCX_HIP(hipEventRecord(c->ev1, c->stream));
  unsigned long long h = 0;
  TCX_HIP(hipMemcpyAsync(&h, acc, 8, hipMemcpyDeviceToHost, c->stream));
  TCX_HIP(hipStreamSynchronize(c->stream));
  TCX_HIP(hipEventElapsedTime(&tm.ms_count, c->ev0, c->ev1));
  tm.ms_dominant = tm.ms_count;
  *count = h;
  if (timing) *timing = tm;
  return TC_OK;
}

}  // extern "C"
