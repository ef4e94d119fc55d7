// Triangle-counting kernels for gfx950 (wave64) and the tc_count dispatcher.
//
// Every kernel computes the forward count over the oriented CSR,
//     T = sum_u sum_{v in N+(u)} |N+(u) & N+(v)|,
// with uint64 per-lane partials, a wave reduction and one atomic per wave.
// Integer sums are order-free, so any schedule gives the same exact count.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "engine.h"
#include "internal.h"

namespace tcx {
namespace {

using namespace dev;

// largest e in [0, E) with pre[e] <= t  (pre ascending, pre[0] = 0 <= t)
__device__ inline uint32_t owner_of(const uint32_t* pre, uint32_t E, uint32_t t) {
  uint32_t lo = 0, hi = E - 1;
  while (lo < hi) {
    uint32_t mid = (lo + hi + 1) >> 1;
    if (pre[mid] <= t) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// ---- work models -----------------------------------------------------------

// Hu/tc.cu:22-35: W(u) = (sum_{v in N+(u)} d+(v)) * (unsigned)log2(d+(u)).
// floor(log2(d)) for d >= 1 is 31 - clz(d) exactly.
__global__ void k_hu_work(const uint32_t* __restrict__ rp, const uint32_t* __restrict__ col,
                          uint32_t n, int64_t* __restrict__ w) {
  for (uint64_t u = gtid(); u < n; u += gstride()) {
    uint32_t b = rp[u], e = rp[u + 1];
    int64_t tot = 0;
    for (uint32_t k = b; k < e; ++k) {
      uint32_t v = col[k];
      tot += rp[v + 1] - rp[v];
    }
    uint32_t d = e - b;
    w[u] = d == 0 ? 0 : tot * (int64_t)(31 - __clz(d));
  }
}

// Hu/tc.cu:36-67 binarySearchValue (direction -1) over an inclusive prefix.
__device__ inline uint32_t hu_bsv(const int64_t* a, int64_t value, uint32_t len) {
  int64_t s = 0, e = (int64_t)len - 1, right = 0;
  bool found = false;
  int64_t mid = (s + e) / 2;
  while (s <= e) {
    int64_t x = a[mid];
    if (x == value) { right = mid; found = true; break; }
    if (x < value) s = mid + 1;
    else { if (e == 0) break; e = mid - 1; }
    mid = (s + e) / 2;
  }
  if (!found) return (uint32_t)s;
  while (right - 1 >= 0 && a[right - 1] == value) --right;
  return (uint32_t)right;
}

__global__ void k_block_starts(const int64_t* __restrict__ pre, uint32_t n, int64_t step,
                               uint32_t nblocks, int32_t* __restrict__ start) {
  for (uint64_t i = gtid(); i <= nblocks; i += gstride()) {
    if (i == 0) start[0] = 0;
    else if (i == nblocks) start[nblocks] = (int32_t)n;
    else start[i] = (int32_t)hu_bsv(pre, (int64_t)i * step + 1, n);
  }
}

// AUTO engine per-vertex work: wedges sum_{v in N+(u)} d+(v) plus d+(u).
__global__ void k_auto_work(const uint32_t* __restrict__ rp, const uint32_t* __restrict__ col,
                            uint32_t n, uint64_t* __restrict__ key) {
  for (uint64_t u = gtid(); u < n; u += gstride()) {
    uint32_t b = rp[u], e = rp[u + 1];
    uint64_t tot = e - b;
    for (uint32_t k = b; k < e; ++k) {
      uint32_t v = col[k];
      tot += rp[v + 1] - rp[v];
    }
    key[u] = ~tot;  // ascending sort of ~work = descending work
  }
}

// Work-list key of every CSR edge u->v: the task that streams the shorter
// out-list takes it ((u << 32 | v) when d+(v) <= d+(u), else (v << 32 | u)).
// Edges that close no triangle become the sentinel (n << 32): v has no
// out-edge, or the owner u has only v as out-neighbour.
__global__ void k_wl_keys(const uint32_t* __restrict__ rp, const uint32_t* __restrict__ col,
                          const uint32_t* __restrict__ esrc, uint64_t m, uint32_t n,
                          uint64_t* __restrict__ key) {
  for (uint64_t e = gtid(); e < m; e += gstride()) {
    const uint32_t u = esrc[e], v = col[e];
    const uint32_t du = rp[u + 1] - rp[u], dv = rp[v + 1] - rp[v];
    uint64_t k;
    if (dv == 0 || (dv <= du && du < 2)) k = (uint64_t)n << 32;
    else if (dv <= du) k = ((uint64_t)u << 32) | v;
    else k = ((uint64_t)v << 32) | u;
    key[e] = k;
  }
}

// Per task x of the work-list schedule: work = d+(x) + sum of d+(y) over
// W(x) (as ~work, for a descending sort), its class (0 none: W(x) empty or
// d+(x) < 2; 1 wave task; 2 block task; 3 block task of the 8192-entry
// kernel; 4 lane task) and the bytes it streams (task id, row bounds of x in
// both CSRs, N+(x), and per y: its row bounds, stored with the entry, and
// N+(y)).  Lane tasks read fixed 32-byte rows instead; the figure is the
// same for every split of the tasks.
__global__ void k_wl_work(const uint32_t* __restrict__ rp, const uint32_t* __restrict__ wrp,
                          const uint32_t* __restrict__ wcol, uint32_t n, uint32_t wmax,
                          uint32_t hmin, uint32_t lmax, uint64_t wave_work,
                          uint64_t* __restrict__ key, uint8_t* __restrict__ cls,
                          unsigned long long* bytes) {
  unsigned long long b = 0;
  for (uint64_t x = gtid(); x < n; x += gstride()) {
    const uint32_t dx = rp[x + 1] - rp[x];
    const uint32_t wb = wrp[x], we = wrp[x + 1];
    uint64_t work = dx, streamed = 0;
    for (uint32_t i = wb; i < we; ++i) {
      const uint32_t y = wcol[i];
      streamed += rp[y + 1] - rp[y];
    }
    work += streamed;
    uint8_t k = 0;
    if (we > wb && dx >= 2) {
      if (dx <= lmax && we - wb <= dev::kLaneMaxList) k = 4;
      else if (dx > hmin) k = 3;
      else if (dx > wmax || work > wave_work) k = 2;
      else k = 1;
      b += 20ull + 4ull * dx + 8ull * (we - wb) + 4ull * streamed;
    }
    key[x] = ~work;
    cls[x] = k;
  }
  b = dev::wave_sum(b);
  if ((threadIdx.x & 63) == 0 && b) atomicAdd(bytes, b);
}

// rearrangeEdge key (Fox/tc.cu:133-139): the longer of the two lists.
__global__ void k_fox_rkey(const uint32_t* __restrict__ rp, const uint32_t* __restrict__ col,
                           const uint32_t* __restrict__ esrc, const uint32_t* __restrict__ eids,
                           uint64_t S, uint32_t* __restrict__ key) {
  for (uint64_t i = gtid(); i < S; i += gstride()) {
    const uint32_t e = eids[i], u = esrc[e], v = col[e];
    const uint32_t a = rp[u + 1] - rp[u], b = rp[v + 1] - rp[v];
    key[i] = a > b ? a : b;
  }
}

// Deals the S sorted edges in groups of G round-robin over B workgroup
// slots: group g goes to slot g % B as its (g / B)-th group, and slot b's
// groups are laid out contiguously, slots in order (Fox/tc.cu:164-174 meant
// this interleave; here it is a permutation for any S).  Only the last group
// can be short, and it is the last of its slot.
__global__ void k_fox_deal(const uint32_t* __restrict__ sorted, uint64_t S, uint64_t G,
                           uint64_t B, uint32_t* __restrict__ out) {
  const uint64_t ng = (S + G - 1) / G, last = ng - 1, short_by = ng * G - S;
  for (uint64_t p = gtid(); p < S; p += gstride()) {
    const uint64_t g = p / G, r = p % G, b = g % B, t = g / B;
    uint64_t before = G * (b * (ng / B) + (b < ng % B ? b : ng % B));  // groups of slots < b
    if (last % B < b) before -= short_by;
    out[before + t * G + r] = sorted[p];
  }
}

__global__ void k_low32(const uint64_t* __restrict__ keys, uint64_t m, uint32_t* __restrict__ out) {
  for (uint64_t i = gtid(); i < m; i += gstride()) out[i] = (uint32_t)keys[i];
}

// Row x of the lane tasks' layout: N+(x) in 8 words, kEmpty after its
// end; rows of longer lists are left all kEmpty (no lane task reads them:
// a lane task and every list it streams hold at most eight entries).
__global__ void k_ell8(const uint32_t* __restrict__ rp, const uint32_t* __restrict__ col,
                       uint32_t n, uint4* __restrict__ ell) {
  for (uint64_t x = gtid(); x < n; x += gstride()) {
    const uint32_t b = rp[x], d = rp[x + 1] - b;
    uint32_t r[8];
#pragma unroll
    for (uint32_t k = 0; k < 8; ++k) r[k] = (d <= 8 && k < d) ? col[b + k] : dev::kEmpty;
    ell[2 * x] = make_uint4(r[0], r[1], r[2], r[3]);
    ell[2 * x + 1] = make_uint4(r[4], r[5], r[6], r[7]);
  }
}

// se[i] = the bounds of N+(y) for the work-list entry y = wl[i].
__global__ void k_wl_bounds(const uint32_t* __restrict__ rp, const uint32_t* __restrict__ wl,
                            uint64_t m, uint2* __restrict__ se) {
  for (uint64_t i = gtid(); i < m; i += gstride()) {
    const uint32_t y = wl[i];
    se[i] = make_uint2(rp[y], rp[y + 1]);
  }
}

__global__ void k_eq_flags(const uint8_t* __restrict__ in, uint64_t n, uint8_t v,
                           uint8_t* __restrict__ out) {
  for (uint64_t i = gtid(); i < n; i += gstride()) out[i] = in[i] == v ? 1 : 0;
}

__global__ void k_gather_u8(const uint8_t* __restrict__ in, const uint32_t* __restrict__ idx,
                            uint32_t n, uint8_t* __restrict__ out) {
  for (uint64_t i = gtid(); i < n; i += gstride()) out[i] = in[idx[i]];
}

__global__ void k_iota(uint32_t* __restrict__ p, uint32_t n) {
  for (uint64_t i = gtid(); i < n; i += gstride()) p[i] = (uint32_t)i;
}

// Fox/tc.cu:18-33: W(u,v) = int(min(d+u, d+v) * log2(max + 2)).
__global__ void k_fox_work(const uint32_t* __restrict__ rp, const uint32_t* __restrict__ col,
                           const uint32_t* __restrict__ esrc, uint64_t m, int32_t* __restrict__ w) {
  for (uint64_t i = gtid(); i < m; i += gstride()) {
    uint32_t u = esrc[i], v = col[i];
    int a = (int)(rp[u + 1] - rp[u]), b = (int)(rp[v + 1] - rp[v]);
    int large = a > b ? a : b;
    w[i] = (int)((a + b - large) * log2((double)large + 2));
  }
}

// Bin of each edge (number of thresholds <= w) plus a block-level histogram,
// so the global counters see one atomic per block and bin.
__global__ void k_fox_bin(const int32_t* __restrict__ w, uint64_t m, const int32_t* thr, int nthr,
                          uint32_t* __restrict__ bin, unsigned long long* counts) {
  __shared__ unsigned long long s_cnt[8];
  if (threadIdx.x < 8) s_cnt[threadIdx.x] = 0;
  __syncthreads();
  for (uint64_t i = gtid(); i < m; i += gstride()) {
    int32_t x = w[i];
    int j = 0;
    while (j < nthr && x >= thr[j]) ++j;
    bin[i] = (uint32_t)j;
    atomicAdd(s_cnt + j, 1ull);
  }
  __syncthreads();
  if (threadIdx.x < 8 && s_cnt[threadIdx.x]) atomicAdd(counts + threadIdx.x, s_cnt[threadIdx.x]);
}

__global__ void k_row_flags(const uint32_t* __restrict__ rp, uint32_t n, uint32_t min_d,
                            uint32_t max_d, uint8_t* __restrict__ f) {
  for (uint64_t u = gtid(); u < n; u += gstride()) {
    uint32_t d = rp[u + 1] - rp[u];
    f[u] = (d >= min_d && d <= max_d) ? 1 : 0;
  }
}

// ---- Hu: workgroup per vertex range, N+(u) slices staged in LDS ----------
//
// Hu/tc.cu:69-201 re-designed for wave64: the range's adjacency slice is
// copied into LDS by all lanes (coalesced), the wedges u->v->w of the slice
// are flattened with an LDS prefix sum so consecutive lanes read consecutive
// w from HBM, and w is binary-searched in the LDS copy of N+(u).  A vertex
// whose list exceeds the slice searches N+(u) in global memory instead
// (the reference's cacheWorked == false path).
template <int SLICE, int ROWCAP>
struct HuLds {
  uint32_t col[SLICE];
  uint32_t vst[SLICE];
  uint32_t pre[SLICE + 1];
  uint16_t row[SLICE];
  uint32_t rof[ROWCAP + 1];
  uint32_t wsum[TB / 64];
  uint32_t u2;
};

// Stage edges [ebase, ebase+E) into LDS: v, start of N+(v), wedge prefix.
template <int SLICE, int ROWCAP>
__device__ inline uint32_t stage_edges(HuLds<SLICE, ROWCAP>& L, const uint32_t* __restrict__ rp,
                                       const uint32_t* __restrict__ col, uint32_t ebase, uint32_t E,
                                       bool by_row = false) {
  const uint32_t tid = threadIdx.x;
  for (uint32_t e = tid; e < E; e += TB) {
    uint32_t v = col[ebase + e];
    uint32_t vs = rp[v];
    L.col[e] = v;
    L.vst[e] = vs;
    uint32_t c = rp[v + 1] - vs;
    if (by_row) {
      const uint32_t r = L.row[e];
      if (L.rof[r + 1] - L.rof[r] < 2) c = 0;
    }
    L.pre[e] = c;
  }
  __syncthreads();
  // block exclusive scan, blocked layout: thread t owns [t*K, t*K+K)
  constexpr int K = SLICE / TB;
  uint32_t loc[K];
  uint32_t sum = 0;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    uint32_t e = tid * K + k;
    uint32_t x = e < E ? L.pre[e] : 0;
    loc[k] = sum;
    sum += x;
  }
  // inclusive wave scan of per-thread sums
  const uint32_t lane = tid & 63, wv = tid >> 6;
  uint32_t inc = sum;
  for (int off = 1; off < 64; off <<= 1) {
    uint32_t o = __shfl_up(inc, off, 64);
    if (lane >= (uint32_t)off) inc += o;
  }
  if (lane == 63) L.wsum[wv] = inc;
  __syncthreads();
  uint32_t wbase = 0;
  for (uint32_t i = 0; i < wv; ++i) wbase += L.wsum[i];
  uint32_t tbase = wbase + inc - sum;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    uint32_t e = tid * K + k;
    if (e < E) L.pre[e] = tbase + loc[k];
  }
  uint32_t total = 0;
  for (uint32_t i = 0; i < TB / 64; ++i) total += L.wsum[i];
  if (tid == 0) L.pre[E] = total;
  __syncthreads();
  return total;
}

template <int SLICE, int ROWCAP>
__global__ __launch_bounds__(TB) void k_hu(const uint32_t* __restrict__ rp,
                                           const uint32_t* __restrict__ col,
                                           const int32_t* __restrict__ start, int skip_hubs,
                                           unsigned long long* total) {
  __shared__ HuLds<SLICE, ROWCAP> L;
  const uint32_t tid = threadIdx.x;
  const uint32_t ub = (uint32_t)start[blockIdx.x], ue = (uint32_t)start[blockIdx.x + 1];
  unsigned long long cnt = 0;
  uint32_t u = ub;
  while (u < ue) {
    if (tid == 0) {
      uint32_t base = rp[u];
      if (rp[u + 1] - base > SLICE) {
        L.u2 = u;  // hub: its list does not fit the slice
      } else {
        uint32_t lo = u + 1, hi = min(ue, u + ROWCAP);
        while (lo < hi) {
          uint32_t mid = (lo + hi + 1) >> 1;
          if (rp[mid] - base <= SLICE) lo = mid;
          else hi = mid - 1;
        }
        L.u2 = lo;
      }
    }
    __syncthreads();
    const uint32_t u2 = L.u2;
    if (u2 == u && skip_hubs) {
      // scheduled by the hub kernel
      __syncthreads();
      u = u + 1;
      continue;
    }
    if (u2 == u) {
      // hub path: N+(u) stays in HBM/L2
      const uint32_t ab = rp[u], du = rp[u + 1] - ab;
      for (uint32_t k0 = 0; k0 < du; k0 += SLICE) {
        uint32_t E = min((uint32_t)SLICE, du - k0);
        uint32_t T = stage_edges(L, rp, col, ab + k0, E);
        for (uint32_t t = tid; t < T; t += TB) {
          uint32_t e = owner_of(L.pre, E, t);
          uint32_t w = col[L.vst[e] + (t - L.pre[e])];
          cnt += contains(col, ab, ab + du, w);
        }
        __syncthreads();
      }
      u = u + 1;
      continue;
    }
    const uint32_t base = rp[u];
    const uint32_t E = rp[u2] - base;
    const uint32_t R = u2 - u;
    for (uint32_t r = tid; r <= R; r += TB) L.rof[r] = rp[u + r] - base;
    __syncthreads();
    for (uint32_t e = tid; e < E; e += TB) {
      uint32_t lo = 0, hi = R - 1;  // row r with rof[r] <= e < rof[r+1]
      while (lo < hi) {
        uint32_t mid = (lo + hi + 1) >> 1;
        if (L.rof[mid] <= e) lo = mid;
        else hi = mid - 1;
      }
      L.row[e] = (uint16_t)lo;
    }
    // stage_edges begins with a barrier-free pass that writes col/vst/pre,
    // then synchronises; rows with d+(u) < 2 get no wedges.
    uint32_t T = stage_edges(L, rp, col, base, E, true);
    for (uint32_t t = tid; t < T; t += TB) {
      uint32_t e = owner_of(L.pre, E, t);
      uint32_t w = col[L.vst[e] + (t - L.pre[e])];
      uint32_t r = L.row[e];
      cnt += contains(L.col, L.rof[r], L.rof[r + 1], w);
    }
    __syncthreads();
    u = u2;
  }
  flush_count(cnt, total);
}

// ---- Bisson: wave per row, bitmap of N+(u) for long rows -----------------
//
// Bisson/tc.cu:18-104 re-designed: rows with d+ > min_deg mark N+(u) in a
// per-workgroup global bitmap (L2-resident) plus a coarse LDS filter, probe
// the flattened wedges, then clear only the words they set (the reference
// memsets the whole n-bit map per row).  Short rows keep N+(u) in LDS and
// binary-search it.
constexpr int kCoarseWords = 2048;  // 64K-bit LDS filter

__global__ __launch_bounds__(64) void k_bisson(const uint32_t* __restrict__ rp,
                                               const uint32_t* __restrict__ col,
                                               const uint32_t* __restrict__ rows, uint64_t nrows,
                                               uint32_t* __restrict__ bitmaps, uint64_t words,
                                               uint32_t min_deg, uint32_t cshift,
                                               unsigned long long* total) {
  __shared__ uint32_t s_coarse[kCoarseWords];
  __shared__ uint32_t s_list[64];
  __shared__ uint32_t s_vs[64];
  __shared__ uint32_t s_pre[65];
  const uint32_t lane = threadIdx.x;
  uint32_t* bm = bitmaps + (uint64_t)blockIdx.x * words;
  for (uint32_t i = lane; i < kCoarseWords; i += 64) s_coarse[i] = 0;
  __syncthreads();
  unsigned long long cnt = 0;
  for (uint64_t ri = blockIdx.x; ri < nrows; ri += gridDim.x) {
    const uint32_t u = rows[ri];
    const uint32_t base = rp[u], du = rp[u + 1] - base;
    const bool big = du > min_deg;
    if (big) {
      for (uint32_t e = lane; e < du; e += 64) {
        uint32_t w = col[base + e];
        atomicOr(bm + (w >> 5), 1u << (w & 31));
        uint32_t cw = w >> cshift;
        atomicOr(s_coarse + (cw >> 5), 1u << (cw & 31));
      }
    } else {
      s_list[lane] = lane < du ? col[base + lane] : 0xFFFFFFFFu;
    }
    __syncthreads();
    for (uint32_t k0 = 0; k0 < du; k0 += 64) {
      const uint32_t E = min(64u, du - k0);
      uint32_t c = 0;
      if (lane < E) {
        uint32_t v = col[base + k0 + lane];
        uint32_t vs = rp[v];
        s_vs[lane] = vs;
        c = rp[v + 1] - vs;
      }
      uint32_t inc = c;
      for (int off = 1; off < 64; off <<= 1) {
        uint32_t o = __shfl_up(inc, off, 64);
        if (lane >= (uint32_t)off) inc += o;
      }
      s_pre[lane] = inc - c;
      if (lane == 63) s_pre[64] = inc;
      __syncthreads();
      const uint32_t T = s_pre[64];
      for (uint32_t t = lane; t < T; t += 64) {
        uint32_t e = owner_of(s_pre, E, t);
        uint32_t w = col[s_vs[e] + (t - s_pre[e])];
        if (big) {
          uint32_t cw = w >> cshift;
          if (s_coarse[cw >> 5] & (1u << (cw & 31))) {
            uint32_t word = __hip_atomic_load(bm + (w >> 5), __ATOMIC_RELAXED,
                                              __HIP_MEMORY_SCOPE_AGENT);
            cnt += (word >> (w & 31)) & 1u;
          }
        } else {
          cnt += contains(s_list, 0, du, w);
        }
      }
      __syncthreads();
    }
    if (big) {
      for (uint32_t e = lane; e < du; e += 64) {
        uint32_t w = col[base + e];
        bm[w >> 5] = 0u;
        s_coarse[(w >> cshift) >> 5] = 0u;
      }
      __syncthreads();
    }
  }
  flush_count(cnt, total);
}

// ---- Fox: lanes-per-edge bins --------------------------------------------
// Fox/tc.cu:64-125: each group of L lanes takes one edge and binary-searches
// its strided share of the shorter list in the longer one.
template <int L>
__global__ __launch_bounds__(TB) void k_fox(const uint32_t* __restrict__ rp,
                                            const uint32_t* __restrict__ col,
                                            const uint32_t* __restrict__ esrc,
                                            const uint32_t* __restrict__ eids, uint64_t count,
                                            unsigned long long* total) {
  const uint64_t g = gtid() / L, ng = gstride() / L;
  const uint32_t lane = threadIdx.x % L;
  unsigned long long cnt = 0;
  for (uint64_t i = g; i < count; i += ng) {
    uint32_t e = eids[i];
    uint32_t u = esrc[e], v = col[e];
    uint32_t a = rp[u], da = rp[u + 1] - a, b = rp[v], db = rp[v + 1] - b;
    if (da < db) {
      uint32_t t = a; a = b; b = t;
      t = da; da = db; db = t;
    }
    for (uint32_t j = lane; j < db; j += L) cnt += contains(col, a, a + da, col[b + j]);
  }
  flush_count(cnt, total);
}

// ---- TriCore: wave per edge, 64 sampled pivots of the longer list ---------
// tricore/src/tricount_gpu.cu:215-290 with 64 samples (one per lane); the
// pivots live in LDS and each lane does the two-phase search for its
// strided elements of the shorter list.  Optional node-range filter on the
// second endpoint implements the 2-D partition passes (:337-406).
__global__ __launch_bounds__(64) void k_tricore(const uint32_t* __restrict__ rp,
                                                const uint32_t* __restrict__ col,
                                                const uint32_t* __restrict__ esrc, uint64_t e_lo,
                                                uint64_t e_hi, uint32_t v_lo, uint32_t v_hi,
                                                unsigned long long* total) {
  __shared__ uint32_t s_samp[64];
  const uint32_t lane = threadIdx.x;
  unsigned long long cnt = 0;
  for (uint64_t e = e_lo + blockIdx.x; e < e_hi; e += gridDim.x) {
    uint32_t u = esrc[e], v = col[e];
    if (v < v_lo || v >= v_hi) continue;
    uint32_t a = rp[u], da = rp[u + 1] - a, b = rp[v], db = rp[v + 1] - b;
    if (da < db) {
      uint32_t t = a; a = b; b = t;
      t = da; da = db; db = t;
    }
    if (db == 0) continue;
    s_samp[lane] = col[a + (uint32_t)(((uint64_t)lane * da) >> 6)];
    __syncthreads();
    for (uint32_t j = lane; j < db; j += 64) {
      uint32_t x = col[b + j];
      if (s_samp[0] > x) continue;
      uint32_t lo = 0, hi = 63;
      while (lo < hi) {
        uint32_t mid = (lo + hi + 1) >> 1;
        if (s_samp[mid] <= x) lo = mid;
        else hi = mid - 1;
      }
      uint32_t seg_lo = (uint32_t)(((uint64_t)lo * da) >> 6);
      uint32_t seg_hi = lo == 63 ? da : (uint32_t)(((uint64_t)(lo + 1) * da) >> 6);
      cnt += contains(col + a, seg_lo, seg_hi, x);
    }
    __syncthreads();
  }
  flush_count(cnt, total);
}

}  // namespace

// ---- schedule builders ------------------------------------------------------


static int ensure_version(tc_graph* g) {
  SchedCache& sc = g->sc;
  if (sc.version != g->sched_version) {
    sc.hu_nblocks = 0;
    sc.auto_ok = false;
    sc.wl_ok = false;
    sc.rows_ok = false;
    sc.fox_lanes = 0;
    sc.version = g->sched_version;
  }
  return TC_OK;
}

static unsigned grid_for(uint64_t n, int tb = TB) {
  uint64_t b = (n + tb - 1) / tb;
  if (b < 1) b = 1;
  if (b > (1u << 20)) b = 1u << 20;
  return (unsigned)b;
}

static int auto_schedule(tc_ctx* c, tc_graph* g, uint32_t wmax, uint32_t hmin);

// tri[u] = the triangles whose first vertex is u: the AUTO kernel's block
// and wave tasks (default thresholds) with one output per vertex.
int per_vertex_counts(tc_ctx* c, tc_graph* g, unsigned long long* tri) {
  if (g->n == 0) return TC_OK;
  TCX_TRY(auto_schedule(c, g, 64u, 2048u));
  const SchedCache& sc = g->sc;
  const uint32_t* rp = g->rowptr.as<uint32_t>();
  const uint32_t* col = g->col.as<uint32_t>();
  DevBuf cur;
  TCX_TRY(cur.reserve(4 * sizeof(uint32_t)));
  uint32_t* cursor = cur.as<uint32_t>();
  TCX_HIP(hipMemsetAsync(cursor, 0, 4 * sizeof(uint32_t), c->stream));
  TCX_HIP(hipMemsetAsync(tri, 0, 8 * (size_t)g->n, c->stream));
  if (sc.n_large || sc.n_small)
    hipLaunchKernelGGL(k_auto_per_vertex<2048>, dim3((uint32_t)c->num_cus * 8u), dim3(TB), 0,
                       c->stream, rp, col, sc.auto_large.as<uint32_t>(), (uint32_t)sc.n_large,
                       sc.auto_small.as<uint32_t>(), (uint32_t)sc.n_small, cursor, tri);
  if (sc.n_huge)
    hipLaunchKernelGGL(k_auto_per_vertex<8192>,
                       dim3(std::min<uint64_t>((uint64_t)c->num_cus * 4, sc.n_huge)), dim3(TB), 0,
                       c->stream, rp, col, sc.auto_huge.as<uint32_t>(), (uint32_t)sc.n_huge,
                       (const uint32_t*)nullptr, 0u, cursor + 2, tri);
  TCX_HIP(hipGetLastError());
  TCX_HIP(hipStreamSynchronize(c->stream));  // before `cur` is freed
  return TC_OK;
}

// start[] over a per-vertex work array with the Hu partition rule.
static int block_starts(tc_ctx* c, tc_graph* g, DevBuf& w, uint32_t nblocks, DevBuf& start) {
  const uint32_t n = g->n;
  DevBuf pre;
  TCX_TRY(pre.reserve(8 * ((size_t)n + 1)));
  TCX_TRY(inclusive_scan_i64(c, w.as<int64_t>(), pre.as<int64_t>(), n));
  int64_t last = 0;
  if (n) {
    TCX_HIP(hipMemcpyAsync(&last, pre.as<int64_t>() + n - 1, 8, hipMemcpyDeviceToHost, c->stream));
    TCX_HIP(hipStreamSynchronize(c->stream));
  }
  int64_t step = (last + nblocks - 1) / nblocks;  // Hu/tc.cu:294
  TCX_TRY(start.reserve(4 * ((size_t)nblocks + 1)));
  hipLaunchKernelGGL(k_block_starts, dim3(grid_for(nblocks + 1ull)), dim3(TB), 0, c->stream,
                     pre.as<int64_t>(), n, step, nblocks, start.as<int32_t>());
  TCX_HIP(hipGetLastError());
  TCX_HIP(hipStreamSynchronize(c->stream));
  return TC_OK;
}

static int hu_schedule(tc_ctx* c, tc_graph* g, uint32_t nblocks) {
  SchedCache& sc = g->sc;
  if (sc.hu_nblocks == nblocks) return TC_OK;
  DevBuf w;
  TCX_TRY(w.reserve(8 * ((size_t)g->n + 1)));
  hipLaunchKernelGGL(k_hu_work, dim3(grid_for(g->n)), dim3(TB), 0, c->stream,
                     g->rowptr.as<uint32_t>(), g->col.as<uint32_t>(), g->n, w.as<int64_t>());
  TCX_HIP(hipGetLastError());
  TCX_TRY(block_starts(c, g, w, nblocks, sc.hu_start));
  sc.hu_nblocks = nblocks;
  return TC_OK;
}

static int rows_schedule(tc_ctx* c, tc_graph* g) {
  SchedCache& sc = g->sc;
  if (sc.rows_ok) return TC_OK;
  DevBuf flags;
  TCX_TRY(flags.reserve((size_t)g->n + 1));
  TCX_TRY(sc.rows.reserve(4 * ((size_t)g->n + 1)));
  hipLaunchKernelGGL(k_row_flags, dim3(grid_for(g->n)), dim3(TB), 0, c->stream,
                     g->rowptr.as<uint32_t>(), g->n, 2u, 0xFFFFFFFFu, flags.as<uint8_t>());
  TCX_HIP(hipGetLastError());
  TCX_TRY(select_indices(c, flags.as<uint8_t>(), g->n, sc.rows.as<uint32_t>(), &sc.nrows));
  sc.rows_ok = true;
  return TC_OK;
}

// Fox thresholds, Fox/tc.cu:316-331, generalised to max_lanes = 64.
static void fox_thresholds(int max_work, uint32_t max_lanes, std::vector<int32_t>& thr) {
  thr.clear();
  int workPerThreadB = 8;
  double maxPowerOf8 = std::log(max_work) / std::log(workPerThreadB);
  int prev = INT32_MIN;
  for (uint32_t i = 2; i <= max_lanes; i *= 2) {
    double index = (double)(maxPowerOf8 - 1) * (i - 1) / (double)max_lanes + 1;
    int p = powl(workPerThreadB, index);
    if (p < prev) p = prev;  // keep bins ordered where the reference's would overlap
    thr.push_back(p);
    prev = p;
  }
}

static int fox_schedule(tc_ctx* c, tc_graph* g, uint32_t max_lanes, bool rearrange) {
  SchedCache& sc = g->sc;
  if (sc.fox_lanes == max_lanes && sc.fox_rearr == rearrange) return TC_OK;
  const uint64_t m = g->m_or;
  DevBuf w, bin, btmp, ids, itmp;
  TCX_TRY(w.reserve(4 * std::max<uint64_t>(m, 1)));
  TCX_TRY(bin.reserve(4 * std::max<uint64_t>(m, 1)));
  TCX_TRY(btmp.reserve(4 * std::max<uint64_t>(m, 1)));
  TCX_TRY(ids.reserve(4 * std::max<uint64_t>(m, 1)));
  TCX_TRY(itmp.reserve(4 * std::max<uint64_t>(m, 1)));
  hipLaunchKernelGGL(k_fox_work, dim3(grid_for(m)), dim3(TB), 0, c->stream,
                     g->rowptr.as<uint32_t>(), g->col.as<uint32_t>(), g->esrc.as<uint32_t>(), m,
                     w.as<int32_t>());
  TCX_HIP(hipGetLastError());
  uint32_t mx = 0;
  TCX_TRY(max_u32(c, w.as<uint32_t>(), m, &mx));
  std::vector<int32_t> thr;
  fox_thresholds((int)mx, max_lanes, thr);
  int nthr = (int)thr.size();
  const int nb = nthr + 1;
  TCX_TRY(c->scratch.reserve(4096));
  int32_t* d_thr = c->scratch.as<int32_t>();
  unsigned long long* d_cnt = reinterpret_cast<unsigned long long*>(c->scratch.as<char>() + 256);
  TCX_HIP(hipMemcpyAsync(d_thr, thr.data(), 4 * nthr, hipMemcpyHostToDevice, c->stream));
  TCX_HIP(hipMemsetAsync(d_cnt, 0, 8 * 8, c->stream));
  hipLaunchKernelGGL(k_fox_bin, dim3(std::min<unsigned>(grid_for(m), 4096)), dim3(TB), 0, c->stream,
                     w.as<int32_t>(), m, d_thr, nthr, bin.as<uint32_t>(), d_cnt);
  hipLaunchKernelGGL(k_iota, dim3(grid_for(m)), dim3(TB), 0, c->stream, ids.as<uint32_t>(),
                     (uint32_t)m);
  TCX_HIP(hipGetLastError());
  std::vector<unsigned long long> cnt(nb);
  TCX_HIP(hipMemcpyAsync(cnt.data(), d_cnt, 8 * nb, hipMemcpyDeviceToHost, c->stream));
  TCX_HIP(hipStreamSynchronize(c->stream));
  sc.fox_off.assign(nb + 1, 0);
  for (int j = 0; j < nb; ++j) sc.fox_off[j + 1] = sc.fox_off[j] + cnt[j];
  // stable grouping of edge ids by bin (keys have at most 3 bits)
  uint32_t *sk, *sv;
  TCX_TRY(sort_pairs_u32_u32(c, bin.as<uint32_t>(), btmp.as<uint32_t>(), ids.as<uint32_t>(),
                             itmp.as<uint32_t>(), m, 3, &sk, &sv));
  TCX_TRY(sc.fox_eids.reserve(4 * std::max<uint64_t>(m, 1)));
  if (!rearrange) {
    TCX_HIP(hipMemcpyAsync(sc.fox_eids.p, sv, 4 * m, hipMemcpyDeviceToDevice, c->stream));
  } else {
    // rearrangeEdge per bin: sort by max(d+(u), d+(v)), then deal groups of
    // up to 32 edges round-robin over the bin launch's workgroups
    uint32_t* ids_in = sv == ids.as<uint32_t>() ? ids.as<uint32_t>() : itmp.as<uint32_t>();
    uint32_t* spare = ids_in == ids.as<uint32_t>() ? itmp.as<uint32_t>() : ids.as<uint32_t>();
    for (int j = 0; j < nb; ++j) {
      const uint64_t lo = sc.fox_off[j], S = sc.fox_off[j + 1] - lo;
      if (!S) continue;
      hipLaunchKernelGGL(k_fox_rkey, dim3(grid_for(S)), dim3(TB), 0, c->stream,
                         g->rowptr.as<uint32_t>(), g->col.as<uint32_t>(), g->esrc.as<uint32_t>(),
                         ids_in + lo, S, bin.as<uint32_t>());
      TCX_HIP(hipGetLastError());
      uint32_t *rk, *rv;
      TCX_TRY(sort_pairs_u32_u32(c, bin.as<uint32_t>(), btmp.as<uint32_t>(), ids_in + lo,
                                 spare + lo, S, 32, &rk, &rv));
      const uint32_t lanes = 1u << j;
      const uint64_t E = TB / lanes;  // edges per workgroup of k_fox<lanes>
      const uint64_t G = std::min<uint64_t>(32, E);
      const uint64_t B = std::max<uint64_t>(1, (S + E - 1) / E);
      hipLaunchKernelGGL(k_fox_deal, dim3(grid_for(S)), dim3(TB), 0, c->stream, rv, S, G, B,
                         sc.fox_eids.as<uint32_t>() + lo);
      TCX_HIP(hipGetLastError());
    }
  }
  TCX_HIP(hipStreamSynchronize(c->stream));
  sc.fox_lanes = max_lanes;
  sc.fox_rearr = rearrange;
  return TC_OK;
}

static int gather_flags(tc_ctx* c, const uint8_t* in, const uint32_t* idx, uint32_t n,
                        uint8_t* out) {
  hipLaunchKernelGGL(k_gather_u8, dim3(grid_for(n)), dim3(TB), 0, c->stream, in, idx, n, out);
  TCX_HIP(hipGetLastError());
  return TC_OK;
}

static int auto_schedule(tc_ctx* c, tc_graph* g, uint32_t wmax, uint32_t hmin) {
  SchedCache& sc = g->sc;
  if (sc.auto_ok && sc.auto_wmax == wmax && sc.auto_hmin == hmin) return TC_OK;
  const uint32_t n = g->n;
  const uint32_t* rp = g->rowptr.as<uint32_t>();
  DevBuf flags, key, ktmp, ids, itmp;
  TCX_TRY(flags.reserve((size_t)n + 1));
  TCX_TRY(sc.auto_small.reserve(4 * ((size_t)n + 1)));
  hipLaunchKernelGGL(k_row_flags, dim3(grid_for(n)), dim3(TB), 0, c->stream, rp, n, 2u, wmax,
                     flags.as<uint8_t>());
  TCX_HIP(hipGetLastError());
  TCX_TRY(select_indices(c, flags.as<uint8_t>(), n, sc.auto_small.as<uint32_t>(), &sc.n_small));
  // large and huge vertices sorted by descending work
  TCX_TRY(key.reserve(8 * ((size_t)n + 1)));
  TCX_TRY(ktmp.reserve(8 * ((size_t)n + 1)));
  TCX_TRY(ids.reserve(4 * ((size_t)n + 1)));
  TCX_TRY(itmp.reserve(4 * ((size_t)n + 1)));
  hipLaunchKernelGGL(k_auto_work, dim3(grid_for(n)), dim3(TB), 0, c->stream, rp,
                     g->col.as<uint32_t>(), n, key.as<uint64_t>());
  hipLaunchKernelGGL(k_iota, dim3(grid_for(n)), dim3(TB), 0, c->stream, ids.as<uint32_t>(), n);
  TCX_HIP(hipGetLastError());
  uint64_t* sk;
  uint32_t* sv;
  TCX_TRY(sort_pairs_u64_u32(c, key.as<uint64_t>(), ktmp.as<uint64_t>(), ids.as<uint32_t>(),
                             itmp.as<uint32_t>(), n, 64, &sk, &sv));
  // sv = vertices by descending work; keep the two degree classes in order
  DevBuf dflag;
  TCX_TRY(dflag.reserve((size_t)n + 1));
  TCX_TRY(sc.auto_large.reserve(4 * ((size_t)n + 1)));
  TCX_TRY(sc.auto_huge.reserve(4 * ((size_t)n + 1)));
  hipLaunchKernelGGL(k_row_flags, dim3(grid_for(n)), dim3(TB), 0, c->stream, rp, n,
                     std::max(wmax + 1, 2u), hmin, flags.as<uint8_t>());
  TCX_HIP(hipGetLastError());
  // flags are indexed by vertex; gather them into sorted order
  TCX_TRY(gather_flags(c, flags.as<uint8_t>(), sv, n, dflag.as<uint8_t>()));
  TCX_TRY(select_flagged_u32(c, sv, dflag.as<uint8_t>(), sc.auto_large.as<uint32_t>(), n,
                             &sc.n_large));
  hipLaunchKernelGGL(k_row_flags, dim3(grid_for(n)), dim3(TB), 0, c->stream, rp, n,
                     std::max(hmin + 1, 2u), 0xFFFFFFFFu, flags.as<uint8_t>());
  TCX_HIP(hipGetLastError());
  TCX_TRY(gather_flags(c, flags.as<uint8_t>(), sv, n, dflag.as<uint8_t>()));
  TCX_TRY(select_flagged_u32(c, sv, dflag.as<uint8_t>(), sc.auto_huge.as<uint32_t>(), n,
                             &sc.n_huge));
  sc.auto_ok = true;
  sc.auto_wmax = wmax;
  sc.auto_hmin = hmin;
  return TC_OK;
}


static int wl_schedule(tc_ctx* c, tc_graph* g, uint32_t wmax, uint32_t hmin, uint32_t lmax) {
  SchedCache& sc = g->sc;
  if (sc.wl_ok && sc.wl_wmax == wmax && sc.wl_hmin == hmin && sc.wl_lmax == lmax) return TC_OK;
  const uint32_t n = g->n;
  const uint64_t m = g->m_or;
  const uint32_t* rp = g->rowptr.as<uint32_t>();
  {
    DevBuf keys, tmp;
    TCX_TRY(keys.reserve(8 * std::max<uint64_t>(m, 1)));
    TCX_TRY(tmp.reserve(8 * std::max<uint64_t>(m, 1)));
    hipLaunchKernelGGL(k_wl_keys, dim3(grid_for(m)), dim3(TB), 0, c->stream, rp,
                       g->col.as<uint32_t>(), g->esrc.as<uint32_t>(), m, n, keys.as<uint64_t>());
    TCX_HIP(hipGetLastError());
    uint64_t* sorted;
    TCX_TRY(sort_u64(c, keys.as<uint64_t>(), tmp.as<uint64_t>(), m, 32 + bit_width_u32(n),
                     &sorted));
    // the sentinels sort last; the kept entries are those below row n
    TCX_TRY(sc.wl_rp.reserve(4 * ((size_t)n + 1)));
    hipLaunchKernelGGL(dev::k_rowptr_search<uint64_t>, dim3(grid_for((uint64_t)n + 1)), dim3(TB), 0,
                       c->stream, sorted, m, n, sc.wl_rp.as<uint32_t>());
    TCX_HIP(hipGetLastError());
    uint32_t kept = 0;
    TCX_HIP(hipMemcpyAsync(&kept, sc.wl_rp.as<uint32_t>() + n, 4, hipMemcpyDeviceToHost, c->stream));
    TCX_HIP(hipStreamSynchronize(c->stream));
    sc.wl_entries = kept;
    TCX_TRY(sc.wl_col.reserve(4 * std::max<uint64_t>(kept, 1)));
    TCX_TRY(sc.wl_se.reserve(8 * std::max<uint64_t>(kept, 1)));
    if (kept) {
      hipLaunchKernelGGL(k_low32, dim3(grid_for(kept)), dim3(TB), 0, c->stream, sorted,
                         (uint64_t)kept, sc.wl_col.as<uint32_t>());
      hipLaunchKernelGGL(k_wl_bounds, dim3(grid_for(kept)), dim3(TB), 0, c->stream, rp,
                         sc.wl_col.as<uint32_t>(), (uint64_t)kept, sc.wl_se.as<uint2>());
    }
    TCX_HIP(hipGetLastError());
  }
  DevBuf key, ktmp, ids, itmp, cls, flags, dflag;
  TCX_TRY(key.reserve(8 * ((size_t)n + 1)));
  TCX_TRY(ktmp.reserve(8 * ((size_t)n + 1)));
  TCX_TRY(ids.reserve(4 * ((size_t)n + 1)));
  TCX_TRY(itmp.reserve(4 * ((size_t)n + 1)));
  TCX_TRY(cls.reserve((size_t)n + 1));
  TCX_TRY(flags.reserve((size_t)n + 1));
  TCX_TRY(dflag.reserve((size_t)n + 1));
  TCX_TRY(c->accum.reserve(64));
  unsigned long long* acc = c->accum.as<unsigned long long>() + 4;
  TCX_HIP(hipMemsetAsync(acc, 0, 8, c->stream));
  hipLaunchKernelGGL(k_wl_work, dim3(grid_for(n)), dim3(TB), 0, c->stream, rp,
                     sc.wl_rp.as<uint32_t>(), sc.wl_col.as<uint32_t>(), n, wmax, hmin, lmax,
                     kWlWaveWork, key.as<uint64_t>(), cls.as<uint8_t>(), acc);
  hipLaunchKernelGGL(k_iota, dim3(grid_for(n)), dim3(TB), 0, c->stream, ids.as<uint32_t>(), n);
  hipLaunchKernelGGL(k_eq_flags, dim3(grid_for(n)), dim3(TB), 0, c->stream, cls.as<uint8_t>(),
                     (uint64_t)n, (uint8_t)1, flags.as<uint8_t>());
  TCX_HIP(hipGetLastError());
  TCX_TRY(sc.wl_small.reserve(4 * ((size_t)n + 1)));
  TCX_TRY(select_indices(c, flags.as<uint8_t>(), n, sc.wl_small.as<uint32_t>(), &sc.nw_small));
  hipLaunchKernelGGL(k_eq_flags, dim3(grid_for(n)), dim3(TB), 0, c->stream, cls.as<uint8_t>(),
                     (uint64_t)n, (uint8_t)4, flags.as<uint8_t>());
  TCX_HIP(hipGetLastError());
  TCX_TRY(sc.wl_lane.reserve(4 * ((size_t)n + 1)));
  TCX_TRY(select_indices(c, flags.as<uint8_t>(), n, sc.wl_lane.as<uint32_t>(), &sc.nw_lane));
  // the lane tasks' 32-byte rows (n of them; a learned schedule may run any
  // task of d+ <= 8 as a lane)
  TCX_TRY(sc.wl_ell.reserve(32 * ((size_t)n + 1)));
  hipLaunchKernelGGL(k_ell8, dim3(grid_for(n)), dim3(TB), 0, c->stream, rp,
                     g->col.as<uint32_t>(), n, sc.wl_ell.as<uint4>());
  TCX_HIP(hipGetLastError());
  uint64_t* sk;
  uint32_t* sv;
  TCX_TRY(sort_pairs_u64_u32(c, key.as<uint64_t>(), ktmp.as<uint64_t>(), ids.as<uint32_t>(),
                             itmp.as<uint32_t>(), n, 64, &sk, &sv));
  TCX_TRY(gather_flags(c, cls.as<uint8_t>(), sv, n, dflag.as<uint8_t>()));
  TCX_TRY(sc.wl_large.reserve(4 * ((size_t)n + 1)));
  TCX_TRY(sc.wl_huge.reserve(4 * ((size_t)n + 1)));
  hipLaunchKernelGGL(k_eq_flags, dim3(grid_for(n)), dim3(TB), 0, c->stream, dflag.as<uint8_t>(),
                     (uint64_t)n, (uint8_t)2, flags.as<uint8_t>());
  TCX_HIP(hipGetLastError());
  TCX_TRY(select_flagged_u32(c, sv, flags.as<uint8_t>(), sc.wl_large.as<uint32_t>(), n,
                             &sc.nw_large));
  hipLaunchKernelGGL(k_eq_flags, dim3(grid_for(n)), dim3(TB), 0, c->stream, dflag.as<uint8_t>(),
                     (uint64_t)n, (uint8_t)3, flags.as<uint8_t>());
  TCX_HIP(hipGetLastError());
  TCX_TRY(select_flagged_u32(c, sv, flags.as<uint8_t>(), sc.wl_huge.as<uint32_t>(), n,
                             &sc.nw_huge));
  unsigned long long bytes = 0;
  TCX_HIP(hipMemcpyAsync(&bytes, acc, 8, hipMemcpyDeviceToHost, c->stream));
  TCX_HIP(hipStreamSynchronize(c->stream));
  sc.wl_bytes = (double)bytes;
  sc.wl_ok = true;
  sc.wl_wmax = wmax;
  sc.wl_hmin = hmin;
  sc.wl_lmax = lmax;
  return TC_OK;
}

int worklists(tc_ctx* c, tc_graph* g) {
  ensure_version(g);
  if (g->m_or == 0 || g->n == 0) return TC_OK;
  return wl_schedule(c, g, 64u, 2048u, kLaneMaxDeg);
}

}  // namespace tcx

using namespace tcx;

namespace {

struct Timer {
  tc_ctx* c;
  float ms = 0;
  explicit Timer(tc_ctx* cc) : c(cc) {}
  int begin() {
    TCX_HIP(hipEventRecord(c->ev0, c->stream));
    return TC_OK;
  }
  int end() {
    TCX_HIP(hipEventRecord(c->ev1, c->stream));
    TCX_HIP(hipEventSynchronize(c->ev1));
    TCX_HIP(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    return TC_OK;
  }
};

// Event marks between consecutive count launches: mark() before each launch,
// finish() after the last; gives the total and the longest launch.
struct LaunchMarks {
  tc_ctx* c;
  int n = 0;
  explicit LaunchMarks(tc_ctx* cc) : c(cc) {}
  int mark() {
    if (n >= 16) return TC_OK;  // later launches fold into the last interval
    if (!c->lev[n]) TCX_HIP(hipEventCreate(&c->lev[n]));
    TCX_HIP(hipEventRecord(c->lev[n], c->stream));
    ++n;
    return TC_OK;
  }
  int finish(tc_timing* tm) {
    if (!c->lev[n]) TCX_HIP(hipEventCreate(&c->lev[n]));
    TCX_HIP(hipEventRecord(c->lev[n], c->stream));
    TCX_HIP(hipEventSynchronize(c->lev[n]));
    tm->ms_count = 0;
    tm->ms_dominant = 0;
    for (int i = 0; i < n; ++i) {
      float ms = 0;
      TCX_HIP(hipEventElapsedTime(&ms, c->lev[i], c->lev[i + 1]));
      tm->ms_count += ms;
      if (ms > tm->ms_dominant) {
        tm->ms_dominant = ms;
        tm->dominant_kernel = (uint32_t)i;
      }
    }
    return TC_OK;
  }
};

int check_ready(tc_ctx* c, tc_graph* g) {
  if (!c || !g) return fail(TC_ERR_ARG, "NULL argument");
  if (g->ctx != c) return fail(TC_ERR_ARG, "graph belongs to another context");
  if (g->orient_mode < 0) return fail(TC_ERR_STATE, "graph must be oriented before counting");
  TCX_HIP(hipSetDevice(c->device));
  return TC_OK;
}

}  // namespace

extern "C" {

int tc_count(tc_ctx* c, tc_graph* g, const tc_sched* sched_in, uint64_t* count,
             tc_timing* timing) {
  TCX_TRY(check_ready(c, g));
  if (!count) return fail(TC_ERR_ARG, "NULL count");
  tc_sched s;
  memset(&s, 0, sizeof(s));
  if (sched_in) s = *sched_in;
  else s.kernel = TC_KERNEL_AUTO;
  if (s.kernel < TC_KERNEL_HU || s.kernel > TC_KERNEL_AUTO)
    return fail(TC_ERR_ARG, "unknown kernel %d", s.kernel);
  ensure_version(g);
  TCX_TRY(c->accum.reserve(64));
  unsigned long long* acc = c->accum.as<unsigned long long>();
  tc_timing tm;
  memset(&tm, 0, sizeof(tm));
  *count = 0;
  if (g->m_or == 0 || g->n == 0) {
    if (timing) *timing = tm;
    return TC_OK;
  }
  const uint32_t* rp = g->rowptr.as<uint32_t>();
  const uint32_t* col = g->col.as<uint32_t>();
  Timer prep(c);
  LaunchMarks lm(c);
  TCX_TRY(prep.begin());
  // schedule (cached per graph)
  switch (s.kernel) {
    case TC_KERNEL_HU:
      TCX_TRY(hu_schedule(c, g, s.grid_size ? s.grid_size : 30000u));  // Hu/tc.cu:7
      break;
    case TC_KERNEL_BISSON:
      TCX_TRY(rows_schedule(c, g));
      break;
    case TC_KERNEL_FOX: {
      uint32_t L = s.fox_max_lanes ? s.fox_max_lanes : 64;
      if (L != 1 && L != 2 && L != 4 && L != 8 && L != 16 && L != 32 && L != 64)
        return fail(TC_ERR_ARG, "fox_max_lanes must be a power of two <= 64");
      if (s.fox_rearrange > 1) return fail(TC_ERR_ARG, "fox_rearrange must be 0 or 1");
      TCX_TRY(fox_schedule(c, g, L, s.fox_rearrange == 1));
      break;
    }
    case TC_KERNEL_AUTO: {
      const uint32_t wmax = s.auto_thresholds[0] ? s.auto_thresholds[0] : 64u;
      const uint32_t hmin = s.auto_thresholds[1] ? s.auto_thresholds[1] : 2048u;
      if (wmax > kWaveMax) return fail(TC_ERR_ARG, "auto_thresholds[0] must be <= %u", kWaveMax);
      if (hmin < wmax) return fail(TC_ERR_ARG, "auto_thresholds[1] must be >= auto_thresholds[0]");
      if (s.auto_thresholds[2] > TC_AUTO_PUSH)
        return fail(TC_ERR_ARG, "auto_thresholds[2] must be TC_AUTO_WORKLIST or TC_AUTO_PUSH");
      if (s.auto_thresholds[3] > kLaneMaxDeg)
        return fail(TC_ERR_ARG, "auto_thresholds[3] must be <= %u", kLaneMaxDeg);
      const uint32_t lmax = s.auto_thresholds[3] ? s.auto_thresholds[3] : kLaneMaxDeg;
      if (s.auto_thresholds[2] == TC_AUTO_PUSH) TCX_TRY(auto_schedule(c, g, wmax, hmin));
      else TCX_TRY(wl_schedule(c, g, wmax, hmin, lmax));
      break;
    }
    default:
      break;
  }
  TCX_TRY(prep.end());
  TCX_HIP(hipMemsetAsync(acc, 0, sizeof(unsigned long long), c->stream));
  switch (s.kernel) {
    case TC_KERNEL_HU: {
      TCX_TRY(lm.mark());
      hipLaunchKernelGGL((k_hu<2048, 1024>), dim3(g->sc.hu_nblocks), dim3(TB), 0, c->stream, rp,
                         col, g->sc.hu_start.as<int32_t>(), 0, acc);
      tm.launches = 1;
      break;
    }
    case TC_KERNEL_BISSON: {
      uint32_t G = s.grid_size ? s.grid_size : (uint32_t)c->num_cus * 8u;
      uint32_t min_deg = s.bitmap_min_deg ? s.bitmap_min_deg : 64;
      if (min_deg > 64) min_deg = 64;  // the short-row path holds at most one list per wave
      uint64_t words = ((uint64_t)g->n + 31) / 32;
      uint64_t need = words * G;
      if (c->bitmap_words < need) {
        TCX_TRY(c->bitmaps.reserve(4 * need));
        TCX_HIP(hipMemsetAsync(c->bitmaps.p, 0, 4 * need, c->stream));
        c->bitmap_words = need;
      }
      uint32_t cshift = 0;
      while (((uint64_t)(g->n - 1) >> cshift) >= (uint64_t)kCoarseWords * 32) ++cshift;
      TCX_TRY(lm.mark());
      hipLaunchKernelGGL(k_bisson, dim3(G), dim3(64), 0, c->stream, rp, col,
                         g->sc.rows.as<uint32_t>(), g->sc.nrows, c->bitmaps.as<uint32_t>(), words,
                         min_deg, cshift, acc);
      tm.launches = 1;
      break;
    }
    case TC_KERNEL_FOX: {
      const auto& off = g->sc.fox_off;
      const uint32_t* eids = g->sc.fox_eids.as<uint32_t>();
      const uint32_t* es = g->esrc.as<uint32_t>();
      for (size_t j = 0; j + 1 < off.size(); ++j) {
        uint64_t cnt = off[j + 1] - off[j];
        if (!cnt) continue;
        const uint32_t* ids = eids + off[j];
        uint32_t lanes = 1u << j;
        unsigned grid = (unsigned)std::min<uint64_t>(20000, (cnt * lanes + TB - 1) / TB);
        TCX_TRY(lm.mark());
        switch (lanes) {
          case 1: hipLaunchKernelGGL(k_fox<1>, dim3(grid), dim3(TB), 0, c->stream, rp, col, es, ids, cnt, acc); break;
          case 2: hipLaunchKernelGGL(k_fox<2>, dim3(grid), dim3(TB), 0, c->stream, rp, col, es, ids, cnt, acc); break;
          case 4: hipLaunchKernelGGL(k_fox<4>, dim3(grid), dim3(TB), 0, c->stream, rp, col, es, ids, cnt, acc); break;
          case 8: hipLaunchKernelGGL(k_fox<8>, dim3(grid), dim3(TB), 0, c->stream, rp, col, es, ids, cnt, acc); break;
          case 16: hipLaunchKernelGGL(k_fox<16>, dim3(grid), dim3(TB), 0, c->stream, rp, col, es, ids, cnt, acc); break;
          case 32: hipLaunchKernelGGL(k_fox<32>, dim3(grid), dim3(TB), 0, c->stream, rp, col, es, ids, cnt, acc); break;
          default: hipLaunchKernelGGL(k_fox<64>, dim3(grid), dim3(TB), 0, c->stream, rp, col, es, ids, cnt, acc); break;
        }
        tm.launches++;
      }
      break;
    }
    case TC_KERNEL_TRICORE: {
      uint32_t parts = s.tricore_parts ? s.tricore_parts : 1;
      uint32_t G = s.grid_size ? s.grid_size : (uint32_t)c->num_cus * 32u;
      const uint32_t* es = g->esrc.as<uint32_t>();
      if (parts > g->n) parts = g->n;
      // node_split as in tricore/src/main.cpp:78-86
      std::vector<uint32_t> split(parts + 1);
      for (uint32_t i = 0; i <= parts; ++i)
        split[i] = i == parts ? g->n : (uint32_t)((uint64_t)(g->n / parts) * i);
      std::vector<uint32_t> ebeg(parts + 1);
      for (uint32_t i = 0; i <= parts; ++i)
        TCX_HIP(hipMemcpyAsync(&ebeg[i], rp + split[i], 4, hipMemcpyDeviceToHost, c->stream));
      TCX_HIP(hipStreamSynchronize(c->stream));
      for (uint32_t pm = 0; pm < parts; ++pm) {
        for (uint32_t pn = pm; pn < parts; ++pn) {
          TCX_TRY(lm.mark());
          hipLaunchKernelGGL(k_tricore, dim3(G), dim3(64), 0, c->stream, rp, col, es,
                             (uint64_t)ebeg[pm], (uint64_t)ebeg[pm + 1], split[pn], split[pn + 1], acc);
          tm.launches++;
          if (pm != pn) {
            TCX_TRY(lm.mark());
            hipLaunchKernelGGL(k_tricore, dim3(G), dim3(64), 0, c->stream, rp, col, es,
                               (uint64_t)ebeg[pn], (uint64_t)ebeg[pn + 1], split[pm], split[pm + 1],
                               acc);
            tm.launches++;
          }
        }
      }
      break;
    }
    case TC_KERNEL_AUTO: {
      const SchedCache& sc = g->sc;
      TCX_TRY(c->scratch.reserve(64));
      uint32_t* cursor = c->scratch.as<uint32_t>();
      TCX_HIP(hipMemsetAsync(cursor, 0, 8 * sizeof(uint32_t), c->stream));
      const uint32_t G = s.grid_size ? s.grid_size : (uint32_t)c->num_cus * 8u;
      if (s.auto_thresholds[2] != TC_AUTO_PUSH) {
        // cursors: [0] block, [1] wave, [2] lane tasks; [4..6] the 8192-entry launch
        const uint32_t* wrp = sc.wl_rp.as<uint32_t>();
        const uint32_t* wcol = sc.wl_col.as<uint32_t>();
        const uint2* wse = sc.wl_se.as<uint2>();
        const uint4* ell = sc.wl_ell.as<uint4>();
        if (sc.nw_large || sc.nw_small || sc.nw_lane) {
          TCX_TRY(lm.mark());
          if (sc.nw_small <= (uint64_t)kStaticWaves * G * (TB / 64))
            hipLaunchKernelGGL((k_auto_wl<2048, true>), dim3(G), dim3(TB), 0, c->stream, rp, col,
                               wrp, wcol, wse, ell, sc.wl_large.as<uint32_t>(), (uint32_t)sc.nw_large,
                               sc.wl_small.as<uint32_t>(), (uint32_t)sc.nw_small,
                               sc.wl_lane.as<uint32_t>(), (uint32_t)sc.nw_lane, cursor, acc);
          else
            hipLaunchKernelGGL((k_auto_wl<2048>), dim3(G), dim3(TB), 0, c->stream, rp, col, wrp,
                               wcol, wse, ell, sc.wl_large.as<uint32_t>(), (uint32_t)sc.nw_large,
                               sc.wl_small.as<uint32_t>(), (uint32_t)sc.nw_small,
                               sc.wl_lane.as<uint32_t>(), (uint32_t)sc.nw_lane, cursor, acc);
          tm.launches++;
        }
        if (sc.nw_huge) {
          TCX_TRY(lm.mark());
          hipLaunchKernelGGL((k_auto_wl<8192>),
                             dim3(std::min<uint64_t>((uint64_t)c->num_cus * 4, sc.nw_huge)),
                             dim3(TB), 0, c->stream, rp, col, wrp, wcol, wse, ell, sc.wl_huge.as<uint32_t>(),
                             (uint32_t)sc.nw_huge, (const uint32_t*)nullptr, 0u,
                             (const uint32_t*)nullptr, 0u, cursor + 4, acc);
          tm.launches++;
        }
        break;
      }
      if (sc.n_large || sc.n_small) {
        TCX_TRY(lm.mark());
        if (sc.n_small <= (uint64_t)kStaticWaves * G * (TB / 64))
          hipLaunchKernelGGL((k_auto<2048, true>), dim3(G), dim3(TB), 0, c->stream, rp, col,
                             sc.auto_large.as<uint32_t>(), (uint32_t)sc.n_large,
                             sc.auto_small.as<uint32_t>(), (uint32_t)sc.n_small, cursor, acc);
        else
          hipLaunchKernelGGL(k_auto<2048>, dim3(G), dim3(TB), 0, c->stream, rp, col,
                             sc.auto_large.as<uint32_t>(), (uint32_t)sc.n_large,
                             sc.auto_small.as<uint32_t>(), (uint32_t)sc.n_small, cursor, acc);
        tm.launches++;
      }
      if (sc.n_huge) {
        TCX_TRY(lm.mark());
        hipLaunchKernelGGL(k_auto<8192>, dim3(std::min<uint64_t>((uint64_t)c->num_cus * 4, sc.n_huge)),
                           dim3(TB), 0, c->stream, rp, col, sc.auto_huge.as<uint32_t>(),
                           (uint32_t)sc.n_huge, (const uint32_t*)nullptr, 0u, cursor + 2, acc);
        tm.launches++;
      }
      break;
    }
  }
  TCX_HIP(hipGetLastError());
  TCX_TRY(lm.finish(&tm));
  unsigned long long h = 0;
  TCX_HIP(hipMemcpyAsync(&h, acc, 8, hipMemcpyDeviceToHost, c->stream));
  TCX_HIP(hipStreamSynchronize(c->stream));
  *count = h;
  tm.ms_prep = prep.ms;
  if (timing) *timing = tm;
  return TC_OK;
}

int tc_auto_stream_bytes(tc_ctx* c, tc_graph* g, const tc_sched* sched, double* bytes) {
  TCX_TRY(check_ready(c, g));
  if (!bytes) return fail(TC_ERR_ARG, "NULL output");
  tc_sched s;
  memset(&s, 0, sizeof(s));
  if (sched) s = *sched;
  const uint32_t wmax = s.auto_thresholds[0] ? s.auto_thresholds[0] : 64u;
  const uint32_t hmin = s.auto_thresholds[1] ? s.auto_thresholds[1] : 2048u;
  if (wmax > kWaveMax || hmin < wmax || s.auto_thresholds[2] > TC_AUTO_PUSH ||
      s.auto_thresholds[3] > kLaneMaxDeg)
    return fail(TC_ERR_ARG, "bad AUTO thresholds");
  const uint32_t lmax = s.auto_thresholds[3] ? s.auto_thresholds[3] : kLaneMaxDeg;
  if (s.auto_thresholds[2] == TC_AUTO_PUSH) {
    *bytes = 4.0 * ((double)g->n + 1) + 12.0 * (double)g->m_or + 4.0 * (double)g->sum_in_out;
    return TC_OK;
  }
  ensure_version(g);
  if (g->m_or == 0 || g->n == 0) {
    *bytes = 0;
    return TC_OK;
  }
  TCX_TRY(wl_schedule(c, g, wmax, hmin, lmax));
  *bytes = g->sc.wl_bytes;
  return TC_OK;
}

int tc_count_per_vertex(tc_ctx* c, tc_graph* g, uint64_t* tri) {
  TCX_TRY(check_ready(c, g));
  if (!tri) return fail(TC_ERR_ARG, "NULL output");
  if (g->n == 0) return TC_OK;
  DevBuf d;
  TCX_TRY(d.reserve(8 * (size_t)g->n));
  TCX_TRY(tcx::per_vertex_counts(c, g, d.as<unsigned long long>()));
  TCX_HIP(hipMemcpyAsync(tri, d.p, 8 * (size_t)g->n, hipMemcpyDeviceToHost, c->stream));
  TCX_HIP(hipStreamSynchronize(c->stream));
  return TC_OK;
}

int tc_hu_workload(tc_ctx* c, tc_graph* g, int64_t* w) {
  TCX_TRY(check_ready(c, g));
  if (!w) return fail(TC_ERR_ARG, "NULL output");
  if (g->n == 0) return TC_OK;
  DevBuf d;
  TCX_TRY(d.reserve(8 * (size_t)g->n));
  hipLaunchKernelGGL(k_hu_work, dim3(grid_for(g->n)), dim3(TB), 0, c->stream,
                     g->rowptr.as<uint32_t>(), g->col.as<uint32_t>(), g->n, d.as<int64_t>());
  TCX_HIP(hipGetLastError());
  TCX_HIP(hipMemcpyAsync(w, d.p, 8 * (size_t)g->n, hipMemcpyDeviceToHost, c->stream));
  TCX_HIP(hipStreamSynchronize(c->stream));
  return TC_OK;
}

int tc_hu_block_starts(tc_ctx* c, tc_graph* g, uint32_t nblocks, int32_t* start) {
  TCX_TRY(check_ready(c, g));
  if (!start || nblocks == 0) return fail(TC_ERR_ARG, "bad arguments");
  ensure_version(g);
  TCX_TRY(hu_schedule(c, g, nblocks));
  TCX_HIP(hipMemcpyAsync(start, g->sc.hu_start.p, 4 * ((size_t)nblocks + 1), hipMemcpyDeviceToHost,
                         c->stream));
  TCX_HIP(hipStreamSynchronize(c->stream));
  return TC_OK;
}

int tc_fox_workload(tc_ctx* c, tc_graph* g, int32_t* w) {
  TCX_TRY(check_ready(c, g));
  if (!w) return fail(TC_ERR_ARG, "NULL output");
  if (g->m_or == 0) return TC_OK;
  DevBuf d;
  TCX_TRY(d.reserve(4 * g->m_or));
  hipLaunchKernelGGL(k_fox_work, dim3(grid_for(g->m_or)), dim3(TB), 0, c->stream,
                     g->rowptr.as<uint32_t>(), g->col.as<uint32_t>(), g->esrc.as<uint32_t>(), g->m_or,
                     d.as<int32_t>());
  TCX_HIP(hipGetLastError());
  TCX_HIP(hipMemcpyAsync(w, d.p, 4 * g->m_or, hipMemcpyDeviceToHost, c->stream));
  TCX_HIP(hipStreamSynchronize(c->stream));
  return TC_OK;
}

int tc_fox_bin_sizes(tc_ctx* c, tc_graph* g, uint32_t max_lanes, uint64_t* count) {
  TCX_TRY(check_ready(c, g));
  if (!count) return fail(TC_ERR_ARG, "NULL output");
  if (max_lanes == 0 || max_lanes > 64 || (max_lanes & (max_lanes - 1)))
    return fail(TC_ERR_ARG, "max_lanes must be a power of two <= 64");
  ensure_version(g);
  uint32_t nb = 1;
  for (uint32_t l = max_lanes; l > 1; l >>= 1) ++nb;
  if (g->m_or == 0) {
    for (uint32_t j = 0; j < nb; ++j) count[j] = 0;
    return TC_OK;
  }
  TCX_TRY(fox_schedule(c, g, max_lanes, g->sc.fox_lanes == max_lanes && g->sc.fox_rearr));
  for (uint32_t j = 0; j < nb; ++j) count[j] = g->sc.fox_off[j + 1] - g->sc.fox_off[j];
  return TC_OK;
}

}  // extern "C"
