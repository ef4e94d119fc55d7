// k-clique counting over an oriented (acyclic) CSR, k-clique/kcl.cu:16-215.
//
// The reference grows embeddings level by level and materialises every
// (k-1)-level in HBM.  Here each root u owns the sub-graph induced by N+(u):
//  * BITMAP path (d+(u) <= 512): the sub-graph's adjacency is built as a
//    d x d bit matrix in LDS; a lane takes one second vertex and walks the
//    remaining levels with AND + popcount on W-word bitsets.
//  * MERGE path (any d+(u)): a wave walks the clique tree depth-first;
//    candidate lists live in per-wave HBM scratch and each intersection is
//    a wave-wide binary-search + ballot compaction.
// Every k-clique has a unique source in the DAG and a unique chain order,
// so it is counted exactly once on either path.
#include <algorithm>
#include <chrono>
#include <cstring>

#include "internal.h"

namespace tcx {
namespace {

constexpr int kMaxK = 12;
constexpr int kKclFeatures = TC_MODEL_FEATURES;
constexpr uint32_t kPivotMax = 1024;  // widest bit matrix in LDS (16 words per row)
constexpr int kCompactWords = 4;      // pivot matrices this wide pack C's members per pass

__device__ inline uint64_t gtid() { return (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; }
__device__ inline uint64_t gstride() { return (uint64_t)gridDim.x * blockDim.x; }

__device__ inline unsigned long long wave_sum(unsigned long long v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}

__device__ inline int find_idx(const uint32_t* a, uint32_t n, uint32_t x) {
  uint32_t lo = 0, hi = n;
  while (lo < hi) {
    uint32_t mid = (lo + hi) >> 1;
    uint32_t v = a[mid];
    if (v == x) return (int)mid;
    if (v < x) lo = mid + 1;
    else hi = mid;
  }
  return -1;
}

__device__ inline bool contains(const uint32_t* a, uint32_t lo, uint32_t hi, uint32_t x) {
  while (lo < hi) {
    uint32_t mid = (lo + hi) >> 1;
    uint32_t v = a[mid];
    if (v == x) return true;
    if (v < x) lo = mid + 1;
    else hi = mid;
  }
  return false;
}

// keys[i] = hi - d+(roots[i]): an ascending sort puts the widest roots first.
__global__ void k_root_keys(const uint32_t* __restrict__ rp, const uint32_t* __restrict__ roots,
                            uint64_t nr, uint32_t hi, uint32_t* __restrict__ keys) {
  for (uint64_t i = gtid(); i < nr; i += gstride()) {
    const uint32_t u = roots[i];
    keys[i] = hi - (rp[u + 1] - rp[u]);
  }
}

// f[u] = lo <= d+(u) <= hi, and (with a plan) the plan sends u to `want`
__global__ void k_deg_flags(const uint32_t* __restrict__ rp, uint32_t n, uint32_t lo, uint32_t hi,
                            const uint8_t* __restrict__ path, uint8_t want,
                            uint8_t* __restrict__ f) {
  for (uint64_t u = gtid(); u < n; u += gstride()) {
    uint32_t d = rp[u + 1] - rp[u];
    f[u] = (d >= lo && d <= hi && (!path || path[u] == want)) ? 1 : 0;
  }
}

__global__ void k_plan_keys(const uint32_t* __restrict__ roots, uint64_t nr,
                            const uint32_t* __restrict__ key, uint32_t* __restrict__ keys) {
  for (uint64_t i = gtid(); i < nr; i += gstride()) keys[i] = key[roots[i]];
}

// Per-root features X[n][8] for k (roots with d+ >= k-1; others zero):
// log2(1+d+), log2(1+m) with m the edges of the sub-graph of N+(u),
// density m / C(d+,2), k, log2(1+sum_v d+(v)), log2 of the k-1 cliques a
// random graph of that density would hold (C(d+,k-1) p^C(k-1,2)), the
// bit-matrix width class (0: d+ <= 64 ... 4: <= 1024, 5: above),
// log2(1+max_v d+(v)).
__global__ void k_kcl_features(const uint32_t* __restrict__ rp, const uint32_t* __restrict__ col,
                               const unsigned long long* __restrict__ tri, uint32_t n, uint32_t k,
                               float* __restrict__ X) {
  for (uint64_t u = gtid(); u < n; u += gstride()) {
    const uint32_t b = rp[u], d = rp[u + 1] - b;
    float* x = X + u * kKclFeatures;
    if (d + 1 < k) {
      for (int j = 0; j < kKclFeatures; ++j) x[j] = 0.f;
      continue;
    }
    uint64_t s2 = 0;
    uint32_t mx = 0;
    for (uint32_t e = b; e < b + d; ++e) {
      const uint32_t v = col[e];
      const uint32_t dv = rp[v + 1] - rp[v];
      s2 += dv;
      mx = dv > mx ? dv : mx;
    }
    const double m = (double)tri[u];
    const double pairs = 0.5 * (double)d * (double)(d - 1);
    const double dens = pairs > 0 ? m / pairs : 0.0;
    const double r = (double)(k - 1);
    double est = -30.0;
    if (dens > 0) {
      const double lc = (lgamma((double)d + 1) - lgamma(r + 1) - lgamma((double)d - r + 1)) / M_LN2;
      est = fmax(-30.0, lc + 0.5 * r * (r - 1) * log2(dens));
    }
    x[0] = (float)log2(1.0 + d);
    x[1] = (float)log2(1.0 + m);
    x[2] = (float)dens;
    x[3] = (float)k;
    x[4] = (float)log2(1.0 + (double)s2);
    x[5] = (float)est;
    x[6] = d <= 64 ? 0.f : d <= 128 ? 1.f : d <= 256 ? 2.f : d <= 512 ? 3.f : d <= 1024 ? 4.f : 5.f;
    x[7] = (float)log2(1.0 + mx);
  }
}

// Learned plan from predicted log2 costs Y[n][2] (pivot, merge).  The fixed
// rule (pivot iff d+ <= 1024) is the prior: a root leaves it only for a
// predicted saving above kPlanMargin.  Each launch takes its roots widest
// first, as kcl_count does (ordering by predicted cost measured slower).
constexpr float kPlanMargin = 0.25f;
__global__ void k_kcl_plan(const uint32_t* __restrict__ rp, uint32_t n, uint32_t need,
                           const float* __restrict__ Y, uint8_t* __restrict__ path,
                           uint32_t* __restrict__ key) {
  const float lm = log2f(1.f - kPlanMargin);
  for (uint64_t u = gtid(); u < n; u += gstride()) {
    const uint32_t d = rp[u + 1] - rp[u];
    if (d < need) {
      path[u] = 255;
      key[u] = 0xFFFFFFFFu;
      continue;
    }
    const float yp = Y[2 * u], ym = Y[2 * u + 1];
    bool pivot = d <= kPivotMax;
    if (pivot && ym < yp + lm) pivot = false;
    path[u] = pivot ? 0 : 1;
    key[u] = 0xFFFFFFFFu - d;
  }
}

// ---- BITMAP path ------------------------------------------------------------
constexpr int kBitmapMaxK = 9;  // LDS DFS stack depth bound (k - 3 levels)

template <int W>
__global__ __launch_bounds__(64) void k_kcl_bitmap(const uint32_t* __restrict__ rp,
                                                   const uint32_t* __restrict__ col,
                                                   const uint32_t* __restrict__ roots,
                                                   uint64_t nroots, uint32_t k,
                                                   unsigned long long* total) {
  constexpr int D = 64 * W;
  constexpr int LV = kBitmapMaxK - 3;
  __shared__ uint32_t s_list[D];
  __shared__ uint64_t s_rows[D * W];
  __shared__ uint64_t s_stack[64 * LV * 2 * W];  // per lane: (set, remaining) per level
  const uint32_t lane = threadIdx.x;
  unsigned long long cnt = 0;
  uint64_t* my = s_stack + (uint64_t)lane * LV * 2 * W;
  const int depth = (int)k - 2;  // vertices still to choose after u and the lane's first vertex
  for (uint64_t ri = blockIdx.x; ri < nroots; ri += gridDim.x) {
    const uint32_t u = roots[ri];
    const uint32_t base = rp[u], d = rp[u + 1] - base;
    for (uint32_t i = lane; i < d; i += 64) s_list[i] = col[base + i];
    __syncthreads();
    // row i = local indices of N+(c_i) & N+(u)
    for (uint32_t i = lane; i < d; i += 64) {
      const uint32_t ci = s_list[i];
      uint64_t row[W];
#pragma unroll
      for (int w = 0; w < W; ++w) row[w] = 0;
      for (uint32_t e = rp[ci]; e < rp[ci + 1]; ++e) {
        int j = find_idx(s_list, d, col[e]);
        if (j >= 0) row[j >> 6] |= 1ull << (j & 63);
      }
#pragma unroll
      for (int w = 0; w < W; ++w) s_rows[i * W + w] = row[w];
    }
    __syncthreads();
    for (uint32_t i = lane; i < d; i += 64) {
      const uint64_t* r0 = s_rows + i * W;
      if (depth == 1) {
        for (int w = 0; w < W; ++w) cnt += __popcll(r0[w]);
        continue;
      }
      // DFS over levels 0..depth-2; level L holds the candidates for the
      // (L+3)-th clique vertex and the part of them not yet tried.
      for (int w = 0; w < W; ++w) {
        my[w] = r0[w];
        my[W + w] = r0[w];
      }
      int L = 0;
      while (L >= 0) {
        uint64_t* set = my + (2 * L) * W;
        uint64_t* rem = set + W;
        int j = -1;
        for (int w = 0; w < W; ++w) {
          if (rem[w]) {
            int b = __ffsll((long long)rem[w]) - 1;
            rem[w] &= rem[w] - 1;
            j = w * 64 + b;
            break;
          }
        }
        if (j < 0) {
          --L;
          continue;
        }
        const uint64_t* rj = s_rows + j * W;
        if (L == depth - 2) {
          for (int w = 0; w < W; ++w) cnt += __popcll(set[w] & rj[w]);
        } else {
          uint64_t* nset = set + 2 * W;
          uint64_t any = 0;
          for (int w = 0; w < W; ++w) {
            uint64_t x = set[w] & rj[w];
            nset[w] = x;
            nset[W + w] = x;
            any |= x;
          }
          if (any) ++L;
        }
      }
    }
    __syncthreads();
  }
  cnt = wave_sum(cnt);
  if (lane == 0 && cnt) atomicAdd(total, cnt);
}

// ---- MERGE path -------------------------------------------------------------
// Level L holds the candidates after choosing u and L-1 more vertices
// (level 1 = N+(u)).  Levels 2..k-2 are materialised in the wave's scratch
// (slot L-2); at level k-2 every candidate x adds |S & N+(x)| directly.
// The DFS control is uniform across the wave: every lane writes the same
// values to s_size / s_idx, so no barrier is needed for them.
__global__ __launch_bounds__(64) void k_kcl_merge(const uint32_t* __restrict__ rp,
                                                  const uint32_t* __restrict__ col,
                                                  const uint32_t* __restrict__ roots,
                                                  uint64_t nroots, uint32_t k, uint32_t max_d,
                                                  uint32_t slots, uint32_t* __restrict__ scratch,
                                                  unsigned long long* total,
                                                  float* __restrict__ cost, float cost_scale) {
  __shared__ uint32_t s_size[kMaxK];
  __shared__ uint32_t s_idx[kMaxK];
  const uint32_t lane = threadIdx.x;
  unsigned long long cnt = 0;
  uint32_t* mine = scratch + (uint64_t)blockIdx.x * slots * max_d;
  const int last = (int)k - 2;
  for (uint64_t ri = blockIdx.x; ri < nroots; ri += gridDim.x) {
    const uint32_t u = roots[ri];
    const uint64_t t_root = cost ? wall_clock64() : 0;
    const uint32_t ub = rp[u], d = rp[u + 1] - ub;
    s_size[1] = d;
    s_idx[1] = 0;
    int L = 1;
    while (L >= 1) {
      const uint32_t* S = L == 1 ? col + ub : mine + (uint64_t)(L - 2) * max_d;
      const uint32_t sz = s_size[L];
      if (L == last) {
        for (uint32_t ix = 0; ix < sz; ++ix) {
          const uint32_t x = S[ix];
          const uint32_t xb = rp[x], xe = rp[x + 1];
          for (uint32_t t = lane; t < sz; t += 64) cnt += contains(col, xb, xe, S[t]);
        }
        --L;
        continue;
      }
      const uint32_t ix = s_idx[L];
      if (ix >= sz) {
        --L;
        continue;
      }
      s_idx[L] = ix + 1;
      const uint32_t x = S[ix];
      const uint32_t xb = rp[x], xe = rp[x + 1];
      uint32_t* out = mine + (uint64_t)(L - 1) * max_d;
      uint32_t nsz = 0;
      for (uint32_t t0 = 0; t0 < sz; t0 += 64) {
        const uint32_t t = t0 + lane;
        const bool keep = t < sz && contains(col, xb, xe, S[t]);
        const uint64_t bal = __ballot(keep);
        const uint32_t pos = __popcll(bal & ((1ull << lane) - 1));
        if (keep) out[nsz + pos] = S[t];
        nsz += __popcll(bal);
      }
      __syncthreads();  // compacted list visible to every lane
      if (nsz > 0) {
        s_size[L + 1] = nsz;
        s_idx[L + 1] = 0;
        ++L;
      }
    }
    __syncthreads();
    if (cost && lane == 0) cost[u] = (float)(wall_clock64() - t_root) * cost_scale;
  }
  cnt = wave_sum(cnt);
  if (lane == 0 && cnt) atomicAdd(total, cnt);
}

// ---- PIVOT path ---------------------------------------------------------------
// Pivoting over the succinct clique tree (the counting scheme of Jain and
// Seshadhri, "The Power of Pivoting for Exact Clique Counting", WSDM 2020),
// one wave per root u.  The sub-graph induced by N+(u) is held as a symmetric
// d x d bit matrix in LDS.  A tree node (C, h, p) has candidate set C, h
// held vertices (u and the non-pivot branch vertices) and p pivots; every
// k-clique of the sub-tree is the held set plus S from the pivots plus a
// clique of C.  A node picks the pivot x of C with most neighbours in C and
// branches on x and on the vertices of C outside N(x), each branch removing
// its vertex from C for the later ones; a leaf (C empty) holds
// binom(p, k - h) k-cliques.  For the one k asked:
//   h > k or h + p + |C| < k  -> nothing below,
//   k - h = 0, 1, 2          -> closed forms 1, p + |C|,
//                               binom(p, 2) + p |C| + edges(C),
// so the tree is only walked where at least three more vertices are needed.
// The walk's stack (C and the branches still to take, per level) lives in
// LDS for one- and two-word matrices and in the wave's slice of `stk`
// otherwise: d + 1 levels of 2W + 1 words.
__device__ inline unsigned long long binom_small(uint32_t n, uint32_t r) {
  if (r > n) return 0;
  unsigned __int128 c = 1;
  for (uint32_t i = 1; i <= r; ++i) c = c * (n - r + i) / i;
  return (unsigned long long)c;
}

__device__ inline uint64_t wave_or_u64(uint64_t v) {
  for (int off = 32; off > 0; off >>= 1) v |= __shfl_xor(v, off, 64);
  return v;
}

__device__ inline uint32_t wave_sum_u32(uint32_t v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

__device__ inline uint32_t wave_max_u32(uint32_t v) {
  for (int off = 32; off > 0; off >>= 1) v = max(v, __shfl_xor(v, off, 64));
  return v;
}

__device__ inline void wave_sync_kcl() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// NW waves share one root: they build its bit matrix together, then wave w
// walks the sub-trees of the root's branches w, w + NW, ... (each branch's
// candidate set is the root's C minus the earlier branch vertices, so the
// branches are independent).  Every wave has its own LDS slot for the
// current C and its own stack slice; a stack word is only read back by the
// lane that wrote it (lane w holds word w of C and of the branch set).
template <int W, int NW>
__global__ __launch_bounds__(64 * NW) void k_kcl_pivot(const uint32_t* __restrict__ rp,
                                                       const uint32_t* __restrict__ col,
                                                       const uint32_t* __restrict__ roots,
                                                       uint64_t nroots, uint32_t copies,
                                                       uint32_t chunk, uint32_t k,
                                                       uint64_t* __restrict__ stk_all,
                                                       unsigned long long* total,
                                                       unsigned long long* claim,
                                                       float* __restrict__ cost, float cost_scale) {
  constexpr uint32_t D = 64 * W;
  constexpr uint32_t LW = 2 * W + 1;  // stack words per level
  static_assert(NW <= W, "each wave's 64-entry member slot lies in s_list[64 W]");
  __shared__ uint32_t s_list[D];
  // bit-matrix row stride in words: rows of 8 and 16 words are padded by
  // two words (16-byte aligned, so row reads stay 128-bit), which spreads
  // the rows that the lanes read at once over the LDS banks
  constexpr uint32_t RS = W >= 8 ? W + 2 : W;
  __shared__ uint64_t s_M[D * RS];
  // Matrices of 2 words or more (d+ > 64) filter the out-list entries
  // through a hashed bitmap of N+(u) (32 bits per possible member): an entry
  // that misses it is not in N+(u) and skips the binary search.  On one-word
  // matrices the search is short and most entries are members, so the
  // filter and its queue cost more than they save (RGG 10^6 k = 4: 3.41
  // against 2.27 ms).
  constexpr bool kFilter = W >= 2;
  constexpr uint32_t kFiltWords = kFilter ? 64 * W : 1;
  constexpr uint32_t kFiltShift = 32u - (uint32_t)__builtin_ctz(2048u * (uint32_t)W);
  __shared__ uint32_t s_filt[kFiltWords];
  __shared__ uint64_t s_Cs[NW][W];
  const uint32_t tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  uint64_t* s_C = s_Cs[wv];
  // one- and two-word matrices keep the walk's stack in LDS (one wave per
  // workgroup), the wider ones in their slice of stk_all
  constexpr bool kLdsStack = W <= 2;
  static_assert(!kLdsStack || NW == 1, "an LDS stack holds one walker");
  __shared__ uint64_t s_stk[kLdsStack ? (D + 1) * LW : 1];
  uint64_t* stk = kLdsStack ? s_stk : stk_all + ((uint64_t)blockIdx.x * NW + wv) * (D + 1) * LW;
  unsigned long long cnt = 0;  // lane 0 of each wave accumulates
  const uint32_t Q = copies * NW;  // walkers per root
  // (root, copy) tasks are claimed `chunk` at a time from a counter, in the
  // order of `roots` (heaviest first), so no workgroup is left holding a
  // fixed share of the heavy roots
  const uint64_t ntask = nroots * copies;
  __shared__ unsigned long long s_ti;
  __shared__ uint32_t s_next;
  for (;;) {
    if (tid == 0) s_ti = atomicAdd(claim, (unsigned long long)chunk);
    __syncthreads();
    const uint64_t tb = s_ti;
    if (tb >= ntask) break;
    const uint64_t te = tb + chunk < ntask ? tb + chunk : ntask;
    for (uint64_t ti = tb; ti < te; ++ti) {
      const uint32_t u = roots[ti / copies];
      const uint64_t t_task = (cost && tid == 0) ? wall_clock64() : 0;
      const uint32_t q = (uint32_t)(ti % copies) * NW + wv;  // this wave's walker number
      const uint32_t base = rp[u], d = rp[u + 1] - base;
      for (uint32_t i = tid; i < d; i += 64 * NW) s_list[i] = col[base + i];
      for (uint32_t i = tid; i < d * RS; i += 64 * NW) s_M[i] = 0;
      if constexpr (kFilter)
        for (uint32_t i = tid; i < kFiltWords; i += 64 * NW) s_filt[i] = 0;
      if (tid == 0) s_next = NW;
      __syncthreads();
      if constexpr (kFilter) {
        for (uint32_t i = tid; i < d; i += 64 * NW) {
          const uint32_t h = (s_list[i] * 0x9E3779B1u) >> kFiltShift;
          atomicOr(&s_filt[h >> 5], 1u << (h & 31));
        }
        __syncthreads();
      }
      // symmetric adjacency of N+(u): bit (i, j) for every arc c_i -> c_j.
      // A wave takes four out-lists per step: when none is longer than 64,
      // each gets 16 lanes; otherwise the whole wave streams them in turn.
      // Entries that hit the hashed bitmap of N+(u) are queued, one per
      // lane, in registers (qi, qx: lane l holds slot l of qn); a full queue
      // resolves 64 binary searches at once, so the searches run on full
      // waves although only a few entries of each step pass.
      auto add_arcs = [&](uint32_t i, uint32_t x) {
        uint32_t lo = 0, len = d;
        while (len > 1) {
          const uint32_t half = len >> 1;
          lo = s_list[lo + half] <= x ? lo + half : lo;
          len -= half;
        }
        if (s_list[lo] == x) {
          atomicOr((unsigned long long*)&s_M[i * RS + (lo >> 6)], 1ull << (lo & 63));
          atomicOr((unsigned long long*)&s_M[lo * RS + (i >> 6)], 1ull << (i & 63));
        }
      };
      uint32_t qi = 0, qx = 0, qn = 0;
      // wave-uniform: every lane calls it; active lanes offer (i, x)
      auto offer = [&](bool active, uint32_t i, uint32_t x) {
        bool pass = false;
        if (active) {
          const uint32_t h = (x * 0x9E3779B1u) >> kFiltShift;
          pass = (s_filt[h >> 5] >> (h & 31)) & 1u;
        }
        const uint64_t mask = __ballot(pass);
        const uint32_t cnt = __popcll(mask);
        if (cnt == 0) return;
        // lane of the r-th set bit of mask: the largest lo with
        // popcount(mask below lo) <= r
        auto nth = [&](uint32_t r) -> uint32_t {
          uint32_t lo = 0;
          for (uint32_t step = 32; step; step >>= 1)
            if ((uint32_t)__popcll(mask & ((1ull << (lo + step)) - 1)) <= r) lo += step;
          return lo;
        };
        const uint32_t take = min(cnt, 64u - qn);
        uint32_t src = nth((lane - qn) & 63u);
        uint32_t vi = __shfl(i, src, 64), vx = __shfl(x, src, 64);
        if (lane >= qn && lane < qn + take) {
          qi = vi;
          qx = vx;
        }
        qn += take;
        if (qn == 64) {
          add_arcs(qi, qx);
          qn = cnt - take;
          src = nth((take + lane) & 63u);
          vi = __shfl(i, src, 64);
          vx = __shfl(x, src, 64);
          if (lane < qn) {
            qi = vi;
            qx = vx;
          }
        }
      };
      for (uint32_t i0 = wv * 4u; i0 < d; i0 += NW * 4u) {
        const uint32_t gi = i0 + (lane >> 4);
        uint32_t eb = 0, ee = 0;
        if (gi < d) {
          const uint32_t ci = s_list[gi];
          eb = rp[ci];
          ee = rp[ci + 1];
        }
        if constexpr (!kFilter) {
          if (__ballot(ee - eb > 64u) == 0) {
            for (uint32_t e = eb + (lane & 15u); e < ee; e += 16u) add_arcs(gi, col[e]);
          } else {
            for (uint32_t g = 0; g < 4u && i0 + g < d; ++g) {
              const uint32_t b = __shfl(eb, 16 * g, 64), e1 = __shfl(ee, 16 * g, 64);
              for (uint32_t e = b + lane; e < e1; e += 64) add_arcs(i0 + g, col[e]);
            }
          }
          continue;
        }
        if (__ballot(ee - eb > 64u) == 0) {
          for (uint32_t k0 = 0;; k0 += 16) {
            const uint32_t e = eb + k0 + (lane & 15u);
            const bool act = e < ee;
            if (__ballot(act) == 0) break;
            offer(act, gi, act ? col[e] : 0u);
          }
        } else {
          for (uint32_t g = 0; g < 4u && i0 + g < d; ++g) {
            const uint32_t b = __shfl(eb, 16 * g, 64), e1 = __shfl(ee, 16 * g, 64);
            for (uint32_t e0 = b; e0 < e1; e0 += 64) {
              const uint32_t e = e0 + lane;
              const bool act = e < e1;
              offer(act, i0 + g, act ? col[e] : 0u);
            }
          }
        }
      }
      if (lane < qn) add_arcs(qi, qx);
      __syncthreads();

      // Node (s_C, h, p): counted in closed form (push = false / no pivot
      // needed), or its pivot is returned for the caller to branch on.
      // Returns true when the node needs branching; piv and its branch set
      // (lanes < W, word `lane`) are written to *piv_out / *todo_out.
      // Nodes with h = 1 are walked by every walker: their single terms are
      // counted by walker 0 (`lead`), and the sums of the r = 2, 3 closed
      // forms are split over the walkers by pass (64 members of C).  Those sums
      // count every edge of C once (at its smaller end) and every triangle
      // once (at its smallest vertex), so each walker's share is an integer.
      auto node = [&](uint32_t h, uint32_t p, uint32_t* piv_out, uint64_t* todo_out) -> bool {
        const bool lead = h != 1 || q == 0;
        const bool part = h == 1 && Q > 1;
        uint64_t cr[W];  // C, the same in every lane
        uint32_t cs = 0;
#pragma unroll
        for (int w = 0; w < W; ++w) {
          cr[w] = s_C[w];
          cs += (uint32_t)__popcll(cr[w]);
        }
        const uint32_t r = k - h;
        if (cs == 0) {
          if (lane == 0 && lead) cnt += binom_small(p, r);
          return false;
        }
        if (h + p + cs < k) return false;
        if (r == 0) {
          if (lane == 0 && lead) cnt += 1;
          return false;
        }
        if (r == 1) {
          if (lane == 0 && lead) cnt += p + cs;
          return false;
        }
        // The members of C are taken 64 at a time, one per lane.  With
        // W >= kCompactWords words they are packed in index order through the
        // wave's 64-entry slot of s_list (free once the bit matrix is built),
        // so a sparse C takes ceil(|C| / 64) passes; narrower matrices take
        // one pass per non-empty word.
        constexpr uint32_t kNone = 0xFFFFFFFFu;
        uint32_t pcw = 0, pex = 0;  // lane w < W: members in word w / before it
        if constexpr (W >= kCompactWords) {
          pcw = lane < (uint32_t)W ? (uint32_t)__popcll(s_C[lane]) : 0u;
          uint32_t inc = pcw;
#pragma unroll
          for (uint32_t off = 1; off < (uint32_t)W; off <<= 1) {
            const uint32_t v = __shfl_up(inc, off, 64);
            if (lane >= off) inc += v;
          }
          pex = inc - pcw;
        }
        auto for_members = [&](bool split, auto&& f) {
          if constexpr (W >= kCompactWords) {
            uint32_t* pl = s_list + 64u * wv;
            for (uint32_t j = 0, lo = 0; lo < cs; ++j, lo += 64) {
              if (split && j % Q != q) continue;
              uint64_t words = __ballot(lane < (uint32_t)W && pex < lo + 64 && pex + pcw > lo);
              while (words) {
                const uint32_t w = (uint32_t)__builtin_ctzll(words);
                words &= words - 1;
                const uint64_t v = s_C[w];
                const uint32_t rk = __shfl(pex, w, 64) +
                                    __builtin_amdgcn_mbcnt_hi((uint32_t)(v >> 32),
                                                              __builtin_amdgcn_mbcnt_lo((uint32_t)v, 0u));
                if (((v >> lane) & 1ull) && rk >= lo && rk < lo + 64) pl[rk - lo] = 64u * w + lane;
              }
              wave_sync_kcl();
              const uint32_t x = lo + lane < cs ? pl[lane] : kNone;
              wave_sync_kcl();
              f(x);
            }
          } else {
#pragma unroll
            for (int t = 0; t < W; ++t) {
              if (cr[t] == 0 || (split && (uint32_t)t % Q != q)) continue;
              f(((cr[t] >> lane) & 1ull) ? 64u * t + lane : kNone);
            }
          }
        };
        if (r == 2) {
          // binom(p,2) + p|C| + edges(C); lane x counts its neighbours in C
          // above x
          uint32_t esum = 0;
          for_members(part, [&](uint32_t x) {
            if (x == kNone) return;
            const uint32_t t = x >> 6;
            const uint64_t lm = ~((2ull << (x & 63)) - 1);
#pragma unroll
            for (int w = 0; w < W; ++w) {
              const uint64_t above = (uint32_t)w < t ? 0ull : ((uint32_t)w > t ? ~0ull : lm);
              esum += (uint32_t)__popcll(s_M[x * RS + w] & cr[w] & above);
            }
          });
          const unsigned long long e = wave_sum_u32(esum);
          if (lane == 0)
            cnt += (lead ? binom_small(p, 2) + (unsigned long long)p * cs : 0ull) + e;
          return false;
        }
        if (r == 3) {
          // binom(p,3) + binom(p,2)|C| + p edges(C) + triangles(C).  Lane x
          // holds A = N(x) & C above x: |A| edges start at x, and the pairs
          // y < z of A with y ~ z count each triangle whose smallest vertex
          // is x once.  z runs over A above y only: the rest of y's word and
          // the words after it.
          uint64_t esum = 0, tsum = 0;
          for_members(part, [&](uint32_t x) {
            if (x == kNone) return;
            const uint32_t t = x >> 6;
            const uint64_t lm = ~((2ull << (x & 63)) - 1);
            uint64_t ax[W];
#pragma unroll
            for (int w = 0; w < W; ++w) {
              const uint64_t above = (uint32_t)w < t ? 0ull : ((uint32_t)w > t ? ~0ull : lm);
              ax[w] = s_M[x * RS + w] & cr[w] & above;
              esum += (uint32_t)__popcll(ax[w]);
            }
#pragma unroll
            for (int w = 0; w < W; ++w) {
              uint64_t ys = ax[w];
              while (ys) {
                const uint32_t y = 64u * w + (uint32_t)__builtin_ctzll(ys);
                ys &= ys - 1;  // now the members of A in word w above y
                uint32_t c3 = (uint32_t)__popcll(ys & s_M[y * RS + w]);
#pragma unroll
                for (int w2 = w + 1; w2 < W; ++w2) c3 += (uint32_t)__popcll(ax[w2] & s_M[y * RS + w2]);
                tsum += c3;
              }
            }
          });
          const unsigned long long e = wave_sum(esum), t2 = wave_sum(tsum);
          if (lane == 0)
            cnt += (lead ? binom_small(p, 3) + binom_small(p, 2) * cs : 0ull) +
                   (unsigned long long)p * e + t2;
          return false;
        }
        // pivot: the member of C with most neighbours in C, the smallest such
        // index on ties (any member gives exact counts; the best one leaves
        // the fewest branches).  Every walker searches all of C.
        uint32_t best = 0;
        auto search = [&](uint32_t x) {
          if (x == kNone) return;
          uint32_t sc = 0;
#pragma unroll
          for (int w = 0; w < W; ++w) sc += (uint32_t)__popcll(s_M[x * RS + w] & cr[w]);
          const uint32_t key = (sc << 16) | (0xFFFFu - x);  // (score, smallest index)
          best = key > best ? key : best;
        };
        for_members(false, search);
        best = wave_max_u32(best);
        const uint32_t piv = 0xFFFFu - (best & 0xFFFFu);
        *piv_out = piv;
        if (lane < (uint32_t)W) {
          uint64_t todo = s_C[lane] & ~s_M[piv * RS + lane];  // C outside N(pivot)
          if (lane == (piv >> 6)) todo |= 1ull << (piv & 63);
          *todo_out = todo;
        }
        return true;
      };

      // Every walker runs the same depth-first walk from the root.  Nodes
      // with h = 1 (the root and its pivot chain) are walked by all of them
      // and counted as node() describes; the sub-trees entered through a
      // first non-pivot branch (h = 2) are numbered in walk order and each
      // is taken by one walker (below).  One heavy root thus spreads over Q
      // waves.
      if (lane < (uint32_t)W) {
        const uint32_t lo = 64u * lane;
        s_C[lane] = d >= lo + 64 ? ~0ull : (d > lo ? ((1ull << (d - lo)) - 1) : 0ull);
      }
      wave_sync_kcl();
      uint32_t seq = 0, mine = q;
      int L = -1;
      uint32_t piv = 0;
      uint64_t todo = 0;
      // one call site of node(): the node in s_C with (nh, np) is evaluated
      // when pending (first the root), then the walk picks the next branch
      uint32_t nh = 1, np = 0;
      bool pending = true;
      for (;;) {
        if (pending) {
          const bool br = node(nh, np, &piv, &todo);
          if (br) {
            ++L;
            uint64_t* nv = stk + (uint64_t)L * LW;
            if (lane < (uint32_t)W) {
              nv[lane] = s_C[lane];
              nv[W + lane] = todo;
            }
            if (lane == 0) nv[2 * W] = (uint64_t)piv | ((uint64_t)nh << 16) | ((uint64_t)np << 32);
          }
          wave_sync_kcl();
          pending = false;
        }
        if (L < 0) break;
        uint64_t* lv = stk + (uint64_t)L * LW;
        const uint64_t todo_w = lane < (uint32_t)W ? lv[W + lane] : 0ull;
        const uint64_t nzl = __ballot(todo_w != 0);
        if (nzl == 0) {
          --L;
          continue;
        }
        const uint32_t fwl = (uint32_t)__builtin_ctzll(nzl);
        const uint64_t twl = __shfl(todo_w, fwl, 64);
        const uint32_t x = fwl * 64u + (uint32_t)__builtin_ctzll(twl);
        const uint64_t meta = __shfl(lane == 0 ? lv[2 * W] : 0ull, 0, 64);
        const uint32_t lpiv = (uint32_t)(meta & 0xFFFFu), h = (uint32_t)((meta >> 16) & 0xFFFFu),
                       p = (uint32_t)(meta >> 32);
        const bool xp = x == lpiv;
        const uint32_t h2 = h + (xp ? 0u : 1u), p2 = p + (xp ? 1u : 0u);
        // take x: child C = C & N(x); x leaves C and the branch set
        bool skip = h2 > k;
        if (h == 1 && h2 == 2) {
          // with one copy per root the walkers of the workgroup claim
          // sub-trees from an LDS counter as they reach them (each holds one
          // claim, always ahead of its own walk); otherwise walker q takes
          // every Q-th
          bool take;
          if (NW > 1 && copies == 1) {
            take = seq == mine;
            if (take) {
              uint32_t cl = 0;
              if (lane == 0) cl = atomicAdd(&s_next, 1u);
              mine = __shfl(cl, 0, 64);
            }
          } else {
            take = (seq % Q) == q;
          }
          ++seq;
          skip = skip || !take;
        }
        if (lane < (uint32_t)W) {
          const uint64_t c = lv[lane];
          const uint64_t xb = lane == fwl ? (1ull << (x & 63)) : 0ull;
          if (!skip) s_C[lane] = c & s_M[x * RS + lane];
          lv[lane] = c & ~xb;
          lv[W + lane] = todo_w & ~xb;
        }
        wave_sync_kcl();
        if (skip) continue;
        nh = h2;
        np = p2;
        pending = true;
      }
      __syncthreads();
      if (cost && tid == 0) atomicAdd(cost + u, (float)(wall_clock64() - t_task) * cost_scale);
    }
  }
  if (lane == 0 && cnt) atomicAdd(total, cnt);
}

}  // namespace
}  // namespace tcx

using namespace tcx;

static unsigned grid_for(uint64_t n, int tb = 256) {
  uint64_t b = (n + tb - 1) / tb;
  if (b < 1) b = 1;
  if (b > (1u << 20)) b = 1u << 20;
  return (unsigned)b;
}

namespace {

struct KclPlan {
  const uint8_t* path;  // 0 pivot, 1 merge, 255 none
  const uint32_t* key;  // ascending = launched first
};

// Fixed plan for cost measurement: every root with need <= d+ <= hi goes to
// `want`, widest first.
__global__ void k_fixed_plan(const uint32_t* __restrict__ rp, uint32_t n, uint32_t need,
                             uint32_t hi, uint8_t want, uint8_t* __restrict__ path,
                             uint32_t* __restrict__ key) {
  for (uint64_t u = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; u < n;
       u += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t d = rp[u + 1] - rp[u];
    path[u] = (d >= need && d <= hi) ? want : 255;
    key[u] = 0xFFFFFFFFu - d;
  }
}

// Counts k-cliques over the roots with d+ >= k-1 (k >= 3, m+ > 0).  plan
// (optional): per-root path and launch order.  cost (optional, device, n
// floats, zeroed by the caller): receives each root's measured
// CU-microseconds on the pivot and merge paths.
int kcl_run(tc_ctx* c, tc_graph* g, uint32_t k, int variant, const KclPlan* plan, float* cost,
            uint64_t* count, tc_timing* timing) {
  tc_timing tm;
  memset(&tm, 0, sizeof(tm));
  const uint32_t* rp = g->rowptr.as<uint32_t>();
  const uint32_t* col = g->col.as<uint32_t>();
  // root buckets: [lo, hi] of d+ and the path that takes them
  struct Bucket { uint32_t lo, hi; int w; bool pivot; };  // w = bitmap words, 0 = merge
  std::vector<Bucket> buckets;
  const uint32_t need = k - 1;  // a root needs k-1 out-neighbours
  if (plan) {
    // learned: pivot buckets hold the roots the plan sends to the pivot
    // path, one merge launch takes the rest
    buckets.push_back({need, 64, 1, true});
    buckets.push_back({65, 128, 2, true});
    buckets.push_back({129, 256, 4, true});
    buckets.push_back({257, 512, 8, true});
    buckets.push_back({513, kPivotMax, 16, true});
    buckets.push_back({need, 0xFFFFFFFFu, 0, false});
  } else if (variant == KCL_VARIANT_AUTO || variant == KCL_VARIANT_PIVOT) {
    buckets.push_back({need, 64, 1, true});
    buckets.push_back({65, 128, 2, true});
    buckets.push_back({129, 256, 4, true});
    buckets.push_back({257, 512, 8, true});
    buckets.push_back({513, kPivotMax, 16, true});
    buckets.push_back({kPivotMax + 1, 0xFFFFFFFFu, 0, false});
  } else if (variant == KCL_VARIANT_MERGE || k > (uint32_t)kBitmapMaxK) {
    buckets.push_back({need, 0xFFFFFFFFu, 0, false});
  } else {
    buckets.push_back({need, 64, 1, false});
    buckets.push_back({65, 128, 2, false});
    buckets.push_back({129, 256, 4, false});
    buckets.push_back({257, 512, 8, false});
    buckets.push_back({513, 0xFFFFFFFFu, 0, false});
  }
  DevBuf flags, roots, scratch, sortbuf, claimbuf;
  const float us_per_tick = 1000.f / (float)c->wall_clock_khz;
  TCX_TRY(flags.reserve((size_t)g->n + 1));
  TCX_TRY(roots.reserve(4 * ((size_t)g->n + 1)));
  TCX_TRY(claimbuf.reserve(64));
  TCX_TRY(c->accum.reserve(64));
  unsigned long long* acc = c->accum.as<unsigned long long>();
  TCX_HIP(hipMemsetAsync(acc, 0, 8, c->stream));
  float total_ms = 0;
  for (const Bucket& b : buckets) {
    uint32_t lo = std::max(b.lo, need);
    if (lo > b.hi) continue;
    const uint8_t* path = plan ? plan->path : nullptr;
    const uint8_t want = b.w == 0 ? 1 : 0;
    hipLaunchKernelGGL(k_deg_flags, dim3(grid_for(g->n)), dim3(256), 0, c->stream, rp, g->n, lo,
                       b.hi, path, want, flags.as<uint8_t>());
    TCX_HIP(hipGetLastError());
    uint64_t nr = 0;
    TCX_TRY(select_indices(c, flags.as<uint8_t>(), g->n, roots.as<uint32_t>(), &nr));
    if (nr == 0) continue;
    // each launch takes its roots in the plan's order, or (pivot launches
    // without a plan) widest first; sorted before the timed region
    const uint32_t* order = roots.as<uint32_t>();
    if (!plan && b.pivot) {
      TCX_TRY(sortbuf.reserve(12 * (nr + 1)));
      uint32_t* keys = sortbuf.as<uint32_t>();
      uint32_t* ktmp = keys + nr + 1;
      uint32_t* vtmp = ktmp + nr + 1;
      hipLaunchKernelGGL(k_root_keys, dim3(grid_for(nr)), dim3(256), 0, c->stream, rp,
                         roots.as<uint32_t>(), nr, b.hi, keys);
      TCX_HIP(hipGetLastError());
      uint32_t *sk = nullptr, *sv = nullptr;
      TCX_TRY(sort_pairs_u32_u32(c, keys, ktmp, roots.as<uint32_t>(), vtmp, nr, 11, &sk, &sv));
      order = sv;
    } else if (plan) {
      TCX_TRY(sortbuf.reserve(12 * (nr + 1)));
      uint32_t* keys = sortbuf.as<uint32_t>();
      uint32_t* ktmp = keys + nr + 1;
      uint32_t* vtmp = ktmp + nr + 1;
      hipLaunchKernelGGL(k_plan_keys, dim3(grid_for(nr)), dim3(256), 0, c->stream,
                         roots.as<uint32_t>(), nr, plan->key, keys);
      TCX_HIP(hipGetLastError());
      uint32_t *sk = nullptr, *sv = nullptr;
      TCX_TRY(sort_pairs_u32_u32(c, keys, ktmp, roots.as<uint32_t>(), vtmp, nr, 32, &sk, &sv));
      order = sv;
    }
    TCX_HIP(hipEventRecord(c->ev0, c->stream));
    if (b.w == 0) {
      uint32_t G = (uint32_t)std::min<uint64_t>(nr, (uint64_t)c->num_cus * 8);
      uint32_t maxd = std::max<uint32_t>(g->max_out, 1);
      uint32_t slots = k > 3 ? k - 3 : 1;
      TCX_TRY(scratch.reserve(4ull * G * slots * maxd));
      // cost: CU-microseconds (one wave of the num_cus * 8 holds 1/8 of a CU)
      hipLaunchKernelGGL(k_kcl_merge, dim3(G), dim3(64), 0, c->stream, rp, col, order, nr, k,
                         maxd, slots, scratch.as<uint32_t>(), acc, cost, us_per_tick / 8.f);
    } else if (b.pivot) {
      // workgroups resident per CU (LDS- and register-bound for the wide
      // matrices); the wide ones run 8 or 16 walkers per root, and a root
      // is repeated `copies` times when few roots would leave CUs idle
      const uint64_t per_cu = b.w == 1 ? 32 : (b.w <= 4 ? 16 : (b.w == 8 ? 2 : 1));
      const uint64_t nw = b.w <= 4 ? 1 : (b.w == 8 ? 8 : 16);
      const uint64_t slots = (uint64_t)c->num_cus * per_cu;
      const uint64_t max_copies = b.w <= 4 ? 1 : (b.w == 8 ? 8 : 32);
      const uint32_t copies =
          (uint32_t)std::min<uint64_t>(max_copies, std::max<uint64_t>(1, (4 * slots + nr - 1) / nr));
      const uint32_t G = (uint32_t)std::min<uint64_t>(nr * copies, slots);
      const uint64_t D = 64ull * b.w;
      TCX_TRY(scratch.reserve(8ull * G * nw * (D + 1) * (2ull * b.w + 1)));
      uint64_t* stk = scratch.as<uint64_t>();
      const uint32_t* rts = order;  // widest first (stable: ties in id order)
      // cost: CU-microseconds (a workgroup holds 1/per_cu of a CU)
      const float cs = us_per_tick / (float)per_cu;
      unsigned long long* claim = claimbuf.as<unsigned long long>();
      TCX_HIP(hipMemsetAsync(claim, 0, 8, c->stream));
      // narrow roots are cheap: claim several per atomic
      const uint32_t chunk = b.w == 1 ? 16 : (b.w == 2 ? 8 : (b.w == 4 ? 4 : 1));
      switch (b.w) {
        case 1: hipLaunchKernelGGL((k_kcl_pivot<1, 1>), dim3(G), dim3(64), 0, c->stream, rp, col, rts, nr, copies, chunk, k, stk, acc, claim, cost, cs); break;
        case 2: hipLaunchKernelGGL((k_kcl_pivot<2, 1>), dim3(G), dim3(64), 0, c->stream, rp, col, rts, nr, copies, chunk, k, stk, acc, claim, cost, cs); break;
        case 4: hipLaunchKernelGGL((k_kcl_pivot<4, 1>), dim3(G), dim3(64), 0, c->stream, rp, col, rts, nr, copies, chunk, k, stk, acc, claim, cost, cs); break;
        case 8: hipLaunchKernelGGL((k_kcl_pivot<8, 8>), dim3(G), dim3(512), 0, c->stream, rp, col, rts, nr, copies, chunk, k, stk, acc, claim, cost, cs); break;
        default: hipLaunchKernelGGL((k_kcl_pivot<16, 16>), dim3(G), dim3(1024), 0, c->stream, rp, col, rts, nr, copies, chunk, k, stk, acc, claim, cost, cs); break;
      }
    } else {
      uint32_t G = (uint32_t)std::min<uint64_t>(nr, (uint64_t)c->num_cus * 16);
      switch (b.w) {
        case 1: hipLaunchKernelGGL(k_kcl_bitmap<1>, dim3(G), dim3(64), 0, c->stream, rp, col, roots.as<uint32_t>(), nr, k, acc); break;
        case 2: hipLaunchKernelGGL(k_kcl_bitmap<2>, dim3(G), dim3(64), 0, c->stream, rp, col, roots.as<uint32_t>(), nr, k, acc); break;
        case 4: hipLaunchKernelGGL(k_kcl_bitmap<4>, dim3(G), dim3(64), 0, c->stream, rp, col, roots.as<uint32_t>(), nr, k, acc); break;
        default: hipLaunchKernelGGL(k_kcl_bitmap<8>, dim3(G), dim3(64), 0, c->stream, rp, col, roots.as<uint32_t>(), nr, k, acc); break;
      }
    }
    TCX_HIP(hipGetLastError());
    TCX_HIP(hipEventRecord(c->ev1, c->stream));
    TCX_HIP(hipEventSynchronize(c->ev1));
    float ms = 0;
    TCX_HIP(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    total_ms += ms;
    if (ms > tm.ms_dominant) {
      tm.ms_dominant = ms;
      tm.dominant_kernel = (uint32_t)b.w;
    }
    tm.launches++;
  }
  unsigned long long h = 0;
  TCX_HIP(hipMemcpyAsync(&h, acc, 8, hipMemcpyDeviceToHost, c->stream));
  TCX_HIP(hipStreamSynchronize(c->stream));
  *count = h;
  tm.ms_count = total_ms;
  if (timing) *timing = tm;
  return TC_OK;
}

int kcl_check(tc_ctx* c, tc_graph* g, uint32_t k) {
  if (!c || !g) return fail(TC_ERR_ARG, "NULL argument");
  if (g->ctx != c) return fail(TC_ERR_ARG, "graph belongs to another context");
  if (g->orient_mode < 0) return fail(TC_ERR_STATE, "graph must be oriented before counting");
  if (k < 3 || k > (uint32_t)kMaxK) return fail(TC_ERR_RANGE, "k must be in 3..%d", kMaxK);
  TCX_HIP(hipSetDevice(c->device));
  return TC_OK;
}

// per-root features (device, n x 8) for k
int root_features(tc_ctx* c, tc_graph* g, uint32_t k, DevBuf& X) {
  const uint32_t n = g->n;
  DevBuf tri;
  TCX_TRY(tri.reserve(8 * ((size_t)n + 1)));
  TCX_TRY(X.reserve(4ull * kKclFeatures * ((size_t)n + 1)));
  TCX_TRY(per_vertex_counts(c, g, tri.as<unsigned long long>()));
  hipLaunchKernelGGL(k_kcl_features, dim3(grid_for(n)), dim3(256), 0, c->stream,
                     g->rowptr.as<uint32_t>(), g->col.as<uint32_t>(),
                     tri.as<unsigned long long>(), n, k, X.as<float>());
  TCX_HIP(hipGetLastError());
  TCX_HIP(hipStreamSynchronize(c->stream));
  return TC_OK;
}

// Plan for an explicit root list: listed roots with d+ >= need take the
// pivot path (d+ <= kPivotMax) or the merge path, widest first; path must
// hold 255 everywhere before the call.
__global__ void k_roots_plan(const uint32_t* __restrict__ rp, const uint32_t* __restrict__ roots,
                             uint64_t nroots, uint32_t n, uint32_t need, uint8_t* __restrict__ path,
                             uint32_t* __restrict__ key) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < nroots;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t u = roots[i];
    if (u >= n) continue;
    const uint32_t d = rp[u + 1] - rp[u];
    if (d < need) continue;
    path[u] = d <= kPivotMax ? 0 : 1;
    key[u] = 0xFFFFFFFFu - d;
  }
}

__global__ void k_fill_u8(uint8_t* __restrict__ p, uint64_t n, uint8_t v) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

// Roots wider than the LDS bit matrix would take the merge path.  Every
// acyclic orientation gives the same count, and the degeneracy order bounds
// each out-degree by the degeneracy: *out is a degeneracy-oriented copy of g
// (built once, kept with g's schedules) when g's orientation leaves such
// roots and the copy is narrower, else g itself.  prep_ms receives the
// copy's build time (0 when kept).
int narrow_graph(tc_ctx* c, tc_graph* g, tc_graph** out, float* prep_ms) {
  *out = g;
  *prep_ms = 0;
  if (g->max_out <= kPivotMax || g->orient_mode == TC_ORIENT_DEGENERACY) return TC_OK;
  SchedCache& sc = g->sc;
  if (sc.kcl_degen_version != g->sched_version || !sc.kcl_degen) {
    const auto t0 = std::chrono::steady_clock::now();
    tc_graph* h = nullptr;
    TCX_TRY(graph_clone_records(c, g, &h));
    std::shared_ptr<tc_graph> keep(h, [](tc_graph* x) { tc_graph_destroy(x); });
    TCX_TRY(tc_graph_orient(c, h, TC_ORIENT_DEGENERACY, 0));
    sc.kcl_degen = keep;
    sc.kcl_degen_version = g->sched_version;
    *prep_ms =
        std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - t0).count();
  }
  if (sc.kcl_degen->max_out < g->max_out) *out = sc.kcl_degen.get();
  return TC_OK;
}

// The learned plan of (graph, k, model), built once and kept with the
// graph's schedules (rebuilt when the CSR, k or the model's parameters
// change).  prep_ms (optional) receives the build time, 0 when cached.
int cached_plan(tc_ctx* c, tc_graph* g, uint32_t k, tc_model* m, float* prep_ms) {
  SchedCache& sc = g->sc;
  if (prep_ms) *prep_ms = 0;
  if (sc.kcl_version == g->sched_version && sc.kcl_k == k && sc.kcl_model == model_id(m))
    return TC_OK;
  const uint32_t n = g->n;
  const auto t0 = std::chrono::steady_clock::now();
  DevBuf X, Y;
  TCX_TRY(root_features(c, g, k, X));
  TCX_TRY(Y.reserve(8 * ((size_t)n + 1)));
  TCX_TRY(model_predict(c, m, X.as<float>(), n, Y.as<float>()));
  TCX_TRY(sc.kcl_path.reserve((size_t)n + 1));
  TCX_TRY(sc.kcl_key.reserve(4 * ((size_t)n + 1)));
  hipLaunchKernelGGL(k_kcl_plan, dim3(grid_for(n)), dim3(256), 0, c->stream,
                     g->rowptr.as<uint32_t>(), n, k - 1, Y.as<float>(),
                     sc.kcl_path.as<uint8_t>(), sc.kcl_key.as<uint32_t>());
  TCX_HIP(hipGetLastError());
  TCX_HIP(hipStreamSynchronize(c->stream));
  sc.kcl_version = g->sched_version;
  sc.kcl_k = k;
  sc.kcl_model = model_id(m);
  if (prep_ms)
    *prep_ms = std::chrono::duration<float, std::milli>(std::chrono::steady_clock::now() - t0)
                   .count();
  return TC_OK;
}

}  // namespace

extern "C" {

int kcl_root_features(tc_ctx* c, tc_graph* g, uint32_t k, float* X) {
  TCX_TRY(kcl_check(c, g, k));
  if (!X) return fail(TC_ERR_ARG, "NULL output");
  if (g->n == 0) return TC_OK;
  DevBuf dX;
  TCX_TRY(root_features(c, g, k, dX));
  TCX_HIP(hipMemcpyAsync(X, dX.p, 4ull * kKclFeatures * g->n, hipMemcpyDeviceToHost, c->stream));
  TCX_HIP(hipStreamSynchronize(c->stream));
  return TC_OK;
}

int kcl_measure_costs(tc_ctx* c, tc_graph* g, uint32_t k, int variant, float* cost_us,
                      uint64_t* count) {
  TCX_TRY(kcl_check(c, g, k));
  if (!cost_us) return fail(TC_ERR_ARG, "NULL output");
  if (variant != KCL_VARIANT_PIVOT && variant != KCL_VARIANT_MERGE)
    return fail(TC_ERR_ARG, "costs are measured for KCL_VARIANT_PIVOT or KCL_VARIANT_MERGE");
  const uint32_t n = g->n;
  if (count) *count = 0;
  if (n == 0) return TC_OK;
  DevBuf path, key, cost;
  TCX_TRY(path.reserve((size_t)n + 1));
  TCX_TRY(key.reserve(4 * ((size_t)n + 1)));
  TCX_TRY(cost.reserve(4 * ((size_t)n + 1)));
  TCX_HIP(hipMemsetAsync(cost.p, 0, 4 * (size_t)n, c->stream));
  const bool pivot = variant == KCL_VARIANT_PIVOT;
  hipLaunchKernelGGL(k_fixed_plan, dim3(grid_for(n)), dim3(256), 0, c->stream,
                     g->rowptr.as<uint32_t>(), n, k - 1, pivot ? kPivotMax : 0xFFFFFFFFu,
                     (uint8_t)(pivot ? 0 : 1), path.as<uint8_t>(), key.as<uint32_t>());
  TCX_HIP(hipGetLastError());
  uint64_t total = 0;
  if (g->m_or) {
    const KclPlan plan{path.as<uint8_t>(), key.as<uint32_t>()};
    TCX_TRY(kcl_run(c, g, k, variant, &plan, cost.as<float>(), &total, nullptr));
  }
  TCX_HIP(hipMemcpyAsync(cost_us, cost.p, 4 * (size_t)n, hipMemcpyDeviceToHost, c->stream));
  TCX_HIP(hipStreamSynchronize(c->stream));
  if (count) *count = total;
  return TC_OK;
}

int kcl_learned_plan(tc_ctx* c, tc_graph* g, uint32_t k, tc_model* m, uint8_t* path) {
  TCX_TRY(kcl_check(c, g, k));
  if (!m || !path) return fail(TC_ERR_ARG, "NULL argument");
  const uint32_t n = g->n;
  if (n == 0) return TC_OK;
  TCX_TRY(cached_plan(c, g, k, m, nullptr));
  TCX_HIP(hipMemcpyAsync(path, g->sc.kcl_path.p, n, hipMemcpyDeviceToHost, c->stream));
  TCX_HIP(hipStreamSynchronize(c->stream));
  return TC_OK;
}

int kcl_count_learned(tc_ctx* c, tc_graph* g, uint32_t k, tc_model* m, uint64_t* count,
                      tc_timing* timing) {
  if (!count) return fail(TC_ERR_ARG, "NULL argument");
  if (!m) return kcl_count(c, g, k, KCL_VARIANT_AUTO, count, timing);
  TCX_TRY(kcl_check(c, g, k));
  *count = 0;
  tc_timing tm;
  memset(&tm, 0, sizeof(tm));
  const uint32_t n = g->n;
  if (g->m_or == 0 || n == 0) {
    if (timing) *timing = tm;
    return TC_OK;
  }
  float copy_ms = 0, prep_ms = 0;
  tc_graph* h = nullptr;
  TCX_TRY(narrow_graph(c, g, &h, &copy_ms));
  TCX_TRY(cached_plan(c, h, k, m, &prep_ms));
  const KclPlan plan{h->sc.kcl_path.as<uint8_t>(), h->sc.kcl_key.as<uint32_t>()};
  TCX_TRY(kcl_run(c, h, k, KCL_VARIANT_AUTO, &plan, nullptr, count, &tm));
  tm.ms_prep = copy_ms + prep_ms;
  if (timing) *timing = tm;
  return TC_OK;
}

int kcl_count(tc_ctx* c, tc_graph* g, uint32_t k, int variant, uint64_t* count,
              tc_timing* timing) {
  if (!c || !g || !count) return fail(TC_ERR_ARG, "NULL argument");
  if (g->ctx != c) return fail(TC_ERR_ARG, "graph belongs to another context");
  if (g->orient_mode < 0) return fail(TC_ERR_STATE, "graph must be oriented before counting");
  if (k == 0 || k > (uint32_t)kMaxK) return fail(TC_ERR_RANGE, "k must be in 1..%d", kMaxK);
  if (variant < KCL_VARIANT_AUTO || variant > KCL_VARIANT_PIVOT)
    return fail(TC_ERR_ARG, "unknown k-clique variant %d", variant);
  TCX_HIP(hipSetDevice(c->device));
  tc_timing tm;
  memset(&tm, 0, sizeof(tm));
  *count = 0;
  if (k == 1) { *count = g->n; if (timing) *timing = tm; return TC_OK; }
  if (k == 2) { *count = g->m_or; if (timing) *timing = tm; return TC_OK; }
  if (g->m_or == 0) { if (timing) *timing = tm; return TC_OK; }
  if (k == 3 && variant == KCL_VARIANT_AUTO) {
    // 3-cliques are triangles: the AUTO triangle engine counts them directly
    tc_sched s;
    memset(&s, 0, sizeof(s));
    s.kernel = TC_KERNEL_AUTO;
    return tc_count(c, g, &s, count, timing);
  }
  if (variant == KCL_VARIANT_AUTO) {
    float prep_ms = 0;
    tc_graph* h = nullptr;
    TCX_TRY(narrow_graph(c, g, &h, &prep_ms));
    if (h != g) {
      TCX_TRY(kcl_run(c, h, k, variant, nullptr, nullptr, count, timing));
      if (timing) timing->ms_prep += prep_ms;
      return TC_OK;
    }
  }
  return kcl_run(c, g, k, variant, nullptr, nullptr, count, timing);
}

int kcl_count_roots(tc_ctx* c, tc_graph* g, uint32_t k, const uint32_t* roots, uint64_t nroots,
                    uint64_t* count, tc_timing* timing) {
  TCX_TRY(kcl_check(c, g, k));
  if (!count || (nroots && !roots)) return fail(TC_ERR_ARG, "NULL argument");
  *count = 0;
  tc_timing tm;
  memset(&tm, 0, sizeof(tm));
  const uint32_t n = g->n;
  if (g->m_or == 0 || n == 0 || nroots == 0) {
    if (timing) *timing = tm;
    return TC_OK;
  }
  for (uint64_t i = 0; i < nroots; ++i)
    if (roots[i] >= n) return fail(TC_ERR_RANGE, "root %u >= n = %u", roots[i], n);
  DevBuf droots, path, key;
  TCX_TRY(droots.reserve(4 * nroots));
  TCX_TRY(path.reserve((size_t)n + 1));
  TCX_TRY(key.reserve(4 * ((size_t)n + 1)));
  TCX_HIP(hipMemcpyAsync(droots.p, roots, 4 * nroots, hipMemcpyHostToDevice, c->stream));
  hipLaunchKernelGGL(k_fill_u8, dim3(grid_for(n)), dim3(256), 0, c->stream, path.as<uint8_t>(),
                     (uint64_t)n, (uint8_t)255);
  hipLaunchKernelGGL(k_roots_plan, dim3(grid_for(nroots)), dim3(256), 0, c->stream,
                     g->rowptr.as<uint32_t>(), droots.as<uint32_t>(), nroots, n, k - 1,
                     path.as<uint8_t>(), key.as<uint32_t>());
  TCX_HIP(hipGetLastError());
  const KclPlan plan{path.as<uint8_t>(), key.as<uint32_t>()};
  TCX_TRY(kcl_run(c, g, k, KCL_VARIANT_AUTO, &plan, nullptr, count, timing));
  return TC_OK;
}

int kcl_gpu_solver(const char* path, uint32_t k, uint64_t* total, size_t n_chunk) {
  if (!path || !total) return fail(TC_ERR_ARG, "NULL argument");
  if (n_chunk == 0) return fail(TC_ERR_ARG, "n_chunk must be at least 1");
  uint64_t m = 0;
  uint32_t n = 0;
  const size_t len = strlen(path);
  const bool gr = len >= 3 && strcmp(path + len - 3, ".gr") == 0;
  std::vector<uint32_t> s, d;
  if (gr) {
    TCX_TRY(tc_read_gr(path, nullptr, nullptr, 0, &m, &n));
    s.resize(std::max<uint64_t>(m, 1));
    d.resize(std::max<uint64_t>(m, 1));
    TCX_TRY(tc_read_gr(path, s.data(), d.data(), m, &m, &n));
  } else {
    TCX_TRY(tc_file_edge_count(path, &m));
    s.resize(std::max<uint64_t>(m, 1));
    d.resize(std::max<uint64_t>(m, 1));
    TCX_TRY(tc_read_edges(path, s.data(), d.data(), m, &m));
  }
  tc_ctx* c = nullptr;
  TCX_TRY(tc_ctx_create(0, &c));
  tc_graph* g = nullptr;
  int st = tc_graph_create(c, s.data(), d.data(), m, n, &g);
  // degeneracy order: every out-degree is at most the degeneracy, which
  // keeps the roots on the LDS pivot path
  if (st == TC_OK) st = tc_graph_orient(c, g, TC_ORIENT_DEGENERACY, 0);
  if (st == TC_OK && (n_chunk == 1 || k < 3)) {
    st = kcl_count(c, g, k, KCL_VARIANT_AUTO, total, nullptr);
  } else if (st == TC_OK) {
    // k-clique/kcl.cu:163-209 splits the DAG's edges into N_CHUNK contiguous
    // ranges and counts the cliques whose first edge lies in each range in
    // turn.  Here a root's out-edges stay together: root u belongs to chunk
    // rowptr[u] / ceil(m / N_CHUNK), and the chunks are counted one after the
    // other through kcl_count_roots.
    const uint32_t nv = g->n;
    std::vector<uint32_t> rp((size_t)nv + 1);
    st = tc_graph_get_csr(g, rp.data(), nullptr);
    const uint64_t mo = g->m_or;
    const uint64_t len = mo ? (mo - 1) / n_chunk + 1 : 1;
    std::vector<std::vector<uint32_t>> chunks(n_chunk);
    for (uint32_t u = 0; st == TC_OK && u < nv; ++u)
      if (rp[u + 1] > rp[u]) chunks[std::min<uint64_t>(rp[u] / len, n_chunk - 1)].push_back(u);
    uint64_t sum = 0;
    for (size_t j = 0; st == TC_OK && j < n_chunk; ++j) {
      uint64_t part = 0;
      if (!chunks[j].empty())
        st = kcl_count_roots(c, g, k, chunks[j].data(), chunks[j].size(), &part, nullptr);
      sum += part;
    }
    if (st == TC_OK) *total = sum;
  }
  tc_graph_destroy(g);
  tc_ctx_destroy(c);
  return st;
}

}  // extern "C"
