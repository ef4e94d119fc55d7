// Device graph: upload, simplification, orientation (ID / degree / A-direction
// / k-core), A-order relabelling and oriented-CSR construction, all in HBM.
//
// Layout in HBM (one graph):
//   src, dst   : uint32[m_raw]   raw records as read from the edge file
//   rowptr     : uint32[n+1]     oriented CSR offsets
//   col        : uint32[m+]      oriented CSR targets, ascending per row
//   esrc       : uint32[m+]      source of each CSR edge (edge-centric kernels)
//   vkey       : int32[n]        A-direction priority or core number
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>

#include "engine.h"
#include "internal.h"

namespace tcx {
namespace {

constexpr int TB = 256;

inline unsigned grid_for(uint64_t n, int tb = TB) {
  uint64_t b = (n + tb - 1) / tb;
  if (b < 1) b = 1;
  if (b > (1u << 20)) b = 1u << 20;
  return (unsigned)b;
}

__device__ inline uint64_t gtid() { return (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; }
__device__ inline uint64_t gstride() { return (uint64_t)gridDim.x * blockDim.x; }

__global__ void k_fill_u32(uint32_t* p, uint32_t v, uint64_t n) {
  for (uint64_t i = gtid(); i < n; i += gstride()) p[i] = v;
}
__global__ void k_fill_i32(int32_t* p, int32_t v, uint64_t n) {
  for (uint64_t i = gtid(); i < n; i += gstride()) p[i] = v;
}

// Degree over the raw records (duplicates and self-loops counted),
// Hu/preTC.cu:119-123, Bisson/preTC.cu:65-76, tricount_gpu.cu:37-47.
// (prune) flag[v] = deg[v] <= 1, Hu/preTC.cu:139-144.
__global__ void k_low_degree_flag(const uint32_t* deg, uint32_t n, uint32_t* flag) {
  for (uint64_t v = gtid(); v < n; v += gstride()) flag[v] = deg[v] <= 1 ? 1u : 0u;
}

// Orient one raw record into a packed (src<<32 | dst) key; self-loops and
// pruned endpoints become the sentinel (n_out << 32), which sorts last.
// tie (optional, TC_ORIENT_DEGREE): equal degrees compare tie[] before ids.
__global__ void k_orient_keys(const uint32_t* __restrict__ s, const uint32_t* __restrict__ d,
                              uint64_t m, const uint32_t* __restrict__ deg, int mode,
                              const uint32_t* __restrict__ pruned_before, uint32_t n_out,
                              const uint32_t* __restrict__ tie, uint64_t* __restrict__ keys) {
  const uint64_t sentinel = (uint64_t)n_out << 32;
  for (uint64_t i = gtid(); i < m; i += gstride()) {
    uint32_t a = s[i], b = d[i];
    bool swap;
    if (mode == TC_ORIENT_ID) {
      swap = a < b;  // src = larger id (Fox/preTC.cu:94-98)
    } else if (mode == TC_ORIENT_DEGREE) {
      uint32_t da = deg[a], db = deg[b];
      if (tie && da == db && tie[a] != tie[b]) swap = tie[a] > tie[b];
      else swap = da > db || (da == db && a < b);  // Hu/preTC.cu:129
    } else {
      uint32_t da = deg[a], db = deg[b];
      swap = da > db || (da == db && a > b);  // tricount_gpu.cu:76
    }
    uint32_t os = swap ? b : a, od = swap ? a : b;
    uint64_t key;
    if (os == od) {
      key = sentinel;
    } else if (pruned_before) {
      if (deg[os] <= 1 || deg[od] <= 1) key = sentinel;  // Hu/preTC.cu:151-160
      else key = ((uint64_t)(os - pruned_before[os]) << 32) | (od - pruned_before[od]);
    } else {
      key = ((uint64_t)os << 32) | od;
    }
    keys[i] = key;
  }
}

// Canonical simple-edge key (min<<32 | max); self-loops -> sentinel.
__global__ void k_simple_keys(const uint32_t* __restrict__ s, const uint32_t* __restrict__ d,
                              uint64_t m, uint32_t n, uint64_t* __restrict__ keys) {
  for (uint64_t i = gtid(); i < m; i += gstride()) {
    uint32_t a = s[i], b = d[i];
    uint32_t lo = a < b ? a : b, hi = a < b ? b : a;
    keys[i] = a == b ? ((uint64_t)n << 32) : (((uint64_t)lo << 32) | hi);
  }
}

// Both directions of every non-loop record (A-direction-seri.cpp:215-221).
__global__ void k_sym_keys(const uint32_t* __restrict__ s, const uint32_t* __restrict__ d,
                           uint64_t m, uint32_t n, uint64_t* __restrict__ keys) {
  for (uint64_t i = gtid(); i < m; i += gstride()) {
    uint32_t a = s[i], b = d[i];
    if (a == b) {
      keys[2 * i] = (uint64_t)n << 32;
      keys[2 * i + 1] = (uint64_t)n << 32;
    } else {
      keys[2 * i] = ((uint64_t)a << 32) | b;
      keys[2 * i + 1] = ((uint64_t)b << 32) | a;
    }
  }
}

__global__ void k_unzip(const uint64_t* __restrict__ keys, uint64_t m, uint32_t* __restrict__ s,
                        uint32_t* __restrict__ d) {
  for (uint64_t i = gtid(); i < m; i += gstride()) {
    uint64_t k = keys[i];
    s[i] = (uint32_t)(k >> 32);
    d[i] = (uint32_t)k;
  }
}

// Row offsets + targets + sources from keys sorted by (src,dst); one thread
// per key plus one for the tail (tricount_gpu.cu:127-137 builds the same
// offsets from a sorted first-endpoint array).
// col (and esrc) from sorted (src << 32 | dst) keys; rowptr comes from
// dev::k_rowptr_search, so runs of empty rows cost nothing extra.
__global__ void k_csr_cols(const uint64_t* __restrict__ keys, uint64_t m,
                           uint32_t* __restrict__ col, uint32_t* __restrict__ esrc) {
  for (uint64_t i = gtid(); i < m; i += gstride()) {
    const uint64_t k = keys[i];
    col[i] = (uint32_t)k;
    if (esrc) esrc[i] = (uint32_t)(k >> 32);
  }
}

__device__ inline unsigned long long wave_sum_u64(unsigned long long v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}

// acc[0] += sum d+^2, acc[1] += sum d- d+, acc[2] = max d+
__global__ void k_csr_stats(const uint32_t* __restrict__ rowptr, const uint32_t* __restrict__ indeg,
                            uint32_t n, unsigned long long* acc) {
  unsigned long long s1 = 0, s2 = 0;
  uint32_t mx = 0;
  for (uint64_t v = gtid(); v < n; v += gstride()) {
    uint64_t d = rowptr[v + 1] - rowptr[v];
    s1 += d * d;
    s2 += d * (uint64_t)indeg[v];
    mx = d > mx ? (uint32_t)d : mx;
  }
  s1 = wave_sum_u64(s1);
  s2 = wave_sum_u64(s2);
  for (int off = 32; off > 0; off >>= 1) {
    uint32_t o = __shfl_down(mx, off, 64);
    mx = o > mx ? o : mx;
  }
  if ((threadIdx.x & 63) == 0) {
    atomicAdd(acc + 0, s1);
    atomicAdd(acc + 1, s2);
    atomicMax(acc + 2, (unsigned long long)mx);
  }
}

// ---- A-direction peel (edge_directing/A-direction-seri.cpp:35-153) --------
//
// The serial peel assigns a vertex the label of the frontier level during
// which its runtime degree first drops to the threshold, so processing one
// whole level in parallel reproduces its labels exactly.  Runtime degrees
// start at twice the simple degree (each neighbour is listed twice, :44-73)
// and each peeled neighbour removes 2.
__global__ void k_adir_init(const uint32_t* __restrict__ rp, uint32_t n, int32_t* rtd,
                            int32_t* prio) {
  for (uint64_t v = gtid(); v < n; v += gstride()) {
    rtd[v] = 2 * (int32_t)(rp[v + 1] - rp[v]);
    prio[v] = -1;
  }
}

__global__ void k_adir_scan(uint32_t n, const int32_t* __restrict__ rtd, int32_t* prio, float t,
                            int32_t label, uint32_t* frontier, uint32_t* fcount) {
  for (uint64_t v = gtid(); v < n; v += gstride()) {
    if (prio[v] == -1 && (float)rtd[v] <= t) {
      prio[v] = label;
      frontier[atomicAdd(fcount, 1u)] = (uint32_t)v;
    }
  }
}

// A peeled neighbour takes 2 from the runtime degree (it is listed twice in
// allEdges); a vertex is peeled when its degree first reaches the threshold,
// which makes the labels independent of the order of the frontier.
__global__ void k_adir_expand(const uint32_t* __restrict__ frontier, uint32_t fsize,
                              const uint32_t* __restrict__ rp, const uint32_t* __restrict__ col,
                              int32_t* rtd, int32_t* prio, float t, int32_t label,
                              uint32_t* next, uint32_t* ncount) {
  const uint32_t lane = threadIdx.x & 63;
  const uint64_t wave = gtid() >> 6, nwaves = gstride() >> 6;
  for (uint64_t f = wave; f < fsize; f += nwaves) {
    uint32_t v = frontier[f];
    for (uint32_t e = rp[v] + lane; e < rp[v + 1]; e += 64) {
      uint32_t y = col[e];
      if (__hip_atomic_load(prio + y, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != -1) continue;
      int32_t old = atomicSub(rtd + y, 2);
      if ((float)old > t && (float)(old - 2) <= t) {
        prio[y] = label;
        next[atomicAdd(ncount, 1u)] = y;
      }
    }
  }
}

// ---- Bisson/A-direction.cpp degree record ----------------------------------
// Bisson's peel (Bisson/A-direction.cpp:43-163) records degreeRecord[v] =
// runTimeDegree[v] when the scan peels v (:96), but a cascade push writes the
// slot of the pushing frontier element's INDEX, not of the pushed vertex
// (:118 writes degreeRecord[i]).  The record therefore depends on the serial
// frontier order, which these kernels rebuild: a round's first frontier is
// ascending ids (:91-98); the next frontier lists pushed vertices in the order
// of their crossing decrement, i.e. by (pusher index p, position e of the
// vertex in the pusher's allEdges list).  allEdges[x] (:75-80 over the (i,
// nbr) pairs of :253-258) is [N(x) below x] + [N(x)] + [N(x) above x], so a
// neighbour y of rank r in N(x), with L neighbours below x, sits at
// e1 = r, e2 = L + r when y < x and at e1 = L + r, e2 = deg(x) + r when y > x.

__global__ void k_adir_scan_flags(uint32_t n, const int32_t* __restrict__ rtd,
                                  const int32_t* __restrict__ prio, float t,
                                  uint8_t* __restrict__ flags) {
  for (uint64_t v = gtid(); v < n; v += gstride())
    flags[v] = (prio[v] == -1 && (float)rtd[v] <= t) ? 1 : 0;
}

__global__ void k_adir_mark_scan(const uint32_t* __restrict__ F, uint32_t fs,
                                 const int32_t* __restrict__ rtd, int32_t* __restrict__ prio,
                                 int32_t label, int32_t* __restrict__ rec) {
  for (uint64_t i = gtid(); i < fs; i += gstride()) {
    const uint32_t v = F[i];
    prio[v] = label;
    rec[v] = rtd[v];
  }
}

// One event (y << 32 | p) per frontier element p and unpeeled neighbour y
// (each stands for y's two entries in allEdges[F[p]]); rtd[y] -= 2.
__global__ void k_adir_events(const uint32_t* __restrict__ F, uint32_t fs,
                              const uint32_t* __restrict__ rp, const uint32_t* __restrict__ col,
                              int32_t* rtd, const int32_t* __restrict__ prio,
                              uint64_t* __restrict__ ev, unsigned long long* ecount) {
  const uint32_t lane = threadIdx.x & 63;
  const uint64_t wave = gtid() >> 6, nwaves = gstride() >> 6;
  for (uint64_t p = wave; p < fs; p += nwaves) {
    const uint32_t x = F[p], b = rp[x], e = rp[x + 1];
    for (uint32_t base = b; base < e; base += 64) {
      const uint32_t i = base + lane;
      const uint32_t y = i < e ? col[i] : 0;
      const bool live = i < e && prio[y] == -1;
      const uint64_t mask = __ballot(live);
      unsigned long long at = 0;
      if (lane == 0 && mask) at = atomicAdd(ecount, (unsigned long long)__popcll(mask));
      at = __shfl(at, 0, 64);
      if (live) {
        atomicSub(rtd + y, 2);
        ev[at + __popcll(mask & ((1ull << lane) - 1))] = ((uint64_t)y << 32) | p;
      }
    }
  }
}

__device__ inline uint32_t lower_bound_u32(const uint32_t* a, uint32_t lo, uint32_t hi,
                                           uint32_t v) {
  while (lo < hi) {
    const uint32_t mid = lo + ((hi - lo) >> 1);
    if (a[mid] < v) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

// Over the events sorted by (y, p): y's k-th decrement (k = its runtime
// degree at the level's start - floor(t)) peels it, if it has that many.
// The j = ceil(k/2)-th pusher p* made it, at e1 (k odd) or e2; slot p* of
// the record takes floor(t) (the runtime degree right after the crossing).
// Pushed vertices go out with the key (p* << 32 | e*) of the next order.
__global__ void k_adir_cross(const uint64_t* __restrict__ ev, uint64_t E,
                             const uint32_t* __restrict__ F, const uint32_t* __restrict__ rp,
                             const uint32_t* __restrict__ col, const int32_t* __restrict__ rtd,
                             int32_t* __restrict__ prio, int32_t tfloor, int32_t label,
                             int32_t* __restrict__ rec, uint64_t* __restrict__ nkey,
                             uint32_t* __restrict__ nid, uint32_t* ncount) {
  for (uint64_t i = gtid(); i < E; i += gstride()) {
    const uint32_t y = (uint32_t)(ev[i] >> 32);
    if (i > 0 && (uint32_t)(ev[i - 1] >> 32) == y) continue;
    uint64_t lo = i, hi = E;  // first event of a larger y
    const uint64_t bound = ((uint64_t)y + 1) << 32;
    while (lo < hi) {
      const uint64_t mid = lo + ((hi - lo) >> 1);
      if (ev[mid] < bound) lo = mid + 1;
      else hi = mid;
    }
    const int64_t cy = (int64_t)(lo - i);
    const int64_t k = (int64_t)rtd[y] + 2 * cy - tfloor;
    if (k < 1 || k > 2 * cy) continue;
    const uint32_t p = (uint32_t)ev[i + (uint64_t)((k + 1) / 2) - 1];
    const uint32_t x = F[p], b = rp[x], e = rp[x + 1];
    const uint32_t r = lower_bound_u32(col, b, e, y) - b;
    const uint32_t L = lower_bound_u32(col, b, e, x) - b;
    const uint32_t e1 = y < x ? r : L + r;
    const uint32_t e2 = y < x ? L + r : (e - b) + r;
    prio[y] = label;
    rec[p] = tfloor;
    const uint32_t at = atomicAdd(ncount, 1u);
    nkey[at] = ((uint64_t)p << 32) | ((k & 1) ? e1 : e2);
    nid[at] = y;
  }
}

__global__ void k_min_unpeeled(uint32_t n, const int32_t* __restrict__ val,
                               const int32_t* __restrict__ state, int32_t* out) {
  int32_t mn = INT32_MAX;
  for (uint64_t v = gtid(); v < n; v += gstride())
    if (state[v] == -1 && val[v] < mn) mn = val[v];
  for (int off = 32; off > 0; off >>= 1) {
    int32_t o = __shfl_down(mn, off, 64);
    mn = o < mn ? o : mn;
  }
  if ((threadIdx.x & 63) == 0 && mn != INT32_MAX) atomicMin(out, mn);
}

// ---- k-core (edge_directing/kcore_direction.cpp:91-126) --------------------
// Level-synchronous peel: at level k every vertex whose remaining degree is
// <= k is removed with core number k; core numbers are order-independent.
__global__ void k_core_init(const uint32_t* __restrict__ rp, uint32_t n, int32_t* deg,
                            int32_t* core) {
  for (uint64_t v = gtid(); v < n; v += gstride()) {
    deg[v] = (int32_t)(rp[v + 1] - rp[v]);
    core[v] = -1;
  }
}

// `step` (optional) receives the index of the frontier that removes each
// vertex: every frontier is one step, numbered over the whole peel.
__global__ void k_core_scan(uint32_t n, const int32_t* __restrict__ deg, int32_t* core, int32_t k,
                            uint32_t* frontier, uint32_t* fcount, int32_t* step, int32_t t) {
  for (uint64_t v = gtid(); v < n; v += gstride()) {
    if (core[v] == -1 && deg[v] <= k) {
      core[v] = k;
      if (step) step[v] = t;
      frontier[atomicAdd(fcount, 1u)] = (uint32_t)v;
    }
  }
}

__global__ void k_core_expand(const uint32_t* __restrict__ frontier, uint32_t fsize,
                              const uint32_t* __restrict__ rp, const uint32_t* __restrict__ col,
                              int32_t* deg, int32_t* core, int32_t k, uint32_t* next,
                              uint32_t* ncount, int32_t* step, int32_t t) {
  const uint32_t lane = threadIdx.x & 63;
  const uint64_t wave = gtid() >> 6, nwaves = gstride() >> 6;
  for (uint64_t f = wave; f < fsize; f += nwaves) {
    uint32_t v = frontier[f];
    for (uint32_t e = rp[v] + lane; e < rp[v + 1]; e += 64) {
      uint32_t y = col[e];
      if (__hip_atomic_load(core + y, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != -1) continue;
      int32_t old = atomicSub(deg + y, 1);
      if (old == k + 1) {
        core[y] = k;
        if (step) step[y] = t;
        next[atomicAdd(ncount, 1u)] = y;
      }
    }
  }
}

// keep[e] for a sorted symmetric key list: ADIR (mode 2), or KCORE (mode 3)
// and DEGENERACY (mode 5), whose keys both point up from the lower (key, id).
// tie (optional, A-direction): equal priorities compare tie[] before ids.
__global__ void k_sym_keep(const uint64_t* __restrict__ keys, uint64_t m2,
                           const int32_t* __restrict__ key, int mode,
                           const uint32_t* __restrict__ tie, uint8_t* keep) {
  for (uint64_t e = gtid(); e < m2; e += gstride()) {
    uint64_t kk = keys[e];
    uint32_t i = (uint32_t)(kk >> 32), j = (uint32_t)kk;
    int32_t pi = key[i], pj = key[j];
    bool k;
    if (tie && pi == pj && tie[i] != tie[j])
      k = tie[i] < tie[j];
    else if (mode == TC_ORIENT_ADIR || mode == TC_ORIENT_ADIR_OMP)
      k = pj > pi || (pj == pi && i < j);  // A-direction-seri.cpp:259-261, A-direction.cpp:266-268
    else k = pi < pj || (pi == pj && i < j);                         // kcore_direction.cpp:137-138
    keep[e] = k ? 1 : 0;
  }
}

// Tie keys of learned orientation ties: a vertex's forced out-degree (the
// neighbours strictly above it in the primary order: higher A-direction
// priority, or higher degree), ascending or (desc) descending.
__global__ void k_forced_sym(const uint32_t* __restrict__ rp, const uint32_t* __restrict__ col,
                             uint32_t n, const int32_t* __restrict__ key, int desc,
                             uint32_t* __restrict__ tie) {
  for (uint64_t v = gtid(); v < n; v += gstride()) {
    uint32_t h = 0;
    for (uint32_t e = rp[v]; e < rp[v + 1]; ++e) h += key[col[e]] > key[v];
    tie[v] = desc ? ~h : h;
  }
}

// over raw records (every record counts, as the degree does)
__global__ void k_forced_raw(const uint32_t* __restrict__ s, const uint32_t* __restrict__ d,
                             uint64_t m, const uint32_t* __restrict__ deg, uint32_t* h) {
  for (uint64_t i = gtid(); i < m; i += gstride()) {
    const uint32_t a = s[i], b = d[i];
    if (a == b) continue;
    if (deg[b] > deg[a]) atomicAdd(h + a, 1u);
    else if (deg[a] > deg[b]) atomicAdd(h + b, 1u);
  }
}

__global__ void k_tie_flip(uint32_t* __restrict__ h, uint32_t n) {
  for (uint64_t v = gtid(); v < n; v += gstride()) h[v] = ~h[v];
}

// A-direction cost function sum_v |d+(v) - avg| (verification.cpp:43-45's
// f_lp over the oriented CSR): per-block partial sums over a fixed grid, so
// the host's ordered sum of the partials is deterministic.
__global__ __launch_bounds__(256) void k_adir_cost(const uint32_t* __restrict__ rp, uint32_t n,
                                                   double avg, double* __restrict__ partial) {
  __shared__ double red[4];
  double acc = 0;
  for (uint64_t v = gtid(); v < n; v += gstride()) acc += fabs((double)(rp[v + 1] - rp[v]) - avg);
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) partial[blockIdx.x] = ((red[0] + red[1]) + red[2]) + red[3];
}

// A-order helpers: oriented out-degree over raw records (A-order.cpp:88-99)
// and the relabelled "reordered" records (:219-228).
__global__ void k_aorder_outdeg(const uint32_t* __restrict__ s, const uint32_t* __restrict__ d,
                                uint64_t m, const uint32_t* __restrict__ deg, uint32_t* rtd) {
  for (uint64_t i = gtid(); i < m; i += gstride()) {
    uint32_t a = s[i], b = d[i];
    bool swap = deg[a] > deg[b] || (deg[a] == deg[b] && a < b);
    atomicAdd(rtd + (swap ? b : a), 1u);
  }
}

// ---- id relabellings of the format tools ----------------------------------
__global__ void k_present_flags(const uint32_t* __restrict__ deg, uint32_t n,
                                uint32_t* __restrict__ f) {
  for (uint64_t v = gtid(); v < n; v += gstride()) f[v] = deg[v] ? 1u : 0u;
}

__global__ void k_fill_u64(uint64_t* p, uint64_t v, uint64_t n) {
  for (uint64_t i = gtid(); i < n; i += gstride()) p[i] = v;
}

// first[v] = smallest record index whose source is v
__global__ void k_first_src(const uint32_t* __restrict__ s, uint64_t m,
                            unsigned long long* first) {
  for (uint64_t i = gtid(); i < m; i += gstride()) atomicMin(first + s[i], (unsigned long long)i);
}

// adjust2tc order: ids by first appearance as a source, then the others by id
__global__ void k_first_keys(const uint64_t* __restrict__ first, uint32_t n, uint64_t m,
                             uint64_t* __restrict__ key, uint32_t* __restrict__ ids) {
  for (uint64_t v = gtid(); v < n; v += gstride()) {
    key[v] = first[v] != ~0ull ? first[v] : m + v;
    ids[v] = (uint32_t)v;
  }
}

// directionByReorder order: (degree, id) ascending
__global__ void k_degree_keys(const uint32_t* __restrict__ deg, uint32_t n,
                              uint64_t* __restrict__ key, uint32_t* __restrict__ ids) {
  for (uint64_t v = gtid(); v < n; v += gstride()) {
    key[v] = ((uint64_t)deg[v] << 32) | v;
    ids[v] = (uint32_t)v;
  }
}

// Bisson/A-direction.cpp:287-293 order: (priority, runtime degree at peel)
// ascending; ties keep ascending ids (stable sort of ids in order)
__global__ void k_adir_relabel_keys(const int32_t* __restrict__ prio,
                                    const int32_t* __restrict__ rec, uint32_t n,
                                    uint64_t* __restrict__ key, uint32_t* __restrict__ ids) {
  for (uint64_t v = gtid(); v < n; v += gstride()) {
    key[v] = ((uint64_t)(uint32_t)prio[v] << 32) | (uint32_t)rec[v];
    ids[v] = (uint32_t)v;
  }
}

// perm[order[j]] = j, or n - 1 - j when reversed
__global__ void k_rank_perm(const uint32_t* __restrict__ order, uint32_t n, bool reversed,
                            uint32_t* __restrict__ perm) {
  for (uint64_t j = gtid(); j < n; j += gstride())
    perm[order[j]] = reversed ? (uint32_t)(n - 1 - j) : (uint32_t)j;
}

__global__ void k_apply_perm(uint32_t* s, uint32_t* d, uint64_t m,
                             const uint32_t* __restrict__ perm) {
  for (uint64_t i = gtid(); i < m; i += gstride()) {
    s[i] = perm[s[i]];
    d[i] = perm[d[i]];
  }
}

__global__ void k_aorder_apply(uint32_t* s, uint32_t* d, uint64_t m, const uint32_t* __restrict__ deg,
                               const uint32_t* __restrict__ perm) {
  for (uint64_t i = gtid(); i < m; i += gstride()) {
    uint32_t a = s[i], b = d[i];
    bool swap = deg[a] > deg[b] || (deg[a] == deg[b] && a < b);
    uint32_t os = swap ? b : a, od = swap ? a : b;
    s[i] = perm[os];
    d[i] = perm[od];
  }
}

// rowptr + col (+ esrc) of the CSR of m sorted, unique keys over n rows.
int csr_launch(tc_ctx* c, const uint64_t* keys, uint64_t m, uint32_t n, uint32_t* rowptr,
               uint32_t* col, uint32_t* esrc) {
  if (m)
    hipLaunchKernelGGL(k_csr_cols, dim3(grid_for(m)), dim3(TB), 0, c->stream, keys, m, col, esrc);
  hipLaunchKernelGGL(dev::k_rowptr_search<uint64_t>, dim3(grid_for((uint64_t)n + 1)), dim3(TB), 0,
                     c->stream, keys, m, n, rowptr);
  TCX_HIP(hipGetLastError());
  return TC_OK;
}

// h[x] += occurrences of x in a[0, m) (and b[0, m) when b is not null); h
// must be zeroed.
int count_ids(tc_ctx* c, const uint32_t* a, const uint32_t* b, uint64_t m, uint32_t n,
              uint32_t* h) {
  if (m == 0) return TC_OK;
  const unsigned G = (unsigned)std::min<uint64_t>(grid_for(m), (uint64_t)c->num_cus * 4u);
  hipLaunchKernelGGL(dev::k_count_ids<8192>, dim3(G), dim3(256), 0, c->stream, a, b, m, n, h);
  TCX_HIP(hipGetLastError());
  return TC_OK;
}

int read_u32(tc_ctx* c, const uint32_t* dptr, uint32_t* out) {
  TCX_HIP(hipMemcpyAsync(out, dptr, sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
  TCX_HIP(hipStreamSynchronize(c->stream));
  return TC_OK;
}

}  // namespace

// Sort keys (in `keys`, scratch of the same size in `tmp`), drop duplicates
// and the sentinel, build the graph's oriented CSR.
static int csr_from_keys(tc_ctx* c, tc_graph* g, DevBuf& keys, DevBuf& tmp, uint64_t nkeys,
                         uint32_t n_out, bool already_sorted_unique) {
  uint64_t* sorted = keys.as<uint64_t>();
  uint64_t nuniq = nkeys;
  uint64_t* uniq = sorted;
  if (!already_sorted_unique) {
    int end_bit = 32 + bit_width_u32(n_out);
    TCX_TRY(sort_u64(c, keys.as<uint64_t>(), tmp.as<uint64_t>(), nkeys, end_bit, &sorted));
    uint64_t* other = sorted == keys.as<uint64_t>() ? tmp.as<uint64_t>() : keys.as<uint64_t>();
    TCX_TRY(unique_u64(c, sorted, other, nkeys, &nuniq));
    uniq = other;
    if (nuniq > 0) {
      uint64_t last = 0;
      TCX_HIP(hipMemcpyAsync(&last, uniq + nuniq - 1, 8, hipMemcpyDeviceToHost, c->stream));
      TCX_HIP(hipStreamSynchronize(c->stream));
      if ((uint32_t)(last >> 32) == n_out) --nuniq;  // sentinel
    }
  }
  if (nuniq > kMaxCsrEntries)
    return fail(TC_ERR_RANGE, "%llu oriented edges exceed the 32-bit CSR offsets",
                (unsigned long long)nuniq);
  g->n = n_out;
  g->m_or = nuniq;
  TCX_TRY(g->rowptr.reserve(sizeof(uint32_t) * ((size_t)n_out + 1)));
  TCX_TRY(g->col.reserve(sizeof(uint32_t) * std::max<uint64_t>(nuniq, 1)));
  TCX_TRY(g->esrc.reserve(sizeof(uint32_t) * std::max<uint64_t>(nuniq, 1)));
  TCX_TRY(csr_launch(c, uniq, nuniq, n_out, g->rowptr.as<uint32_t>(), g->col.as<uint32_t>(),
                     g->esrc.as<uint32_t>()));
  g->sched_version++;
  return compute_csr_stats(c, g);
}

int compute_csr_stats(tc_ctx* c, tc_graph* g) {
  DevBuf indeg;
  TCX_TRY(indeg.reserve(sizeof(uint32_t) * ((size_t)g->n + 1)));
  TCX_TRY(c->accum.reserve(64));
  unsigned long long* acc = c->accum.as<unsigned long long>();
  TCX_HIP(hipMemsetAsync(acc, 0, 3 * sizeof(unsigned long long), c->stream));
  hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(g->n)), dim3(TB), 0, c->stream,
                     indeg.as<uint32_t>(), 0u, (uint64_t)g->n);
  TCX_TRY(count_ids(c, g->col.as<uint32_t>(), nullptr, g->m_or, g->n, indeg.as<uint32_t>()));
  hipLaunchKernelGGL(k_csr_stats, dim3(grid_for(g->n)), dim3(TB), 0, c->stream,
                     g->rowptr.as<uint32_t>(), indeg.as<uint32_t>(), g->n, acc);
  TCX_HIP(hipGetLastError());
  unsigned long long h[3];
  TCX_HIP(hipMemcpyAsync(h, acc, sizeof(h), hipMemcpyDeviceToHost, c->stream));
  TCX_HIP(hipStreamSynchronize(c->stream));
  g->sum_out_sq = h[0];
  g->sum_in_out = h[1];
  g->max_out = (uint32_t)h[2];
  return TC_OK;
}

// Simple symmetric adjacency as sorted unique keys (both directions).
static int sym_keys(tc_ctx* c, tc_graph* g, DevBuf& keys, DevBuf& tmp, uint64_t** out,
                    uint64_t* m2) {
  uint64_t nk = 2 * g->m_raw;
  TCX_TRY(keys.reserve(8 * std::max<uint64_t>(nk, 1)));
  TCX_TRY(tmp.reserve(8 * std::max<uint64_t>(nk, 1)));
  hipLaunchKernelGGL(k_sym_keys, dim3(grid_for(g->m_raw)), dim3(TB), 0, c->stream,
                     g->src.as<uint32_t>(), g->dst.as<uint32_t>(), g->m_raw, g->n_raw,
                     keys.as<uint64_t>());
  TCX_HIP(hipGetLastError());
  uint64_t* sorted;
  TCX_TRY(sort_u64(c, keys.as<uint64_t>(), tmp.as<uint64_t>(), nk,
                   32 + bit_width_u32(g->n_raw), &sorted));
  uint64_t* other = sorted == keys.as<uint64_t>() ? tmp.as<uint64_t>() : keys.as<uint64_t>();
  uint64_t nu = 0;
  TCX_TRY(unique_u64(c, sorted, other, nk, &nu));
  if (nu > 0) {
    uint64_t last = 0;
    TCX_HIP(hipMemcpyAsync(&last, other + nu - 1, 8, hipMemcpyDeviceToHost, c->stream));
    TCX_HIP(hipStreamSynchronize(c->stream));
    if ((uint32_t)(last >> 32) == g->n_raw) --nu;
  }
  *out = other;
  *m2 = nu;
  return TC_OK;
}

static int sym_csr(tc_ctx* c, const uint64_t* keys, uint64_t m2, uint32_t n, DevBuf& rp,
                   DevBuf& col) {
  TCX_TRY(rp.reserve(sizeof(uint32_t) * ((size_t)n + 1)));
  TCX_TRY(col.reserve(sizeof(uint32_t) * std::max<uint64_t>(m2, 1)));
  TCX_TRY(csr_launch(c, keys, m2, n, rp.as<uint32_t>(), col.as<uint32_t>(), nullptr));
  return TC_OK;
}

// One round-level of Bisson's peel with the serial frontier order (see the
// degree-record kernels): frontier F (fs, in serial order) -> next frontier
// N in serial order; *ns its size.
struct AdirOrdered {
  DevBuf ev, evtmp, nkey, nktmp, nid, nitmp, flags;
};

static int adir_level_ordered(tc_ctx* c, AdirOrdered& o, const uint32_t* F, uint32_t fs,
                              const uint32_t* rp, const uint32_t* col, int32_t* rtd,
                              int32_t* prio, float t, int32_t label, int32_t* rec, uint32_t* N,
                              uint32_t* ns) {
  unsigned long long* ecount = c->accum.as<unsigned long long>();
  uint32_t* cnt = c->scratch.as<uint32_t>();
  TCX_HIP(hipMemsetAsync(ecount, 0, 8, c->stream));
  hipLaunchKernelGGL(k_adir_events, dim3(std::min<unsigned>(4096, grid_for(fs * 64ull))),
                     dim3(TB), 0, c->stream, F, fs, rp, col, rtd, prio, o.ev.as<uint64_t>(),
                     ecount);
  TCX_HIP(hipGetLastError());
  unsigned long long E = 0;
  TCX_HIP(hipMemcpyAsync(&E, ecount, 8, hipMemcpyDeviceToHost, c->stream));
  TCX_HIP(hipStreamSynchronize(c->stream));
  *ns = 0;
  if (E == 0) return TC_OK;
  uint64_t* sorted;
  TCX_TRY(sort_u64(c, o.ev.as<uint64_t>(), o.evtmp.as<uint64_t>(), E, 64, &sorted));
  TCX_HIP(hipMemsetAsync(cnt, 0, 4, c->stream));
  hipLaunchKernelGGL(k_adir_cross, dim3(grid_for(E)), dim3(TB), 0, c->stream, sorted,
                     (uint64_t)E, F, rp, col, rtd, prio, (int32_t)floorf(t), label, rec,
                     o.nkey.as<uint64_t>(), o.nid.as<uint32_t>(), cnt);
  TCX_HIP(hipGetLastError());
  uint32_t pushed = 0;
  TCX_TRY(read_u32(c, cnt, &pushed));
  if (pushed) {
    uint64_t* sk;
    uint32_t* order;
    TCX_TRY(sort_pairs_u64_u32(c, o.nkey.as<uint64_t>(), o.nktmp.as<uint64_t>(),
                               o.nid.as<uint32_t>(), o.nitmp.as<uint32_t>(), pushed, 64, &sk,
                               &order));
    TCX_HIP(hipMemcpyAsync(N, order, 4ull * pushed, hipMemcpyDeviceToDevice, c->stream));
  }
  *ns = pushed;
  return TC_OK;
}

// A-direction peel.  With rec non-NULL it also fills Bisson's degree record
// (zero where nothing is written, as the reference's memset leaves it), which
// needs the serial frontier order; without it the frontiers are unordered.
// thre_param > 0 selects the OpenMP program's schedule
// (edge_directing/A-direction.cpp:36-161): threshold0 = 2 totalEdge / n, a
// round's cascade runs only while its frontier holds more than n / 5000 and
// more than 200 vertices, for at most six levels, and the threshold grows by
// (int)threshold0 * thre_param / 5 per round.
static int adir_peel(tc_ctx* c, uint32_t n, uint64_t m2, const uint32_t* rp, const uint32_t* col,
                     int32_t* prio, int32_t* rec = nullptr, double thre_param = 0) {
  const bool omp = thre_param > 0;
  DevBuf rtd, fa, fb;
  TCX_TRY(rtd.reserve(sizeof(int32_t) * ((size_t)n + 1)));
  TCX_TRY(fa.reserve(sizeof(uint32_t) * ((size_t)n + 1)));
  TCX_TRY(fb.reserve(sizeof(uint32_t) * ((size_t)n + 1)));
  TCX_TRY(c->scratch.reserve(64));
  TCX_TRY(c->accum.reserve(64));
  uint32_t* cnt = c->scratch.as<uint32_t>();
  AdirOrdered ord;
  if (rec) {
    TCX_TRY(ord.ev.reserve(8 * std::max<uint64_t>(m2, 1)));
    TCX_TRY(ord.evtmp.reserve(8 * std::max<uint64_t>(m2, 1)));
    TCX_TRY(ord.nkey.reserve(8 * ((size_t)n + 1)));
    TCX_TRY(ord.nktmp.reserve(8 * ((size_t)n + 1)));
    TCX_TRY(ord.nid.reserve(4 * ((size_t)n + 1)));
    TCX_TRY(ord.nitmp.reserve(4 * ((size_t)n + 1)));
    TCX_TRY(ord.flags.reserve((size_t)n + 1));
    TCX_HIP(hipMemsetAsync(rec, 0, 4ull * n, c->stream));
  }
  hipLaunchKernelGGL(k_adir_init, dim3(grid_for(n)), dim3(TB), 0, c->stream, rp, n,
                     rtd.as<int32_t>(), prio);
  TCX_HIP(hipGetLastError());
  // A-direction-seri.cpp:37 and :77, with edgeNum = 2m (totalEdge, :228,:252);
  // A-direction.cpp:38,78
  unsigned int edgeNum = (unsigned int)m2;
  int nodeNum = (int)n;
  float threshold = omp ? 2 * ((float)edgeNum) / nodeNum : ((float)edgeNum) / nodeNum + 1;
  int threshold_old = threshold;
  auto step = [&]() {
    if (omp) threshold += (double)threshold_old * thre_param / 5;  // A-direction.cpp:139
    else threshold += threshold_old;                               // A-direction-seri.cpp:127
  };
  int32_t label = 0;
  uint64_t peeled = 0;
  uint32_t* F = fa.as<uint32_t>();
  uint32_t* N = fb.as<uint32_t>();
  const unsigned expand_grid = 4096;
  for (;;) {
    uint32_t fs = 0;
    if (rec) {  // ascending ids, as the serial scan appends them
      hipLaunchKernelGGL(k_adir_scan_flags, dim3(grid_for(n)), dim3(TB), 0, c->stream, n,
                         rtd.as<int32_t>(), prio, threshold, ord.flags.as<uint8_t>());
      TCX_HIP(hipGetLastError());
      uint64_t f64 = 0;
      TCX_TRY(select_indices(c, ord.flags.as<uint8_t>(), n, F, &f64));
      fs = (uint32_t)f64;
      if (fs)
        hipLaunchKernelGGL(k_adir_mark_scan, dim3(grid_for(fs)), dim3(TB), 0, c->stream, F, fs,
                           rtd.as<int32_t>(), prio, label, rec);
      TCX_HIP(hipGetLastError());
    } else {
      TCX_HIP(hipMemsetAsync(cnt, 0, sizeof(uint32_t), c->stream));
      hipLaunchKernelGGL(k_adir_scan, dim3(grid_for(n)), dim3(TB), 0, c->stream, n,
                         rtd.as<int32_t>(), prio, threshold, label, F, cnt);
      TCX_HIP(hipGetLastError());
      TCX_TRY(read_u32(c, cnt, &fs));
    }
    peeled += fs;
    label++;
    int inner = 0;
    while (fs > 0 && (!omp || (fs > (uint32_t)(nodeNum / 5000) && fs > 200 && inner <= 5))) {
      inner++;
      uint32_t ns = 0;
      if (rec) {
        TCX_TRY(adir_level_ordered(c, ord, F, fs, rp, col, rtd.as<int32_t>(), prio, threshold,
                                   label, rec, N, &ns));
      } else {
        TCX_HIP(hipMemsetAsync(cnt, 0, sizeof(uint32_t), c->stream));
        hipLaunchKernelGGL(k_adir_expand,
                           dim3(std::min<unsigned>(expand_grid, grid_for(fs * 64ull))), dim3(TB),
                           0, c->stream, F, fs, rp, col, rtd.as<int32_t>(), prio, threshold,
                           label, N, cnt);
        TCX_HIP(hipGetLastError());
        TCX_TRY(read_u32(c, cnt, &ns));
      }
      peeled += ns;
      std::swap(F, N);
      fs = ns;
      label++;
    }
    if (peeled >= n) break;
    step();
    // Rounds whose scan finds nothing only advance the label and threshold.
    TCX_HIP(hipMemsetAsync(cnt, 0x7f, sizeof(int32_t), c->stream));
    hipLaunchKernelGGL(k_min_unpeeled, dim3(grid_for(n)), dim3(TB), 0, c->stream, n,
                       rtd.as<int32_t>(), prio, (int32_t*)cnt);
    TCX_HIP(hipGetLastError());
    int32_t minr = 0;
    TCX_TRY(read_u32(c, cnt, (uint32_t*)&minr));
    while (!((float)minr <= threshold)) {
      label++;
      step();
    }
  }
  return TC_OK;
}

// Core numbers into `core`; with `step` non-NULL, also the removal step of
// every vertex.  A vertex removed at step t has at most core[v] neighbours
// removed at step t or later (its remaining degree when its frontier formed),
// so orienting every edge from the lower (step, id) bounds d+(v) by core[v].
static int kcore_peel(tc_ctx* c, uint32_t n, const uint32_t* rp, const uint32_t* col,
                      int32_t* core, int32_t* step = nullptr) {
  DevBuf deg, fa, fb;
  TCX_TRY(deg.reserve(sizeof(int32_t) * ((size_t)n + 1)));
  TCX_TRY(fa.reserve(sizeof(uint32_t) * ((size_t)n + 1)));
  TCX_TRY(fb.reserve(sizeof(uint32_t) * ((size_t)n + 1)));
  TCX_TRY(c->scratch.reserve(64));
  uint32_t* cnt = c->scratch.as<uint32_t>();
  hipLaunchKernelGGL(k_core_init, dim3(grid_for(n)), dim3(TB), 0, c->stream, rp, n,
                     deg.as<int32_t>(), core);
  TCX_HIP(hipGetLastError());
  uint64_t removed = 0;
  int32_t k = 0, t = 0;
  uint32_t* F = fa.as<uint32_t>();
  uint32_t* N = fb.as<uint32_t>();
  while (removed < n) {
    TCX_HIP(hipMemsetAsync(cnt, 0, sizeof(uint32_t), c->stream));
    hipLaunchKernelGGL(k_core_scan, dim3(grid_for(n)), dim3(TB), 0, c->stream, n,
                       deg.as<int32_t>(), core, k, F, cnt, step, t);
    TCX_HIP(hipGetLastError());
    uint32_t fs = 0;
    TCX_TRY(read_u32(c, cnt, &fs));
    removed += fs;
    if (fs) ++t;
    while (fs > 0) {
      TCX_HIP(hipMemsetAsync(cnt, 0, sizeof(uint32_t), c->stream));
      hipLaunchKernelGGL(k_core_expand, dim3(std::min<unsigned>(4096, grid_for(fs * 64ull))),
                         dim3(TB), 0, c->stream, F, fs, rp, col, deg.as<int32_t>(), core, k, N,
                         cnt, step, t);
      TCX_HIP(hipGetLastError());
      uint32_t ns = 0;
      TCX_TRY(read_u32(c, cnt, &ns));
      removed += ns;
      if (ns) ++t;
      std::swap(F, N);
      fs = ns;
    }
    if (removed >= n) break;
    TCX_HIP(hipMemsetAsync(cnt, 0x7f, sizeof(int32_t), c->stream));
    hipLaunchKernelGGL(k_min_unpeeled, dim3(grid_for(n)), dim3(TB), 0, c->stream, n,
                       deg.as<int32_t>(), core, (int32_t*)cnt);
    TCX_HIP(hipGetLastError());
    int32_t mind = 0;
    TCX_TRY(read_u32(c, cnt, (uint32_t*)&mind));
    k = std::max(k + 1, mind);
  }
  return TC_OK;
}

int graph_clone_records(tc_ctx* c, const tc_graph* g, tc_graph** out) {
  *out = nullptr;
  const bool from_csr = g->csr_only;
  const uint64_t m = from_csr ? (g->orient_mode >= 0 ? g->m_or : 0) : g->m_raw;
  const void* ps = from_csr ? g->esrc.p : g->src.p;
  const void* pd = from_csr ? g->col.p : g->dst.p;
  tc_graph* h = new tc_graph();
  h->ctx = c;
  h->m_raw = m;
  int st = h->src.reserve(4 * std::max<uint64_t>(m, 1));
  if (st == TC_OK) st = h->dst.reserve(4 * std::max<uint64_t>(m, 1));
  if (st == TC_OK && m &&
      (hipMemcpyAsync(h->src.p, ps, 4 * m, hipMemcpyDeviceToDevice, c->stream) != hipSuccess ||
       hipMemcpyAsync(h->dst.p, pd, 4 * m, hipMemcpyDeviceToDevice, c->stream) != hipSuccess))
    st = fail(TC_ERR_HIP, "record copy failed");
  uint32_t ms = 0, md = 0;
  if (st == TC_OK && m) st = max_u32(c, h->src.as<uint32_t>(), m, &ms);
  if (st == TC_OK && m) st = max_u32(c, h->dst.as<uint32_t>(), m, &md);
  if (st != TC_OK) {
    delete h;
    return st;
  }
  h->n = m ? std::max(ms, md) + 1 : 0;
  h->n_raw = h->n;
  *out = h;
  return TC_OK;
}

}  // namespace tcx

using namespace tcx;

static int check_graph(tc_graph* g) {
  if (!g || !g->ctx) return fail(TC_ERR_ARG, "NULL graph");
  TCX_HIP(hipSetDevice(g->ctx->device));
  return TC_OK;
}

// Passes over the raw records need them: a graph loaded from a CSR has none.
static int check_records(tc_graph* g) {
  TCX_TRY(check_graph(g));
  if (g->csr_only)
    return fail(TC_ERR_STATE, "the graph was built from an oriented CSR and holds no raw records");
  return TC_OK;
}

extern "C" {

int tc_graph_create(tc_ctx* c, const uint32_t* src, const uint32_t* dst, uint64_t m,
                    uint32_t n_hint, tc_graph** out) {
  if (!c || !out) return fail(TC_ERR_ARG, "NULL argument");
  if (m && (!src || !dst)) return fail(TC_ERR_ARG, "NULL edge arrays");
  *out = nullptr;
  TCX_HIP(hipSetDevice(c->device));
  tc_graph* g = new tc_graph();
  g->ctx = c;
  g->m_raw = m;
  int st = g->src.reserve(4 * std::max<uint64_t>(m, 1));
  if (st == TC_OK) st = g->dst.reserve(4 * std::max<uint64_t>(m, 1));
  if (st != TC_OK) { delete g; return st; }
  if (m) {
    hipError_t e1 = hipMemcpyAsync(g->src.p, src, 4 * m, hipMemcpyHostToDevice, c->stream);
    hipError_t e2 = hipMemcpyAsync(g->dst.p, dst, 4 * m, hipMemcpyHostToDevice, c->stream);
    if (e1 != hipSuccess || e2 != hipSuccess) {
      delete g;
      return fail(TC_ERR_HIP, "edge upload failed");
    }
    uint32_t ms = 0, md = 0;
    st = max_u32(c, g->src.as<uint32_t>(), m, &ms);
    if (st == TC_OK) st = max_u32(c, g->dst.as<uint32_t>(), m, &md);
    if (st != TC_OK) { delete g; return st; }
    uint32_t mx = std::max(ms, md);
    if (mx >= 0x7FFFFFFFu) {
      delete g;
      return fail(TC_ERR_RANGE, "vertex id %u exceeds 2^31-2", mx);
    }
    if (n_hint && mx >= n_hint) {
      delete g;
      return fail(TC_ERR_RANGE, "vertex id %u >= n_hint %u", mx, n_hint);
    }
    g->n = n_hint ? n_hint : mx + 1;
  } else {
    g->n = n_hint;
  }
  g->n_raw = g->n;
  *out = g;
  return TC_OK;
}

int tc_graph_destroy(tc_graph* g) {
  if (!g) return TC_OK;
  if (g->ctx) {
    (void)hipSetDevice(g->ctx->device);
    (void)hipStreamSynchronize(g->ctx->stream);
  }
  delete g;
  return TC_OK;
}

int tc_graph_simplify(tc_ctx* c, tc_graph* g) {
  TCX_TRY(check_records(g));
  if (c != g->ctx) return fail(TC_ERR_ARG, "graph belongs to another context");
  if (g->m_raw == 0) return TC_OK;
  DevBuf keys, tmp;
  TCX_TRY(keys.reserve(8 * g->m_raw));
  TCX_TRY(tmp.reserve(8 * g->m_raw));
  hipLaunchKernelGGL(k_simple_keys, dim3(grid_for(g->m_raw)), dim3(TB), 0, c->stream,
                     g->src.as<uint32_t>(), g->dst.as<uint32_t>(), g->m_raw, g->n_raw,
                     keys.as<uint64_t>());
  TCX_HIP(hipGetLastError());
  uint64_t* sorted;
  TCX_TRY(sort_u64(c, keys.as<uint64_t>(), tmp.as<uint64_t>(), g->m_raw,
                   32 + bit_width_u32(g->n_raw), &sorted));
  uint64_t* other = sorted == keys.as<uint64_t>() ? tmp.as<uint64_t>() : keys.as<uint64_t>();
  uint64_t nu = 0;
  TCX_TRY(unique_u64(c, sorted, other, g->m_raw, &nu));
  if (nu > 0) {
    uint64_t last = 0;
    TCX_HIP(hipMemcpyAsync(&last, other + nu - 1, 8, hipMemcpyDeviceToHost, c->stream));
    TCX_HIP(hipStreamSynchronize(c->stream));
    if ((uint32_t)(last >> 32) == g->n_raw) --nu;
  }
  hipLaunchKernelGGL(k_unzip, dim3(grid_for(nu)), dim3(TB), 0, c->stream, other, nu,
                     g->src.as<uint32_t>(), g->dst.as<uint32_t>());
  TCX_HIP(hipGetLastError());
  TCX_HIP(hipStreamSynchronize(c->stream));
  g->m_raw = nu;
  return TC_OK;
}

int tc_graph_relabel_aorder(tc_ctx* c, tc_graph* g, uint32_t* perm_out) {
  tc_bw_model ref;
  bw_model_reference(&ref);
  return tc_graph_relabel_aorder_model(c, g, &ref, perm_out);
}

int tc_graph_relabel_aorder_model(tc_ctx* c, tc_graph* g, const tc_bw_model* model,
                                  uint32_t* perm_out) {
  if (!model) return fail(TC_ERR_ARG, "NULL bandwidth model");
  // the bucket heaps need a strict weak order: finite, positive bandwidths
  for (int l = 2; l < 10; ++l)
    if (!(std::isfinite(model->table[l]) && model->table[l] > 0 &&
          std::isfinite(model->lambda[l]) && model->lambda[l] >= 0))
      return fail(TC_ERR_ARG, "bandwidth model entry %d is not finite and positive", l);
  if (!(std::isfinite(model->a) && model->a > 0 && std::isfinite(model->b) &&
        std::isfinite(model->lambda_tail) && model->lambda_tail >= 0))
    return fail(TC_ERR_ARG, "bandwidth model power law is not finite and positive");
  TCX_TRY(check_records(g));
  if (c != g->ctx) return fail(TC_ERR_ARG, "graph belongs to another context");
  uint32_t n = g->n_raw;
  if (n == 0) return fail(TC_ERR_UNSUPPORTED, "A-order on an empty graph");
  DevBuf deg, rtd, perm;
  TCX_TRY(deg.reserve(4 * (size_t)n));
  TCX_TRY(rtd.reserve(4 * (size_t)n));
  TCX_TRY(perm.reserve(4 * (size_t)n));
  hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(n)), dim3(TB), 0, c->stream, deg.as<uint32_t>(),
                     0u, (uint64_t)n);
  hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(n)), dim3(TB), 0, c->stream, rtd.as<uint32_t>(),
                     0u, (uint64_t)n);
  TCX_TRY(count_ids(c, g->src.as<uint32_t>(), g->dst.as<uint32_t>(), g->m_raw, n,
                    deg.as<uint32_t>()));
  hipLaunchKernelGGL(k_aorder_outdeg, dim3(grid_for(g->m_raw)), dim3(TB), 0, c->stream,
                     g->src.as<uint32_t>(), g->dst.as<uint32_t>(), g->m_raw, deg.as<uint32_t>(),
                     rtd.as<uint32_t>());
  TCX_HIP(hipGetLastError());
  std::vector<uint32_t> h_rtd(n), h_perm(n);
  TCX_HIP(hipMemcpyAsync(h_rtd.data(), rtd.p, 4 * (size_t)n, hipMemcpyDeviceToHost, c->stream));
  TCX_HIP(hipStreamSynchronize(c->stream));
  TCX_TRY(aorder_perm_from_outdeg(h_rtd.data(), n, h_perm.data(), *model));
  TCX_HIP(hipMemcpyAsync(perm.p, h_perm.data(), 4 * (size_t)n, hipMemcpyHostToDevice, c->stream));
  hipLaunchKernelGGL(k_aorder_apply, dim3(grid_for(g->m_raw)), dim3(TB), 0, c->stream,
                     g->src.as<uint32_t>(), g->dst.as<uint32_t>(), g->m_raw, deg.as<uint32_t>(),
                     perm.as<uint32_t>());
  TCX_HIP(hipGetLastError());
  // The relabelled records are what A-order writes to its "reordered" file
  // (A-order.cpp:206-228); a program reading them back takes the vertex count
  // as the largest id + 1 (A-direction-seri.cpp:187-193), not the old n.
  uint32_t new_n = 0;
  if (g->m_raw) {
    uint32_t ms = 0, md = 0;
    TCX_TRY(max_u32(c, g->src.as<uint32_t>(), g->m_raw, &ms));
    TCX_TRY(max_u32(c, g->dst.as<uint32_t>(), g->m_raw, &md));
    new_n = std::max(ms, md) + 1;
  }
  TCX_HIP(hipStreamSynchronize(c->stream));
  if (perm_out) std::copy(h_perm.begin(), h_perm.end(), perm_out);
  g->n = g->n_raw = new_n;
  g->orient_mode = -1;  // any CSR built before is stale
  return TC_OK;
}

int tc_graph_relabel(tc_ctx* c, tc_graph* g, int mode, uint32_t* perm_out) {
  TCX_TRY(check_records(g));
  if (c != g->ctx) return fail(TC_ERR_ARG, "graph belongs to another context");
  if (mode < TC_RELABEL_COMPACT || mode > TC_RELABEL_ADIR)
    return fail(TC_ERR_ARG, "unknown relabel mode %d", mode);
  const uint64_t m = g->m_raw;
  uint32_t* s = g->src.as<uint32_t>();
  uint32_t* d = g->dst.as<uint32_t>();
  uint32_t n = 1;  // max id + 1 over the records, as the tools compute it
  if (m) {
    uint32_t ms = 0, md = 0;
    TCX_TRY(max_u32(c, s, m, &ms));
    TCX_TRY(max_u32(c, d, m, &md));
    n = std::max(ms, md) + 1;
  }
  DevBuf deg, perm, key, ktmp, ids, itmp;
  TCX_TRY(perm.reserve(4ull * n));
  uint32_t n_new = n;
  if (mode == TC_RELABEL_COMPACT || mode == TC_RELABEL_DEGREE) {
    TCX_TRY(deg.reserve(4ull * n));
    hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(n)), dim3(TB), 0, c->stream,
                       deg.as<uint32_t>(), 0u, (uint64_t)n);
    TCX_TRY(count_ids(c, s, d, m, n, deg.as<uint32_t>()));
  }
  if (mode == TC_RELABEL_COMPACT) {
    // new id = number of present ids below (tools/serialize.cpp:32-44)
    DevBuf flags;
    TCX_TRY(flags.reserve(4ull * n));
    hipLaunchKernelGGL(k_present_flags, dim3(grid_for(n)), dim3(TB), 0, c->stream,
                       deg.as<uint32_t>(), n, flags.as<uint32_t>());
    TCX_TRY(exclusive_scan_u32(c, flags.as<uint32_t>(), perm.as<uint32_t>(), n));
    uint32_t last_perm = 0, last_flag = 0;
    TCX_TRY(read_u32(c, perm.as<uint32_t>() + n - 1, &last_perm));
    TCX_TRY(read_u32(c, flags.as<uint32_t>() + n - 1, &last_flag));
    n_new = m ? last_perm + last_flag : 1;
  } else {
    TCX_TRY(key.reserve(8ull * n));
    TCX_TRY(ktmp.reserve(8ull * n));
    TCX_TRY(ids.reserve(4ull * n));
    TCX_TRY(itmp.reserve(4ull * n));
    if (mode == TC_RELABEL_ADIR) {
      // peel the simple symmetric graph over the tools' n = max id + 1
      DevBuf skeys_buf, stmp, srp, scol, prio, rec;
      uint64_t* skeys = nullptr;
      uint64_t m2 = 0;
      TCX_TRY(sym_keys(c, g, skeys_buf, stmp, &skeys, &m2));
      TCX_TRY(sym_csr(c, skeys, m2, n, srp, scol));
      TCX_TRY(prio.reserve(4ull * n));
      TCX_TRY(rec.reserve(4ull * n));
      TCX_TRY(adir_peel(c, n, m2, srp.as<uint32_t>(), scol.as<uint32_t>(), prio.as<int32_t>(),
                        rec.as<int32_t>()));
      hipLaunchKernelGGL(k_adir_relabel_keys, dim3(grid_for(n)), dim3(TB), 0, c->stream,
                         prio.as<int32_t>(), rec.as<int32_t>(), n, key.as<uint64_t>(),
                         ids.as<uint32_t>());
    } else if (mode == TC_RELABEL_FIRST_SRC) {
      DevBuf first;
      TCX_TRY(first.reserve(8ull * n));
      hipLaunchKernelGGL(k_fill_u64, dim3(grid_for(n)), dim3(TB), 0, c->stream,
                         first.as<uint64_t>(), ~0ull, (uint64_t)n);
      if (m)
        hipLaunchKernelGGL(k_first_src, dim3(grid_for(m)), dim3(TB), 0, c->stream, s, m,
                           first.as<unsigned long long>());
      hipLaunchKernelGGL(k_first_keys, dim3(grid_for(n)), dim3(TB), 0, c->stream,
                         first.as<uint64_t>(), n, m, key.as<uint64_t>(), ids.as<uint32_t>());
    } else {
      hipLaunchKernelGGL(k_degree_keys, dim3(grid_for(n)), dim3(TB), 0, c->stream,
                         deg.as<uint32_t>(), n, key.as<uint64_t>(), ids.as<uint32_t>());
    }
    TCX_HIP(hipGetLastError());
    uint64_t* sk;
    uint32_t* order;
    TCX_TRY(sort_pairs_u64_u32(c, key.as<uint64_t>(), ktmp.as<uint64_t>(), ids.as<uint32_t>(),
                               itmp.as<uint32_t>(), n, 64, &sk, &order));
    hipLaunchKernelGGL(k_rank_perm, dim3(grid_for(n)), dim3(TB), 0, c->stream, order, n,
                       mode == TC_RELABEL_DEGREE || mode == TC_RELABEL_ADIR, perm.as<uint32_t>());
  }
  if (m)
    hipLaunchKernelGGL(k_apply_perm, dim3(grid_for(m)), dim3(TB), 0, c->stream, s, d, m,
                       perm.as<uint32_t>());
  TCX_HIP(hipGetLastError());
  if (perm_out)
    TCX_HIP(hipMemcpyAsync(perm_out, perm.p, 4ull * n, hipMemcpyDeviceToHost, c->stream));
  TCX_HIP(hipStreamSynchronize(c->stream));
  g->n = g->n_raw = n_new;
  g->orient_mode = -1;  // any CSR built before is stale
  return TC_OK;
}

static int orient_peeled(tc_ctx* c, tc_graph* g, int mode, double thre_param, int tie);
static int orient_impl(tc_ctx* c, tc_graph* g, int mode, uint32_t flags, int tie);

int tc_graph_orient(tc_ctx* c, tc_graph* g, int mode, uint32_t flags) {
  TCX_TRY(check_records(g));
  if (c != g->ctx) return fail(TC_ERR_ARG, "graph belongs to another context");
  if (mode < TC_ORIENT_ID || mode > TC_ORIENT_DEGENERACY)
    return fail(TC_ERR_ARG, "unknown orientation mode %d", mode);
  if ((flags & TC_FLAG_PRUNE_DEG1) && mode != TC_ORIENT_DEGREE)
    return fail(TC_ERR_ARG, "degree<=1 pruning applies to TC_ORIENT_DEGREE only");
  return orient_impl(c, g, mode, flags, TC_TIE_ID);
}

int tc_graph_orient_ties(tc_ctx* c, tc_graph* g, int mode, int tie) {
  TCX_TRY(check_records(g));
  if (c != g->ctx) return fail(TC_ERR_ARG, "graph belongs to another context");
  if (mode != TC_ORIENT_ADIR && mode != TC_ORIENT_DEGREE)
    return fail(TC_ERR_ARG, "tie rules apply to TC_ORIENT_ADIR and TC_ORIENT_DEGREE");
  if (tie < TC_TIE_ID || tie > TC_TIE_FORCED_DESC) return fail(TC_ERR_ARG, "unknown tie rule %d", tie);
  return orient_impl(c, g, mode, 0, tie);
}

int tc_graph_adir_cost(tc_ctx* c, tc_graph* g, double* cost) {
  if (!c || !g || !cost) return fail(TC_ERR_ARG, "NULL argument");
  if (c != g->ctx) return fail(TC_ERR_ARG, "graph belongs to another context");
  if (g->orient_mode < 0) return fail(TC_ERR_STATE, "graph must be oriented");
  TCX_HIP(hipSetDevice(c->device));
  *cost = 0;
  if (g->n == 0) return TC_OK;
  constexpr unsigned kGrid = 1024;
  DevBuf part;
  TCX_TRY(part.reserve(8 * kGrid));
  const double avg = (double)g->m_or / (double)g->n;
  hipLaunchKernelGGL(k_adir_cost, dim3(kGrid), dim3(256), 0, c->stream, g->rowptr.as<uint32_t>(),
                     g->n, avg, part.as<double>());
  TCX_HIP(hipGetLastError());
  std::vector<double> h(kGrid);
  TCX_HIP(hipMemcpyAsync(h.data(), part.p, 8 * kGrid, hipMemcpyDeviceToHost, c->stream));
  TCX_HIP(hipStreamSynchronize(c->stream));
  double sum = 0;
  for (double x : h) sum += x;
  *cost = sum;
  return TC_OK;
}

static int orient_impl(tc_ctx* c, tc_graph* g, int mode, uint32_t flags, int tie) {
  // Every pass over the raw records uses their id range; only a pruned CSR
  // has fewer rows.
  const uint32_t n = g->n_raw;
  const uint64_t m = g->m_raw;
  if (mode == TC_ORIENT_ID || mode == TC_ORIENT_DEGREE || mode == TC_ORIENT_DEGREE_TRICORE) {
    DevBuf deg, pruned, keys, tmp;
    TCX_TRY(deg.reserve(4 * ((size_t)n + 1)));
    TCX_TRY(keys.reserve(8 * std::max<uint64_t>(m, 1)));
    TCX_TRY(tmp.reserve(8 * std::max<uint64_t>(m, 1)));
    hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(n + 1)), dim3(TB), 0, c->stream,
                       deg.as<uint32_t>(), 0u, (uint64_t)n + 1);
    if (mode != TC_ORIENT_ID)
      TCX_TRY(count_ids(c, g->src.as<uint32_t>(), g->dst.as<uint32_t>(), m, n, deg.as<uint32_t>()));
    TCX_HIP(hipGetLastError());
    uint32_t n_out = n;
    const uint32_t* pruned_before = nullptr;
    DevBuf tkey;
    if (tie != TC_TIE_ID) {
      TCX_TRY(tkey.reserve(4 * ((size_t)n + 1)));
      hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(n + 1)), dim3(TB), 0, c->stream,
                         tkey.as<uint32_t>(), 0u, (uint64_t)n + 1);
      hipLaunchKernelGGL(k_forced_raw, dim3(grid_for(m)), dim3(TB), 0, c->stream,
                         g->src.as<uint32_t>(), g->dst.as<uint32_t>(), m, deg.as<uint32_t>(),
                         tkey.as<uint32_t>());
      if (tie == TC_TIE_FORCED_DESC)
        hipLaunchKernelGGL(k_tie_flip, dim3(grid_for(n)), dim3(TB), 0, c->stream,
                           tkey.as<uint32_t>(), n);
      TCX_HIP(hipGetLastError());
    }
    if (flags & TC_FLAG_PRUNE_DEG1) {
      DevBuf flag;
      TCX_TRY(flag.reserve(4 * ((size_t)n + 1)));
      TCX_TRY(pruned.reserve(4 * ((size_t)n + 1)));
      hipLaunchKernelGGL(k_low_degree_flag, dim3(grid_for(n)), dim3(TB), 0, c->stream,
                         deg.as<uint32_t>(), n, flag.as<uint32_t>());
      TCX_HIP(hipGetLastError());
      TCX_TRY(exclusive_scan_u32(c, flag.as<uint32_t>(), pruned.as<uint32_t>(), n));
      uint32_t lastp = 0, lastf = 0;
      if (n) {
        TCX_TRY(read_u32(c, pruned.as<uint32_t>() + n - 1, &lastp));
        TCX_TRY(read_u32(c, flag.as<uint32_t>() + n - 1, &lastf));
      }
      n_out = n - (lastp + lastf);
      pruned_before = pruned.as<uint32_t>();
      TCX_HIP(hipStreamSynchronize(c->stream));
    }
    hipLaunchKernelGGL(k_orient_keys, dim3(grid_for(m)), dim3(TB), 0, c->stream,
                       g->src.as<uint32_t>(), g->dst.as<uint32_t>(), m, deg.as<uint32_t>(), mode,
                       pruned_before, n_out, tie != TC_TIE_ID ? tkey.as<uint32_t>() : nullptr,
                       keys.as<uint64_t>());
    TCX_HIP(hipGetLastError());
    TCX_TRY(csr_from_keys(c, g, keys, tmp, m, n_out, false));
    g->orient_mode = mode;
    return TC_OK;
  }
  return orient_peeled(c, g, mode, 0, tie);
}

// A-direction / k-core: peel the simple symmetric graph, then keep the
// symmetric entries that point up the peel order.
static int orient_peeled(tc_ctx* c, tc_graph* g, int mode, double thre_param, int tie) {
  const uint32_t n = g->n_raw;
  DevBuf keys, tmp, srp, scol, keep;
  uint64_t* skeys = nullptr;
  uint64_t m2 = 0;
  TCX_TRY(sym_keys(c, g, keys, tmp, &skeys, &m2));
  TCX_TRY(sym_csr(c, skeys, m2, n, srp, scol));
  TCX_TRY(g->vkey.reserve(4 * ((size_t)n + 1)));
  if (n > 0) {
    if (mode == TC_ORIENT_ADIR || mode == TC_ORIENT_ADIR_OMP) {
      TCX_TRY(adir_peel(c, n, m2, srp.as<uint32_t>(), scol.as<uint32_t>(), g->vkey.as<int32_t>(),
                        nullptr, mode == TC_ORIENT_ADIR_OMP ? thre_param : 0));
    } else if (mode == TC_ORIENT_KCORE) {
      TCX_TRY(kcore_peel(c, n, srp.as<uint32_t>(), scol.as<uint32_t>(), g->vkey.as<int32_t>()));
    } else {
      DevBuf core;
      TCX_TRY(core.reserve(4 * ((size_t)n + 1)));
      TCX_TRY(kcore_peel(c, n, srp.as<uint32_t>(), scol.as<uint32_t>(), core.as<int32_t>(),
                         g->vkey.as<int32_t>()));
    }
  }
  DevBuf tkey;
  if (tie != TC_TIE_ID && n > 0) {
    TCX_TRY(tkey.reserve(4 * ((size_t)n + 1)));
    hipLaunchKernelGGL(k_forced_sym, dim3(grid_for(n)), dim3(TB), 0, c->stream,
                       srp.as<uint32_t>(), scol.as<uint32_t>(), n, g->vkey.as<int32_t>(),
                       tie == TC_TIE_FORCED_DESC ? 1 : 0, tkey.as<uint32_t>());
    TCX_HIP(hipGetLastError());
  }
  TCX_TRY(keep.reserve(std::max<uint64_t>(m2, 1)));
  hipLaunchKernelGGL(k_sym_keep, dim3(grid_for(m2)), dim3(TB), 0, c->stream, skeys, m2,
                     g->vkey.as<int32_t>(), mode, tkey.p ? tkey.as<uint32_t>() : nullptr,
                     keep.as<uint8_t>());
  TCX_HIP(hipGetLastError());
  uint64_t* dst_keys = skeys == keys.as<uint64_t>() ? tmp.as<uint64_t>() : keys.as<uint64_t>();
  uint64_t mo = 0;
  TCX_TRY(select_flagged_u64(c, skeys, keep.as<uint8_t>(), dst_keys, m2, &mo));
  DevBuf* holder = dst_keys == keys.as<uint64_t>() ? &keys : &tmp;
  DevBuf* spare = holder == &keys ? &tmp : &keys;
  TCX_TRY(csr_from_keys(c, g, *holder, *spare, mo, n, true));
  g->orient_mode = mode;
  return TC_OK;
}

int tc_graph_orient_adir_omp(tc_ctx* c, tc_graph* g, double thre_param) {
  TCX_TRY(check_records(g));
  if (c != g->ctx) return fail(TC_ERR_ARG, "graph belongs to another context");
  if (!(thre_param > 0) || !std::isfinite(thre_param))
    return fail(TC_ERR_ARG, "thre_param must be finite and positive (the threshold never grows "
                            "otherwise)");
  return orient_peeled(c, g, TC_ORIENT_ADIR_OMP, thre_param, TC_TIE_ID);
}

// Stage file of an oriented CSR: "TCAMDCSR", uint32 version 1, uint32 n,
// uint64 m, int32 orientation, uint32 has_key, rowptr[n+1], col[m] (uint32),
// then key[n] (int32) when has_key; little endian.
namespace {
struct CsrFileHeader {
  char magic[8];
  uint32_t version, n;
  uint64_t m;
  int32_t orient_mode;
  uint32_t has_key;
};
static_assert(sizeof(CsrFileHeader) == 32, "stage file header is 32 bytes");
constexpr char kCsrMagic[8] = {'T', 'C', 'A', 'M', 'D', 'C', 'S', 'R'};

bool has_vertex_key(int mode) {
  return mode == TC_ORIENT_ADIR || mode == TC_ORIENT_KCORE || mode == TC_ORIENT_DEGENERACY ||
         mode == TC_ORIENT_ADIR_OMP;
}
}  // namespace

int tc_graph_save_csr(const tc_graph* g, const char* path) {
  if (!g || !g->ctx || !path) return fail(TC_ERR_ARG, "NULL argument");
  if (g->orient_mode < 0) return fail(TC_ERR_STATE, "graph not oriented");
  const uint32_t n = g->n;
  const uint64_t m = g->m_or;
  std::vector<uint32_t> rp((size_t)n + 1), col(std::max<uint64_t>(m, 1));
  std::vector<int32_t> key;
  const bool wk = has_vertex_key(g->orient_mode);
  TCX_TRY(tc_graph_get_csr(g, rp.data(), col.data()));
  if (wk && n) {
    key.resize(n);
    TCX_TRY(tc_graph_get_vertex_key(g, key.data()));
  }
  CsrFileHeader h;
  memcpy(h.magic, kCsrMagic, 8);
  h.version = 1;
  h.n = n;
  h.m = m;
  h.orient_mode = g->orient_mode;
  h.has_key = wk ? 1u : 0u;
  FILE* f = fopen(path, "wb");
  if (!f) return fail(TC_ERR_IO, "cannot open %s for writing", path);
  bool ok = fwrite(&h, sizeof(h), 1, f) == 1 &&
            fwrite(rp.data(), 4, (size_t)n + 1, f) == (size_t)n + 1 &&
            (m == 0 || fwrite(col.data(), 4, m, f) == m) &&
            (!wk || n == 0 || fwrite(key.data(), 4, n, f) == n);
  ok = (fclose(f) == 0) && ok;
  if (!ok) return fail(TC_ERR_IO, "write to %s failed", path);
  return TC_OK;
}

int tc_graph_load_csr(tc_ctx* c, const char* path, tc_graph** out) {
  if (!c || !path || !out) return fail(TC_ERR_ARG, "NULL argument");
  *out = nullptr;
  FILE* f = fopen(path, "rb");
  if (!f) return fail(TC_ERR_IO, "cannot open %s", path);
  CsrFileHeader h;
  if (fread(&h, sizeof(h), 1, f) != 1 || memcmp(h.magic, kCsrMagic, 8) != 0 || h.version != 1 ||
      h.orient_mode < TC_ORIENT_ID || h.orient_mode > TC_ORIENT_ADIR_OMP ||
      h.has_key != (has_vertex_key(h.orient_mode) ? 1u : 0u) || h.n > 0x7FFFFFFEu ||
      h.m > kMaxCsrEntries) {
    fclose(f);
    return fail(TC_ERR_FORMAT, "%s is not an oriented-CSR stage file", path);
  }
  const uint32_t n = h.n;
  const uint64_t m = h.m;
  std::vector<uint32_t> rp((size_t)n + 1), col(std::max<uint64_t>(m, 1));
  std::vector<int32_t> key(h.has_key ? std::max<uint32_t>(n, 1) : 0);
  bool ok = fread(rp.data(), 4, (size_t)n + 1, f) == (size_t)n + 1 &&
            (m == 0 || fread(col.data(), 4, m, f) == m) &&
            (!h.has_key || n == 0 || fread(key.data(), 4, n, f) == n);
  char extra;
  ok = ok && fread(&extra, 1, 1, f) == 0;  // nothing after the arrays
  fclose(f);
  if (!ok) return fail(TC_ERR_FORMAT, "%s is truncated or has trailing bytes", path);
  // the kernels index rp[col[e] + 1]: every structural property is checked
  if (rp[0] != 0 || rp[n] != m) return fail(TC_ERR_FORMAT, "%s: row offsets do not span m", path);
  std::vector<uint32_t> esrc(std::max<uint64_t>(m, 1));
  for (uint32_t v = 0; v < n; ++v) {
    if (rp[v + 1] < rp[v]) return fail(TC_ERR_FORMAT, "%s: row offsets decrease at %u", path, v);
    for (uint32_t e = rp[v]; e < rp[v + 1]; ++e) {
      if (col[e] >= n || col[e] == v || (e > rp[v] && col[e] <= col[e - 1]))
        return fail(TC_ERR_FORMAT, "%s: row %u is not ascending, loop-free and in range", path, v);
      esrc[e] = v;
    }
  }
  TCX_HIP(hipSetDevice(c->device));
  tc_graph* g = new tc_graph();
  g->ctx = c;
  g->n = g->n_raw = n;
  g->csr_only = true;
  g->m_or = m;
  int st = g->rowptr.reserve(4 * ((size_t)n + 1));
  if (st == TC_OK) st = g->col.reserve(4 * std::max<uint64_t>(m, 1));
  if (st == TC_OK) st = g->esrc.reserve(4 * std::max<uint64_t>(m, 1));
  if (st == TC_OK && h.has_key) st = g->vkey.reserve(4 * ((size_t)n + 1));
  if (st == TC_OK &&
      (hipMemcpyAsync(g->rowptr.p, rp.data(), 4 * ((size_t)n + 1), hipMemcpyHostToDevice,
                      c->stream) != hipSuccess ||
       (m && hipMemcpyAsync(g->col.p, col.data(), 4 * m, hipMemcpyHostToDevice, c->stream) !=
                 hipSuccess) ||
       (m && hipMemcpyAsync(g->esrc.p, esrc.data(), 4 * m, hipMemcpyHostToDevice, c->stream) !=
                 hipSuccess) ||
       (h.has_key && n &&
        hipMemcpyAsync(g->vkey.p, key.data(), 4 * (size_t)n, hipMemcpyHostToDevice, c->stream) !=
            hipSuccess)))
    st = fail(TC_ERR_HIP, "stage file upload failed");
  if (st == TC_OK && hipStreamSynchronize(c->stream) != hipSuccess)
    st = fail(TC_ERR_HIP, "stage file upload failed");
  if (st == TC_OK) {
    g->orient_mode = h.orient_mode;
    g->sched_version++;
    st = compute_csr_stats(c, g);
  }
  if (st != TC_OK) {
    delete g;
    return st;
  }
  *out = g;
  return TC_OK;
}

int tc_graph_info_get(const tc_graph* g, tc_graph_info* info) {
  if (!g || !info) return fail(TC_ERR_ARG, "NULL argument");
  info->n = g->n;
  info->max_out_degree = g->max_out;
  info->m_raw = g->m_raw;
  info->m_oriented = g->orient_mode >= 0 ? g->m_or : 0;
  info->orient_mode = g->orient_mode;
  info->n_raw = g->n_raw;
  info->sum_out_sq = g->orient_mode >= 0 ? g->sum_out_sq : 0;
  info->sum_in_out = g->orient_mode >= 0 ? g->sum_in_out : 0;
  double n1 = (double)g->n + 1, mo = (double)info->m_oriented;
  info->bytes_alg = 4 * n1 + 8 * mo + 4.0 * (double)info->sum_out_sq + 4.0 * (double)info->sum_in_out;
  info->bytes_min = 4 * n1 + 4 * mo;
  return TC_OK;
}

int tc_graph_get_csr(const tc_graph* g, uint32_t* rowptr, uint32_t* col) {
  if (!g || !g->ctx) return fail(TC_ERR_ARG, "NULL graph");
  if (g->orient_mode < 0) return fail(TC_ERR_STATE, "graph not oriented");
  TCX_HIP(hipSetDevice(g->ctx->device));
  hipStream_t s = g->ctx->stream;
  if (rowptr)
    TCX_HIP(hipMemcpyAsync(rowptr, g->rowptr.p, 4 * ((size_t)g->n + 1), hipMemcpyDeviceToHost, s));
  if (col && g->m_or)
    TCX_HIP(hipMemcpyAsync(col, g->col.p, 4 * g->m_or, hipMemcpyDeviceToHost, s));
  TCX_HIP(hipStreamSynchronize(s));
  return TC_OK;
}

int tc_graph_get_edges(const tc_graph* g, uint32_t* src, uint32_t* dst) {
  if (!g || !g->ctx) return fail(TC_ERR_ARG, "NULL graph");
  TCX_HIP(hipSetDevice(g->ctx->device));
  hipStream_t s = g->ctx->stream;
  if (g->m_raw) {
    if (src) TCX_HIP(hipMemcpyAsync(src, g->src.p, 4 * g->m_raw, hipMemcpyDeviceToHost, s));
    if (dst) TCX_HIP(hipMemcpyAsync(dst, g->dst.p, 4 * g->m_raw, hipMemcpyDeviceToHost, s));
  }
  TCX_HIP(hipStreamSynchronize(s));
  return TC_OK;
}

int tc_graph_get_vertex_key(const tc_graph* g, int32_t* key) {
  if (!g || !g->ctx || !key) return fail(TC_ERR_ARG, "NULL argument");
  if (g->orient_mode != TC_ORIENT_ADIR && g->orient_mode != TC_ORIENT_KCORE &&
      g->orient_mode != TC_ORIENT_DEGENERACY && g->orient_mode != TC_ORIENT_ADIR_OMP)
    return fail(TC_ERR_STATE,
                "vertex keys exist after ADIR, ADIR_OMP, KCORE or DEGENERACY orientation only");
  TCX_HIP(hipSetDevice(g->ctx->device));
  if (g->n)
    TCX_HIP(hipMemcpyAsync(key, g->vkey.p, 4 * (size_t)g->n, hipMemcpyDeviceToHost, g->ctx->stream));
  TCX_HIP(hipStreamSynchronize(g->ctx->stream));
  return TC_OK;
}

int tc_graph_device_csr(const tc_graph* g, const uint32_t** rowptr, const uint32_t** col) {
  if (!g || !rowptr || !col) return fail(TC_ERR_ARG, "NULL argument");
  if (g->orient_mode < 0) return fail(TC_ERR_STATE, "graph not oriented");
  *rowptr = g->rowptr.as<uint32_t>();
  *col = g->col.as<uint32_t>();
  return TC_OK;
}

}  // extern "C"
