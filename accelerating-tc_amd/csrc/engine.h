// Device building blocks of the AUTO engine (vertex-centric, N+(u) in LDS),
// shared by the count kernels, the batched kernel and the cost-measurement
// kernels.
//
// For a vertex u the work is the multiset X = concatenation of N+(v) over
// v in N+(u); the count is |{x in X : x in N+(u)}|.  A wave
//  1. generates the entries of X 64 at a time (whole lists streamed by the
//     wave, or short lists concatenated across lanes),
//  2. tests each entry against a hashed bitmap of N+(u) in LDS (one LDS read;
//     most entries are not in N+(u) and stop here),
//  3. appends the survivors to a per-wave queue in LDS (ballot + mbcnt), and
//  4. once 64 survivors are queued, resolves them with full lanes by a
//     branchless binary search of N+(u).
// Step 2 replaces a log2(d+(u))-deep chain of dependent LDS reads per entry
// by one read; step 3 keeps the searches at full lane occupancy.  The
// per-vertex kernel and the push forms work this way.  The work-list
// kernels (k_auto_wl, the batched work-list form) replace steps 2-4 by one
// lookup per entry in a rank-indexed table of N+(u) (see RankLds below).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace tcx {
namespace dev {

constexpr int TB = 256;                 // threads per workgroup (4 waves)
constexpr uint32_t kEmpty = 0xFFFFFFFFu;
constexpr uint32_t kWaveMax = 512;      // longest N+(u) the wave variant holds in LDS
constexpr uint32_t kWaveChunk = 16;     // wave tasks taken per cursor bump
constexpr uint32_t kStaticWaves = 16;   // wave tasks per wave up to which they are strided
constexpr uint32_t kQueue = 320;        // per-wave survivor queue (< 64 + 4 * 64 held)
constexpr uint32_t kFlatMin = 64;       // lists this long or longer are streamed whole
constexpr uint32_t kFilterMin = 1;      // wave variant: bitmap for every d+(u) above this
constexpr uint32_t kLaneMaxDeg = 8;     // lane tasks: d+(x) up to this ...
constexpr uint32_t kLaneMaxList = 16;   // ... and at most this many work-list entries
constexpr uint32_t kLaneClaim = 256;    // lane tasks per wave claim (4 per lane)
// Wave tasks whose work (d+(x) + sum of d+(y) over W(x)) exceeds this many
// list entries run as block tasks (a hub reached by many in-edges can carry
// millions).
constexpr uint64_t kWlWaveWork = 1ull << 16;

__device__ inline uint64_t gtid() { return (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; }
__device__ inline uint64_t gstride() { return (uint64_t)gridDim.x * blockDim.x; }

__device__ inline unsigned long long wave_sum(unsigned long long v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}

__device__ inline void flush_count(unsigned long long cnt, unsigned long long* total) {
  cnt = wave_sum(cnt);
  if ((threadIdx.x & 63) == 0 && cnt) atomicAdd(total, cnt);
}

// The workgroup's count in one atomic; every thread of the workgroup calls
// it (red: TB / 64 words of LDS).
__device__ inline void flush_block(unsigned long long cnt, unsigned long long* total,
                                   unsigned long long* red) {
  cnt = wave_sum(cnt);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = cnt;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long s = 0;
    for (uint32_t w = 0; w < blockDim.x / 64; ++w) s += red[w];
    if (s) atomicAdd(total, s);
  }
}

// Claims `amount` entries of a task list of n from its cursor and returns
// the first (n or more when the list is exhausted).  A worker's first claim
// in a phase (first = true) reads the cursor before it adds: atomics on one
// address are serialised (about 10-20 ns each), and without the read every
// wave of a persistent grid would pay one per phase whose list ran out
// before it got there (about 0.1 ms a launch).  Later claims add at once:
// the read costs a memory round trip that short tasks would feel.  An
// empty list costs nothing.
__device__ inline uint32_t claim(uint32_t* cursor, uint32_t n, uint32_t amount, bool first) {
  if (n == 0) return 0;
  if (first && __hip_atomic_load(cursor, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= n)
    return n;
  return atomicAdd(cursor, amount);
}

// Histogram of ids: h[x] += 1 for every x of a[0, m) (and of b[0, m) when b
// is not null).  Ids below HOT are counted in LDS first: skewed inputs
// (R-MAT, Barabasi-Albert) put their hubs at small ids, whose global atomics
// would otherwise serialise in one L2 channel.
template <int HOT>
__global__ __launch_bounds__(256) void k_count_ids(const uint32_t* __restrict__ a,
                                                  const uint32_t* __restrict__ b, uint64_t m,
                                                  uint32_t n, uint32_t* __restrict__ h) {
  __shared__ uint32_t hot[HOT];
  for (uint32_t i = threadIdx.x; i < (uint32_t)HOT; i += blockDim.x) hot[i] = 0;
  __syncthreads();
  for (uint64_t i = gtid(); i < m; i += gstride()) {
    const uint32_t x = a[i];
    if (x < (uint32_t)HOT) atomicAdd(hot + x, 1u);
    else atomicAdd(h + x, 1u);
    if (b) {
      const uint32_t y = b[i];
      if (y < (uint32_t)HOT) atomicAdd(hot + y, 1u);
      else atomicAdd(h + y, 1u);
    }
  }
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < (uint32_t)HOT && i < n; i += blockDim.x)
    if (hot[i]) atomicAdd(h + i, hot[i]);
}

// rowptr[r] = first position of the ascending (src << 32 | dst) keys whose
// src is >= r, for r in [0, n]: one independent search per row.
template <typename Key>
__global__ void k_rowptr_search(const Key* __restrict__ keys, uint64_t m, uint32_t n,
                                uint32_t* __restrict__ rowptr) {
  for (uint64_t r = gtid(); r <= n; r += gstride()) {
    const Key x = (Key)(r << 32);
    uint64_t lo = 0, hi = m;
    while (lo < hi) {
      const uint64_t mid = (lo + hi) >> 1;
      if (keys[mid] < x) lo = mid + 1;
      else hi = mid;
    }
    rowptr[r] = (uint32_t)lo;
  }
}

// x in a[lo, hi) for ascending a (global memory, early exit)
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

// Branchless membership test in an ascending LDS list a[0, n): the loop
// count depends on n only, so the lanes of a wave never diverge.  n must be
// wave-uniform: it is read from the first lane, so the loop's counters live
// in scalar registers and a step costs one add, one compare and one select.
// UNROLL (the work-list kernels): four halvings per trip while at least four
// remain (len > 15), a quarter of the loop's scalar bookkeeping.  The
// per-vertex and push-form kernels keep the one-halving loop: with the
// unrolled search the compiler moved their block variant out of line, and
// the per-vertex kernel in that form once stopped returning on R-MAT
// graphs (the cause was not found; the inlined form has always finished).
template <bool UNROLL = false>
__device__ inline uint32_t lds_has(const uint32_t* a, uint32_t n, uint32_t x) {
  n = (uint32_t)__builtin_amdgcn_readfirstlane((int)n);
  const uint32_t* p = a;
  uint32_t len = n;
  if constexpr (UNROLL) {
    while (len > 15) {
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const uint32_t half = len >> 1;
        p = p[half] <= x ? p + half : p;
        len -= half;
      }
    }
  }
  while (len > 1) {
    const uint32_t half = len >> 1;
    p = p[half] <= x ? p + half : p;
    len -= half;
  }
  return (n > 0 && *p == x) ? 1u : 0u;
}

__device__ inline void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Inclusive prefix sum over the 64 lanes of a wave.
__device__ inline uint32_t wave_incl_scan(uint32_t x, uint32_t lane) {
  for (uint32_t off = 1; off < 64; off <<= 1) {
    const uint32_t y = __shfl_up(x, off, 64);
    if (lane >= off) x += y;
  }
  return x;
}

// Inclusive prefix sum over the 64 lanes of a wave in six DPP adds (no LDS
// round trips): shifts by 1 and 2 within each 16-lane row give every lane
// the sum of its last four; shifts by 4 and 8 (on the banks above the
// first one and two) complete each row; the last lane of rows 0 and 2 is
// then added to rows 1 and 3, and that of row 1 to rows 2 and 3.  Lanes
// whose source is disabled or outside the row add 0.
__device__ inline uint32_t wave_incl_scan_dpp(uint32_t x) {
  x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x111, 0xf, 0xf, false);  // row_shr:1
  x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x112, 0xf, 0xf, false);  // row_shr:2
  x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x114, 0xf, 0xe, false);  // row_shr:4
  x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x118, 0xf, 0xc, false);  // row_shr:8
  x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x142, 0xa, 0xf, false);  // row_bcast:15
  x += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)x, 0x143, 0xc, 0xf, false);  // row_bcast:31
  return x;
}

// Number of set bits of m below this lane, plus q.
__device__ inline uint32_t lanes_below(uint64_t m, uint32_t q = 0u) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                   __builtin_amdgcn_mbcnt_lo((uint32_t)m, q));
}

// A wave-uniform value held in a scalar register.
__device__ inline uint32_t uniform(uint32_t v) {
  return (uint32_t)__builtin_amdgcn_readfirstlane((int)v);
}

// The lanes where b holds.  A predicate that is already a lane mask stays
// one (HIP's int-valued ballot turns it into a value and back: two vector
// instructions).
__device__ inline uint64_t lanes(bool b) { return __builtin_amdgcn_ballot_w64(b); }

// Largest j in [0, 64) with F[j] <= f, for ascending F with F[0] <= f
// (branchless, six LDS reads).
__device__ inline uint32_t seg64(const uint32_t* F, uint32_t f) {
  const uint32_t* p = F;
  uint32_t len = 64;
  while (len > 1) {
    const uint32_t half = len >> 1;
    p = p[half] <= f ? p + half : p;
    len -= half;
  }
  return (uint32_t)(p - F);
}

// Hash of a vertex id to (32 - shift) bits: the low bits with the bits above
// bit 15 folded in (neighbouring ids, common after A-order, stay spread).
__device__ inline uint32_t fhash(uint32_t x, uint32_t shift) {
  return (x ^ (x >> 15)) & (0xFFFFFFFFu >> shift);
}

__device__ inline void bm_set(uint32_t* bm, uint32_t x, uint32_t shift) {
  const uint32_t h = fhash(x, shift);
  atomicOr(bm + (h >> 5), 1u << (h & 31));
}

// Per-wave probe state for one vertex u (all fields wave-uniform).
struct Probe {
  const uint32_t* a;   // N+(u), ascending: LDS, or global when `global`
  uint32_t n;          // d+(u)
  const uint32_t* bm;  // hashed bitmap of N+(u) (LDS), used when `filt`
  uint32_t shift;      // 32 - log2(bitmap bits)
  uint32_t* Q;         // kQueue-entry survivor queue (LDS)
  uint32_t qn;         // queued survivors
  bool filt;
  bool global;
};

template <bool UNROLL = false>
__device__ inline uint32_t probe_search(const Probe& P, uint32_t x) {
  if (!P.global) return lds_has<UNROLL>(P.a, P.n, x);
  return (x != kEmpty && contains(P.a, 0, P.n, x)) ? 1u : 0u;
}

// Bitmap bit of x, 0 or 1 (any x: the word index is masked to the bitmap).
// The bit extract reads only the low five bits of its offset, which the
// mask keeps, so it takes the hash before the mask.
__device__ inline uint32_t bm_bit(const Probe& P, uint32_t x) {
  const uint32_t f = x ^ (x >> 15);
  // word (fhash(x) >> 5): bits 5 .. 31 - shift of f
  const uint32_t w = P.bm[__builtin_amdgcn_ubfe(f, 5u, 27u - P.shift)];
  return __builtin_amdgcn_ubfe(w, f, 1u);
}

__device__ inline bool bm_hit(const Probe& P, uint32_t x) { return bm_bit(P, x) != 0u; }

// Appends x (kEmpty: nothing) to the queue when it passes the filter.  The
// bitmap is read whether or not the filter is on (where it is off the probe
// decides nothing), so the predicate is a plain lane mask.  The queue count
// stays wave-uniform (a sum of ballot counts) and is kept in a scalar
// register.
__device__ inline void push(Probe& P, uint32_t x) {
  const bool pass = (x != kEmpty) & (bm_hit(P, x) | !P.filt);
  const uint64_t m = lanes(pass);
  if (pass) P.Q[lanes_below(m, P.qn)] = x;
  P.qn = uniform(P.qn + (uint32_t)__popcll(m));
}

// push() of four entries, with the four bitmap probes issued before any
// queue write so that their LDS reads overlap.  Faster for long streamed
// lists; short-list kernels keep the one-entry form, which needs fewer
// registers.  Entries past a list's end (kEmpty) are not told apart: one
// whose bitmap bit is set (or any, with the filter off) is queued, and the
// search rejects it, as no vertex id is kEmpty.  Each lane mask is the
// compare of the extracted bit.  FILT: the caller guarantees P.filt (the
// work-list kernels build the bitmap for every task), which drops the
// filter-off masks.
template <bool FILT = false>
__device__ inline void push4(Probe& P, uint32_t x0, uint32_t x1, uint32_t x2, uint32_t x3) {
  const bool nf = !FILT && !P.filt;
  const uint64_t all = nf ? ~0ull : 0ull;
  const uint32_t b0 = bm_bit(P, x0), b1 = bm_bit(P, x1), b2 = bm_bit(P, x2), b3 = bm_bit(P, x3);
  const bool c0 = b0 != 0u, c1 = b1 != 0u, c2 = b2 != 0u, c3 = b3 != 0u;
  const uint64_t m0 = lanes(c0) | all, m1 = lanes(c1) | all;
  const uint64_t m2 = lanes(c2) | all, m3 = lanes(c3) | all;
  uint32_t q = P.qn;
  if (c0 || nf) P.Q[lanes_below(m0, q)] = x0;
  q += (uint32_t)__popcll(m0);
  if (c1 || nf) P.Q[lanes_below(m1, q)] = x1;
  q += (uint32_t)__popcll(m1);
  if (c2 || nf) P.Q[lanes_below(m2, q)] = x2;
  q += (uint32_t)__popcll(m2);
  if (c3 || nf) P.Q[lanes_below(m3, q)] = x3;
  P.qn = uniform(q + (uint32_t)__popcll(m3));
}

// Resolves queued survivors 64 at a time; `all` also resolves a partial tail.
// WL: the work-list kernels' search (see lds_has).
template <bool WL = false>
__device__ inline uint32_t drain(Probe& P, uint32_t lane, bool all) {
  uint32_t c = 0;
  if (P.qn >= 64 || (all && P.qn)) {
    wave_sync();
    while (P.qn >= 64) {
      P.qn -= 64;
      c += probe_search<WL>(P, P.Q[P.qn + lane]);
    }
    if (all && P.qn) {
      c += probe_search<WL>(P, lane < P.qn ? P.Q[lane] : kEmpty);
      P.qn = 0;
    }
    wave_sync();
  }
  return c;
}

// Streams col[vb, ve) with the whole wave, 4 * U 64-entry chunks in flight
// (the queue is drained after every four).  U = 2 suits long lists (R-MAT
// hubs); U = 1 wastes fewer lanes on short ones (many small graphs).
template <int U = 2, bool WL = false>
__device__ inline uint32_t stream_list(const uint32_t* __restrict__ col, uint32_t vb, uint32_t ve,
                                       uint32_t lane, Probe& P) {
  uint32_t c = 0;
  for (uint32_t k0 = vb; k0 < ve; k0 += 256 * U) {
    const uint32_t k = k0 + lane;
    uint32_t w[4 * U];
#pragma unroll
    for (int j = 0; j < 4 * U; ++j) w[j] = k + 64 * j < ve ? col[k + 64 * j] : kEmpty;
#pragma unroll
    for (int h = 0; h < U; ++h) {
      if constexpr (U >= 2) {
        push4(P, w[4 * h], w[4 * h + 1], w[4 * h + 2], w[4 * h + 3]);
      } else {
#pragma unroll
        for (int j = 0; j < 4; ++j) push(P, w[4 * h + j]);
      }
      c += drain<WL>(P, lane, false);
    }
  }
  return c;
}

// The long lists of a wave's 64 lists (the lanes set in `mask`; lane j
// holds col[vb, ve)) streamed back to back by the whole wave: 4 * U chunks
// of 64 entries in flight, taken in list order across list boundaries, so
// that lists of a few chunks each (R-MAT's work lists average about 270
// entries) still keep 4 * U chunks of loads in flight per wave.
template <int U>
__device__ inline uint32_t stream_many(const uint32_t* __restrict__ col, uint64_t mask,
                                       uint32_t vb, uint32_t ve, uint32_t lane, Probe& P) {
  uint32_t c = 0;
  uint32_t pos = 0, end = 0;  // wave-uniform: the current list's next chunk and its end
  auto next_list = [&]() {
    const uint32_t j = (uint32_t)__builtin_ctzll(mask);
    mask &= mask - 1;
    pos = (uint32_t)__builtin_amdgcn_readlane((int)vb, (int)j);
    end = (uint32_t)__builtin_amdgcn_readlane((int)ve, (int)j);
  };
  if (mask) next_list();
  while (pos < end) {
    uint32_t w[4 * U];
#pragma unroll
    for (int s = 0; s < 4 * U; ++s) {
      const uint32_t k = pos + lane;  // no wrap: CSRs hold fewer than 2^32 - 512 entries
      w[s] = k < end ? col[k] : kEmpty;
      pos += 64;
      if (pos >= end) {
        if (mask)
          next_list();
        else
          pos = end = 0;  // all lists done: the remaining chunks load nothing
      }
    }
#pragma unroll
    for (int h = 0; h < U; ++h) {
      push4<true>(P, w[4 * h], w[4 * h + 1], w[4 * h + 2], w[4 * h + 3]);
      c += drain<true>(P, lane, false);
    }
  }
  return c;
}

// Each lane holds one list col[vb, ve) (vb == ve for none).  Lists of
// kFlatMin entries or more are streamed by the whole wave (with MANY, back to
// back through stream_many; else one at a time); the shorter ones are
// concatenated and walked 64 entries per step (two steps in flight), each
// lane finding the list of its entry by a search over the exclusive prefix
// of the lengths, kept with the list starts in the wave's LDS slot F[128].
// WL (the work-list kernels) streams through stream_many and searches with
// the unrolled loop; the per-vertex and push-form kernels keep the
// one-list-at-a-time stream and the one-halving search.
template <int U = 2, bool WL = false>
__device__ inline uint32_t lists64(const uint32_t* __restrict__ col, uint32_t vb, uint32_t ve,
                                   uint32_t lane, uint32_t* F, Probe& P) {
  const uint32_t len = ve - vb;
  const bool lng = len >= kFlatMin;
  uint32_t c = 0;
  uint64_t mask = lanes(lng);
  if constexpr (WL && U >= 2) {
    if (mask) c += stream_many<U>(col, mask, vb, ve, lane, P);
  } else {
    while (mask) {
      const uint32_t j = (uint32_t)__builtin_ctzll(mask);
      mask &= mask - 1;
      c += stream_list<U, WL>(col, __shfl(vb, j, 64), __shfl(ve, j, 64), lane, P);
    }
  }
  const uint32_t sl = lng ? 0u : len;
  const uint32_t incl = wave_incl_scan(sl, lane);
  const uint32_t total = __shfl(incl, 63, 64);
  if (total == 0) return c;
  F[lane] = incl - sl;
  F[64 + lane] = vb;
  wave_sync();
  for (uint32_t f0 = 0; f0 < total; f0 += 128) {
    const uint32_t fa = f0 + lane, fb = fa + 64;
    uint32_t xa = kEmpty, xb = kEmpty;
    if (fa < total) {
      const uint32_t j = seg64(F, fa);
      xa = col[F[64 + j] + (fa - F[j])];
    }
    if (fb < total) {
      const uint32_t j = seg64(F, fb);
      xb = col[F[64 + j] + (fb - F[j])];
    }
    push(P, xa);
    push(P, xb);
    c += drain<WL>(P, lane, false);
  }
  wave_sync();  // F is rewritten by the next call
  return c;
}

// Bounds col[vb, ve) of list j of a work list: `wl` holds vertex ids (the
// bounds read from rp: a random row-offset line per list) or, as uint2, the
// bounds themselves (read in sequence with the list).
__device__ inline void wl_bounds(const uint32_t* __restrict__ rp, const uint32_t* __restrict__ wl,
                                 uint32_t j, uint32_t& vb, uint32_t& ve) {
  const uint32_t v = wl[j];
  vb = rp[v];
  ve = rp[v + 1];
}
__device__ inline void wl_bounds(const uint32_t* __restrict__, const uint2* __restrict__ wl,
                                 uint32_t j, uint32_t& vb, uint32_t& ve) {
  const uint2 b = wl[j];
  vb = b.x;
  ve = b.y;
}

// LDS of one wave for the wave variant.
struct WaveSlot {
  uint32_t* L;   // N+(u), kWaveMax entries
  uint32_t* bm;  // hashed bitmap, 2^(32 - shift) bits, all zero between vertices
  uint32_t shift;
  uint32_t* F;   // 128 words
  uint32_t* Q;   // kQueue words
};

// Wave variant: one wave counts the wedges of u (d+(u) <= kWaveMax).  Lane j
// takes the j-th out-neighbour of each 64-neighbour group.  The bitmap is
// built for d+(u) > FMIN and its words are cleared again at the end (below
// that, every entry goes straight to the search).
// With WL, the lists streamed against N+(u) are N+(y) for y in the work list
// wl[0, wn) (global memory; ids or bounds, see wl_bounds) instead of N+(v)
// for v in N+(u).
template <uint32_t FMIN = kFilterMin, int U = 2, bool WL = false, typename WLE = uint32_t>
__device__ inline uint32_t wave_vertex(const uint32_t* __restrict__ rp,
                                       const uint32_t* __restrict__ col, uint32_t u,
                                       uint32_t lane, const WaveSlot& W,
                                       const WLE* __restrict__ wl = nullptr, uint32_t wn = 0) {
  const uint32_t ub = rp[u], du = uniform(rp[u + 1] - ub);
  // the work-list stream (stream_many) relies on the filter being on
  const bool filt = (WL && U >= 2) || du > FMIN;
  for (uint32_t i = lane; i < du; i += 64) {
    const uint32_t x = col[ub + i];
    W.L[i] = x;
    if (filt) bm_set(W.bm, x, W.shift);
  }
  wave_sync();
  Probe P{W.L, du, W.bm, W.shift, W.Q, 0u, filt, false};
  uint32_t c = 0;
  const uint32_t nl = WL ? wn : du;
  for (uint32_t j0 = 0; j0 < nl; j0 += 64) {
    uint32_t mvb = 0, mve = 0;
    if (j0 + lane < nl) {
      if constexpr (WL) {
        wl_bounds(rp, wl, j0 + lane, mvb, mve);
      } else {
        const uint32_t v = W.L[j0 + lane];
        mvb = rp[v];
        mve = rp[v + 1];
      }
    }
    c += lists64<U, WL>(col, mvb, mve, lane, W.F, P);
  }
  c += drain<WL>(P, lane, true);
  if (filt)
    for (uint32_t i = lane; i < du; i += 64) W.bm[fhash(W.L[i], W.shift) >> 5] = 0;
  wave_sync();
  return c;
}

// Workgroup LDS: N+(u) and its bitmap (CAP * 16 bits) for block tasks; in
// wave phases both split into four per-wave slots.
template <int CAP>
struct BlockLds {
  uint32_t u[CAP];
  uint32_t bm[CAP / 2];
  uint32_t f[TB / 64][128];
  uint32_t q[TB / 64][kQueue];
  uint32_t t;
  static constexpr uint32_t kBlockShift = 32u - (uint32_t)__builtin_ctz((unsigned)CAP * 16u);
  static constexpr uint32_t kWaveShift = 32u - (uint32_t)__builtin_ctz((unsigned)CAP * 4u);
  static_assert(CAP / 4 >= (int)kWaveMax, "wave slots must hold kWaveMax entries");
};

template <int CAP>
__device__ inline WaveSlot wave_slot(BlockLds<CAP>& S, uint32_t wv) {
  return WaveSlot{S.u + wv * (CAP / 4), S.bm + wv * (CAP / 8), BlockLds<CAP>::kWaveShift,
                  S.f[wv], S.q[wv]};
}

// Zeroes the bitmap region (once per workgroup, before the first vertex).
template <int CAP>
__device__ inline void clear_bitmap(BlockLds<CAP>& S) {
  for (uint32_t i = threadIdx.x; i < (uint32_t)CAP / 2; i += TB) S.bm[i] = 0;
  __syncthreads();
}

// Block variant: the whole workgroup counts the wedges of u.  N+(u) is copied
// into LDS (when it fits; longer lists are searched in global memory) and
// hashed into the bitmap; wave w takes the out-neighbours w, w+4, w+8, ...
// 64 at a time.  Every thread must call it; it ends with the bitmap clear
// and a barrier.  With WL, the lists streamed are N+(y) for y in the work
// list wl[0, wn) (global memory; ids or bounds, see wl_bounds), wave w
// taking entries w, w+4, w+8, ...
template <int CAP, int U = 2, bool WL = false, typename WLE = uint32_t>
__device__ inline uint32_t block_vertex(const uint32_t* __restrict__ rp,
                                        const uint32_t* __restrict__ col, uint32_t u,
                                        BlockLds<CAP>& S, const WLE* __restrict__ wl = nullptr,
                                        uint32_t wn = 0) {
  constexpr uint32_t shift = BlockLds<CAP>::kBlockShift;
  const uint32_t tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const uint32_t ub = rp[u], du = uniform(rp[u + 1] - ub);
  const bool in_lds = du <= (uint32_t)CAP;
  for (uint32_t i = tid; i < du; i += TB) {
    const uint32_t x = col[ub + i];
    if (in_lds) S.u[i] = x;
    bm_set(S.bm, x, shift);
  }
  __syncthreads();
  const uint32_t* a = in_lds ? S.u : col + ub;
  Probe P{a, du, S.bm, shift, S.q[wv], 0u, true, !in_lds};
  uint32_t c = 0;
  const uint32_t nl = WL ? wn : du;
  const uint32_t mine = nl > wv ? (nl - wv + 3) / 4 : 0;
  for (uint32_t g0 = 0; g0 < mine; g0 += 64) {
    uint32_t vb = 0, ve = 0;
    if (g0 + lane < mine) {
      if constexpr (WL) {
        wl_bounds(rp, wl, 4 * (g0 + lane) + wv, vb, ve);
      } else {
        const uint32_t v = a[4 * (g0 + lane) + wv];
        vb = rp[v];
        ve = rp[v + 1];
      }
    }
    c += lists64<U, WL>(col, vb, ve, lane, S.f[wv], P);
  }
  c += drain<WL>(P, lane, true);
  __syncthreads();
  for (uint32_t i = tid; i < du; i += TB) S.bm[fhash(a[i], shift) >> 5] = 0;
  __syncthreads();
  return c;
}

// The AUTO count kernel (one launch, persistent grid):
//  phase 1 - block tasks (vertex list `blk`, in the schedule's order), one
//            workgroup each via cursor[0];
//  phase 2 - wave tasks (list `wav`, d+ <= kWaveMax), one wave each,
//            kWaveChunk at a time via cursor[1]; they fill the tail left by
//            phase 1.
template <int CAP, bool STATIC_WAVES = false>
__global__ __launch_bounds__(TB) void k_auto(const uint32_t* __restrict__ rp,
                                             const uint32_t* __restrict__ col,
                                             const uint32_t* __restrict__ blk, uint32_t nblk,
                                             const uint32_t* __restrict__ wav, uint32_t nwav,
                                             uint32_t* cursor, unsigned long long* total) {
  __shared__ BlockLds<CAP> S;
  __shared__ unsigned long long red[TB / 64];
  const uint32_t tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  clear_bitmap(S);
  unsigned long long cnt = 0;
  for (bool first = true;; first = false) {
    if (tid == 0) S.t = claim(cursor, nblk, 1u, first);
    __syncthreads();
    const uint32_t t = S.t;
    if (t >= nblk) break;
    cnt += block_vertex<CAP>(rp, col, blk[t], S);
  }
  const WaveSlot W = wave_slot(S, wv);
  if constexpr (STATIC_WAVES) {
    // few wave tasks: a fixed stride over them, no counter (claims on one
    // atomic counter are throughput-bound, about 20 ns each)
    const uint32_t nw = gridDim.x * (TB / 64);
    for (uint32_t i = blockIdx.x * (TB / 64) + wv; i < nwav; i += nw)
      cnt += wave_vertex(rp, col, wav[i], lane, W);
    flush_block(cnt, total, red);
    return;
  }
  for (bool first = true;; first = false) {
    uint32_t i0 = 0;
    if (lane == 0) i0 = claim(cursor + 1, nwav, kWaveChunk, first);
    i0 = __shfl(i0, 0, 64);
    if (i0 >= nwav) break;
    const uint32_t i1 = min(nwav, i0 + kWaveChunk);
    for (uint32_t i = i0; i < i1; ++i) cnt += wave_vertex(rp, col, wav[i], lane, W);
  }
  flush_block(cnt, total, red);
}

// Lane task: one lane counts task x of the work-list schedule when
// d+(x) <= kLaneMaxDeg.  Out-lists come from `ell`, the out-lists of up to
// eight entries as fixed 32-byte rows padded with kEmpty: a row is two
// 16-byte loads from one cache line and needs no row-offset read (with
// lists of a few entries at random rows, the offsets' line and a list that
// straddles two lines made up most of the traffic).  N+(x) is held in
// registers; every list N+(y), y in W(x), has at most d+(x) entries (the
// work-list rule streams the shorter list), and each of its entries is
// compared with all of N+(x).  The padding of N+(x) is turned into
// kEmpty - 1 so that it matches none of a streamed row's (vertex ids are
// below 2^31 - 1).
static_assert(kLaneMaxDeg == 8, "lane rows are 32 bytes");
__device__ inline uint32_t lane_task(const uint4* __restrict__ ell,
                                     const uint32_t* __restrict__ wrp,
                                     const uint32_t* __restrict__ wcol, uint32_t x) {
  const uint4 a0 = ell[2 * (uint64_t)x], a1 = ell[2 * (uint64_t)x + 1];
  uint32_t s[kLaneMaxDeg] = {a0.x, a0.y, a0.z, a0.w, a1.x, a1.y, a1.z, a1.w};
#pragma unroll
  for (uint32_t k = 0; k < kLaneMaxDeg; ++k) s[k] = s[k] == kEmpty ? kEmpty - 1 : s[k];
  const uint32_t wb = wrp[x], we = wrp[x + 1];
  uint32_t c = 0;
  for (uint32_t i = wb; i < we; ++i) {
    const uint32_t y = wcol[i];
    const uint4 b0 = ell[2 * (uint64_t)y], b1 = ell[2 * (uint64_t)y + 1];
    const uint32_t e[kLaneMaxDeg] = {b0.x, b0.y, b0.z, b0.w, b1.x, b1.y, b1.z, b1.w};
#pragma unroll
    for (uint32_t k = 0; k < kLaneMaxDeg; ++k) {
#pragma unroll
      for (uint32_t j = 0; j < kLaneMaxDeg; ++j) c += e[k] == s[j] ? 1u : 0u;
    }
  }
  return c;
}

// ---- Rank-indexed membership (the work-list kernels) -----------------------
// N+(x) is hashed (fhash) into a bitmap whose words sit in `bw` as pairs
// (bits, prefix): .y holds the number of set bits in the words before
// (bits 0..30) and, in bit 31, a flag for a word in which two members share
// a bit.  The member whose hash is the r-th set bit is stored at H[r], and
// H[total] = kEmpty - 1 (no id, and not the padding kEmpty).  An entry y is
// a member iff H[popcount(bits below y's bit) + prefix] == y: one 8-byte and
// one 4-byte LDS read and a handful of vector instructions, with no queue
// and no search.  A colliding member may have lost its H slot, so a miss in
// a flagged word is queued and resolved by a binary search of N+(x) in
// global memory (sorted), 64 at a time.  The flag bit drops out of the H
// address (r << 2).

// LDS of the work-list kernel: one table region, seen as the workgroup's
// (block tasks: CAP members, CAP * 16 bits) or as four wave slots (wave
// tasks: CAP / 4 members, CAP * 4 bits each).
template <int CAP>
struct RankLds {
  static constexpr uint32_t kWords = CAP / 2;                  // block bitmap words
  static constexpr uint32_t kWaveWords = CAP / 8;              // wave slot bitmap words
  static constexpr uint32_t kWaveH = CAP / 4 + 2;              // >= CAP / 4 + 1, even
  static constexpr uint32_t kSlot = 2 * kWaveWords + kWaveH;   // words per wave slot
  static constexpr uint32_t kRegion = 4 * kSlot;               // words
  static_assert(kRegion >= 2 * kWords + CAP + 1, "the block view needs CAP + 1 H entries");
  static_assert(CAP / 4 >= (int)kWaveMax, "wave slots must hold kWaveMax members");
  static constexpr uint32_t kQ = 96;                           // secondary queue per wave
  alignas(16) uint32_t region[kRegion];
  uint32_t f[TB / 64][128];
  uint32_t q[TB / 64][kQ];
  uint32_t wsum[TB / 64];
  uint32_t t;
};

struct RankSlot {
  uint2* bw;
  uint32_t* H;
  uint32_t nwm;   // bitmap words - 1
  uint32_t* F;
  uint32_t* Q;
};

template <int CAP>
__device__ inline RankSlot rank_block(RankLds<CAP>& S, uint32_t wv) {
  return RankSlot{reinterpret_cast<uint2*>(S.region), S.region + 2 * RankLds<CAP>::kWords,
                  RankLds<CAP>::kWords - 1, S.f[wv], S.q[wv]};
}

template <int CAP>
__device__ inline RankSlot rank_wave(RankLds<CAP>& S, uint32_t wv) {
  uint32_t* b = S.region + wv * RankLds<CAP>::kSlot;
  return RankSlot{reinterpret_cast<uint2*>(b), b + 2 * RankLds<CAP>::kWaveWords,
                  RankLds<CAP>::kWaveWords - 1, S.f[wv], S.q[wv]};
}

__device__ inline void rank_insert(uint2* bw, uint32_t nwm, uint32_t x) {
  const uint32_t f = x ^ (x >> 15);
  const uint32_t wi = (f >> 5) & nwm, bit = 1u << (f & 31);
  const uint32_t old = atomicOr(&bw[wi].x, bit);
  if (old & bit) atomicOr(&bw[wi].y, 0x80000000u);
}

__device__ inline void rank_place(const uint2* bw, uint32_t* H, uint32_t nwm, uint32_t x) {
  const uint32_t f = x ^ (x >> 15);
  const uint2 e = bw[(f >> 5) & nwm];
  H[__builtin_popcount(__builtin_amdgcn_ubfe(e.x, 0u, f)) + (e.y & 0x7FFFFFFFu)] = x;
}

__device__ inline void rank_clear(uint2* bw, uint32_t nwm, uint32_t x) {
  const uint32_t f = x ^ (x >> 15);
  bw[(f >> 5) & nwm] = make_uint2(0u, 0u);
}

// Prefix of the set bits over the NW words of bw (the whole workgroup; K
// consecutive words per thread); returns the total.  Ends with a barrier.
template <uint32_t NW>
__device__ inline uint32_t rank_scan_block(uint2* bw, uint32_t* wsum) {
  constexpr uint32_t K = NW / TB;
  static_assert(K * TB == NW, "words per thread");
  const uint32_t tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  uint32_t c[K], s = 0;
#pragma unroll
  for (uint32_t k = 0; k < K; ++k) {
    c[k] = (uint32_t)__builtin_popcount(bw[tid * K + k].x);
    s += c[k];
  }
  const uint32_t incl = wave_incl_scan_dpp(s);
  if (lane == 63) wsum[wv] = incl;
  __syncthreads();
  uint32_t before = 0, total = 0;
#pragma unroll
  for (uint32_t w = 0; w < TB / 64; ++w) {
    const uint32_t t = wsum[w];
    before += w < wv ? t : 0u;
    total += t;
  }
  uint32_t p = before + incl - s;
#pragma unroll
  for (uint32_t k = 0; k < K; ++k) {
    bw[tid * K + k].y = (bw[tid * K + k].y & 0x80000000u) | p;
    p += c[k];
  }
  __syncthreads();
  return total;
}

// Per-wave probe state (all fields wave-uniform).
struct RProbe {
  const uint2* bw;
  const uint32_t* H;
  uint32_t amask;            // byte offsets of bw words: (nwm << 3)
  uint32_t* Q;               // secondary queue (RankLds::kQ entries)
  uint32_t qn;
  const uint32_t* sa;        // N+(x) in global memory, ascending
  uint32_t sn;
  uint32_t hits;
};

// Resolves up to 64 queued entries by binary search in global memory.
__device__ inline void rank_resolve(RProbe& P, uint32_t lane) {
  wave_sync();
  const uint32_t take = P.qn < 64u ? P.qn : 64u;
  const uint32_t y = lane < take ? P.Q[P.qn - take + lane] : kEmpty;
  const bool hit = y != kEmpty && contains(P.sa, 0, P.sn, y);
  P.hits = uniform(P.hits + (uint32_t)__popcll(lanes(hit)));
  P.qn = uniform(P.qn - take);
  wave_sync();
}

// N entries (kEmpty: none) probed together: the table reads of all N are
// issued before their compares.
template <int N>
__device__ inline void rank_probe(RProbe& P, const uint32_t (&x)[N], uint32_t lane) {
  uint32_t f[N], h[N];
  uint2 e[N];
#pragma unroll
  for (int i = 0; i < N; ++i) {
    f[i] = x[i] ^ (x[i] >> 15);
    e[i] = *reinterpret_cast<const uint2*>(reinterpret_cast<const char*>(P.bw) +
                                           ((f[i] >> 2) & P.amask));
  }
#pragma unroll
  for (int i = 0; i < N; ++i) {
    const uint32_t r = (uint32_t)__builtin_popcount(__builtin_amdgcn_ubfe(e[i].x, 0u, f[i])) + e[i].y;
    h[i] = *reinterpret_cast<const uint32_t*>(reinterpret_cast<const char*>(P.H) + (r << 2));
  }
  uint32_t hs = 0;  // scalar: the hits of the N entries
#pragma unroll
  for (int i = 0; i < N; ++i) {
    const uint64_t hm = lanes(h[i] == x[i]);
    // misses on a set bit of a flagged word (an unset bit holds no member)
    const uint64_t sm = lanes((int)e[i].y < 0) & lanes(__builtin_amdgcn_ubfe(e[i].x, f[i], 1u) != 0u) & ~hm;
    hs += (uint32_t)__popcll(hm);
    if (sm) {
      if ((sm >> lane) & 1ull) P.Q[lanes_below(sm, P.qn)] = x[i];
      P.qn = uniform(P.qn + (uint32_t)__popcll(sm));
      if (P.qn >= 32) rank_resolve(P, lane);
    }
  }
  P.hits = uniform(P.hits + hs);
}

__device__ inline void rank_finish(RProbe& P, uint32_t lane) {
  while (P.qn) rank_resolve(P, lane);
}

// stream_many with rank probes.
__device__ inline void rank_stream_many(const uint32_t* __restrict__ col, uint64_t mask,
                                        uint32_t vb, uint32_t ve, uint32_t lane, RProbe& P) {
  uint32_t pos = 0, end = 0;
  auto next_list = [&]() {
    const uint32_t j = (uint32_t)__builtin_ctzll(mask);
    mask &= mask - 1;
    pos = (uint32_t)__builtin_amdgcn_readlane((int)vb, (int)j);
    end = (uint32_t)__builtin_amdgcn_readlane((int)ve, (int)j);
  };
  if (mask) next_list();
  while (pos < end) {
    uint32_t w[8];
#pragma unroll
    for (int s = 0; s < 8; ++s) {
      const uint32_t k = pos + lane;  // no wrap: CSRs hold fewer than 2^32 - 512 entries
      w[s] = k < end ? col[k] : kEmpty;
      pos += 64;
      if (pos >= end) {
        if (mask)
          next_list();
        else
          pos = end = 0;
      }
    }
    const uint32_t a[4] = {w[0], w[1], w[2], w[3]};
    rank_probe<4>(P, a, lane);
    const uint32_t b[4] = {w[4], w[5], w[6], w[7]};
    rank_probe<4>(P, b, lane);
  }
}

// lists64 with rank probes (long lists back to back, short ones
// concatenated across lanes through F).
__device__ inline void rank_lists64(const uint32_t* __restrict__ col, uint32_t vb, uint32_t ve,
                                    uint32_t lane, uint32_t* F, RProbe& P) {
  const uint32_t len = ve - vb;
  const bool lng = len >= kFlatMin;
  const uint64_t mask = lanes(lng);
  if (mask) rank_stream_many(col, mask, vb, ve, lane, P);
  const uint32_t sl = lng ? 0u : len;
  const uint32_t incl = wave_incl_scan_dpp(sl);
  const uint32_t total = __shfl(incl, 63, 64);
  if (total == 0) return;
  F[lane] = incl - sl;
  F[64 + lane] = vb;
  wave_sync();
  for (uint32_t f0 = 0; f0 < total; f0 += 128) {
    const uint32_t fa = f0 + lane, fb = fa + 64;
    uint32_t xa = kEmpty, xb = kEmpty;
    if (fa < total) {
      const uint32_t j = seg64(F, fa);
      xa = col[F[64 + j] + (fa - F[j])];
    }
    if (fb < total) {
      const uint32_t j = seg64(F, fb);
      xb = col[F[64 + j] + (fb - F[j])];
    }
    const uint32_t a[2] = {xa, xb};
    rank_probe<2>(P, a, lane);
  }
  wave_sync();
}

// Block task x (d+(x) <= CAP): the workgroup builds the table of N+(x) and
// wave w streams the lists of work-list entries w, w+4, ...  Every thread
// must call it; it ends with the table cleared and a barrier.  Returns the
// count in lane 0 of each wave (0 elsewhere).
template <int CAP>
__device__ inline uint32_t block_rank(const uint32_t* __restrict__ rp,
                                      const uint32_t* __restrict__ col, uint32_t x,
                                      RankLds<CAP>& S, const uint2* __restrict__ wl,
                                      uint32_t wn) {
  const uint32_t tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const RankSlot T = rank_block(S, wv);
  const uint32_t ub = rp[x], du = uniform(rp[x + 1] - ub);
  for (uint32_t i = tid; i < du; i += TB) rank_insert(T.bw, T.nwm, col[ub + i]);
  __syncthreads();
  const uint32_t total = rank_scan_block<RankLds<CAP>::kWords>(T.bw, S.wsum);
  if (tid == 0) T.H[total] = kEmpty - 1;
  for (uint32_t i = tid; i < du; i += TB) rank_place(T.bw, T.H, T.nwm, col[ub + i]);
  __syncthreads();
  RProbe P{T.bw, T.H, T.nwm << 3, T.Q, 0u, col + ub, du, 0u};
  const uint32_t mine = wn > wv ? (wn - wv + 3) / 4 : 0;
  for (uint32_t g0 = 0; g0 < mine; g0 += 64) {
    uint32_t vb = 0, ve = 0;
    if (g0 + lane < mine) {
      const uint2 b = wl[4 * (g0 + lane) + wv];
      vb = b.x;
      ve = b.y;
    }
    rank_lists64(col, vb, ve, lane, T.F, P);
  }
  rank_finish(P, lane);
  __syncthreads();
  for (uint32_t i = tid; i < total; i += TB) rank_clear(T.bw, T.nwm, T.H[i]);
  __syncthreads();
  return lane == 0 ? P.hits : 0u;
}

// Block task whose N+(x) exceeds the table (d+(x) > CAP): every streamed
// entry is searched in N+(x) in global memory.  Per-lane counts.
__device__ inline uint32_t block_global(const uint32_t* __restrict__ rp,
                                        const uint32_t* __restrict__ col, uint32_t x,
                                        const uint2* __restrict__ wl, uint32_t wn) {
  const uint32_t tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const uint32_t ub = rp[x], ue = rp[x + 1];
  uint32_t c = 0;
  for (uint32_t j = wv; j < wn; j += TB / 64) {
    const uint2 b = wl[j];
    for (uint32_t k = b.x + lane; k < b.y; k += 64) c += contains(col, ub, ue, col[k]) ? 1u : 0u;
  }
  return c;
}

// Prefix of the set bits over the first nw words of a wave slot (nw a
// multiple of 64, wave-uniform); returns the total (wave-uniform).
__device__ inline uint32_t rank_scan_wave(uint2* bw, uint32_t nw, uint32_t lane) {
  const uint32_t K = nw >> 6;
  uint32_t s = 0;
  for (uint32_t k = 0; k < K; ++k) s += (uint32_t)__builtin_popcount(bw[lane * K + k].x);
  const uint32_t incl = wave_incl_scan_dpp(s);
  const uint32_t total = uniform(__shfl(incl, 63, 64));
  uint32_t p = incl - s;
  for (uint32_t k = 0; k < K; ++k) {
    const uint2 e = bw[lane * K + k];
    bw[lane * K + k].y = (e.y & 0x80000000u) | p;
    p += (uint32_t)__builtin_popcount(e.x);
  }
  return total;
}

// Wave task x (d+(x) <= kWaveMax) in the wave's slot.  The table takes
// about 16 bits per member (64 words and up, the slot's at most), so the
// prefix pass of a short task is one word per lane.  Returns the count in
// lane 0 (0 elsewhere).
template <int CAP>
__device__ inline uint32_t wave_rank(const uint32_t* __restrict__ rp,
                                     const uint32_t* __restrict__ col, uint32_t x, uint32_t lane,
                                     const RankSlot& T, const uint2* __restrict__ wl,
                                     uint32_t wn) {
  static_assert(RankLds<CAP>::kWaveWords >= 64, "at least one word per lane");
  const uint32_t ub = rp[x], du = uniform(rp[x + 1] - ub);
  uint32_t nw = 64;
  while (nw < RankLds<CAP>::kWaveWords && 2 * nw < du) nw <<= 1;
  const uint32_t nwm = nw - 1;
  // the first 128 members are kept in F (free until the stream) for the
  // place pass: a short task reads N+(x) from memory once
  for (uint32_t i = lane; i < du; i += 64) {
    const uint32_t v = col[ub + i];
    if (i < 128) T.F[i] = v;
    rank_insert(T.bw, nwm, v);
  }
  wave_sync();
  const uint32_t total = rank_scan_wave(T.bw, nw, lane);
  if (lane == 0) T.H[total] = kEmpty - 1;
  wave_sync();
  for (uint32_t i = lane; i < du; i += 64) rank_place(T.bw, T.H, nwm, i < 128 ? T.F[i] : col[ub + i]);
  wave_sync();
  RProbe P{T.bw, T.H, nwm << 3, T.Q, 0u, col + ub, du, 0u};
  for (uint32_t j0 = 0; j0 < wn; j0 += 64) {
    uint32_t vb = 0, ve = 0;
    if (j0 + lane < wn) {
      const uint2 b = wl[j0 + lane];
      vb = b.x;
      ve = b.y;
    }
    rank_lists64(col, vb, ve, lane, T.F, P);
  }
  rank_finish(P, lane);
  wave_sync();
  for (uint32_t i = lane; i < total; i += 64) rank_clear(T.bw, nwm, T.H[i]);
  wave_sync();
  return lane == 0 ? P.hits : 0u;
}

// The AUTO count kernel over work lists (the default schedule): task x
// builds the rank-indexed table of N+(x) in LDS (block tasks: the
// workgroup's, d+(x) <= CAP; wave tasks: the wave's slot) and streams N+(y)
// for y in W(x) = wcol[wrp[x], wrp[x+1]), each entry decided by one table
// lookup.  A block task whose N+(x) exceeds CAP searches N+(x) in global
// memory instead.  Each oriented edge u->v sits in exactly one work
// list, at the endpoint whose task streams the shorter out-list (u when
// d+(v) <= d+(u), else v), so the sum over tasks of |N+(x) & N+(y)| is the
// triangle count; edges into a vertex with no out-edge are left out.
// Block and wave tasks read each entry's list bounds from wse (wse[i] =
// (rp[y], rp[y+1]) for y = wcol[i]), in sequence, instead of a random
// row-offset line per list.
// Phase 3 counts the lane tasks (`lan`), kLaneClaim per wave claim from
// cursor[2], from the 32-byte rows `ell` (see lane_task; unused when nlan is
// zero).
// Eight waves per SIMD (the persistent grid's eight workgroups per CU): the
// register budget is held at 64 VGPRs and the LDS (RankLds<2048>) at 20 KB
// per workgroup so the occupancy survives the unrolled streams.
template <int CAP, bool STATIC_WAVES = false>
__global__ __launch_bounds__(TB) __attribute__((amdgpu_waves_per_eu(8, 8)))
void k_auto_wl(const uint32_t* __restrict__ rp, const uint32_t* __restrict__ col,
               const uint32_t* __restrict__ wrp, const uint32_t* __restrict__ wcol,
               const uint2* __restrict__ wse, const uint4* __restrict__ ell,
               const uint32_t* __restrict__ blk, uint32_t nblk,
               const uint32_t* __restrict__ wav, uint32_t nwav,
               const uint32_t* __restrict__ lan, uint32_t nlan,
               uint32_t* cursor, unsigned long long* total) {
  __shared__ RankLds<CAP> S;
  __shared__ unsigned long long red[TB / 64];
  const uint32_t tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  for (uint32_t i = tid; i < RankLds<CAP>::kRegion; i += TB) S.region[i] = 0;
  __syncthreads();
  unsigned long long cnt = 0;
  for (bool first = true;; first = false) {
    if (tid == 0) S.t = claim(cursor, nblk, 1u, first);
    __syncthreads();
    const uint32_t t = S.t;
    if (t >= nblk) break;
    const uint32_t x = blk[t];
    const uint32_t wb = wrp[x], wn = wrp[x + 1] - wb;
    if (uniform(rp[x + 1] - rp[x]) <= (uint32_t)CAP) {
      [[clang::always_inline]] cnt += block_rank<CAP>(rp, col, x, S, wse + wb, wn);
    } else {
      cnt += block_global(rp, col, x, wse + wb, wn);
      __syncthreads();  // every wave has read S.t before it is claimed again
    }
  }
  // the wave slots overlap the block view's table: start them from zero
  for (uint32_t i = tid; i < RankLds<CAP>::kRegion; i += TB) S.region[i] = 0;
  __syncthreads();
  const RankSlot W = rank_wave(S, wv);
  if constexpr (STATIC_WAVES) {
    const uint32_t nw = gridDim.x * (TB / 64);
    for (uint32_t i = blockIdx.x * (TB / 64) + wv; i < nwav; i += nw) {
      const uint32_t x = wav[i];
      const uint32_t wb = wrp[x];
      cnt += wave_rank<CAP>(rp, col, x, lane, W, wse + wb, wrp[x + 1] - wb);
    }
  } else {
    for (bool first = true;; first = false) {
      uint32_t i0 = 0;
      if (lane == 0) i0 = claim(cursor + 1, nwav, kWaveChunk, first);
      i0 = __shfl(i0, 0, 64);
      if (i0 >= nwav) break;
      const uint32_t i1 = min(nwav, i0 + kWaveChunk);
      for (uint32_t i = i0; i < i1; ++i) {
        const uint32_t x = wav[i];
        const uint32_t wb = wrp[x];
        cnt += wave_rank<CAP>(rp, col, x, lane, W, wse + wb, wrp[x + 1] - wb);
      }
    }
  }
  for (bool first = true;; first = false) {
    uint32_t i0 = 0;
    if (lane == 0) i0 = claim(cursor + 2, nlan, kLaneClaim, first);
    i0 = __shfl(i0, 0, 64);
    if (i0 >= nlan) break;
    for (uint32_t i = i0 + lane; i < min(nlan, i0 + kLaneClaim); i += 64)
      cnt += lane_task(ell, wrp, wcol, lan[i]);
  }
  flush_block(cnt, total, red);
}

// Per-vertex form of k_auto: tri[u] = the triangles whose first vertex (the
// tail of the two other arcs) is u, for every vertex of the two task lists.
// A block task adds one partial per wave, so tri must start at zero.
template <int CAP>
__global__ __launch_bounds__(TB) void k_auto_per_vertex(const uint32_t* __restrict__ rp,
                                                        const uint32_t* __restrict__ col,
                                                        const uint32_t* __restrict__ blk,
                                                        uint32_t nblk,
                                                        const uint32_t* __restrict__ wav,
                                                        uint32_t nwav, uint32_t* cursor,
                                                        unsigned long long* __restrict__ tri) {
  __shared__ BlockLds<CAP> S;
  const uint32_t tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  clear_bitmap(S);
  for (bool first = true;; first = false) {
    if (tid == 0) S.t = claim(cursor, nblk, 1u, first);
    __syncthreads();
    const uint32_t t = S.t;
    if (t >= nblk) break;
    const uint32_t u = blk[t];
    const unsigned long long c = wave_sum(block_vertex<CAP>(rp, col, u, S));
    if (lane == 0 && c) atomicAdd(tri + u, c);
  }
  const WaveSlot W = wave_slot(S, wv);
  for (bool first = true;; first = false) {
    uint32_t i0 = 0;
    if (lane == 0) i0 = claim(cursor + 1, nwav, kWaveChunk, first);
    i0 = __shfl(i0, 0, 64);
    if (i0 >= nwav) break;
    const uint32_t i1 = min(nwav, i0 + kWaveChunk);
    for (uint32_t i = i0; i < i1; ++i) {
      const uint32_t u = wav[i];
      const unsigned long long c = wave_sum(wave_vertex(rp, col, u, lane, W));
      if (lane == 0) tri[u] = c;
    }
  }
}

}  // namespace dev
}  // namespace tcx
