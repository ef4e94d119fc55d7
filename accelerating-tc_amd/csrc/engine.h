// Device building blocks of the AUTO engine (vertex-centric, N+(u) in LDS),
// shared by the count kernels and the cost-measurement kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace tcx {
namespace dev {

constexpr int TB = 256;                 // threads per workgroup (4 waves)
constexpr uint32_t kEmpty = 0xFFFFFFFFu;
constexpr int kStage = 1024;            // out-neighbour row bounds staged per pass
constexpr uint32_t kWaveMax = 512;      // longest N+(u) the wave variant holds in LDS
constexpr uint32_t kWaveChunk = 16;     // wave tasks taken per cursor bump

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
// count depends on n only, so the lanes of a wave never diverge.
__device__ inline uint32_t lds_has(const uint32_t* a, uint32_t n, uint32_t x) {
  const uint32_t* p = a;
  uint32_t len = n;
  while (len > 1) {
    const uint32_t half = len >> 1;
    p = p[half] <= x ? p + half : p;
    len -= half;
  }
  return (n > 0 && *p == x) ? 1u : 0u;
}

// Streams N+(v) = col[vb, ve) with the wave's lanes, four 64-entry chunks in
// flight per lane, and counts the entries present in the LDS list a[0, n).
__device__ inline uint32_t count_list(const uint32_t* __restrict__ col, uint32_t vb, uint32_t ve,
                                      uint32_t lane, const uint32_t* a, uint32_t n) {
  uint32_t c = 0;
  for (uint32_t k = vb + lane; k < ve; k += 256) {
    const uint32_t w0 = col[k];
    const uint32_t w1 = k + 64 < ve ? col[k + 64] : kEmpty;
    const uint32_t w2 = k + 128 < ve ? col[k + 128] : kEmpty;
    const uint32_t w3 = k + 192 < ve ? col[k + 192] : kEmpty;
    c += lds_has(a, n, w0);
    c += lds_has(a, n, w1);
    c += lds_has(a, n, w2);
    c += lds_has(a, n, w3);
  }
  return c;
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

// Each lane holds one list col[vb, ve) (vb == ve for none).  Counts the
// entries of all 64 lists that are present in the ascending LDS list a[0, n).
// Lists of kFlatMin entries or more are streamed one at a time by the whole
// wave (count_list).  The shorter ones are concatenated: the wave walks the
// concatenation 64 entries per step (two steps in flight), and each lane
// finds the list its entry belongs to by a search over the exclusive prefix
// of the lengths, kept with the list starts in the wave's LDS slot F[128].
// Short lists thus fill every lane instead of len/64 of them.
constexpr uint32_t kFlatMin = 128;

__device__ inline uint32_t count_lists64(const uint32_t* __restrict__ col, uint32_t vb,
                                         uint32_t ve, uint32_t lane, const uint32_t* a,
                                         uint32_t n, uint32_t* F) {
  const uint32_t len = ve - vb;
  const bool lng = len >= kFlatMin;
  uint32_t c = 0;
  uint64_t mask = __ballot(lng);
  while (mask) {
    const uint32_t j = (uint32_t)__builtin_ctzll(mask);
    mask &= mask - 1;
    c += count_list(col, __shfl(vb, j, 64), __shfl(ve, j, 64), lane, a, n);
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
    c += lds_has(a, n, xa);
    c += lds_has(a, n, xb);
  }
  wave_sync();  // F is rewritten by the next call
  return c;
}

// Wave variant: one wave counts the wedges of u (d+(u) <= kWaveMax) with
// N+(u) in the wave's LDS slot L; lane j takes the j-th out-neighbour of
// each 64-neighbour group, whose lists go through count_lists64 (F: the
// wave's 128-word slot).
__device__ inline uint32_t wave_vertex(const uint32_t* __restrict__ rp,
                                       const uint32_t* __restrict__ col, uint32_t u,
                                       uint32_t lane, uint32_t* L, uint32_t* F) {
  const uint32_t ub = rp[u], du = rp[u + 1] - ub;
  for (uint32_t i = lane; i < du; i += 64) L[i] = col[ub + i];
  wave_sync();
  uint32_t c = 0;
  for (uint32_t j0 = 0; j0 < du; j0 += 64) {
    uint32_t mvb = 0, mve = 0;
    if (j0 + lane < du) {
      const uint32_t v = L[j0 + lane];
      mvb = rp[v];
      mve = rp[v + 1];
    }
    c += count_lists64(col, mvb, mve, lane, L, du, F);
  }
  wave_sync();
  return c;
}

template <int CAP>
struct BlockLds {
  uint32_t u[CAP];        // N+(u); reused as 4 wave slots of CAP/4 in wave phases
  uint32_t vb[kStage];
  uint32_t ve[kStage];
  uint32_t f[TB / 64][128];  // per-wave slots of count_lists64
  uint32_t t;
};

// Block variant: the whole workgroup counts the wedges of u.  N+(u) is copied
// into LDS (when it fits), the row bounds of the out-neighbours are staged in
// LDS, and wave w takes the staged neighbours w, w+4, w+8, ... 64 at a time
// (count_lists64).  Lists longer than CAP are searched in global memory.
// Every thread must call it.
template <int CAP>
__device__ inline uint32_t block_vertex(const uint32_t* __restrict__ rp,
                                        const uint32_t* __restrict__ col, uint32_t u,
                                        BlockLds<CAP>& S) {
  const uint32_t tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const uint32_t ub = rp[u], du = rp[u + 1] - ub;
  const bool in_lds = du <= (uint32_t)CAP;
  uint32_t c = 0;
  if (in_lds)
    for (uint32_t i = tid; i < du; i += TB) S.u[i] = col[ub + i];
  for (uint32_t e0 = 0; e0 < du; e0 += kStage) {
    const uint32_t E = min((uint32_t)kStage, du - e0);
    for (uint32_t i = tid; i < E; i += TB) {
      const uint32_t v = col[ub + e0 + i];
      S.vb[i] = rp[v];
      S.ve[i] = rp[v + 1];
    }
    __syncthreads();
    if (in_lds) {
      // wave wv owns staged neighbours e = 4 * i + wv; 64 of them per group
      const uint32_t mine = E > wv ? (E - wv + 3) / 4 : 0;
      for (uint32_t g0 = 0; g0 < mine; g0 += 64) {
        uint32_t vb = 0, ve = 0;
        if (g0 + lane < mine) {
          const uint32_t e = 4 * (g0 + lane) + wv;
          vb = S.vb[e];
          ve = S.ve[e];
        }
        c += count_lists64(col, vb, ve, lane, S.u, du, S.f[wv]);
      }
    } else {
      for (uint32_t e = wv; e < E; e += TB / 64) {
        const uint32_t vb = S.vb[e], ve = S.ve[e];
        for (uint32_t k = vb + lane; k < ve; k += 64) c += contains(col, ub, ub + du, col[k]);
      }
    }
    __syncthreads();
  }
  return c;
}

// The AUTO count kernel (one launch, persistent grid):
//  phase 1 - block tasks (vertex list `blk`, in the schedule's order), one
//            workgroup each via cursor[0];
//  phase 2 - wave tasks (list `wav`, d+ <= min(kWaveMax, CAP/4)), one wave
//            each, kWaveChunk at a time via cursor[1]; they fill the tail
//            left by phase 1.
template <int CAP>
__global__ __launch_bounds__(TB) void k_auto(const uint32_t* __restrict__ rp,
                                             const uint32_t* __restrict__ col,
                                             const uint32_t* __restrict__ blk, uint32_t nblk,
                                             const uint32_t* __restrict__ wav, uint32_t nwav,
                                             uint32_t* cursor, unsigned long long* total) {
  __shared__ BlockLds<CAP> S;
  const uint32_t tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  unsigned long long cnt = 0;
  for (;;) {
    if (tid == 0) S.t = atomicAdd(cursor, 1u);
    __syncthreads();
    const uint32_t t = S.t;
    if (t >= nblk) break;
    cnt += block_vertex<CAP>(rp, col, blk[t], S);
  }
  uint32_t* L = S.u + wv * (CAP / 4);
  for (;;) {
    uint32_t i0 = 0;
    if (lane == 0) i0 = atomicAdd(cursor + 1, kWaveChunk);
    i0 = __shfl(i0, 0, 64);
    if (i0 >= nwav) break;
    const uint32_t i1 = min(nwav, i0 + kWaveChunk);
    for (uint32_t i = i0; i < i1; ++i) cnt += wave_vertex(rp, col, wav[i], lane, L, S.f[wv]);
  }
  flush_count(cnt, total);
}

}  // namespace dev
}  // namespace tcx
