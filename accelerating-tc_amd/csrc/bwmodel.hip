// List-read bandwidth of this GPU, for A-order's bandwidth model
// (vertex_ordering/A-order/A-order.cpp:20-36 ships numbers measured on the
// GPU the reference targets), and the vector-ALU issue rate that the
// issue-bound kernels are measured against.  One wave reads one list of `len` uint32
// entries starting at a random 4-byte offset of a 1 GiB array (as the
// counting kernels read N+(v) rows of a CSR far larger than the caches); the
// whole GPU is kept busy; the rate is list entries read per second.
#include <algorithm>
#include <cmath>
#include <vector>

#include "internal.h"

namespace tcx {
namespace {

constexpr uint64_t kWords = 1ull << 28;  // 1 GiB of uint32

__global__ void k_fill_words(uint32_t* __restrict__ a, uint64_t n) {
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (uint64_t)gridDim.x * blockDim.x)
    a[i] = (uint32_t)(i * 2654435761u);
}

// splitmix-style hash of the list number -> start offset in [0, span)
__device__ inline uint64_t list_start(uint64_t i, uint64_t span) {
  uint64_t z = i + 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z ^= z >> 31;
  return z % span;
}

// Each wave reads lists w, w + nwaves, ... of `len` entries; the sum of what
// it read goes to out[wave] so that no load is dropped.
__global__ __launch_bounds__(256) void k_list_read(const uint32_t* __restrict__ a, uint64_t span,
                                                   uint32_t len, uint64_t nlists,
                                                   uint32_t* __restrict__ out) {
  const uint32_t lane = threadIdx.x & 63;
  const uint64_t wave = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const uint64_t nwaves = ((uint64_t)gridDim.x * blockDim.x) >> 6;
  uint32_t acc = 0;
  for (uint64_t i = wave; i < nlists; i += nwaves) {
    const uint64_t b = list_start(i, span);
    for (uint32_t j = lane; j < len; j += 64) acc += a[b + j];
  }
  for (int off = 32; off > 0; off >>= 1) acc += __shfl_down(acc, off, 64);
  if (lane == 0) out[wave] = acc;
}

// Issue-rate probe of the vector ALU: every lane runs `iters` rounds of
// eight independent chains of 32-bit integer operations (shift, xor, add,
// and/or: the mix of the counting kernels), so nothing waits on memory or on
// a result; the xor of the chains goes to out so that none is dropped.  The
// grid fills every SIMD with eight waves.  The instructions issued are read
// from the SQ_INSTS_VALU counter of a profiled run.
__global__ __launch_bounds__(256) void k_valu_probe(uint32_t iters, uint32_t seed,
                                                    uint32_t* __restrict__ out) {
  uint32_t x[8];
#pragma unroll
  for (int i = 0; i < 8; ++i) x[i] = seed + threadIdx.x * 8u + (uint32_t)i;
  for (uint32_t r = 0; r < iters; ++r) {
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      x[i] ^= x[i] >> 13;
      x[i] += 0x9E3779B9u;
      x[i] ^= x[i] << 7;
      x[i] = (x[i] & 0x0F0F0F0Fu) | (x[i] >> 4);
    }
  }
  uint32_t acc = 0;
#pragma unroll
  for (int i = 0; i < 8; ++i) acc ^= x[i];
  out[(uint64_t)blockIdx.x * blockDim.x + threadIdx.x] = acc;
}

unsigned grid_for(uint64_t n) {
  return (unsigned)std::max<uint64_t>(1, std::min<uint64_t>((n + 255) / 256, 1u << 20));
}

}  // namespace
}  // namespace tcx

using namespace tcx;

extern "C" {

int tc_bw_model_reference(tc_bw_model* out) {
  if (!out) return fail(TC_ERR_ARG, "NULL output");
  bw_model_reference(out);
  return TC_OK;
}

int tc_measure_list_bandwidth(tc_ctx* c, const uint32_t* lens, uint32_t count, float* gbps) {
  if (!c || (count && (!lens || !gbps))) return fail(TC_ERR_ARG, "NULL argument");
  for (uint32_t i = 0; i < count; ++i)
    if (lens[i] == 0 || lens[i] > (1u << 20))
      return fail(TC_ERR_RANGE, "list length %u outside 1 .. 2^20", lens[i]);
  if (count == 0) return TC_OK;
  TCX_HIP(hipSetDevice(c->device));
  DevBuf arr, out;
  TCX_TRY(arr.reserve(4 * kWords));
  hipLaunchKernelGGL(k_fill_words, dim3(grid_for(kWords)), dim3(256), 0, c->stream,
                     arr.as<uint32_t>(), kWords);
  TCX_HIP(hipGetLastError());
  const uint32_t blocks = (uint32_t)c->num_cus * 8;  // 32 waves per CU
  TCX_TRY(out.reserve(4ull * blocks * 4));
  hipEvent_t e0, e1;
  TCX_HIP(hipEventCreate(&e0));
  TCX_HIP(hipEventCreate(&e1));
  int st = TC_OK;
  for (uint32_t i = 0; i < count && st == TC_OK; ++i) {
    const uint32_t len = lens[i];
    // about 256 MiB of list entries per timed launch, at least 2^22 lists
    const uint64_t nlists = std::max<uint64_t>(1ull << 22, (1ull << 26) / len);
    const uint64_t span = kWords - len;
    float best = 0;
    for (int rep = 0; rep < 4 && st == TC_OK; ++rep) {  // first launch warms up
      if (hipEventRecord(e0, c->stream) != hipSuccess) st = fail(TC_ERR_HIP, "event record");
      hipLaunchKernelGGL(k_list_read, dim3(blocks), dim3(256), 0, c->stream, arr.as<uint32_t>(),
                         span, len, nlists, out.as<uint32_t>());
      if (st == TC_OK && (hipGetLastError() != hipSuccess ||
                          hipEventRecord(e1, c->stream) != hipSuccess ||
                          hipEventSynchronize(e1) != hipSuccess))
        st = fail(TC_ERR_HIP, "list-read launch");
      float ms = 0;
      if (st == TC_OK && hipEventElapsedTime(&ms, e0, e1) != hipSuccess)
        st = fail(TC_ERR_HIP, "event timing");
      if (st == TC_OK && rep > 0 && ms > 0) {
        const float r = (float)(4.0 * len * (double)nlists / (ms * 1e-3) / 1e9);
        best = std::max(best, r);
      }
    }
    gbps[i] = best;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return st;
}

int tc_measure_valu_rate(tc_ctx* c, uint32_t iters, float* ms) {
  if (!c || !ms) return fail(TC_ERR_ARG, "NULL argument");
  if (iters == 0) return fail(TC_ERR_RANGE, "iters must be positive");
  TCX_HIP(hipSetDevice(c->device));
  const uint32_t G = (uint32_t)c->num_cus * 8u;  // 8 waves per SIMD
  DevBuf out;
  TCX_TRY(out.reserve(4ull * G * 256));
  hipLaunchKernelGGL(k_valu_probe, dim3(G), dim3(256), 0, c->stream, iters, 1u,
                     out.as<uint32_t>());  // warm-up
  TCX_HIP(hipGetLastError());
  TCX_HIP(hipEventRecord(c->ev0, c->stream));
  hipLaunchKernelGGL(k_valu_probe, dim3(G), dim3(256), 0, c->stream, iters, 2u,
                     out.as<uint32_t>());
  TCX_HIP(hipGetLastError());
  TCX_HIP(hipEventRecord(c->ev1, c->stream));
  TCX_HIP(hipEventSynchronize(c->ev1));
  TCX_HIP(hipEventElapsedTime(ms, c->ev0, c->ev1));
  return TC_OK;
}

int tc_bw_model_fit(const uint32_t* lens, const float* gbps, uint32_t count, tc_bw_model* out) {
  if (!out || (count && (!lens || !gbps))) return fail(TC_ERR_ARG, "NULL argument");
  tc_bw_model m;
  bw_model_reference(&m);
  bool have[10] = {false};
  double sx = 0, sy = 0, sxx = 0, sxy = 0;
  uint32_t nfit = 0;
  for (uint32_t i = 0; i < count; ++i) {
    if (!(gbps[i] > 0)) return fail(TC_ERR_ARG, "bandwidth %u is not positive", i);
    if (lens[i] < 10) {
      if (lens[i] >= 2) {
        m.table[lens[i]] = gbps[i];
        have[lens[i]] = true;
      }
      continue;
    }
    const double x = std::log((double)lens[i]), y = std::log((double)gbps[i]);
    sx += x;
    sy += y;
    sxx += x * x;
    sxy += x * y;
    ++nfit;
  }
  for (int l = 2; l < 10; ++l)
    if (!have[l]) return fail(TC_ERR_ARG, "no measurement for list length %d", l);
  const double den = nfit * sxx - sx * sx;
  if (nfit < 2 || !(den > 0)) return fail(TC_ERR_ARG, "need two or more lengths >= 10");
  m.b = (nfit * sxy - sx * sy) / den;
  m.a = std::exp((sy - m.b * sx) / nfit);
  *out = m;
  return TC_OK;
}

}  // extern "C"
