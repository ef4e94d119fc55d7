// Counting without device allocation: the exact triangle count of an
// oriented CSR that the caller already holds in device memory, for callers
// that cannot allocate (code under HIP graph capture, tight serving loops).
//
// The AUTO engine's push form without precomputed task lists: workgroups
// claim 1024 vertex ids at a time and count the rows with d+ > 64 of each
// claim as block tasks, then waves claim 256 ids at a time and count the rows
// with 2 <= d+ <= 64 as wave tasks.  The only memory beyond the caller's
// inputs and output is `scratch`, a file-scope device array: two claim
// cursors and one partial count per workgroup.  Each device holds its own
// copy; its address is looked up once per device when a context is created.
#include <mutex>

#include "engine.h"
#include "internal.h"

namespace tcx {
namespace {

using dev::BlockLds;
using dev::TB;
using dev::WaveSlot;

constexpr uint32_t kBlockClaim = 1024;  // vertex ids per workgroup claim (block tasks)
constexpr uint32_t kWaveClaim = 256;    // vertex ids per wave claim (wave tasks)
constexpr uint32_t kWaveMaxDeg = 64;    // rows up to this d+ are wave tasks
constexpr size_t kCursorBytes = 256;    // scratch[0, 256): claim cursors (uint32)
static_assert(kCursorBytes + 8ull * TC_NOALLOC_MAX_GRID <= TC_NOALLOC_SCRATCH_BYTES,
              "scratch holds the cursors and one partial per workgroup");

// cursors at offset 0, partial counts (uint64) from offset kCursorBytes
static __device__ __align__(256) unsigned char scratch[TC_NOALLOC_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void* g_scratch[kMaxDevices] = {};  // device address of `scratch`, per device

__global__ void k_noalloc_clear(uint32_t* cursor) {
  if (threadIdx.x < 4) cursor[threadIdx.x] = 0;
}

template <int CAP>
__global__ __launch_bounds__(TB) void k_auto_scan(const uint32_t* __restrict__ rp,
                                                  const uint32_t* __restrict__ col, uint32_t n,
                                                  uint32_t* cursor,
                                                  unsigned long long* __restrict__ partial) {
  __shared__ BlockLds<CAP> S;
  __shared__ unsigned long long red[TB / 64];
  const uint32_t tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  dev::clear_bitmap(S);
  unsigned long long cnt = 0;
  for (;;) {
    if (tid == 0) S.t = atomicAdd(cursor, kBlockClaim);
    __syncthreads();
    const uint32_t base = S.t;
    __syncthreads();  // every thread has read S.t before thread 0 claims again
    if (base >= n) break;
    const uint32_t end = min(n, base + kBlockClaim);
    for (uint32_t b = base; b < end; b += 64) {
      // every wave reads the same 64 degrees, so the ballot is uniform
      const uint32_t u = b + lane;
      const uint32_t d = u < end ? rp[u + 1] - rp[u] : 0u;
      uint64_t big = __ballot(d > kWaveMaxDeg);
      while (big) {
        const uint32_t j = (uint32_t)__builtin_ctzll(big);
        big &= big - 1;
        cnt += dev::block_vertex<CAP>(rp, col, b + j, S);
      }
    }
  }
  const WaveSlot W = dev::wave_slot(S, wv);
  for (;;) {
    uint32_t base = 0;
    if (lane == 0) base = atomicAdd(cursor + 1, kWaveClaim);
    base = __shfl(base, 0, 64);
    if (base >= n) break;
    const uint32_t end = min(n, base + kWaveClaim);
    for (uint32_t b = base; b < end; b += 64) {
      const uint32_t u = b + lane;
      const uint32_t d = u < end ? rp[u + 1] - rp[u] : 0u;
      uint64_t small = __ballot(d >= 2 && d <= kWaveMaxDeg);
      while (small) {
        const uint32_t j = (uint32_t)__builtin_ctzll(small);
        small &= small - 1;
        cnt += dev::wave_vertex(rp, col, b + j, lane, W);
      }
    }
  }
  cnt = dev::wave_sum(cnt);
  if (lane == 0) red[wv] = cnt;
  __syncthreads();
  if (tid == 0) partial[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

// out[0] = sum of the G partial counts (one workgroup of TB threads)
__global__ __launch_bounds__(TB) void k_noalloc_sum(const unsigned long long* __restrict__ partial,
                                                    uint32_t G, unsigned long long* out) {
  __shared__ unsigned long long red[TB / 64];
  unsigned long long s = 0;
  for (uint32_t i = threadIdx.x; i < G; i += TB) s += partial[i];
  s = dev::wave_sum(s);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) out[0] = red[0] + red[1] + red[2] + red[3];
}

}  // namespace

// Looks up this device's copy of `scratch` (device `device` current).
int noalloc_setup(int device) {
  if (device < 0 || device >= kMaxDevices)
    return fail(TC_ERR_RANGE, "device %d beyond the %d the scratch table holds", device,
                kMaxDevices);
  std::lock_guard<std::mutex> lock(g_mu);
  if (g_scratch[device]) return TC_OK;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p)
    return fail(TC_ERR_HIP, "hipGetSymbolAddress(scratch) on device %d failed: %s", device,
                hipGetErrorString(e));
  g_scratch[device] = p;
  return TC_OK;
}

}  // namespace tcx

using namespace tcx;

extern "C" int tc_count_csr_noalloc(tc_ctx* c, const uint32_t* rowptr, const uint32_t* col,
                                    uint32_t n, uint32_t grid, unsigned long long* count) {
  if (!c || !count) return fail(TC_ERR_ARG, "NULL context or output");
  if (!rowptr || !col) return fail(TC_ERR_ARG, "NULL CSR array");
  if (n > 0x7FFFFFFEu) return fail(TC_ERR_RANGE, "n = %u exceeds 2^31 - 2", n);
  const uint32_t G = grid ? grid : std::min<uint32_t>((uint32_t)c->num_cus * 8u,
                                                      TC_NOALLOC_MAX_GRID);
  if (G > TC_NOALLOC_MAX_GRID)
    return fail(TC_ERR_RANGE, "grid %u exceeds the %u partial counts the static scratch holds",
                G, (unsigned)TC_NOALLOC_MAX_GRID);
  TCX_HIP(hipSetDevice(c->device));
  void* base = nullptr;
  {
    std::lock_guard<std::mutex> lock(g_mu);
    if (c->device >= 0 && c->device < kMaxDevices) base = g_scratch[c->device];
  }
  if (!base) return fail(TC_ERR_STATE, "static scratch of device %d not set up", c->device);
  uint32_t* cursor = static_cast<uint32_t*>(base);
  unsigned long long* partial =
      reinterpret_cast<unsigned long long*>(static_cast<unsigned char*>(base) + kCursorBytes);
  hipLaunchKernelGGL(k_noalloc_clear, dim3(1), dim3(64), 0, c->stream, cursor);
  hipLaunchKernelGGL(k_auto_scan<2048>, dim3(G), dim3(TB), 0, c->stream, rowptr, col, n, cursor,
                     partial);
  hipLaunchKernelGGL(k_noalloc_sum, dim3(1), dim3(TB), 0, c->stream, partial, G, count);
  TCX_HIP(hipGetLastError());
  return TC_OK;
}
