// Function-level TriCore API (tricore/include/tricount.h:7-17) on the engine.
//
// Same signatures and meaning as the reference; the edge array is the
// uint64 view of the record with FIRST = high word (tricore/include/util.h:14-16).
// Calls share one lazily created context on device 0 (the reference also
// uses device 0, tricore/src/tricount_gpu.cu:31).  Failures leave a message
// in tc_last_error() and return 0 / leave outputs untouched.
#include <algorithm>
#include <cstring>
#include <mutex>

#include "internal.h"

namespace tcx {
namespace {

std::mutex g_mu;
tc_ctx* g_ctx = nullptr;

tc_ctx* shared_ctx() {
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_ctx && tc_ctx_create(0, &g_ctx) != TC_OK) g_ctx = nullptr;
  return g_ctx;
}

__device__ inline uint64_t gtid() { return (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; }
__device__ inline uint64_t gstride() { return (uint64_t)gridDim.x * blockDim.x; }

unsigned grid_for(uint64_t n) {
  uint64_t b = (n + 255) / 256;
  return (unsigned)std::max<uint64_t>(1, std::min<uint64_t>(b, 1u << 20));
}

// tricount_gpu.cu:37-47 (deg is zeroed first; the reference leaves it as allocated)
__global__ void k_tri_degree(const uint64_t* __restrict__ e, uint64_t m, uint32_t* deg) {
  for (uint64_t i = gtid(); i < m; i += gstride()) {
    uint64_t x = e[i];
    atomicAdd(deg + (uint32_t)(x >> 32), 1u);
    atomicAdd(deg + (uint32_t)x, 1u);
  }
}

// tricount_gpu.cu:69-80
__global__ void k_tri_redirect(uint64_t* e, uint64_t m, const uint32_t* __restrict__ deg) {
  for (uint64_t i = gtid(); i < m; i += gstride()) {
    uint64_t x = e[i];
    uint32_t f = (uint32_t)(x >> 32), s = (uint32_t)x;
    if (deg[f] > deg[s] || (deg[f] == deg[s] && f > s)) e[i] = ((uint64_t)s << 32) | f;
  }
}

// keys of one partition: FIRST in [lo, hi), self-loops -> sentinel
__global__ void k_tri_part_keys(const uint64_t* __restrict__ e, uint64_t m, uint32_t lo,
                                uint32_t hi, uint64_t* __restrict__ out, unsigned long long* cnt) {
  for (uint64_t i = gtid(); i < m; i += gstride()) {
    uint64_t x = e[i];
    uint32_t f = (uint32_t)(x >> 32), s = (uint32_t)x;
    if (f >= lo && f < hi && f != s) out[atomicAdd(cnt, 1ull)] = x;
  }
}

// node_index[j] for j in [0, count] from the sorted keys (tricount_gpu.cu:127-137):
// the first position whose FIRST is >= start + j, one search per entry.
__global__ void k_tri_node_index(const uint64_t* __restrict__ keys, uint64_t m, uint32_t start,
                                 uint32_t count, uint32_t* __restrict__ index) {
  for (uint64_t j = gtid(); j <= count; j += gstride()) {
    const uint64_t x = ((uint64_t)start + j) << 32;
    uint64_t lo = 0, hi = m;
    while (lo < hi) {
      const uint64_t mid = (lo + hi) >> 1;
      if (keys[mid] < x) lo = mid + 1;
      else hi = mid;
    }
    index[j] = (uint32_t)lo;
  }
}

__global__ void k_tri_adj(const uint64_t* __restrict__ keys, uint64_t m, uint32_t* __restrict__ adj) {
  for (uint64_t i = gtid(); i < m; i += gstride()) adj[i] = (uint32_t)keys[i];
}

}  // namespace
}  // namespace tcx

using namespace tcx;

extern "C" {

uint64_t tc_tricore_init_gpu(void) {
  tc_ctx* c = shared_ctx();
  return c ? c->total_mem : 0;
}

void tc_tricore_cal_degree(const uint64_t* edges, uint64_t m, uint32_t* deg, uint32_t n) {
  tc_ctx* c = shared_ctx();
  if (!c || (m && !edges) || !deg) { set_error("cal_degree: bad arguments"); return; }
  DevBuf de, dd;
  if (de.reserve(8 * std::max<uint64_t>(m, 1)) || dd.reserve(4 * std::max<uint64_t>(n, 1))) return;
  if (hipMemsetAsync(dd.p, 0, 4 * (size_t)n, c->stream) != hipSuccess ||
      (m && hipMemcpyAsync(de.p, edges, 8 * m, hipMemcpyHostToDevice, c->stream) != hipSuccess)) {
    set_error("cal_degree: HIP copy failed");
    return;
  }
  hipLaunchKernelGGL(k_tri_degree, dim3(grid_for(m)), dim3(256), 0, c->stream, de.as<uint64_t>(),
                     m, dd.as<uint32_t>());
  if (hipMemcpyAsync(deg, dd.p, 4 * (size_t)n, hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
      hipStreamSynchronize(c->stream) != hipSuccess)
    set_error("cal_degree: HIP failure");
}

void tc_tricore_redirect_edges(uint64_t* edges, uint64_t m, const uint32_t* deg, uint32_t n) {
  tc_ctx* c = shared_ctx();
  if (!c || (m && !edges) || !deg) { set_error("redirect_edges: bad arguments"); return; }
  if (m == 0) return;
  DevBuf de, dd;
  if (de.reserve(8 * m) || dd.reserve(4 * std::max<uint64_t>(n, 1))) return;
  if (hipMemcpyAsync(de.p, edges, 8 * m, hipMemcpyHostToDevice, c->stream) != hipSuccess ||
      hipMemcpyAsync(dd.p, deg, 4 * (size_t)n, hipMemcpyHostToDevice, c->stream) != hipSuccess) {
    set_error("redirect_edges: HIP copy failed");
    return;
  }
  hipLaunchKernelGGL(k_tri_redirect, dim3(grid_for(m)), dim3(256), 0, c->stream,
                     de.as<uint64_t>(), m, dd.as<uint32_t>());
  if (hipMemcpyAsync(edges, de.p, 8 * m, hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
      hipStreamSynchronize(c->stream) != hipSuccess)
    set_error("redirect_edges: HIP failure");
}

// tricount_gpu.cu:103-115 with this engine's footprint: sorted keys plus a
// scratch copy and the part's offsets (no per-thread result array).
uint32_t tc_tricore_cal_part_num(const uint64_t* edges, uint64_t m, uint32_t n) {
  (void)edges;
  tc_ctx* c = shared_ctx();
  if (!c) return 0;
  uint64_t mem = c->total_mem ? c->total_mem : (1ull << 34);
  uint64_t part = 1;
  while (true) {
    uint64_t pe = m / part + 1, pn = n / part + 1;
    uint64_t use = pe * sizeof(uint64_t) * 2 * 115 / 100 + (pn + 1) * sizeof(uint32_t) * 2;
    if (use < mem) break;
    ++part;
  }
  return (uint32_t)part;
}

// tricount_gpu.cu:146-203: per node-range partition, sort + drop self-loops +
// dedupe, write the partition's adjacency (SECOND values) packed into the
// front of `edges` and its offsets into a malloc'd node_index[i].
void tc_tricore_split(uint64_t* edges, uint64_t m, uint64_t part_num, uint32_t** node_index,
                      const uint32_t* node_split, uint64_t* adj_count) {
  tc_ctx* c = shared_ctx();
  if (!c || !node_index || !node_split || !adj_count || part_num == 0 || (m && !edges)) {
    set_error("split: bad arguments");
    return;
  }
  DevBuf de, keys, tmp, idx, adj;
  if (de.reserve(8 * std::max<uint64_t>(m, 1)) || keys.reserve(8 * std::max<uint64_t>(m, 1)) ||
      tmp.reserve(8 * std::max<uint64_t>(m, 1)) || adj.reserve(4 * std::max<uint64_t>(m, 1)))
    return;
  if (c->accum.reserve(64)) return;
  if (m && hipMemcpyAsync(de.p, edges, 8 * m, hipMemcpyHostToDevice, c->stream) != hipSuccess) {
    set_error("split: upload failed");
    return;
  }
  std::vector<uint64_t> counts(part_num);
  uint32_t* adj_out = reinterpret_cast<uint32_t*>(edges);
  uint64_t written = 0;
  for (uint64_t p = 0; p < part_num; ++p) {
    uint32_t lo = node_split[p], hi = node_split[p + 1];
    unsigned long long* cnt = c->accum.as<unsigned long long>();
    (void)hipMemsetAsync(cnt, 0, 8, c->stream);
    hipLaunchKernelGGL(k_tri_part_keys, dim3(grid_for(m)), dim3(256), 0, c->stream,
                       de.as<uint64_t>(), m, lo, hi, keys.as<uint64_t>(), cnt);
    unsigned long long np = 0;
    (void)hipMemcpyAsync(&np, cnt, 8, hipMemcpyDeviceToHost, c->stream);
    if (hipStreamSynchronize(c->stream) != hipSuccess) { set_error("split: HIP failure"); return; }
    uint64_t* sorted = keys.as<uint64_t>();
    if (sort_u64(c, keys.as<uint64_t>(), tmp.as<uint64_t>(), np, 64, &sorted) != TC_OK) return;
    uint64_t* other = sorted == keys.as<uint64_t>() ? tmp.as<uint64_t>() : keys.as<uint64_t>();
    uint64_t nu = 0;
    if (unique_u64(c, sorted, other, np, &nu) != TC_OK) return;
    uint32_t count = hi - lo + 1;  // tricount_gpu.cu:182
    if (idx.reserve(4 * ((size_t)count + 1))) return;
    hipLaunchKernelGGL(k_tri_node_index, dim3(grid_for((uint64_t)count + 1)), dim3(256), 0,
                       c->stream, other, nu, lo, count, idx.as<uint32_t>());
    if (nu)
      hipLaunchKernelGGL(k_tri_adj, dim3(grid_for(nu)), dim3(256), 0, c->stream, other, nu,
                         adj.as<uint32_t>());
    node_index[p] = (uint32_t*)malloc(4 * ((size_t)count + 1));
    if (!node_index[p]) { set_error("split: host allocation failed"); return; }
    (void)hipMemcpyAsync(node_index[p], idx.p, 4 * ((size_t)count + 1), hipMemcpyDeviceToHost, c->stream);
    if (nu) (void)hipMemcpyAsync(adj_out + written, adj.p, 4 * nu, hipMemcpyDeviceToHost, c->stream);
    if (hipStreamSynchronize(c->stream) != hipSuccess) { set_error("split: HIP failure"); return; }
    counts[p] = nu;
    written += nu;
  }
  adj_count[0] = 0;
  for (uint64_t p = 1; p <= part_num; ++p) adj_count[p] = adj_count[p - 1] + counts[p - 1];
}

// tricount_gpu.cu:337-406: assembles the partitions into one CSR in HBM and
// runs the TriCore kernel over every (m, n) partition pair.
uint64_t tc_tricore_tricount_gpu(uint64_t part_num, uint32_t* adj, const uint64_t* adj_count,
                                 const uint32_t* node_split, uint32_t** node_index) {
  tc_ctx* c = shared_ctx();
  if (!c || !adj_count || !node_split || !node_index || part_num == 0) {
    set_error("tricount_gpu: bad arguments");
    return 0;
  }
  const uint32_t n = node_split[part_num];
  const uint64_t m = adj_count[part_num];
  std::vector<uint32_t> rowptr((size_t)n + 1, 0);
  for (uint64_t p = 0; p < part_num; ++p) {
    uint32_t lo = node_split[p], hi = node_split[p + 1];
    for (uint32_t v = lo; v <= hi; ++v)
      rowptr[v] = (uint32_t)(adj_count[p] + node_index[p][v - lo]);
  }
  // An all-device graph built from the split CSR (rows already sorted/unique).
  tc_graph* g = new tc_graph();
  g->ctx = c;
  g->n = g->n_raw = n;
  g->m_raw = 0;
  g->csr_only = true;
  int st = g->rowptr.reserve(4 * ((size_t)n + 1));
  if (!st) st = g->col.reserve(4 * std::max<uint64_t>(m, 1));
  if (!st) st = g->esrc.reserve(4 * std::max<uint64_t>(m, 1));
  std::vector<uint32_t> esrc(m);
  for (uint32_t v = 0; v < n; ++v)
    for (uint32_t e = rowptr[v]; e < rowptr[v + 1]; ++e) esrc[e] = v;
  if (!st && hipMemcpyAsync(g->rowptr.p, rowptr.data(), 4 * ((size_t)n + 1), hipMemcpyHostToDevice,
                            c->stream) != hipSuccess)
    st = TC_ERR_HIP;
  if (!st && m &&
      (hipMemcpyAsync(g->col.p, adj, 4 * m, hipMemcpyHostToDevice, c->stream) != hipSuccess ||
       hipMemcpyAsync(g->esrc.p, esrc.data(), 4 * m, hipMemcpyHostToDevice, c->stream) != hipSuccess))
    st = TC_ERR_HIP;
  if (!st) {
    (void)hipStreamSynchronize(c->stream);
    g->m_or = m;
    g->orient_mode = TC_ORIENT_DEGREE_TRICORE;
    g->sched_version = 1;
    st = compute_csr_stats(c, g);
  }
  uint64_t total = 0;
  if (!st) {
    tc_sched s;
    memset(&s, 0, sizeof(s));
    s.kernel = TC_KERNEL_TRICORE;
    s.tricore_parts = (uint32_t)part_num;
    st = tc_count(c, g, &s, &total, nullptr);
  }
  delete g;
  if (st) {
    set_error("tricount_gpu failed: %s", tc_last_error());
    return 0;
  }
  return total;
}

}  // extern "C"
