// Orientation cost functions of the raw records: the quantities the
// reference's analysis tools print (tools/costFuncCal.cpp,
// tools/verification.cpp, tools/approximateRatio.cpp).  The per-vertex
// integer counts (degree, out-degree under each orientation rule, degree
// histogram) are computed on the GPU; the final float / double sums run on
// the host in the tools' own summation order, so the printed values are
// reproduced bit for bit.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "engine.h"
#include "internal.h"

namespace tcx {
namespace {

using dev::gtid;
using dev::gstride;

// Per record (s, d), with raw degrees deg:
//   idB[min(s, d)]++                                   (costFuncCal.cpp:47-52)
//   degB[s]++ if deg[s] < deg[d] || (== && s < d), else degB[d]++   (:53-60)
//   rtd[d]++ if deg[s] > deg[d] || (== && s < d), else rtd[s]++  (verification.cpp:50-56)
__global__ void k_rule_counts(const uint32_t* __restrict__ s, const uint32_t* __restrict__ d,
                              uint64_t m, const uint32_t* __restrict__ deg,
                              uint32_t* __restrict__ idb, uint32_t* __restrict__ degb,
                              uint32_t* __restrict__ rtd) {
  for (uint64_t i = gtid(); i < m; i += gstride()) {
    const uint32_t a = s[i], b = d[i];
    const uint32_t da = deg[a], db = deg[b];
    atomicAdd(idb + ((int)a < (int)b ? a : b), 1u);
    atomicAdd(degb + ((da < db || (da == db && (int)a < (int)b)) ? a : b), 1u);
    atomicAdd(rtd + ((da > db || (da == db && (int)a < (int)b)) ? b : a), 1u);
  }
}

__global__ void k_zero(uint32_t* __restrict__ p, uint64_t n) {
  for (uint64_t i = gtid(); i < n; i += gstride()) p[i] = 0;
}

unsigned grid_for(uint64_t n) {
  return (unsigned)std::max<uint64_t>(1, std::min<uint64_t>((n + 255) / 256, 1u << 20));
}

struct RawCounts {
  uint32_t nodes = 1;  // max id + 1 over the records (1 for none)
  std::vector<uint32_t> deg, lp, idb, degb, rtd;
};

int raw_counts(tc_ctx* c, tc_graph* g, RawCounts& rc) {
  const uint64_t m = g->m_raw;
  const uint32_t* s = g->src.as<uint32_t>();
  const uint32_t* d = g->dst.as<uint32_t>();
  uint32_t mx = 0;
  if (m) {
    uint32_t ms = 0, md = 0;
    TCX_TRY(max_u32(c, s, m, &ms));
    TCX_TRY(max_u32(c, d, m, &md));
    mx = std::max(ms, md);
  }
  const uint32_t nodes = mx + 1;
  rc.nodes = nodes;
  DevBuf buf;
  TCX_TRY(buf.reserve(5ull * 4 * nodes));
  uint32_t* deg = buf.as<uint32_t>();
  uint32_t* lp = deg + nodes;
  uint32_t* idb = lp + nodes;
  uint32_t* degb = idb + nodes;
  uint32_t* rtd = degb + nodes;
  hipLaunchKernelGGL(k_zero, dim3(grid_for(5ull * nodes)), dim3(256), 0, c->stream, deg,
                     5ull * nodes);
  if (m) {
    const unsigned G = (unsigned)std::min<uint64_t>(grid_for(m), (uint64_t)c->num_cus * 4u);
    hipLaunchKernelGGL(dev::k_count_ids<8192>, dim3(G), dim3(256), 0, c->stream, s, d, m, nodes,
                       deg);
    hipLaunchKernelGGL(dev::k_count_ids<8192>, dim3(G), dim3(256), 0, c->stream, s,
                       (const uint32_t*)nullptr, m, nodes, lp);
    hipLaunchKernelGGL(k_rule_counts, dim3(grid_for(m)), dim3(256), 0, c->stream, s, d, m, deg,
                       idb, degb, rtd);
  }
  TCX_HIP(hipGetLastError());
  rc.deg.resize(nodes);
  rc.lp.resize(nodes);
  rc.idb.resize(nodes);
  rc.degb.resize(nodes);
  rc.rtd.resize(nodes);
  TCX_HIP(hipMemcpyAsync(rc.deg.data(), deg, 4ull * nodes, hipMemcpyDeviceToHost, c->stream));
  TCX_HIP(hipMemcpyAsync(rc.lp.data(), lp, 4ull * nodes, hipMemcpyDeviceToHost, c->stream));
  TCX_HIP(hipMemcpyAsync(rc.idb.data(), idb, 4ull * nodes, hipMemcpyDeviceToHost, c->stream));
  TCX_HIP(hipMemcpyAsync(rc.degb.data(), degb, 4ull * nodes, hipMemcpyDeviceToHost, c->stream));
  TCX_HIP(hipMemcpyAsync(rc.rtd.data(), rtd, 4ull * nodes, hipMemcpyDeviceToHost, c->stream));
  TCX_HIP(hipStreamSynchronize(c->stream));
  return TC_OK;
}

int check(tc_ctx* c, tc_graph* g) {
  if (!c || !g) return fail(TC_ERR_ARG, "NULL argument");
  if (g->ctx != c) return fail(TC_ERR_ARG, "graph belongs to another context");
  if (g->m_raw >= 0x80000000ull)
    return fail(TC_ERR_RANGE, "the tools count records in 32-bit ints");
  TCX_HIP(hipSetDevice(c->device));
  return TC_OK;
}

}  // namespace
}  // namespace tcx

using namespace tcx;

extern "C" {

int tc_cost_functions(tc_ctx* c, tc_graph* g, int k, float* out) {
  TCX_TRY(check(c, g));
  if (!out) return fail(TC_ERR_ARG, "NULL output");
  RawCounts rc;
  TCX_TRY(raw_counts(c, g, rc));
  const unsigned edge_num = (unsigned)g->m_raw;
  // tools/costFuncCal.cpp:40-88: float average, threshold k * average,
  // one float accumulator per rule, vertices in id order
  const float avg = (float)edge_num / rc.nodes;
  const float thr = k * avg;
  float id_part = 0, deg_part = 0, lp_part = 0;
  for (uint32_t i = 0; i < rc.nodes; ++i) {
    const int ib = (int)rc.idb[i], db = (int)rc.degb[i], lb = (int)rc.lp[i];
    if (ib - avg > thr) id_part += std::fabs(ib - avg);
    if (db - avg > thr) deg_part += std::fabs(db - avg);
    if (lb - avg > thr) lp_part += std::fabs(lb - avg);
  }
  // tools/verification.cpp:36-60: f_lp over every id, f_degree over ids
  // below the largest (the tool's loop bound)
  const float d_ave = (float)g->m_raw / rc.nodes;
  float f_lp = 0, f_degree = 0;
  for (uint32_t i = 0; i < rc.nodes; ++i) f_lp += std::fabs(d_ave - (int)rc.lp[i]);
  for (uint32_t i = 0; i + 1 < rc.nodes; ++i) f_degree += std::fabs((int)rc.rtd[i] - d_ave);
  out[0] = id_part;
  out[1] = deg_part;
  out[2] = lp_part;
  out[3] = f_lp;
  out[4] = f_degree;
  return TC_OK;
}

int tc_approx_ratio(tc_ctx* c, tc_graph* g, double* ratio) {
  TCX_TRY(check(c, g));
  if (!ratio) return fail(TC_ERR_ARG, "NULL output");
  RawCounts rc;
  TCX_TRY(raw_counts(c, g, rc));
  // tools/approximateRatio.cpp:37-94 on the degree histogram
  const double avg_outd = (double)g->m_raw / rc.nodes;
  uint32_t dmax = 0;
  for (uint32_t v : rc.deg) dmax = std::max(dmax, v);
  std::vector<int64_t> percent((size_t)dmax + 2, 0);
  double vlow = 0, vhigh = 0, sum_low = 0, sum_high = 0;
  for (uint32_t i = 0; i < rc.nodes; ++i) {
    const int dv = (int)rc.deg[i];
    if (dv <= avg_outd) {
      vlow++;
      sum_low += dv;
    } else {
      vhigh++;
      sum_high += dv;
    }
    percent[dv]++;
  }
  (void)vlow;
  double lb_opt = 0;
  for (int dd = 1; dd <= avg_outd; ++dd) lb_opt += percent[dd] * (avg_outd - dd);
  double lb_opt2;
  const double avg_vhigh = sum_high / vhigh / 2;
  if (avg_vhigh >= avg_outd) {
    const double t = (sum_high - sum_low) / 2 - avg_outd * vhigh;
    lb_opt2 = t >= 0 ? t : 0;
  } else {
    lb_opt2 = avg_outd * vhigh - sum_high / 2;
  }
  double ub = 0, remain = 0;
  for (int dd = (int)avg_outd + 1; dd <= (int)dmax; ++dd) remain += percent[dd] * dd / 2;
  int dx = (int)avg_outd + 1;
  while (remain > 0) {
    if (remain > (double)dx * percent[dx]) {
      ub += avg_outd * percent[dx];
      remain -= (double)dx * percent[dx];
      dx++;
    } else {
      ub += avg_outd * remain / dx;
      break;
    }
  }
  *ratio = 1 + ub / (lb_opt + lb_opt2);
  return TC_OK;
}

}  // extern "C"
