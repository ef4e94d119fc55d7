// A-order bucket assignment (host part of tc_graph_relabel_aorder).
//
// Follows vertex_ordering/A-order/A-order.cpp:100-210.  The GPU computes the
// oriented out-degree (runTimeDegree, :86-99); this file runs the greedy
// bucket fill, which is a strictly sequential heap walk, and returns the
// permutation old id -> new id.  The heap steps are kept exactly as in the
// reference because they define the permutation; the bandwidth model is a
// parameter whose default is the reference's (bit for bit), and which can be
// replaced by one measured on this GPU (tc_measure_list_bandwidth).
#include <algorithm>
#include <cmath>

#include "internal.h"

namespace {

constexpr int kBucket = 40;  // A-order.cpp:14

// The model's arithmetic follows A-order.cpp:23-36: the power law in double,
// rounded to float; the equal-bandwidth term in float.
float bw_of(const tc_bw_model& m, int len) {
  if (len < 10) return m.table[len];
  return (float)(m.a * std::pow((double)len, m.b));
}
float equal_bw_of(const tc_bw_model& m, int len) {
  if (len < 10) return m.lambda[len] * std::sqrt(bw_of(m, len)) / std::sqrt((float)len);
  return m.lambda_tail * std::sqrt(bw_of(m, len)) / std::sqrt((float)len);
}

struct Slot {
  int id;
  float score;  // "memory superiority"
  int list;     // index into the bucket member lists
};
struct ByScoreMin {  // min-heap order, A-order.cpp:50-55
  bool operator()(const Slot& a, const Slot& b) const {
    return a.score > b.score || (a.score == b.score && a.id > b.id);
  }
};
struct ByScoreMax {  // max-heap order, A-order.cpp:44-49
  bool operator()(const Slot& a, const Slot& b) const {
    return a.score < b.score || (a.score == b.score && a.id < b.id);
  }
};
struct ByIdDesc {  // A-order.cpp:56-60
  bool operator()(const Slot& a, const Slot& b) const { return a.id > b.id; }
};

}  // namespace

namespace tcx {

// The reference's bandwidth lookup model, A-order.cpp:20-36.
void bw_model_reference(tc_bw_model* m) {
  const float bw[10] = {0, 0, 271, 310, 399, 419, 458, 478, 524, 539};
  const float lambda[10] = {0, 0, 145.51f, 121.98f, 53.26f, 23.92f, 16, 14, 12.94f, 10.24f};
  for (int i = 0; i < 10; ++i) {
    m->table[i] = bw[i];
    m->lambda[i] = lambda[i];
  }
  m->a = 384.1;
  m->b = 0.1491;
  m->lambda_tail = 12;
}

// rtd: oriented out-degree per vertex (n entries).  perm: old -> new.
int aorder_perm_from_outdeg(const uint32_t* rtd, uint32_t n, uint32_t* perm,
                            const tc_bw_model& model) {
  uint32_t maxd = 0;
  for (uint32_t v = 0; v < n; ++v) maxd = std::max(maxd, rtd[v]);
  // vertices grouped by out-degree, ascending id inside a group
  std::vector<uint64_t> start(maxd + 2, 0);
  for (uint32_t v = 0; v < n; ++v) start[rtd[v] + 1]++;
  for (uint32_t d = 0; d <= maxd; ++d) start[d + 1] += start[d];
  std::vector<uint32_t> by_deg(n);
  {
    std::vector<uint64_t> pos(start.begin(), start.end() - 1);
    for (uint32_t v = 0; v < n; ++v) by_deg[pos[rtd[v]]++] = v;
  }
  // candidate order: out-degree 2..max (A-order.cpp:113-119); `split` marks
  // where out-degree 5 begins (memoryDominatePos).
  uint64_t first2 = maxd >= 2 ? start[2] : n;
  uint64_t cand = n - first2;
  uint64_t split = maxd >= 5 ? start[5] - first2 : 0;
  uint64_t buckets = cand / kBucket;
  if (buckets < 1 || split < 1 || split > (uint64_t)kBucket * buckets)
    return fail(TC_ERR_UNSUPPORTED,
                "A-order: input outside the reference's defined range (%llu candidates, "
                "%llu of out-degree 2..4)", (unsigned long long)cand,
                (unsigned long long)split);
  auto cand_vertex = [&](uint64_t i) { return by_deg[first2 + i]; };

  std::vector<float> bw(maxd + 1), ebw(maxd + 1);
  for (uint32_t d = 0; d <= maxd; ++d) {
    bw[d] = bw_of(model, (int)d);
    ebw[d] = equal_bw_of(model, (int)d);
  }
  std::vector<Slot> heap(buckets);
  std::vector<std::vector<uint32_t>> members(buckets);
  for (uint64_t i = 0; i < buckets; ++i) heap[i] = Slot{(int)i, 0.0f, (int)i};
  std::make_heap(heap.begin(), heap.end(), ByScoreMin());
  uint64_t done = 0;
  // memory-dominated vertices, highest out-degree first (A-order.cpp:144-157)
  for (int64_t i = (int64_t)(kBucket * buckets) - 1; i >= (int64_t)split; --i) {
    std::pop_heap(heap.begin(), heap.end() - done, ByScoreMin());
    Slot& s = *(heap.rbegin() + done);
    uint32_t v = cand_vertex((uint64_t)i);
    uint32_t d = rtd[v];
    members[s.list].push_back(v);
    s.score += (bw[d] - ebw[d]) / kBucket;
    if (members[s.list].size() >= (size_t)kBucket) ++done;
    else std::push_heap(heap.begin(), heap.end() - done, ByScoreMin());
  }
  std::make_heap(heap.begin(), heap.end() - done, ByScoreMax());
  // out-degree 2..4 vertices (A-order.cpp:165-177)
  for (uint64_t i = 0; i < split; ++i) {
    std::pop_heap(heap.begin(), heap.end() - done, ByScoreMax());
    Slot& s = *(heap.rbegin() + done);
    uint32_t v = cand_vertex(i);
    uint32_t d = rtd[v];
    members[s.list].push_back(v);
    s.score -= (ebw[d] - bw[d]) / kBucket;
    if (members[s.list].size() >= (size_t)kBucket) ++done;
    else std::push_heap(heap.begin(), heap.end() - done, ByScoreMax());
  }
  // A-order.cpp:181 applies sort_heap to a range that is not a heap for this
  // comparator; the bucket order is whatever that call leaves.
  std::sort_heap(heap.begin(), heap.end(), ByIdDesc());
  uint32_t next = 0;
  for (uint64_t b = 0; b < buckets; ++b)
    for (uint32_t v : members[heap[b].list]) perm[v] = next++;
  for (uint64_t i = (uint64_t)kBucket * buckets; i < cand; ++i) perm[cand_vertex(i)] = next++;
  if (maxd >= 1)
    for (uint64_t k = start[1]; k < start[2 <= maxd ? 2 : maxd + 1]; ++k) perm[by_deg[k]] = next++;
  for (uint64_t k = start[0]; k < start[1]; ++k) perm[by_deg[k]] = next++;
  return TC_OK;
}

}  // namespace tcx
