// CPU oracle for the triangle-counting / k-clique path.
//
// TEST INFRASTRUCTURE ONLY: a serial restatement of the reference's CPU-side
// algorithms, used by tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg as the checker.  The product library never links or loads
// it.  Each function cites the reference file:line it restates; oracle.h
// lists how each one is pinned.  Functions pinned only by invariants are
// marked "parity unpinned beyond invariants" below.
#include "oracle.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <utility>
#include <vector>

namespace {

const uint32_t kWiped = 0xFFFFFFFFu;

// Sort + unique a list of packed (hi,lo) pairs and emit CSR over n rows.
uint64_t pairs_to_csr(std::vector<uint64_t>& keys, uint32_t n, uint32_t* rowptr,
                      uint32_t* col) {
  std::sort(keys.begin(), keys.end());
  keys.erase(std::unique(keys.begin(), keys.end()), keys.end());
  for (uint32_t i = 0; i <= n; ++i) rowptr[i] = 0;
  for (uint64_t k : keys) rowptr[(uint32_t)(k >> 32) + 1]++;
  for (uint32_t i = 0; i < n; ++i) rowptr[i + 1] += rowptr[i];
  for (uint64_t i = 0; i < keys.size(); ++i) col[i] = (uint32_t)keys[i];
  return keys.size();
}

inline uint64_t merge_count(const uint32_t* a, uint32_t na, const uint32_t* b,
                            uint32_t nb) {
  uint64_t c = 0;
  uint32_t i = 0, j = 0;
  while (i < na && j < nb) {
    if (a[i] == b[j]) {
      ++c; ++i; ++j;
    } else if (a[i] < b[j]) {
      ++i;
    } else {
      ++j;
    }
  }
  return c;
}

uint64_t vertex_tri(const uint32_t* rowptr, const uint32_t* col, uint32_t u) {
  uint64_t c = 0;
  const uint32_t* nu = col + rowptr[u];
  uint32_t du = rowptr[u + 1] - rowptr[u];
  for (uint32_t e = 0; e < du; ++e) {
    uint32_t v = nu[e];
    c += merge_count(nu, du, col + rowptr[v], rowptr[v + 1] - rowptr[v]);
  }
  return c;
}

// number of r-cliques inside candidate set S (sorted), DAG edges only.
uint64_t clique_rec(const uint32_t* rowptr, const uint32_t* col,
                    const std::vector<uint32_t>& S, uint32_t r) {
  if (r == 1) return S.size();
  uint64_t c = 0;
  std::vector<uint32_t> next;
  for (uint32_t v : S) {
    const uint32_t* nv = col + rowptr[v];
    uint32_t dv = rowptr[v + 1] - rowptr[v];
    next.clear();
    std::set_intersection(S.begin(), S.end(), nv, nv + dv, std::back_inserter(next));
    if (next.size() + 1 < r) continue;
    if (r == 2) c += next.size();
    else c += clique_rec(rowptr, col, next, r - 1);
  }
  return c;
}

// Hu/tc.cu:36-67 (long array, with the e==0 guard of the reference).
uint32_t hu_binary_search_value(const int64_t* array, int64_t value, uint32_t len,
                                int direction) {
  int64_t s = 0, e = (int64_t)len - 1;
  int64_t rightPos = 0;
  bool find = false;
  int64_t mid = (s + e) / 2;
  while (s <= e) {
    if (array[mid] == value) {
      rightPos = mid;
      find = true;
      break;
    } else if (array[mid] < value) {
      s = mid + 1;
    } else {
      if (e == 0) break;
      e = mid - 1;
    }
    mid = (s + e) / 2;
  }
  if (!find) return (uint32_t)s;
  int64_t tmp = rightPos + direction;
  while (tmp >= 0 && tmp < (int64_t)len && array[tmp] == value) {
    rightPos += direction;
    tmp = rightPos + direction;
  }
  return (uint32_t)rightPos;
}

// Fox/tc.cu:34-63 (int array).  The reference's unsigned `e = mid - 1` wraps
// when mid == 0 and reads out of bounds; the restatement stops there and
// returns the insertion point 0 instead (the only in-bounds reading).
uint32_t fox_binary_search_value(const int32_t* array, int32_t value, uint32_t len,
                                 int direction) {
  if (len == 0) return 0;
  uint32_t s = 0, e = len - 1;
  uint32_t rightPos = 0;
  bool find = false;
  uint32_t mid = (s + e) / 2;
  while (s <= e) {
    if (array[mid] == value) {
      rightPos = mid;
      find = true;
      break;
    } else if (array[mid] < value) {
      s = mid + 1;
    } else {
      if (mid == 0) break;
      e = mid - 1;
    }
    mid = (s + e) / 2;
  }
  if (!find) return s;
  int64_t tmp = (int64_t)rightPos + direction;
  while (tmp >= 0 && tmp < (int64_t)len && array[tmp] == value) {
    rightPos += direction;
    tmp = (int64_t)rightPos + direction;
  }
  return rightPos;
}

// vertex_ordering/A-order/A-order.cpp:20-36.
const float kBandwidthTable[10] = {0, 0, 271, 310, 399, 419, 458, 478, 524, 539};
const float kLambdaTable[10] = {0, 0, 145.51, 121.98, 53.26, 23.92, 16, 14, 12.94, 10.24};
float aorder_bandwidth(int length) {
  if (length < 10) return kBandwidthTable[length];
  return 384.1 * std::pow(length, 0.1491);
}
float aorder_equal_bandwidth(int length) {
  if (length < 10)
    return kLambdaTable[length] * std::sqrt(aorder_bandwidth(length)) /
           std::sqrt((float)length);
  return 12 * std::sqrt(aorder_bandwidth(length)) / std::sqrt((float)length);
}

struct Bucket {  // A-order.cpp:38-42
  int id;
  float ms;
  int slot;  // index of this bucket's node list
};
struct LessCmp {  // A-order.cpp:44-49
  bool operator()(const Bucket& a, const Bucket& b) const {
    return a.ms < b.ms || (a.ms == b.ms && a.id < b.id);
  }
};
struct GreatCmp {  // A-order.cpp:50-55
  bool operator()(const Bucket& a, const Bucket& b) const {
    return a.ms > b.ms || (a.ms == b.ms && a.id > b.id);
  }
};
struct IdSort {  // A-order.cpp:56-60
  bool operator()(const Bucket& a, const Bucket& b) const { return a.id > b.id; }
};

}  // namespace

extern "C" {

uint32_t orc_num_nodes(const uint32_t* src, const uint32_t* dst, uint64_t m) {
  if (m == 0) return 0;
  uint32_t mx = 0;
  for (uint64_t i = 0; i < m; ++i) {
    if (src[i] > mx) mx = src[i];
    if (dst[i] > mx) mx = dst[i];
  }
  return mx + 1;
}

void orc_degree_raw(const uint32_t* src, const uint32_t* dst, uint64_t m,
                    uint32_t n, uint32_t* deg) {
  std::memset(deg, 0, sizeof(uint32_t) * n);
  for (uint64_t i = 0; i < m; ++i) {
    deg[src[i]]++;
    deg[dst[i]]++;
  }
}

void orc_orient_simple(const uint32_t* src, const uint32_t* dst, uint64_t m,
                       const uint32_t* deg, int mode, uint32_t* osrc,
                       uint32_t* odst) {
  for (uint64_t i = 0; i < m; ++i) {
    uint32_t s = src[i], d = dst[i];
    bool swap = false;
    if (mode == 0) {
      swap = s < d;  // Fox/preTC.cu:94-98
    } else if (mode == 1) {
      swap = deg[s] > deg[d] || (deg[s] == deg[d] && s < d);  // Hu/preTC.cu:129
    } else {
      swap = deg[s] > deg[d] || (deg[s] == deg[d] && s > d);  // tricount_gpu.cu:76
    }
    osrc[i] = swap ? d : s;
    odst[i] = swap ? s : d;
  }
}

uint32_t orc_prune_deg1(uint32_t* src, uint32_t* dst, uint64_t m, uint32_t n,
                        const uint32_t* deg) {
  std::vector<uint32_t> mini(n);
  uint32_t total = 0;
  for (uint32_t i = 0; i < n; ++i) {  // Hu/preTC.cu:139-144
    if (deg[i] <= 1) total++;
    mini[i] = total;
  }
  for (uint64_t i = 0; i < m; ++i) {  // Hu/preTC.cu:148-166
    uint32_t s = src[i], d = dst[i];
    if (deg[s] <= 1) { src[i] = kWiped; continue; }
    if (deg[d] <= 1) { dst[i] = kWiped; continue; }
    if (s > 0) src[i] = s - mini[s - 1];
    if (d > 0) dst[i] = d - mini[d - 1];
  }
  return n - total;
}

uint64_t orc_build_csr(const uint32_t* src, const uint32_t* dst, uint64_t m,
                       uint32_t n, uint32_t* rowptr, uint32_t* col) {
  std::vector<uint64_t> keys;
  keys.reserve(m);
  for (uint64_t i = 0; i < m; ++i) {
    if (src[i] == kWiped || dst[i] == kWiped) continue;
    if (src[i] == dst[i]) continue;
    keys.push_back(((uint64_t)src[i] << 32) | dst[i]);
  }
  return pairs_to_csr(keys, n, rowptr, col);
}

uint64_t orc_build_sym(const uint32_t* src, const uint32_t* dst, uint64_t m,
                       uint32_t n, uint32_t* rowptr, uint32_t* col) {
  std::vector<uint64_t> keys;
  keys.reserve(2 * m);
  for (uint64_t i = 0; i < m; ++i) {
    if (src[i] == dst[i]) continue;  // A-direction-seri.cpp:217-218
    keys.push_back(((uint64_t)src[i] << 32) | dst[i]);
    keys.push_back(((uint64_t)dst[i] << 32) | src[i]);
  }
  return pairs_to_csr(keys, n, rowptr, col);
}

// Literal restatement of edge_directing/A-direction-seri.cpp:35-153 with the
// originEdges of :242-252 (pairs (i, nbr) for i ascending, nbr ascending).
// The reference needs Boost to build; the restatement is pinned by
// hand-traced graphs (tests/test_oracle.py, tests/test_oracle_pins.py).
// rec (optional): Bisson/A-direction.cpp's degreeRecord: the scan writes the
// peeled vertex's runtime degree (:96); a cascade push writes the pushed
// vertex's runtime degree into the slot of the pushing element's frontier
// index f (:118 writes degreeRecord[i], i the frontier loop index).
// thre_param > 0: edge_directing/A-direction.cpp:36-161 (the OpenMP program)
// in one-thread loop order, with each frontier holding exactly the vertices
// peeled at its step: threshold0 = 2 edgeNum / nodeNum (:38), a cascade
// level runs only while the frontier exceeds nodeNum / 5000 and 200 and for
// at most six levels (:96), the threshold grows by threshold_old *
// thre_param / 5 (:139).
static void adir_peel_impl(uint32_t n, const uint32_t* sym_rowptr, const uint32_t* sym_col,
                           int32_t* prio, int32_t* rec, double thre_param = 0) {
  if (n == 0) return;
  const bool omp = thre_param > 0;
  unsigned int edgeNum = (unsigned int)sym_rowptr[n];  // totalEdge = 2m
  int nodeNum = (int)n;
  float threshold = omp ? 2 * ((float)edgeNum) / nodeNum : ((float)edgeNum) / nodeNum + 1;
  for (uint32_t i = 0; i < n; ++i) prio[i] = -1;
  std::vector<int> degree(n, 0);
  // every (i,nbr) pair bumps both ends: degree = 2*|N(v)|  (:44-47)
  for (uint32_t i = 0; i < n; ++i) {
    for (uint32_t e = sym_rowptr[i]; e < sym_rowptr[i + 1]; ++e) {
      degree[i]++;
      degree[sym_col[e]]++;
    }
  }
  std::vector<int> rtd(degree);
  std::vector<long long> off(n, 0);
  for (uint32_t i = 0; i + 1 < n; ++i) off[i + 1] = off[i] + degree[i];
  std::vector<int> allEdges((size_t)edgeNum * 2);
  {
    std::vector<long long> cur(off);
    for (uint32_t i = 0; i < n; ++i) {  // :66-73
      for (uint32_t e = sym_rowptr[i]; e < sym_rowptr[i + 1]; ++e) {
        int s = (int)i, d = (int)sym_col[e];
        allEdges[cur[s]++] = d;
        allEdges[cur[d]++] = s;
      }
    }
  }
  int priorityNum = 0;
  int threshold_old = threshold;
  int peeled = 0;
  std::vector<int> frontier, next;
  for (;;) {  // :79-128
    frontier.clear();
    next.clear();
    for (int i = 0; i < nodeNum; ++i) {
      if (rtd[i] <= threshold && prio[i] == -1) {
        frontier.push_back(i);
        peeled++;
        prio[i] = priorityNum;
        if (rec) rec[i] = rtd[i];
      }
    }
    priorityNum++;
    int innerLoop = 0;
    while (!frontier.empty() &&
           (!omp || ((int)frontier.size() > nodeNum / 5000 && (int)frontier.size() > 200 &&
                     innerLoop <= 5))) {
      innerLoop++;
      for (size_t f = 0; f < frontier.size(); ++f) {
        int cur = frontier[f];
        const int* lst = allEdges.data() + off[cur];
        for (int j = 0; j < degree[cur]; ++j) {
          int node = lst[j];
          if (prio[node] != -1) {
            if (prio[node] == prio[cur] &&
                (rtd[node] > rtd[cur] || (rtd[node] == rtd[cur] && node > cur)))
              rtd[node]--;
          } else {
            rtd[node]--;
            if (rtd[node] <= threshold) {
              next.push_back(node);
              peeled++;
              prio[node] = priorityNum;
              if (rec) rec[f] = rtd[node];
            }
          }
        }
      }
      std::swap(frontier, next);
      next.clear();
      priorityNum++;
    }
    if (peeled >= nodeNum) break;
    if (omp) threshold += (double)threshold_old * thre_param / 5;
    else threshold += threshold_old;
  }
}

void orc_adir_omp_priority(uint32_t n, const uint32_t* sym_rowptr, const uint32_t* sym_col,
                           double thre_param, int32_t* prio) {
  adir_peel_impl(n, sym_rowptr, sym_col, prio, nullptr, thre_param);
}

void orc_adir_priority(uint32_t n, const uint32_t* sym_rowptr, const uint32_t* sym_col,
                       int32_t* prio) {
  adir_peel_impl(n, sym_rowptr, sym_col, prio, nullptr);
}

// Bisson/A-direction.cpp:180-304: the peel over the records' simple
// symmetric graph (nodeNum = max id + 1, :196-202; self-loops dropped and
// lists deduplicated, :223-238), then units sorted by (priority, degree
// record, id) ascending (:34-42,287-293) and lookUpTable[id] = nodeNum - 1 -
// rank (:294-297).  perm has max id + 1 entries.
void orc_relabel_adir(const uint32_t* src, const uint32_t* dst, uint64_t m, uint32_t* perm) {
  const uint32_t n = orc_num_nodes(src, dst, m);
  std::vector<uint32_t> rp((size_t)n + 1), col(std::max<uint64_t>(2 * m, 1));
  const uint64_t m2 = orc_build_sym(src, dst, m, n, rp.data(), col.data());
  (void)m2;
  std::vector<int32_t> prio(n), rec(n, 0);
  adir_peel_impl(n, rp.data(), col.data(), prio.data(), rec.data());
  struct Unit {
    int priority, degree, id;
    bool operator<(const Unit& b) const {
      return priority < b.priority || (priority == b.priority && degree < b.degree) ||
             (priority == b.priority && degree == b.degree && id < b.id);
    }
  };
  std::vector<Unit> units(n);
  for (uint32_t i = 0; i < n; ++i) units[i] = Unit{prio[i], rec[i], (int)i};
  std::sort(units.begin(), units.end());
  for (uint32_t i = 0; i < n; ++i) perm[units[i].id] = n - 1 - i;
}

uint64_t orc_adir_orient(uint32_t n, const uint32_t* sym_rowptr,
                         const uint32_t* sym_col, const int32_t* prio,
                         uint32_t* osrc, uint32_t* odst) {
  uint64_t k = 0;
  for (uint32_t i = 0; i < n; ++i) {  // A-direction-seri.cpp:254-273
    for (uint32_t e = sym_rowptr[i]; e < sym_rowptr[i + 1]; ++e) {
      uint32_t j = sym_col[e];
      if (prio[j] > prio[i] || (prio[j] == prio[i] && i < j)) {
        osrc[k] = i;
        odst[k] = j;
        ++k;
      }
    }
  }
  return k;
}

void orc_kcore(uint32_t n, const uint32_t* sym_rowptr, const uint32_t* sym_col,
               uint32_t* core) {
  // Batagelj-Zaversnik bucket peel, kcore_direction.cpp:91-126.
  if (n == 0) return;
  std::vector<uint32_t> deg(n), D(n), p(n);
  uint32_t maxd = 0;
  for (uint32_t v = 0; v < n; ++v) {
    deg[v] = sym_rowptr[v + 1] - sym_rowptr[v];
    maxd = std::max(maxd, deg[v]);
  }
  std::vector<uint32_t> b(maxd + 2, 0);
  // stable counting sort of vertices by degree (kcore_direction.cpp:79)
  std::vector<uint32_t> cnt(maxd + 1, 0);
  for (uint32_t v = 0; v < n; ++v) cnt[deg[v]]++;
  uint32_t acc = 0;
  for (uint32_t d = 0; d <= maxd; ++d) {
    b[d] = acc;
    acc += cnt[d];
  }
  b[maxd + 1] = n;
  {
    std::vector<uint32_t> pos(b.begin(), b.begin() + maxd + 1);
    for (uint32_t v = 0; v < n; ++v) {
      D[pos[deg[v]]] = v;
      p[v] = pos[deg[v]]++;
    }
  }
  for (uint32_t i = 0; i < n; ++i) {
    uint32_t v = D[i];
    for (uint32_t e = sym_rowptr[v]; e < sym_rowptr[v + 1]; ++e) {
      uint32_t u = sym_col[e];
      if (deg[u] > deg[v]) {
        uint32_t du = deg[u], pu = p[u], pw = b[du], w = D[pw];
        if (u != w) {
          D[pu] = w; D[pw] = u; p[u] = pw; p[w] = pu;
        }
        b[du]++;
        deg[u]--;
      }
    }
  }
  for (uint32_t v = 0; v < n; ++v) core[v] = deg[v];
}

uint64_t orc_kcore_orient(uint32_t n, const uint32_t* sym_rowptr,
                          const uint32_t* sym_col, const uint32_t* core,
                          uint32_t* osrc, uint32_t* odst) {
  uint64_t k = 0;
  for (uint32_t i = 0; i < n; ++i) {
    for (uint32_t e = sym_rowptr[i]; e < sym_rowptr[i + 1]; ++e) {
      uint32_t j = sym_col[e];
      if (j <= i) continue;
      uint32_t s = i, d = j;  // kcore_direction.cpp:137-143
      if (core[s] > core[d] || (core[s] == core[d] && s > d)) std::swap(s, d);
      osrc[k] = s;
      odst[k] = d;
      ++k;
    }
  }
  return k;
}

void orc_degeneracy(uint32_t n, const uint32_t* sym_rowptr, const uint32_t* sym_col,
                    uint32_t* step, uint32_t* core) {
  const uint32_t kNone = 0xFFFFFFFFu;
  std::vector<uint32_t> deg(n);
  for (uint32_t v = 0; v < n; ++v) {
    deg[v] = sym_rowptr[v + 1] - sym_rowptr[v];
    step[v] = kNone;
  }
  uint32_t removed = 0, k = 0, t = 0;
  std::vector<uint32_t> frontier;
  while (removed < n) {
    // every remaining vertex of degree <= k, then their remaining
    // neighbours that drop to <= k, one frontier per step
    frontier.clear();
    for (uint32_t v = 0; v < n; ++v)
      if (step[v] == kNone && deg[v] <= k) frontier.push_back(v);
    while (!frontier.empty()) {
      for (uint32_t v : frontier) {
        step[v] = t;
        core[v] = k;
      }
      removed += (uint32_t)frontier.size();
      std::vector<uint32_t> next;
      for (uint32_t v : frontier)
        for (uint32_t e = sym_rowptr[v]; e < sym_rowptr[v + 1]; ++e) {
          const uint32_t y = sym_col[e];
          if (step[y] != kNone) continue;
          if (--deg[y] == k) next.push_back(y);
        }
      ++t;
      frontier.swap(next);
    }
    if (removed == n) break;
    uint32_t mind = kNone;
    for (uint32_t v = 0; v < n; ++v)
      if (step[v] == kNone) mind = std::min(mind, deg[v]);
    k = std::max(k + 1, mind);
  }
}

int orc_aorder(const uint32_t* src, const uint32_t* dst, uint64_t m, uint32_t n,
               uint32_t* perm, uint32_t* osrc, uint32_t* odst) {
  const int BUCKETSIZE = 40;  // A-order.cpp:14
  int maxNodeId = (int)n - 1;
  std::vector<int> degree(n, 0), rtd(n, 0);
  for (uint64_t i = 0; i < m; ++i) {  // :80-83
    degree[src[i]]++;
    degree[dst[i]]++;
  }
  for (uint64_t i = 0; i < m; ++i) {  // :88-99
    int s = (int)src[i], d = (int)dst[i];
    if (degree[s] > degree[d] || (degree[s] == degree[d] && s < d)) {
      osrc[i] = d; odst[i] = s;
      rtd[d]++;
    } else {
      osrc[i] = s; odst[i] = d;
      rtd[s]++;
    }
  }
  int maxRtd = 0;
  for (int i = 0; i <= maxNodeId; ++i) maxRtd = std::max(maxRtd, rtd[i]);
  std::vector<std::vector<int>> bucket(maxRtd + 1);
  for (int i = 0; i <= maxNodeId; ++i) bucket[rtd[i]].push_back(i);
  std::vector<std::pair<int, int>> idr;
  size_t mdp = 0;  // memoryDominatePos
  for (int i = 2; i <= maxRtd; ++i) {  // :113-119
    if (i == 5) mdp = idr.size();
    for (int v : bucket[i]) idr.push_back({v, i});
  }
  int B = (int)(idr.size() / BUCKETSIZE);
  // The reference's loops (:144, :165) stay in bounds only when
  // 1 <= memoryDominatePos <= BUCKETSIZE*B.
  if (B < 1 || mdp < 1 || mdp > (size_t)BUCKETSIZE * B) return -1;
  std::vector<Bucket> all(B);
  std::vector<std::vector<int>> nodes(B);
  for (int i = 0; i < B; ++i) all[i] = Bucket{i, 0.0f, i};
  std::make_heap(all.begin(), all.end(), GreatCmp());
  std::vector<float> bw(maxRtd + 1), ebw(maxRtd + 1);
  for (int i = 0; i < maxRtd + 1; ++i) {
    bw[i] = aorder_bandwidth(i);
    ebw[i] = aorder_equal_bandwidth(i);
  }
  int finished = 0;
  for (long long i = (long long)BUCKETSIZE * B - 1; i >= (long long)mdp; --i) {  // :144-157
    std::pop_heap(all.begin(), all.end() - finished, GreatCmp());
    Bucket& cb = *(all.rbegin() + finished);
    int node = idr[i].first, d = idr[i].second;
    nodes[cb.slot].push_back(node);
    cb.ms += (bw[d] - ebw[d]) / BUCKETSIZE;
    if ((int)nodes[cb.slot].size() >= BUCKETSIZE) finished++;
    else std::push_heap(all.begin(), all.end() - finished, GreatCmp());
  }
  std::make_heap(all.begin(), all.end() - finished, LessCmp());
  for (size_t i = 0; i < mdp; ++i) {  // :165-177
    std::pop_heap(all.begin(), all.end() - finished, LessCmp());
    Bucket& cb = *(all.rbegin() + finished);
    int node = idr[i].first, d = idr[i].second;
    nodes[cb.slot].push_back(node);
    cb.ms -= (ebw[d] - bw[d]) / BUCKETSIZE;
    if ((int)nodes[cb.slot].size() >= BUCKETSIZE) finished++;
    else std::push_heap(all.begin(), all.end() - finished, LessCmp());
  }
  // :181 sorts a range that is not a heap under IdSort; the order that
  // results is whatever the standard library's sort_heap produces.
  std::sort_heap(all.begin(), all.end(), IdSort());
  std::vector<int> order;
  order.reserve(n);
  for (int i = 0; i < B; ++i)
    for (int v : nodes[all[i].slot]) order.push_back(v);
  for (size_t i = (size_t)B * BUCKETSIZE; i < idr.size(); ++i) order.push_back(idr[i].first);
  if (maxRtd >= 1) for (int v : bucket[1]) order.push_back(v);
  for (int v : bucket[0]) order.push_back(v);
  for (int i = 0; i <= maxNodeId; ++i) perm[order[i]] = (uint32_t)i;  // :206-210
  return 0;
}

uint64_t orc_count(uint32_t n, const uint32_t* rowptr, const uint32_t* col) {
  uint64_t total = 0;
  for (uint32_t u = 0; u < n; ++u) total += vertex_tri(rowptr, col, u);
  return total;
}

uint64_t orc_count_omp(uint32_t n, const uint32_t* rowptr, const uint32_t* col) {
  uint64_t total = 0;
#pragma omp parallel for schedule(dynamic, 256) reduction(+ : total)
  for (long long u = 0; u < (long long)n; ++u) total += vertex_tri(rowptr, col, (uint32_t)u);
  return total;
}

uint64_t orc_count_rows(uint32_t n, const uint32_t* rowptr, const uint32_t* col,
                        uint32_t u_begin, uint32_t u_end) {
  uint64_t total = 0;
  if (u_end > n) u_end = n;
  for (uint32_t u = u_begin; u < u_end; ++u) total += vertex_tri(rowptr, col, u);
  return total;
}

uint64_t orc_count_rows_strided(uint32_t n, const uint32_t* rowptr, const uint32_t* col,
                                uint32_t start, uint32_t stride) {
  uint64_t total = 0;
  if (stride == 0) stride = 1;
  for (uint64_t u = start; u < n; u += stride) total += vertex_tri(rowptr, col, (uint32_t)u);
  return total;
}

uint64_t orc_count_rows_strided_omp(uint32_t n, const uint32_t* rowptr, const uint32_t* col,
                                    uint32_t start, uint32_t stride, int threads) {
  uint64_t total = 0;
  if (stride == 0) stride = 1;
  const long long rows = start < n ? ((long long)n - start + stride - 1) / stride : 0;
#pragma omp parallel for num_threads(threads) schedule(dynamic, 64) reduction(+ : total)
  for (long long i = 0; i < rows; ++i)
    total += vertex_tri(rowptr, col, (uint32_t)(start + (uint64_t)i * stride));
  return total;
}

// Triangles at the listed rows (ids >= n are skipped); OpenMP over the list
// when threads > 1.  Used to time a sample of rows chosen by the caller.
uint64_t orc_count_row_list(uint32_t n, const uint32_t* rowptr, const uint32_t* col,
                            const uint32_t* rows, uint64_t nrows, int threads) {
  uint64_t total = 0;
  if (threads <= 1) {
    for (uint64_t i = 0; i < nrows; ++i)
      if (rows[i] < n) total += vertex_tri(rowptr, col, rows[i]);
    return total;
  }
#pragma omp parallel for num_threads(threads) schedule(dynamic, 64) reduction(+ : total)
  for (long long i = 0; i < (long long)nrows; ++i)
    if (rows[i] < n) total += vertex_tri(rowptr, col, rows[i]);
  return total;
}

void orc_count_per_vertex(uint32_t n, const uint32_t* rowptr, const uint32_t* col,
                          uint64_t* tri) {
  for (uint32_t u = 0; u < n; ++u) tri[u] = vertex_tri(rowptr, col, u);
}

uint64_t orc_kclique(uint32_t n, const uint32_t* rowptr, const uint32_t* col,
                     uint32_t k) {
  if (k == 0) return 0;
  if (k == 1) return n;
  uint64_t total = 0;
  std::vector<uint32_t> S;
  for (uint32_t u = 0; u < n; ++u) {
    S.assign(col + rowptr[u], col + rowptr[u + 1]);
    if (S.size() + 1 < k) continue;
    total += clique_rec(rowptr, col, S, k - 1);
  }
  return total;
}

uint64_t orc_kclique_rows_strided(uint32_t n, const uint32_t* rowptr, const uint32_t* col,
                                  uint32_t k, uint32_t start, uint32_t stride) {
  if (k < 2 || stride == 0) return 0;
  uint64_t total = 0;
  std::vector<uint32_t> S;
  for (uint64_t u = start; u < n; u += stride) {
    S.assign(col + rowptr[u], col + rowptr[u + 1]);
    if (S.size() + 1 < k) continue;
    total += clique_rec(rowptr, col, S, k - 1);
  }
  return total;
}

// k-cliques whose lowest vertex is rows[i], one count per listed row (ids >= n
// give 0); OpenMP over the list when threads > 1.  The same depth-first
// intersection as orc_kclique, restricted to chosen roots.
void orc_kclique_row_list(uint32_t n, const uint32_t* rowptr, const uint32_t* col, uint32_t k,
                          const uint32_t* rows, uint64_t nrows, uint64_t* counts, int threads) {
#pragma omp parallel for num_threads(threads > 1 ? threads : 1) schedule(dynamic, 1)
  for (long long i = 0; i < (long long)nrows; ++i) {
    const uint32_t u = rows[i];
    uint64_t c = 0;
    if (u < n && k >= 2) {
      std::vector<uint32_t> S(col + rowptr[u], col + rowptr[u + 1]);
      if (S.size() + 1 >= k) c = clique_rec(rowptr, col, S, k - 1);
    } else if (u < n && k == 1) {
      c = 1;
    }
    counts[i] = c;
  }
}

void orc_hu_workload(uint32_t n, const uint32_t* rowptr, const uint32_t* col,
                     int64_t* w) {
  for (uint32_t u = 0; u < n; ++u) {  // Hu/tc.cu:22-35
    uint32_t du = rowptr[u + 1] - rowptr[u];
    int64_t total = 0;
    for (uint32_t e = rowptr[u]; e < rowptr[u + 1]; ++e) {
      uint32_t v = col[e];
      total += rowptr[v + 1] - rowptr[v];
    }
    // log2(0) cast to unsigned is undefined in the reference; it is always
    // multiplied by total == 0 there, so W = 0.
    w[u] = du == 0 ? 0 : total * (int64_t)(unsigned)std::log2((double)du);
  }
}

void orc_hu_block_starts(uint32_t n, const int64_t* w, uint32_t nblocks,
                         int32_t* start) {
  std::vector<int64_t> pre(w, w + n);
  for (uint32_t i = 1; i < n; ++i) pre[i] += pre[i - 1];  // Hu/tc.cu:286-287
  int64_t last = n ? pre[n - 1] : 0;
  int64_t step = (last + nblocks - 1) / nblocks;         // :294
  start[0] = 0;
  for (uint32_t i = 1; i < nblocks; ++i)                 // :298-300
    start[i] = (int32_t)hu_binary_search_value(pre.data(), (int64_t)i * step + 1, n, -1);
  start[nblocks] = (int32_t)n;
}

void orc_fox_workload(uint64_t m, const uint32_t* esrc, const uint32_t* edst,
                      const uint32_t* adjlen, int32_t* w) {
  for (uint64_t i = 0; i < m; ++i) {  // Fox/tc.cu:22-31
    int srcLen = (int)adjlen[esrc[i]];
    int dstLen = (int)adjlen[edst[i]];
    int large = srcLen > dstLen ? srcLen : dstLen;
    int bWork = (srcLen + dstLen - large) * std::log2((double)large + 2);
    w[i] = bWork;
  }
}

void orc_fox_bins(const int32_t* w, uint32_t size, uint32_t max_lanes,
                  uint32_t* bin_start) {
  // Fox/tc.cu:316-331
  int binMaxWork = size ? w[size - 1] : 0;
  int workPerThreadB = 8;
  double maxPowerOf8 = std::log(binMaxWork) / std::log(workPerThreadB);
  int curBinB = (int)max_lanes;
  for (int i = 0; i < curBinB + 2; ++i) bin_start[i] = 0;
  bin_start[1] = 0;
  for (int i = 2; i <= curBinB; i *= 2) {
    double index = (double)(maxPowerOf8 - 1) * (i - 1) / (double)curBinB + 1;
    int partitionPoint = powl(workPerThreadB, index);
    bin_start[i] = fox_binary_search_value(w, partitionPoint, size, -1);
    bin_start[i / 2 + 1] = bin_start[i];
  }
  bin_start[curBinB + 1] = size;
}

void orc_cost_func(const uint32_t* src, const uint32_t* dst, uint64_t m, int k,
                   float* out) {
  // tools/costFuncCal.cpp:21-88
  unsigned edgeNum = (unsigned)m;
  unsigned maxNode = 0;
  for (unsigned i = 0; i < edgeNum; ++i) {
    maxNode = src[i] > maxNode ? src[i] : maxNode;
    maxNode = dst[i] > maxNode ? dst[i] : maxNode;
  }
  unsigned nodeNum = maxNode + 1;
  float avgDegree = (float)edgeNum / nodeNum;
  std::vector<int> degree(nodeNum, 0), idB(nodeNum, 0), degB(nodeNum, 0), lpB(nodeNum, 0);
  for (unsigned i = 0; i < edgeNum; ++i) {
    degree[src[i]]++;
    degree[dst[i]]++;
    lpB[src[i]]++;
  }
  for (unsigned i = 0; i < edgeNum; ++i) {
    int s = (int)src[i], d = (int)dst[i];
    if (s < d) idB[s]++;
    else idB[d]++;
    if (degree[s] < degree[d] || (degree[s] == degree[d] && s < d)) degB[s]++;
    else degB[d]++;
  }
  float idPart = 0, degPart = 0, lpPart = 0;
  float countThreshold = k * avgDegree;
  for (int i = 0; i < (int)nodeNum; ++i) {
    float cv = std::fabs(idB[i] - avgDegree);
    if (idB[i] - avgDegree > countThreshold) idPart += cv;
    cv = std::fabs(degB[i] - avgDegree);
    if (degB[i] - avgDegree > countThreshold) degPart += cv;
    cv = std::fabs(lpB[i] - avgDegree);
    if (lpB[i] - avgDegree > countThreshold) lpPart += cv;
  }
  out[0] = idPart;
  out[1] = degPart;
  out[2] = lpPart;
}

void orc_verification(const uint32_t* src, const uint32_t* dst, uint64_t m,
                      float* out) {
  // tools/verification.cpp:14-60
  int maxNodeId = 0;
  for (uint64_t i = 0; i < m; ++i) {
    maxNodeId = (int)src[i] > maxNodeId ? (int)src[i] : maxNodeId;
    maxNodeId = (int)dst[i] > maxNodeId ? (int)dst[i] : maxNodeId;
  }
  float dAve = (float)m / (maxNodeId + 1);
  std::vector<int> lp(maxNodeId + 1, 0), degree(maxNodeId + 1, 0), rtd(maxNodeId + 1, 0);
  for (uint64_t i = 0; i < m; ++i) {
    lp[src[i]]++;
    degree[src[i]]++;
    degree[dst[i]]++;
  }
  float f_lp = 0;
  for (int i = 0; i <= maxNodeId; ++i) f_lp += std::fabs(dAve - lp[i]);
  for (uint64_t i = 0; i < m; ++i) {
    int s = (int)src[i], d = (int)dst[i];
    if (degree[s] > degree[d] || (degree[s] == degree[d] && s < d)) rtd[d]++;
    else rtd[s]++;
  }
  float f_degree = 0;
  for (int i = 0; i < maxNodeId; ++i) f_degree += std::fabs(rtd[i] - dAve);  // :58 stops before the last id
  out[0] = f_lp;
  out[1] = f_degree;
}

void orc_relabel_compact(const uint32_t* src, const uint32_t* dst, uint64_t m, uint32_t* perm) {
  // tools/serialize.cpp:17-47: present ids keep their order, absent ids are
  // squeezed out (new = id - number of absent ids below it)
  uint32_t mx = 0;
  for (uint64_t i = 0; i < m; ++i) mx = std::max(mx, std::max(src[i], dst[i]));
  std::vector<uint8_t> present((size_t)mx + 1, 0);
  for (uint64_t i = 0; i < m; ++i) present[src[i]] = present[dst[i]] = 1;
  uint32_t absent = 0;
  for (uint32_t i = 0; i <= mx; ++i) {
    perm[i] = i - absent;
    if (!present[i]) ++absent;
  }
}

void orc_relabel_first_src(const uint32_t* src, const uint32_t* dst, uint64_t m,
                           uint32_t* perm) {
  // tools/adjust2tc.cpp:13-35: ids numbered in order of first appearance as
  // a source, then every other id up to the largest in id order
  uint32_t mx = 0;
  for (uint64_t i = 0; i < m; ++i) mx = std::max(mx, std::max(src[i], dst[i]));
  std::vector<int64_t> f((size_t)mx + 1, -1);
  int64_t count = 0;
  for (uint64_t i = 0; i < m; ++i)
    if (f[src[i]] == -1) f[src[i]] = count++;
  for (uint32_t i = 0; i <= mx; ++i)
    if (f[i] == -1) f[i] = count++;
  for (uint32_t i = 0; i <= mx; ++i) perm[i] = (uint32_t)f[i];
}

void orc_relabel_degree(const uint32_t* src, const uint32_t* dst, uint64_t m, uint32_t* perm) {
  // tools/directionByReorder.cpp:38-70: sort (degree, id) ascending; the
  // i-th gets id n - 1 - i, so the highest degree becomes 0
  uint32_t mx = 0;
  for (uint64_t i = 0; i < m; ++i) mx = std::max(mx, std::max(src[i], dst[i]));
  const uint64_t n = (uint64_t)mx + 1;
  std::vector<uint32_t> deg(n, 0);
  for (uint64_t i = 0; i < m; ++i) {
    deg[src[i]]++;
    deg[dst[i]]++;
  }
  std::vector<std::pair<uint32_t, uint32_t>> u(n);
  for (uint64_t i = 0; i < n; ++i) u[i] = {deg[i], (uint32_t)i};
  std::sort(u.begin(), u.end());
  for (uint64_t i = 0; i < n; ++i) perm[u[i].second] = (uint32_t)(n - 1 - i);
}

double orc_approx_ratio(const uint32_t* src, const uint32_t* dst, uint64_t m) {
  // tools/approximateRatio.cpp:16-94
  int maxNodeId = 0;
  for (uint64_t i = 0; i < m; ++i) {
    maxNodeId = (int)src[i] > maxNodeId ? (int)src[i] : maxNodeId;
    maxNodeId = (int)dst[i] > maxNodeId ? (int)dst[i] : maxNodeId;
  }
  std::vector<int> degree(maxNodeId + 1, 0);
  for (uint64_t i = 0; i < m; ++i) {
    degree[src[i]]++;
    degree[dst[i]]++;
  }
  double avg_outd = (double)m / (maxNodeId + 1);
  int dmax = 0;
  for (int i = 0; i <= maxNodeId; ++i) dmax = std::max(dmax, degree[i]);
  std::vector<int> percent(dmax + 2, 0);
  double vlow = 0, vhigh = 0, sumLow = 0, sumHigh = 0;
  for (int i = 0; i <= maxNodeId; ++i) {
    if (degree[i] <= avg_outd) { vlow++; sumLow += degree[i]; }
    else { vhigh++; sumHigh += degree[i]; }
    percent[degree[i]]++;
  }
  double lb_opt = 0;
  for (int d = 1; d <= avg_outd; ++d) lb_opt += percent[d] * (avg_outd - d);
  double lb_opt2 = 0;
  double avg_vhigh = sumHigh / vhigh / 2;
  if (avg_vhigh >= avg_outd) {
    if ((sumHigh - sumLow) / 2 - avg_outd * vhigh >= 0)
      lb_opt2 = (sumHigh - sumLow) / 2 - avg_outd * vhigh;
    else
      lb_opt2 = 0;
  } else {
    lb_opt2 = avg_outd * vhigh - sumHigh / 2;
  }
  double ub = 0, remain = 0;
  for (int d = (int)avg_outd + 1; d <= dmax; ++d) remain += percent[d] * d / 2;
  int dx = (int)avg_outd + 1;
  while (remain > 0) {
    if (remain > dx * percent[dx]) {
      ub += avg_outd * percent[dx];
      remain -= dx * percent[dx];
      dx++;
    } else {
      ub += avg_outd * remain / dx;
      break;
    }
  }
  return 1 + ub / (lb_opt + lb_opt2);
}

}  // extern "C"
