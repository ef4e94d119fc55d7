// Seeded synthetic graph generators (ER, R-MAT, Barabasi-Albert, RGG).
//
// Every random draw comes from splitmix64 keyed by (seed, item index), so the
// output depends only on the arguments, never on the thread count.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "internal.h"

namespace {

inline uint64_t mix64(uint64_t z) {
  z += 0x9e3779b97f4a7c15ull;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

struct Rng {  // sequential splitmix64 stream
  uint64_t s;
  explicit Rng(uint64_t seed) : s(seed) {}
  uint64_t next() {
    uint64_t z = (s += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
  }
  uint32_t below(uint32_t n) { return (uint32_t)(((next() >> 32) * (uint64_t)n) >> 32); }
  double unit() { return (next() >> 11) * (1.0 / 9007199254740992.0); }
};

inline uint64_t stream_seed(uint64_t seed, uint64_t idx) {
  return mix64(seed ^ mix64(idx + 0x632be59bd9b4e019ull));
}

// Open-addressing set of undirected pair keys.
struct PairSet {
  std::vector<uint64_t> slots;
  uint64_t mask;
  explicit PairSet(uint64_t want) {
    uint64_t cap = 16;
    while (cap < 2 * want + 16) cap <<= 1;
    slots.assign(cap, ~0ull);
    mask = cap - 1;
  }
  bool insert(uint64_t key) {
    uint64_t h = mix64(key) & mask;
    while (true) {
      uint64_t s = slots[h];
      if (s == key) return false;
      if (s == ~0ull) {
        slots[h] = key;
        return true;
      }
      h = (h + 1) & mask;
    }
  }
};

struct RmatArgs {
  uint32_t scale;
  uint64_t seed;
  uint64_t ta, tab, tabc;
  uint32_t* src;
  uint32_t* dst;
};

void rmat_range(uint64_t b, uint64_t e, void* p) {
  const RmatArgs& A = *static_cast<const RmatArgs*>(p);
  for (uint64_t i = b; i < e; ++i) {
    Rng r(stream_seed(A.seed, i));
    uint32_t u = 0, v = 0;
    uint64_t bits = 0;
    for (uint32_t lvl = 0; lvl < A.scale; ++lvl) {
      if ((lvl & 1) == 0) bits = r.next();
      uint64_t x = (lvl & 1) ? (bits >> 32) : (bits & 0xffffffffull);
      uint32_t bu, bv;
      if (x < A.ta) { bu = 0; bv = 0; }
      else if (x < A.tab) { bu = 0; bv = 1; }
      else if (x < A.tabc) { bu = 1; bv = 0; }
      else { bu = 1; bv = 1; }
      u = (u << 1) | bu;
      v = (v << 1) | bv;
    }
    A.src[i] = u;
    A.dst[i] = v;
  }
}

struct RggArgs {
  uint32_t n;
  double r;
  uint32_t g;  // grid cells per side
  const double* x;
  const double* y;
  const uint32_t* cell_start;  // g*g+1
  const uint32_t* cell_pts;    // points sorted by cell
  const uint32_t* pt_cell;
  uint64_t* cnt;               // per point (count pass) / offsets (write pass)
  uint32_t* src;
  uint32_t* dst;
};

template <bool kWrite>
void rgg_range(uint64_t b, uint64_t e, void* p) {
  const RggArgs& A = *static_cast<const RggArgs*>(p);
  const double r2 = A.r * A.r;
  std::vector<uint32_t> nb;
  for (uint64_t i = b; i < e; ++i) {
    uint32_t c = A.pt_cell[i];
    int cx = (int)(c % A.g), cy = (int)(c / A.g);
    nb.clear();
    for (int dy = -1; dy <= 1; ++dy) {
      int yy = cy + dy;
      if (yy < 0 || yy >= (int)A.g) continue;
      for (int dx = -1; dx <= 1; ++dx) {
        int xx = cx + dx;
        if (xx < 0 || xx >= (int)A.g) continue;
        uint32_t cc = (uint32_t)yy * A.g + (uint32_t)xx;
        for (uint32_t k = A.cell_start[cc]; k < A.cell_start[cc + 1]; ++k) {
          uint32_t j = A.cell_pts[k];
          if (j <= i) continue;
          double ddx = A.x[i] - A.x[j], ddy = A.y[i] - A.y[j];
          if (ddx * ddx + ddy * ddy <= r2) nb.push_back(j);
        }
      }
    }
    if (kWrite) {
      std::sort(nb.begin(), nb.end());
      uint64_t o = A.cnt[i];
      for (uint32_t j : nb) {
        A.src[o] = (uint32_t)i;
        A.dst[o] = j;
        ++o;
      }
    } else {
      A.cnt[i] = nb.size();
    }
  }
}

}  // namespace

using namespace tcx;

extern "C" {

int tc_gen_er(uint32_t n, uint64_t m, uint64_t seed, uint32_t* src, uint32_t* dst) {
  if (m && (!src || !dst)) return fail(TC_ERR_ARG, "NULL output arrays");
  if (n < 2 && m) return fail(TC_ERR_ARG, "ER needs n >= 2");
  double maxm = (double)n * (n - 1) / 2;
  if ((double)m > maxm) return fail(TC_ERR_ARG, "ER: m exceeds n(n-1)/2");
  Rng r(stream_seed(seed, 0));
  PairSet set(m);
  uint64_t k = 0;
  while (k < m) {
    uint32_t u = r.below(n), v = r.below(n);
    if (u == v) continue;
    uint64_t key = u < v ? ((uint64_t)u << 32 | v) : ((uint64_t)v << 32 | u);
    if (!set.insert(key)) continue;
    src[k] = u;
    dst[k] = v;
    ++k;
  }
  return TC_OK;
}

int tc_gen_rmat(uint32_t scale, uint64_t nedges, double a, double b, double c,
                uint64_t seed, uint32_t* src, uint32_t* dst) {
  if (nedges && (!src || !dst)) return fail(TC_ERR_ARG, "NULL output arrays");
  if (scale == 0 || scale > 31) return fail(TC_ERR_RANGE, "R-MAT scale must be 1..31");
  if (a < 0 || b < 0 || c < 0 || a + b + c > 1.0 + 1e-12)
    return fail(TC_ERR_ARG, "R-MAT probabilities invalid");
  RmatArgs A;
  A.scale = scale;
  A.seed = seed;
  const double two32 = 4294967296.0;
  A.ta = (uint64_t)(a * two32);
  A.tab = (uint64_t)((a + b) * two32);
  A.tabc = (uint64_t)((a + b + c) * two32);
  A.src = src;
  A.dst = dst;
  parallel_for(nedges, rmat_range, &A);
  return TC_OK;
}

int tc_gen_ba(uint32_t n, uint32_t k, uint64_t seed, uint32_t* src, uint32_t* dst,
              uint64_t* m_out) {
  if (!m_out) return fail(TC_ERR_ARG, "NULL m_out");
  if (k == 0 || n <= k) return fail(TC_ERR_ARG, "BA needs 0 < k < n");
  uint64_t m = (uint64_t)(n - k) * k;
  *m_out = m;
  if (!src || !dst) return fail(TC_ERR_ARG, "NULL output arrays");
  // Preferential attachment by repeated-endpoint sampling: vertex t >= k
  // attaches to k distinct targets drawn from the endpoint multiset (the
  // first new vertex attaches to the k seed vertices).
  Rng r(stream_seed(seed, 1));
  std::vector<uint32_t> rep;
  rep.reserve(2 * m);
  std::vector<uint32_t> targets(k);
  for (uint32_t i = 0; i < k; ++i) targets[i] = i;
  uint64_t e = 0;
  for (uint32_t t = k; t < n; ++t) {
    for (uint32_t i = 0; i < k; ++i) {
      src[e] = t;
      dst[e] = targets[i];
      ++e;
      rep.push_back(t);
      rep.push_back(targets[i]);
    }
    // pick the next vertex's targets
    uint32_t got = 0;
    while (got < k) {
      uint32_t cand = rep[(size_t)(((unsigned __int128)r.next() * rep.size()) >> 64)];
      bool dup = false;
      for (uint32_t j = 0; j < got; ++j) dup |= targets[j] == cand;
      if (!dup) targets[got++] = cand;
    }
  }
  return TC_OK;
}

int tc_gen_rgg(uint32_t n, double r, uint64_t seed, uint32_t* src, uint32_t* dst,
               uint64_t* m_out) {
  if (!m_out) return fail(TC_ERR_ARG, "NULL m_out");
  if (!(r > 0) || r > 1.5) return fail(TC_ERR_ARG, "RGG radius must be in (0, 1.5]");
  std::vector<double> x(n), y(n);
  for (uint32_t i = 0; i < n; ++i) {
    Rng q(stream_seed(seed, i));
    x[i] = q.unit();
    y[i] = q.unit();
  }
  uint32_t g = (uint32_t)std::max(1.0, std::floor(1.0 / r));
  if (g > 65536) g = 65536;
  std::vector<uint32_t> pt_cell(n), cell_start((size_t)g * g + 1, 0), cell_pts(n);
  for (uint32_t i = 0; i < n; ++i) {
    uint32_t cx = std::min<uint32_t>(g - 1, (uint32_t)(x[i] * g));
    uint32_t cy = std::min<uint32_t>(g - 1, (uint32_t)(y[i] * g));
    pt_cell[i] = cy * g + cx;
    cell_start[pt_cell[i] + 1]++;
  }
  for (size_t c = 0; c < (size_t)g * g; ++c) cell_start[c + 1] += cell_start[c];
  {
    std::vector<uint32_t> pos(cell_start.begin(), cell_start.end() - 1);
    for (uint32_t i = 0; i < n; ++i) cell_pts[pos[pt_cell[i]]++] = i;
  }
  std::vector<uint64_t> cnt(n + 1, 0);
  RggArgs A{n, r, g, x.data(), y.data(), cell_start.data(), cell_pts.data(),
            pt_cell.data(), cnt.data(), src, dst};
  parallel_for(n, rgg_range<false>, &A);
  uint64_t total = 0;
  for (uint32_t i = 0; i < n; ++i) {
    uint64_t c = cnt[i];
    cnt[i] = total;
    total += c;
  }
  cnt[n] = total;
  *m_out = total;
  if (!src || !dst) return TC_OK;
  parallel_for(n, rgg_range<true>, &A);
  return TC_OK;
}

}  // extern "C"
