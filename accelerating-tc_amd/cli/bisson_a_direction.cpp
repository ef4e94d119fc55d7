// Bisson's A-direction relabel: bisson_a_direction <edges.bin>
// (Bisson/A-direction.cpp:180-315).  The peel and the (priority, degree
// record, id) ranking run on the GPU (TC_RELABEL_ADIR); lpBasedGraph.bin in
// the current directory receives both directions of every simple edge,
// (lookUpTable[i], lookUpTable[nbr]) for old id i ascending and its sorted
// neighbours (:298-304), so that ID order follows A-direction order.
#include <algorithm>

#include "common.h"

int main(int argc, const char* argv[]) {
  if (argc < 2) {
    std::printf("Usage: ./A-direction inputFileName\n");
    return 0;
  }
  const double t0 = cli::now_s();
  cli::Records r = cli::load(argv[1]);
  uint32_t mx = 0;
  for (uint64_t i = 0; i < r.m; ++i) mx = std::max(mx, std::max(r.src[i], r.dst[i]));
  const uint32_t n = r.m ? mx + 1 : 1;
  std::printf("the max nodeNum is %u\n", n);
  // neighbours[i], sorted and unique, self-loops dropped (:223-238)
  std::vector<uint64_t> pairs;
  pairs.reserve(2 * r.m);
  for (uint64_t i = 0; i < r.m; ++i) {
    const uint64_t s = r.src[i], d = r.dst[i];
    if (s == d) continue;
    pairs.push_back(s << 32 | d);
    pairs.push_back(d << 32 | s);
  }
  std::sort(pairs.begin(), pairs.end());
  pairs.erase(std::unique(pairs.begin(), pairs.end()), pairs.end());
  std::printf("the total edges is %llu\n", (unsigned long long)(pairs.size() / 2));
  std::vector<uint32_t> perm(n);
  tc_ctx* ctx = nullptr;
  cli::check(tc_ctx_create(0, &ctx), "tc_ctx_create");
  tc_graph* g = nullptr;
  cli::check(tc_graph_create(ctx, r.src.data(), r.dst.data(), r.m, 0, &g), "tc_graph_create");
  cli::check(tc_graph_relabel(ctx, g, TC_RELABEL_ADIR, perm.data()), "tc_graph_relabel");
  tc_graph_destroy(g);
  tc_ctx_destroy(ctx);
  std::vector<uint32_t> s(pairs.size() ? pairs.size() : 1), d(s.size());
  for (size_t k = 0; k < pairs.size(); ++k) {
    s[k] = perm[pairs[k] >> 32];
    d[k] = perm[pairs[k] & 0xFFFFFFFFu];
  }
  cli::check(tc_write_edges("lpBasedGraph.bin", s.data(), d.data(), pairs.size()),
             "tc_write_edges");
  std::printf("the total optTime is %g s.\n", cli::now_s() - t0);
  return 0;
}
