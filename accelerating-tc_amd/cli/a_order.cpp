// A-order driver: A-order <edges.bin>  (vertex_ordering/A-order/A-order.cpp:62-230).
// Writes `normal` (each record degree-oriented: the end of larger degree,
// ties the smaller id, goes to dst) and `reordered` (the same records with
// the A-order ids) in the current directory.
#include "common.h"

int main(int argc, const char* argv[]) {
  if (argc < 2) {
    std::printf("Usage: ./A-order inputFileName\n");
    return 0;
  }
  cli::Records r = cli::load(argv[1]);
  uint32_t n = 0;
  for (uint64_t i = 0; i < r.m; ++i) n = std::max(n, std::max(r.src[i], r.dst[i]) + 1);
  std::vector<uint64_t> deg(n ? n : 1, 0);
  for (uint64_t i = 0; i < r.m; ++i) {
    deg[r.src[i]]++;
    deg[r.dst[i]]++;
  }
  for (uint64_t i = 0; i < r.m; ++i) {
    const uint32_t s = r.src[i], d = r.dst[i];
    if (deg[s] > deg[d] || (deg[s] == deg[d] && s < d)) {
      r.src[i] = d;
      r.dst[i] = s;
    }
  }
  tc_ctx* ctx = nullptr;
  cli::check(tc_ctx_create(0, &ctx), "tc_ctx_create");
  tc_graph* g = nullptr;
  cli::check(tc_graph_create(ctx, r.src.data(), r.dst.data(), r.m, 0, &g), "tc_graph_create");
  std::vector<uint32_t> perm(n ? n : 1);
  cli::check(tc_graph_relabel_aorder(ctx, g, perm.data()), "tc_graph_relabel_aorder");
  cli::check(tc_write_edges("normal", r.src.data(), r.dst.data(), r.m), "tc_write_edges");
  for (uint64_t i = 0; i < r.m; ++i) {
    r.src[i] = perm[r.src[i]];
    r.dst[i] = perm[r.dst[i]];
  }
  cli::check(tc_write_edges("reordered", r.src.data(), r.dst.data(), r.m), "tc_write_edges");
  tc_graph_destroy(g);
  tc_ctx_destroy(ctx);
  return 0;
}
