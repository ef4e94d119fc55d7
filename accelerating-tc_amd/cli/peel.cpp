// Parallel-peel driver: Peel <in.bin> <out.bin>
// (edge_directing/A-direction-parallel/peel.cu:130-300).  Writes every
// input record oriented by the A-direction priorities (serial numbering of
// A-direction-seri.cpp; the reference's GPU peel uses a relaxed numbering):
// the lower-priority end (ties: smaller id) goes to the record's dst field,
// as outPutResults (peel.cu:40-62) stores it.  Self-loops are kept as read.
#include "common.h"

int main(int argc, const char* argv[]) {
  if (argc != 3) {
    std::printf("Usage: ./Peel inputFileName outFileName\n");
    return 0;
  }
  const double t0 = cli::now_s();
  cli::Records r = cli::load(argv[1]);
  tc_ctx* ctx = nullptr;
  cli::check(tc_ctx_create(0, &ctx), "tc_ctx_create");
  tc_graph* g = nullptr;
  cli::check(tc_graph_create(ctx, r.src.data(), r.dst.data(), r.m, 0, &g), "tc_graph_create");
  cli::check(tc_graph_orient(ctx, g, TC_ORIENT_ADIR, 0), "tc_graph_orient");
  tc_graph_info info;
  cli::check(tc_graph_info_get(g, &info), "tc_graph_info_get");
  std::printf("the node num is %u, and the edgeNum is %llu\n", info.n, (unsigned long long)r.m);
  std::vector<int32_t> prio(info.n ? info.n : 1);
  cli::check(tc_graph_get_vertex_key(g, prio.data()), "tc_graph_get_vertex_key");
  for (uint64_t i = 0; i < r.m; ++i) {
    uint32_t a = r.src[i], b = r.dst[i];
    // a = higher-priority end, b = lower
    const bool b_higher = prio[b] > prio[a] || (prio[b] == prio[a] && b > a);
    if (b_higher) std::swap(a, b);
    r.src[i] = a;
    r.dst[i] = b;
  }
  cli::check(tc_write_edges(argv[2], r.src.data(), r.dst.data(), r.m), "tc_write_edges");
  std::printf("the peel time is %g s.\n", cli::now_s() - t0);
  tc_graph_destroy(g);
  tc_ctx_destroy(ctx);
  return 0;
}
