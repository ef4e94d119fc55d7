// General driver: tc -f <edges.bin> [--kernel hu|bisson|fox|tricore|auto]
//                    [--orient id|degree|adir|kcore|degeneracy|tricore] [--aorder] [--simplify]
// Any method under any orientation; prints the reference's triangle line.
#include <cstring>

#include "common.h"

int main(int argc, const char* argv[]) {
  const char* path = nullptr;
  int kernel = TC_KERNEL_AUTO, orient = TC_ORIENT_ADIR;
  bool aorder = false, simplify = false;
  for (int i = 1; i < argc; ++i) {
    const std::string a = argv[i];
    if (a == "-f" && i + 1 < argc) path = argv[++i];
    else if (a == "--kernel" && i + 1 < argc) {
      const std::string v = argv[++i];
      kernel = v == "hu" ? TC_KERNEL_HU : v == "bisson" ? TC_KERNEL_BISSON
             : v == "fox" ? TC_KERNEL_FOX : v == "tricore" ? TC_KERNEL_TRICORE : TC_KERNEL_AUTO;
    } else if (a == "--orient" && i + 1 < argc) {
      const std::string v = argv[++i];
      orient = v == "id" ? TC_ORIENT_ID : v == "degree" ? TC_ORIENT_DEGREE
             : v == "kcore" ? TC_ORIENT_KCORE : v == "tricore" ? TC_ORIENT_DEGREE_TRICORE
             : v == "degeneracy" ? TC_ORIENT_DEGENERACY : TC_ORIENT_ADIR;
    } else if (a == "--aorder") aorder = true;
    else if (a == "--simplify") simplify = true;
  }
  if (!path) {
    std::printf("Usage: ./tc -f inputFileName [--kernel K] [--orient O] [--aorder] [--simplify]\n");
    return 1;
  }
  const double t0 = cli::now_s();
  cli::Records r = cli::load(path);
  tc_ctx* ctx = nullptr;
  cli::check(tc_ctx_create(0, &ctx), "tc_ctx_create");
  tc_graph* g = nullptr;
  cli::check(tc_graph_create(ctx, r.src.data(), r.dst.data(), r.m, 0, &g), "tc_graph_create");
  if (simplify) cli::check(tc_graph_simplify(ctx, g), "tc_graph_simplify");
  if (aorder) cli::check(tc_graph_relabel_aorder(ctx, g, nullptr), "tc_graph_relabel_aorder");
  cli::check(tc_graph_orient(ctx, g, orient, 0), "tc_graph_orient");
  tc_graph_info info;
  cli::check(tc_graph_info_get(g, &info), "tc_graph_info_get");
  std::printf("the node num is %u, and the edgeNum is %llu\n", info.n,
              (unsigned long long)info.m_oriented);
  tc_sched sched = {};
  sched.kernel = kernel;
  uint64_t count = 0;
  tc_timing tm;
  cli::check(tc_count(ctx, g, &sched, &count, &tm), "tc_count");
  std::printf("There are %llu triangles in the input graph.\n", (unsigned long long)count);
  std::printf("kernel use %g s.\n", tm.ms_count * 1e-3);
  std::printf("Total use time %g s.\n", cli::now_s() - t0);
  tc_graph_destroy(g);
  tc_ctx_destroy(ctx);
  return 0;
}
