// OpenMP A-direction driver: A-direction <edges.bin> <thre_param>
// (edge_directing/A-direction.cpp:186-294).  Peels on the GPU with the
// program's schedule (threshold0 = 2 totalEdge / n, a cascade only for
// frontiers above max(n / 5000, 200) and at most six levels, threshold step
// (int)threshold0 * thre_param / 5; tc_graph_orient_adir_omp) and writes
// lpBasedGraph.bin in the current directory: one record (i, j) per
// undirected edge, with priority[j] > priority[i] or (equal and i < j).
#include "common.h"

int main(int argc, const char* argv[]) {
  if (argc < 3) {
    std::printf("Usage: ./A-direction inputFileName thre_param\n");
    return 0;
  }
  const double t0 = cli::now_s();
  const double thre_param = (double)std::atoi(argv[2]);  // :197
  cli::Records r = cli::load(argv[1]);
  tc_ctx* ctx = nullptr;
  cli::check(tc_ctx_create(0, &ctx), "tc_ctx_create");
  tc_graph* g = nullptr;
  cli::check(tc_graph_create(ctx, r.src.data(), r.dst.data(), r.m, 0, &g), "tc_graph_create");
  cli::check(tc_graph_orient_adir_omp(ctx, g, thre_param), "tc_graph_orient_adir_omp");
  tc_graph_info info;
  cli::check(tc_graph_info_get(g, &info), "tc_graph_info_get");
  std::printf("the max nodeNum is %u\n", info.n);
  std::printf("the total edges is %llu\n", (unsigned long long)info.m_oriented);
  std::vector<uint32_t> rp(info.n + 1), col(info.m_oriented ? info.m_oriented : 1);
  cli::check(tc_graph_get_csr(g, rp.data(), col.data()), "tc_graph_get_csr");
  std::vector<uint32_t> s(col.size()), d(col.size());
  for (uint32_t u = 0; u < info.n; ++u)
    for (uint32_t k = rp[u]; k < rp[u + 1]; ++k) {
      s[k] = u;
      d[k] = col[k];
    }
  cli::check(tc_write_edges("lpBasedGraph.bin", s.data(), d.data(), info.m_oriented),
             "tc_write_edges");
  std::printf("the total optTime is %g s.\n", cli::now_s() - t0);
  std::printf("we have dealed with %llu edges up to now\n", (unsigned long long)info.m_oriented);
  tc_graph_destroy(g);
  tc_ctx_destroy(ctx);
  return 0;
}
