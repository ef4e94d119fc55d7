// kcore_direction <in.bin> <out.bin>  (edge_directing/kcore_direction.cpp):
// records as uint64 (FIRST = high word, :9-11), core numbers by the GPU
// k-core peel, each record written as (lower, higher) core number, ties by
// the smaller id first (:131-143).  The degree loop of the reference runs over
// node_num instead of edge_num (:35); the intended degrees are used here.
#include "common.h"

int main(int argc, const char* argv[]) {
  if (argc < 3) {
    std::printf("Usage: ./kcore_direction in.bin out.bin\n");
    return 1;
  }
  cli::Records r = cli::load(argv[1]);  // src = low word (SECOND), dst = high word (FIRST)
  std::printf("the edge num is %llu\n", (unsigned long long)r.m);
  tc_ctx* ctx = nullptr;
  cli::check(tc_ctx_create(0, &ctx), "tc_ctx_create");
  tc_graph* g = nullptr;
  cli::check(tc_graph_create(ctx, r.src.data(), r.dst.data(), r.m, 0, &g), "tc_graph_create");
  cli::check(tc_graph_orient(ctx, g, TC_ORIENT_KCORE, 0), "tc_graph_orient");
  tc_graph_info info;
  cli::check(tc_graph_info_get(g, &info), "tc_graph_info_get");
  std::printf("the node num is %u\n", info.n);
  std::vector<int32_t> core(info.n ? info.n : 1);
  cli::check(tc_graph_get_vertex_key(g, core.data()), "tc_graph_get_vertex_key");
  FILE* o = std::fopen(argv[2], "wb");
  if (!o) {
    std::fprintf(stderr, "cannot open %s\n", argv[2]);
    return 1;
  }
  std::vector<uint64_t> out(r.m ? r.m : 1);
  for (uint64_t i = 0; i < r.m; ++i) {
    uint32_t a = r.dst[i], b = r.src[i];  // FIRST, SECOND
    if (core[a] > core[b] || (core[a] == core[b] && a > b)) std::swap(a, b);
    out[i] = ((uint64_t)a << 32) | b;
  }
  if (r.m && std::fwrite(out.data(), 8, r.m, o) != r.m) {
    std::fprintf(stderr, "short write on %s\n", argv[2]);
    return 1;
  }
  std::fclose(o);
  tc_graph_destroy(g);
  tc_ctx_destroy(ctx);
  return 0;
}
