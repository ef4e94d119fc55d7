// TriCore driver: tric_gpu -f <edges.bin> [debug]  (tricore/src/main.cpp:15-101,
// tricore/include/util.h:18-24).  The file is read as uint64 edges and goes
// through the function-level TriCore interface (init_gpu, cal_degree,
// redirect_edges, cal_part_num, split, tricount_gpu).
#include <cstring>

#include "common.h"

int main(int argc, char* argv[]) {
  const bool debug = std::string(argv[argc - 1]) == "debug";
  if (argc < 3) {
    std::printf("usage: %s -f graph_name\n", argv[0]);
    return 255;
  }
  const char* path = argv[2];
  FILE* f = std::fopen(path, "rb");
  if (!f) {
    std::fprintf(stderr, "cannot open %s\n", path);
    return 255;
  }
  std::fseek(f, 0, SEEK_END);
  const long len = std::ftell(f);
  std::fseek(f, 0, SEEK_SET);
  if (len < 0 || len % 8) {
    std::fprintf(stderr, "file size error: cannot be divided by 8\n");
    return 255;
  }
  const uint64_t edge_count = (uint64_t)len / 8;
  std::vector<uint64_t> edges(edge_count ? edge_count : 1);
  if (std::fread(edges.data(), 8, edge_count, f) != edge_count) {
    std::fprintf(stderr, "short read on %s\n", path);
    return 255;
  }
  std::fclose(f);
  const uint64_t mem = tc_tricore_init_gpu();
  if (debug) std::fprintf(stderr, "GPU memory: %fMB\n", (double)mem / 1024 / 1024);
  uint32_t node_count = 0;
  for (uint64_t i = 0; i < edge_count; ++i) {
    const uint32_t a = (uint32_t)(edges[i] >> 32), b = (uint32_t)edges[i];
    node_count = std::max(node_count, std::max(a, b));
  }
  node_count++;
  std::vector<uint32_t> deg(node_count);
  tc_tricore_cal_degree(edges.data(), edge_count, deg.data(), node_count);
  tc_tricore_redirect_edges(edges.data(), edge_count, deg.data(), node_count);
  const uint32_t part_num = tc_tricore_cal_part_num(edges.data(), edge_count, node_count);
  if (debug) std::fprintf(stderr, "split into part_num: %u\n", part_num);
  std::vector<uint32_t> node_split(part_num + 1);
  for (uint32_t i = 0; i <= part_num; ++i)
    node_split[i] = i == part_num ? node_count : node_count / part_num * i;
  std::vector<uint32_t*> node_index(part_num);
  std::vector<uint64_t> adj_count(part_num + 1);
  tc_tricore_split(edges.data(), edge_count, part_num, node_index.data(), node_split.data(),
                   adj_count.data());
  for (uint32_t i = 0; i < part_num; ++i)
    if (!node_index[i]) cli::die("tc_tricore_split", TC_ERR_STATE);
  const uint64_t total =
      tc_tricore_tricount_gpu(part_num, reinterpret_cast<uint32_t*>(edges.data()),
                              adj_count.data(), node_split.data(), node_index.data());
  for (uint32_t i = 0; i < part_num; ++i) std::free(node_index[i]);
  std::printf("There are %lu triangles in the input graph.\n", (unsigned long)total);
  return 0;
}
