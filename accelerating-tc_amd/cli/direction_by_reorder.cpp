// directionByReorder <in.bin> <out.bin>  (tools/directionByReorder.cpp):
// ids sorted by (degree, id), the highest degree becomes id 0.
#include "relabel_common.h"

int main(int argc, const char* argv[]) {
  if (argc < 3) {
    std::printf("Usage: ./directionByReorder in.bin out.bin\n");
    return 1;
  }
  cli::Records r = cli::load(argv[1]);
  uint32_t mx = 0;
  for (uint64_t i = 0; i < r.m; ++i) mx = std::max(mx, std::max(r.src[i], r.dst[i]));
  std::printf("the nodeNum is %u\n", mx + 1);
  cli::relabel_records(r, TC_RELABEL_DEGREE);
  cli::check(tc_write_edges(argv[2], r.src.data(), r.dst.data(), r.m), "tc_write_edges");
  return 0;
}
