// serialize <in.bin> <out.bin>  (tools/serialize.cpp): ids that occur keep
// their order, ids that never occur are squeezed out.
#include "relabel_common.h"

int main(int argc, const char* argv[]) {
  if (argc < 3) {
    std::printf("Usage: ./serialize in.bin out.bin\n");
    return 1;
  }
  cli::Records r = cli::load(argv[1]);
  uint32_t mx = 0;
  for (uint64_t i = 0; i < r.m; ++i) mx = std::max(mx, std::max(r.src[i], r.dst[i]));
  std::printf("the maxvalue is %u\n", mx);
  cli::relabel_records(r, TC_RELABEL_COMPACT);
  cli::check(tc_write_edges(argv[2], r.src.data(), r.dst.data(), r.m), "tc_write_edges");
  return 0;
}
