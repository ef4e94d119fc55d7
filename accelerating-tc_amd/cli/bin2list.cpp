// bin2list <in.bin> <out.txt>  (tools/bin2list.cpp).
#include "common.h"

int main(int argc, const char* argv[]) {
  if (argc < 3) {
    std::printf("Usage: ./bin2list in.bin out.txt\n");
    return 1;
  }
  uint64_t m = 0;
  cli::check(tc_bin2list(argv[1], argv[2], &m), "tc_bin2list");
  return 0;
}
