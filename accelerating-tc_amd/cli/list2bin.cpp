// list2bin <in.txt> <out.bin>  (tools/list2bin.cpp:7-24).
#include "common.h"

int main(int argc, const char* argv[]) {
  if (argc < 3) {
    std::printf("Usage: ./list2bin in.txt out.bin\n");
    return 1;
  }
  std::printf("sizeof Edge is 8\n");
  uint64_t m = 0;
  cli::check(tc_list2bin(argv[1], argv[2], &m), "tc_list2bin");
  return 0;
}
