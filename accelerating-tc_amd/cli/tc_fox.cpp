// Fox-method driver: TC -f <edges.bin> <rearrange 0|1>  (Fox/tc.cu:184-189,345).
// ID orientation (Fox/preTC.cu:92-101), degree-binned Fox kernel; a nonzero
// rearrange argument applies rearrangeEdge to every bin (Fox/tc.cu:127-183).
#include "common.h"

int main(int argc, const char* argv[]) {
  if (argc != 4) {
    std::printf("Usage: ./TC -f inputFileName\n");
    return 0;
  }
  return cli::count_file(argv[2], TC_ORIENT_ID, 0, TC_KERNEL_FOX, std::atoi(argv[3]) ? 1u : 0u);
}
