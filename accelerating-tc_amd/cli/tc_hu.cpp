// Hu-method driver: TC -f <edges.bin>  (Hu/tc.cu:202-206,239-337).
// Degree orientation with degree <= 1 pruning, then the Hu kernel.
#include "common.h"

int main(int argc, const char* argv[]) {
  if (argc != 3) {
    std::printf("Usage: ./TC -f inputFileName\n");
    return 0;
  }
  return cli::count_file(argv[2], TC_ORIENT_DEGREE, TC_FLAG_PRUNE_DEG1, TC_KERNEL_HU);
}
