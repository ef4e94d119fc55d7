// Bisson-method driver: TC -f <edges.bin> <chooseIndex>  (Bisson/tc.cu:105-143).
// chooseIndex 0 orients by ID (Bisson/preTC.cu:135-144); any other value by
// degree (the reference relabels by degree, Bisson/preTC.cu:145-161, which
// gives another acyclic order and the same count).
#include "common.h"

int main(int argc, const char* argv[]) {
  if (argc != 4) {
    std::printf("Usage: ./TC -f inputFileName chooseindex\n");
    return 0;
  }
  const int orient = std::atoi(argv[3]) == 0 ? TC_ORIENT_ID : TC_ORIENT_DEGREE;
  return cli::count_file(argv[2], orient, 0, TC_KERNEL_BISSON);
}
