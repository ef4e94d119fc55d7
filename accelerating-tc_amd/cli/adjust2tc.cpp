// adjust2tc <in.txt> <out.txt>  (tools/adjust2tc.cpp): "src dst" lines, ids
// renumbered by first appearance as a source, then the rest by id.
#include <cinttypes>

#include "relabel_common.h"

int main(int argc, const char* argv[]) {
  if (argc < 3) {
    std::printf("Usage: ./adjust2tc in.txt out.txt\n");
    return 1;
  }
  FILE* f = std::fopen(argv[1], "r");
  if (!f) {
    std::fprintf(stderr, "cannot open %s\n", argv[1]);
    return 1;
  }
  cli::Records r;
  long long a, b;
  uint32_t mx = 0;
  while (std::fscanf(f, "%lld %lld", &a, &b) == 2) {
    if (a < 0 || b < 0 || a >= 0x7FFFFFFF || b >= 0x7FFFFFFF) {
      std::fprintf(stderr, "id out of range on line %llu\n", (unsigned long long)r.m + 1);
      return 1;
    }
    r.src.push_back((uint32_t)a);
    r.dst.push_back((uint32_t)b);
    mx = std::max(mx, (uint32_t)std::max(a, b));
    ++r.m;
  }
  std::fclose(f);
  std::printf("the maxvalue is %u\n", mx);
  if (r.m == 0) {
    r.src.resize(1);
    r.dst.resize(1);
  }
  cli::relabel_records(r, TC_RELABEL_FIRST_SRC);
  FILE* o = std::fopen(argv[2], "w");
  if (!o) {
    std::fprintf(stderr, "cannot open %s\n", argv[2]);
    return 1;
  }
  for (uint64_t i = 0; i < r.m; ++i) std::fprintf(o, "%u %u\n", r.src[i], r.dst[i]);
  std::fclose(o);
  return 0;
}
