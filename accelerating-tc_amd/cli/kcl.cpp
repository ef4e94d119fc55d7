// k-clique driver: k-clique-listing-gpu -symmetricGraph -simpleGraph <g.gr> -k=K
// (k-clique/README.md:28, k-clique/kcl_gpu.cpp:12-31).  Flags are accepted in
// any order; -ft=gr is implied by a ".gr" file name, other files are read as
// list2bin records.
#include <cstring>

#include "common.h"

int main(int argc, char** argv) {
  const char* path = nullptr;
  unsigned k = 3;
  bool sym = false, simple = false;
  for (int i = 1; i < argc; ++i) {
    const std::string a = argv[i];
    if (a == "-symmetricGraph") sym = true;
    else if (a == "-simpleGraph") simple = true;
    else if (a.rfind("-k=", 0) == 0) k = (unsigned)std::atoi(a.c_str() + 3);
    else if (a.rfind("-ft=", 0) == 0 || a.rfind("-t=", 0) == 0) continue;
    else if (a[0] != '-') path = argv[i];
  }
  if (!path) {
    std::fprintf(stderr, "usage: %s -symmetricGraph -simpleGraph <graph.gr> -k=K\n", argv[0]);
    return 1;
  }
  if (!simple || !sym) {
    std::fprintf(stderr,
                 "This application requires a symmetric simple graph input  which is symmetric "
                 "and has no multiple edges or self-loops; please use both -symmetricGraph and "
                 "-simpleGraph flag  to indicate the input is a symmetric simple graph\n");
    return 1;
  }
  uint64_t total = 0;
  const double t0 = cli::now_s();
  cli::check(kcl_gpu_solver(path, k, &total, 1), "kcl_gpu_solver");
  std::printf("runtime = %g ms.\n", 1e3 * (cli::now_s() - t0));
  std::printf("\n\ttotal_num_cliques = %llu\n\n", (unsigned long long)total);
  return 0;
}
