// Shared helpers of the command-line drivers: file loading, one-line error
// exits and wall-clock timers.  The drivers mirror the reference programs'
// argv forms and print the same result line.
#pragma once

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "tc_amd.h"

namespace cli {

inline double now_s() {
  using namespace std::chrono;
  return duration<double>(steady_clock::now().time_since_epoch()).count();
}

[[noreturn]] inline void die(const char* what, int st) {
  std::fprintf(stderr, "%s failed: %s (%d): %s\n", what, tc_status_str(st), st, tc_last_error());
  std::exit(1);
}

inline void check(int st, const char* what) {
  if (st != TC_OK) die(what, st);
}

struct Records {
  std::vector<uint32_t> src, dst;
  uint64_t m = 0;
};

inline Records load(const char* path) {
  Records r;
  check(tc_file_edge_count(path, &r.m), "tc_file_edge_count");
  r.src.resize(r.m ? r.m : 1);
  r.dst.resize(r.m ? r.m : 1);
  check(tc_read_edges(path, r.src.data(), r.dst.data(), r.m, &r.m), "tc_read_edges");
  return r;
}

// Loads the file, orients it and counts with one method; prints the
// reference's lines (node/edge numbers, the triangle line, kernel time).
inline int count_file(const char* path, int orient, uint32_t flags, int kernel,
                      uint32_t fox_rearrange = 0) {
  const double t0 = now_s();
  Records r = load(path);
  std::printf("the size of input file is %llu Byte. \n", (unsigned long long)(r.m * 8));
  tc_ctx* ctx = nullptr;
  check(tc_ctx_create(0, &ctx), "tc_ctx_create");
  tc_graph* g = nullptr;
  check(tc_graph_create(ctx, r.src.data(), r.dst.data(), r.m, 0, &g), "tc_graph_create");
  check(tc_graph_orient(ctx, g, orient, flags), "tc_graph_orient");
  tc_graph_info info;
  check(tc_graph_info_get(g, &info), "tc_graph_info_get");
  const double t1 = now_s();
  std::printf("preWork cost %g ms.\n", 1e3 * (t1 - t0));
  std::printf("the node num is %u, and the edgeNum is %llu\n", info.n,
              (unsigned long long)info.m_oriented);
  tc_sched sched = {};
  sched.kernel = kernel;
  sched.fox_rearrange = fox_rearrange;
  uint64_t count = 0;
  tc_timing tm;
  check(tc_count(ctx, g, &sched, &count, &tm), "tc_count");
  const double t2 = now_s();
  std::printf("There are %llu triangles in the input graph.\n", (unsigned long long)count);
  std::printf("kernel use %g s.\n", tm.ms_count * 1e-3);
  std::printf("Total use time %g s.\n", t2 - t0);
  tc_graph_destroy(g);
  tc_ctx_destroy(ctx);
  return 0;
}

}  // namespace cli
