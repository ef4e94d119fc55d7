// Shared body of the relabelling drivers: records in, tc_graph_relabel on the
// GPU, records out in the same order.
#pragma once

#include "common.h"

namespace cli {

inline void relabel_records(Records& r, int mode) {
  tc_ctx* ctx = nullptr;
  check(tc_ctx_create(0, &ctx), "tc_ctx_create");
  tc_graph* g = nullptr;
  check(tc_graph_create(ctx, r.src.data(), r.dst.data(), r.m, 0, &g), "tc_graph_create");
  check(tc_graph_relabel(ctx, g, mode, nullptr), "tc_graph_relabel");
  if (r.m) check(tc_graph_get_edges(g, r.src.data(), r.dst.data()), "tc_graph_get_edges");
  tc_graph_destroy(g);
  tc_ctx_destroy(ctx);
}

}  // namespace cli
