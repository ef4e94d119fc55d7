// Internal declarations shared by the engine's translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <memory>
#include <string>
#include <utility>
#include <vector>

#include "../../include/tc_amd.h"

namespace tcx {

// Most oriented edges a CSR may hold: 32-bit offsets, with the last list's
// 64-entry chunks (up to 512 entries past a list's first) still below 2^32.
constexpr uint64_t kMaxCsrEntries = 0xFFFFFFFFull - 512;

// ---- error plumbing ------------------------------------------------------
void set_error(const char* fmt, ...);
int fail(int status, const char* fmt, ...);

#define TCX_HIP(call)                                                        \
  do {                                                                       \
    hipError_t _e = (call);                                                  \
    if (_e != hipSuccess)                                                    \
      return ::tcx::fail(_e == hipErrorOutOfMemory ? TC_ERR_NOMEM : TC_ERR_HIP, \
                         "%s failed: %s (%s:%d)", #call, hipGetErrorString(_e), \
                         __FILE__, __LINE__);                                \
  } while (0)

#define TCX_TRY(expr)          \
  do {                         \
    int _s = (expr);           \
    if (_s != TC_OK) return _s; \
  } while (0)

// ---- device buffer ------------------------------------------------------
// Owns one hipMalloc'd region; grows on demand (contents are not kept).
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() { release(); }
  int reserve(size_t want);
  void release();
  template <class T> T* as() const { return static_cast<T*>(p); }
};

inline void swap(DevBuf& a, DevBuf& b) {
  std::swap(a.p, b.p);
  std::swap(a.bytes, b.bytes);
}

}  // namespace tcx

struct tc_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  hipEvent_t lev[17] = {};  // per-launch timing marks (created on first use)
  uint64_t total_mem = 0;
  int num_cus = 256;
  int wall_clock_khz = 100000;  // device wall_clock64() rate
  tcx::DevBuf temp;     // rocprim temp storage
  tcx::DevBuf scratch;  // small per-call scratch (counters, bitmaps)
  tcx::DevBuf accum;    // uint64 accumulators
  tcx::DevBuf prim_out; // 64-byte result slot of the device-wide primitives
  tcx::DevBuf bitmaps;  // Bisson per-workgroup bitmaps, all-zero between calls
  uint64_t bitmap_words = 0;
};

namespace tcx {
// Per-graph schedules, rebuilt when the oriented CSR changes.
struct SchedCache {
  uint64_t version = ~0ull;
  uint32_t hu_nblocks = 0;
  DevBuf hu_start;                 // int32[nblocks+1]
  bool auto_ok = false;
  uint32_t auto_wmax = 0, auto_hmin = 0;  // thresholds the AUTO lists were built for
  DevBuf auto_small;               // vertices with 2 <= d+ <= 64, id order
  DevBuf auto_large;               // d+ in (64, 2048], by descending work
  DevBuf auto_huge;                // d+ > 2048, by descending work
  uint64_t n_small = 0, n_large = 0, n_huge = 0;
  // AUTO over work lists: W(x) CSR (wl_rp, wl_col) and the task lists
  bool wl_ok = false;
  uint32_t wl_wmax = 0, wl_hmin = 0, wl_lmax = 0;
  DevBuf wl_rp, wl_col;            // uint32[n+1], uint32[entries]
  DevBuf wl_se;                    // uint2[entries]: (rp[y], rp[y+1]) of wl_col's y
  DevBuf wl_small, wl_large, wl_huge, wl_lane;
  DevBuf wl_ell;                   // uint32[8 n]: out-lists of <= 8 entries, kEmpty-padded rows
  uint64_t nw_small = 0, nw_large = 0, nw_huge = 0, nw_lane = 0;
  uint64_t wl_entries = 0;         // edges kept in the work lists
  double wl_bytes = 0;             // bytes one count streams (see count.hip)
  bool rows_ok = false;
  DevBuf rows;                     // rows with d+ >= 2
  uint64_t nrows = 0;
  uint32_t fox_lanes = 0;
  bool fox_rearr = false;
  DevBuf fox_eids;                 // edge ids grouped by lanes-per-edge bin
  std::vector<uint64_t> fox_off;   // bin offsets (bins+1)
  float ms_prep = 0;
  // learned schedule (per model id)
  uint64_t learned_version = ~0ull, learned_model = ~0ull;
  DevBuf l_blk, l_wav, l_huge, l_lan;
  uint64_t nl_blk = 0, nl_wav = 0, nl_huge = 0, nl_lan = 0;
  uint32_t learned_mask = 0;  // cells whose tasks run narrow
  // learned k-clique plan (per k and model id)
  uint64_t kcl_version = ~0ull, kcl_model = ~0ull;
  uint32_t kcl_k = 0;
  DevBuf kcl_path, kcl_key;
  // degeneracy-oriented copy of the graph for k-clique counting when the
  // caller's orientation leaves roots wider than the LDS bit matrix
  uint64_t kcl_degen_version = ~0ull;
  std::shared_ptr<tc_graph> kcl_degen;
};
}  // namespace tcx

struct tc_graph {
  tc_ctx* ctx = nullptr;
  uint32_t n = 0;           // rows of the oriented CSR (after pruning, if any)
  uint32_t n_raw = 0;       // every raw record id is below this
  uint64_t m_raw = 0;
  bool csr_only = false;    // built from a CSR (stage file, TriCore parts): no raw records
  tcx::DevBuf src, dst;     // raw records (uint32 each)
  // oriented CSR
  int orient_mode = -1;
  uint64_t m_or = 0;
  uint32_t max_out = 0;
  tcx::DevBuf rowptr;       // uint32[n+1]
  tcx::DevBuf col;          // uint32[m_or]
  tcx::DevBuf esrc;         // uint32[m_or]  source of each CSR edge
  tcx::DevBuf vkey;         // int32[n] ADIR priority / KCORE core number
  uint64_t sum_out_sq = 0, sum_in_out = 0;
  // cached schedules (valid for the current CSR)
  uint64_t sched_version = 0;
  tcx::SchedCache sc;
};

namespace tcx {

// ---- primitives (prims.hip) ---------------------------------------------
int sort_u64(tc_ctx* c, uint64_t* keys, uint64_t* tmp, uint64_t n, int end_bit,
             uint64_t** sorted);
int sort_pairs_u64_u32(tc_ctx* c, uint64_t* keys, uint64_t* ktmp, uint32_t* vals,
                       uint32_t* vtmp, uint64_t n, int end_bit, uint64_t** skeys,
                       uint32_t** svals);
int sort_pairs_u32_u32(tc_ctx* c, uint32_t* keys, uint32_t* ktmp, uint32_t* vals,
                       uint32_t* vtmp, uint64_t n, int end_bit, uint32_t** skeys,
                       uint32_t** svals);
int unique_u64(tc_ctx* c, const uint64_t* in, uint64_t* out, uint64_t n,
               uint64_t* n_out);
int exclusive_scan_u32(tc_ctx* c, const uint32_t* in, uint32_t* out, uint64_t n);
int exclusive_scan_u64(tc_ctx* c, const uint64_t* in, uint64_t* out, uint64_t n);
int inclusive_scan_i64(tc_ctx* c, const int64_t* in, int64_t* out, uint64_t n);
int max_u32(tc_ctx* c, const uint32_t* in, uint64_t n, uint32_t* result);
int sum_u64(tc_ctx* c, const uint64_t* in, uint64_t n, uint64_t* result);
int min_i32(tc_ctx* c, const int32_t* in, uint64_t n, int32_t* result);
int select_flagged_u32(tc_ctx* c, const uint32_t* in, const uint8_t* flags,
                       uint32_t* out, uint64_t n, uint64_t* n_out);
int select_indices(tc_ctx* c, const uint8_t* flags, uint64_t n, uint32_t* out,
                   uint64_t* n_out);
int select_flagged_u64(tc_ctx* c, const uint64_t* in, const uint8_t* flags,
                       uint64_t* out, uint64_t n, uint64_t* n_out);

// ---- graph helpers (graph.hip) -------------------------------------------
int build_csr_from_keys(tc_ctx* c, tc_graph* g, uint64_t* keys, uint64_t nkeys,
                        int end_bit);
int compute_csr_stats(tc_ctx* c, tc_graph* g);
int build_sym(tc_ctx* c, tc_graph* g, DevBuf& rowptr, DevBuf& col, uint64_t* m2);
// A new graph holding a device copy of g's raw records (not oriented); its
// vertex count is the records' largest id + 1.  A graph without raw records
// (csr_only) gives its oriented CSR's edges instead.
int graph_clone_records(tc_ctx* c, const tc_graph* g, tc_graph** out);

// ---- work lists of the AUTO engine (count.hip) ---------------------------
// Builds (or keeps) the default work-list schedule of g: sc.wl_rp / wl_col
// and the fixed rule's task lists.
int worklists(tc_ctx* c, tc_graph* g);

// ---- cost model and per-vertex counts (mlp.hip, count.hip) ---------------
// Predicted log2 costs Y[n][2] for device features X[n][8]; fails when the
// model belongs to another context.
int model_predict(tc_ctx* c, tc_model* m, const float* dX, uint32_t n, float* dY);
// Identity of the model's current parameters (changes with every update).
uint64_t model_id(const tc_model* m);
// tri[u] = |{v -> w : v, w in N+(u)}|, the edges of the sub-graph of N+(u)
// (uint64 per vertex, device array of n entries).
int per_vertex_counts(tc_ctx* c, tc_graph* g, unsigned long long* tri);

// ---- allocation-free counting (noalloc.hip) --------------------------------
// Looks up the device address of the static scratch array on `device` (the
// device must be current); called once per device by tc_ctx_create.
int noalloc_setup(int device);

// ---- A-order (aorder.cpp, host heap) -------------------------------------
int aorder_perm_from_outdeg(const uint32_t* rtd, uint32_t n, uint32_t* perm,
                            const tc_bw_model& model);
void bw_model_reference(tc_bw_model* m);

// ---- host parallel-for ----------------------------------------------------
void parallel_for(uint64_t n, void (*fn)(uint64_t begin, uint64_t end, void* arg),
                  void* arg, unsigned max_threads = 0);

inline int bit_width_u32(uint32_t x) {
  int b = 0;
  while (x) { ++b; x >>= 1; }
  return b;
}

constexpr int kWave = 64;

}  // namespace tcx
