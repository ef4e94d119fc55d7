// Device-wide primitives (sort, unique, scan, reduce, select) on rocPRIM.
// All run on the context stream; temp storage is the context's `temp` buffer.
#include <cstring>

#include <rocprim/rocprim.hpp>

#include "internal.h"

namespace tcx {

namespace {
template <class F>
int run_prim(tc_ctx* c, F&& f) {
  size_t bytes = 0;
  TCX_HIP(f(nullptr, bytes));
  TCX_TRY(c->temp.reserve(bytes));
  TCX_HIP(f(c->temp.p, bytes));
  return TC_OK;
}
}  // namespace

int sort_u64(tc_ctx* c, uint64_t* keys, uint64_t* tmp, uint64_t n, int end_bit,
             uint64_t** sorted) {
  *sorted = keys;
  if (n <= 1) return TC_OK;
  rocprim::double_buffer<uint64_t> db(keys, tmp);
  TCX_TRY(run_prim(c, [&](void* t, size_t& b) {
    return rocprim::radix_sort_keys(t, b, db, (size_t)n, 0u, (unsigned)end_bit, c->stream);
  }));
  *sorted = db.current();
  return TC_OK;
}

int sort_pairs_u64_u32(tc_ctx* c, uint64_t* keys, uint64_t* ktmp, uint32_t* vals,
                       uint32_t* vtmp, uint64_t n, int end_bit, uint64_t** skeys,
                       uint32_t** svals) {
  *skeys = keys;
  *svals = vals;
  if (n <= 1) return TC_OK;
  rocprim::double_buffer<uint64_t> kb(keys, ktmp);
  rocprim::double_buffer<uint32_t> vb(vals, vtmp);
  TCX_TRY(run_prim(c, [&](void* t, size_t& b) {
    return rocprim::radix_sort_pairs(t, b, kb, vb, (size_t)n, 0u, (unsigned)end_bit, c->stream);
  }));
  *skeys = kb.current();
  *svals = vb.current();
  return TC_OK;
}

int sort_pairs_u32_u32(tc_ctx* c, uint32_t* keys, uint32_t* ktmp, uint32_t* vals,
                       uint32_t* vtmp, uint64_t n, int end_bit, uint32_t** skeys,
                       uint32_t** svals) {
  *skeys = keys;
  *svals = vals;
  if (n <= 1) return TC_OK;
  rocprim::double_buffer<uint32_t> kb(keys, ktmp);
  rocprim::double_buffer<uint32_t> vb(vals, vtmp);
  TCX_TRY(run_prim(c, [&](void* t, size_t& b) {
    return rocprim::radix_sort_pairs(t, b, kb, vb, (size_t)n, 0u, (unsigned)end_bit, c->stream);
  }));
  *skeys = kb.current();
  *svals = vb.current();
  return TC_OK;
}

int unique_u64(tc_ctx* c, const uint64_t* in, uint64_t* out, uint64_t n,
               uint64_t* n_out) {
  if (n == 0) { *n_out = 0; return TC_OK; }
  TCX_TRY(c->prim_out.reserve(64));
  uint64_t* d_cnt = c->prim_out.as<uint64_t>();
  TCX_TRY(run_prim(c, [&](void* t, size_t& b) {
    return rocprim::unique(t, b, in, out, d_cnt, (size_t)n, rocprim::equal_to<uint64_t>(),
                           c->stream);
  }));
  TCX_HIP(hipMemcpyAsync(n_out, d_cnt, sizeof(uint64_t), hipMemcpyDeviceToHost, c->stream));
  TCX_HIP(hipStreamSynchronize(c->stream));
  return TC_OK;
}

int exclusive_scan_u32(tc_ctx* c, const uint32_t* in, uint32_t* out, uint64_t n) {
  if (n == 0) return TC_OK;
  return run_prim(c, [&](void* t, size_t& b) {
    return rocprim::exclusive_scan(t, b, in, out, 0u, (size_t)n, rocprim::plus<uint32_t>(),
                                   c->stream);
  });
}

int exclusive_scan_u64(tc_ctx* c, const uint64_t* in, uint64_t* out, uint64_t n) {
  if (n == 0) return TC_OK;
  return run_prim(c, [&](void* t, size_t& b) {
    return rocprim::exclusive_scan(t, b, in, out, (uint64_t)0, (size_t)n,
                                   rocprim::plus<uint64_t>(), c->stream);
  });
}

int inclusive_scan_i64(tc_ctx* c, const int64_t* in, int64_t* out, uint64_t n) {
  if (n == 0) return TC_OK;
  return run_prim(c, [&](void* t, size_t& b) {
    return rocprim::inclusive_scan(t, b, in, out, (size_t)n, rocprim::plus<int64_t>(),
                                   c->stream);
  });
}

int max_u32(tc_ctx* c, const uint32_t* in, uint64_t n, uint32_t* result) {
  *result = 0;
  if (n == 0) return TC_OK;
  TCX_TRY(c->prim_out.reserve(64));
  uint32_t* d = c->prim_out.as<uint32_t>();
  TCX_TRY(run_prim(c, [&](void* t, size_t& b) {
    return rocprim::reduce(t, b, in, d, 0u, (size_t)n, rocprim::maximum<uint32_t>(), c->stream);
  }));
  TCX_HIP(hipMemcpyAsync(result, d, sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
  TCX_HIP(hipStreamSynchronize(c->stream));
  return TC_OK;
}

int min_i32(tc_ctx* c, const int32_t* in, uint64_t n, int32_t* result) {
  *result = INT32_MAX;
  if (n == 0) return TC_OK;
  TCX_TRY(c->prim_out.reserve(64));
  int32_t* d = c->prim_out.as<int32_t>();
  TCX_TRY(run_prim(c, [&](void* t, size_t& b) {
    return rocprim::reduce(t, b, in, d, (int32_t)INT32_MAX, (size_t)n,
                           rocprim::minimum<int32_t>(), c->stream);
  }));
  TCX_HIP(hipMemcpyAsync(result, d, sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
  TCX_HIP(hipStreamSynchronize(c->stream));
  return TC_OK;
}

int sum_u64(tc_ctx* c, const uint64_t* in, uint64_t n, uint64_t* result) {
  *result = 0;
  if (n == 0) return TC_OK;
  TCX_TRY(c->prim_out.reserve(64));
  uint64_t* d = c->prim_out.as<uint64_t>();
  TCX_TRY(run_prim(c, [&](void* t, size_t& b) {
    return rocprim::reduce(t, b, in, d, (uint64_t)0, (size_t)n, rocprim::plus<uint64_t>(),
                           c->stream);
  }));
  TCX_HIP(hipMemcpyAsync(result, d, sizeof(uint64_t), hipMemcpyDeviceToHost, c->stream));
  TCX_HIP(hipStreamSynchronize(c->stream));
  return TC_OK;
}

int select_flagged_u32(tc_ctx* c, const uint32_t* in, const uint8_t* flags,
                       uint32_t* out, uint64_t n, uint64_t* n_out) {
  if (n == 0) { *n_out = 0; return TC_OK; }
  TCX_TRY(c->prim_out.reserve(64));
  uint64_t* d_cnt = c->prim_out.as<uint64_t>();
  TCX_TRY(run_prim(c, [&](void* t, size_t& b) {
    return rocprim::select(t, b, in, flags, out, d_cnt, (size_t)n, c->stream);
  }));
  TCX_HIP(hipMemcpyAsync(n_out, d_cnt, sizeof(uint64_t), hipMemcpyDeviceToHost, c->stream));
  TCX_HIP(hipStreamSynchronize(c->stream));
  return TC_OK;
}

int select_flagged_u64(tc_ctx* c, const uint64_t* in, const uint8_t* flags,
                       uint64_t* out, uint64_t n, uint64_t* n_out) {
  if (n == 0) { *n_out = 0; return TC_OK; }
  TCX_TRY(c->prim_out.reserve(64));
  uint64_t* d_cnt = c->prim_out.as<uint64_t>();
  TCX_TRY(run_prim(c, [&](void* t, size_t& b) {
    return rocprim::select(t, b, in, flags, out, d_cnt, (size_t)n, c->stream);
  }));
  TCX_HIP(hipMemcpyAsync(n_out, d_cnt, sizeof(uint64_t), hipMemcpyDeviceToHost, c->stream));
  TCX_HIP(hipStreamSynchronize(c->stream));
  return TC_OK;
}

int select_indices(tc_ctx* c, const uint8_t* flags, uint64_t n, uint32_t* out,
                   uint64_t* n_out) {
  if (n == 0) { *n_out = 0; return TC_OK; }
  TCX_TRY(c->prim_out.reserve(64));
  uint64_t* d_cnt = c->prim_out.as<uint64_t>();
  rocprim::counting_iterator<uint32_t> it(0u);
  TCX_TRY(run_prim(c, [&](void* t, size_t& b) {
    return rocprim::select(t, b, it, flags, out, d_cnt, (size_t)n, c->stream);
  }));
  TCX_HIP(hipMemcpyAsync(n_out, d_cnt, sizeof(uint64_t), hipMemcpyDeviceToHost, c->stream));
  TCX_HIP(hipStreamSynchronize(c->stream));
  return TC_OK;
}

}  // namespace tcx
