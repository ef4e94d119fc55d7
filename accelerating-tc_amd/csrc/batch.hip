// Batched counting of many small graphs in one persistent launch.
//
// The batch is the disjoint union of its graphs, degree-oriented in one CSR
// (each graph keeps its own relative order, so the union's orientation is
// every graph's own).  The AUTO engine's work-list tasks count it (an edge
// never leaves its graph, so neither does a work list); block and wave
// tasks' counts are folded into per-graph uint64 accumulators.
// The reset + count sequence is captured once in a hipGraph and replayed.
#include <algorithm>
#include <atomic>
#include <cstring>
#include <vector>

#include "engine.h"
#include "internal.h"

namespace tcx {
namespace {

using namespace dev;

__global__ void k_batch_reset(uint32_t* __restrict__ cursor, uint32_t ncur,
                              unsigned long long* __restrict__ counts, uint32_t ngraphs) {
  for (uint64_t i = gtid(); i < (uint64_t)ncur + ngraphs; i += gstride()) {
    if (i < ncur) cursor[i] = 0;
    else counts[i - ncur] = 0;
  }
}

__device__ inline void flush_graph(unsigned long long& acc, uint32_t g,
                                   unsigned long long* counts) {
  const unsigned long long w = wave_sum(acc);
  if ((threadIdx.x & 63) == 0 && w) atomicAdd(counts + g, w);
  acc = 0;
}

// The push form: task u streams N+(v) for v in N+(u), each entry tested
// against the hashed bitmap of N+(u) and the survivors searched (wrp and
// wse unused; they keep the work-list kernel's signature).
template <int CAP, uint32_t FMIN>
__global__ __launch_bounds__(TB) void k_batch(const uint32_t* __restrict__ rp,
                                              const uint32_t* __restrict__ col,
                                              const uint32_t* __restrict__ wrp,
                                              const uint2* __restrict__ wse,
                                              const uint32_t* __restrict__ gid,
                                              const uint32_t* __restrict__ blk, uint32_t nblk,
                                              const uint32_t* __restrict__ wav, uint32_t nwav,
                                              uint32_t wave_chunk, uint32_t* cursor,
                                              unsigned long long* counts) {
  __shared__ BlockLds<CAP> S;
  const uint32_t tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  clear_bitmap(S);
  for (bool first = true;; first = false) {
    if (tid == 0) S.t = claim(cursor, nblk, 1u, first);
    __syncthreads();
    const uint32_t t = S.t;
    if (t >= nblk) break;
    const uint32_t u = blk[t];
    unsigned long long c = block_vertex<CAP, 1>(rp, col, u, S);
    flush_graph(c, gid[u], counts);
  }
  const WaveSlot W = wave_slot(S, wv);
  unsigned long long acc = 0;
  uint32_t cur = 0xFFFFFFFFu;
  for (bool first = true;; first = false) {
    uint32_t i0 = 0;
    if (lane == 0) i0 = claim(cursor + 1, nwav, wave_chunk, first);
    i0 = __shfl(i0, 0, 64);
    if (i0 >= nwav) break;
    const uint32_t i1 = min(nwav, i0 + wave_chunk);
    for (uint32_t i = i0; i < i1; ++i) {
      const uint32_t u = wav[i];
      const uint32_t g = gid[u];
      if (g != cur) {
        if (cur != 0xFFFFFFFFu) flush_graph(acc, cur, counts);
        cur = g;
      }
      acc += wave_vertex<FMIN, 1>(rp, col, u, lane, W);
    }
  }
  if (cur != 0xFFFFFFFFu) flush_graph(acc, cur, counts);
}

// The work-list form (AUTO's default schedule; wse holds each entry's list
// bounds): each task builds the rank-indexed table of N+(u) and decides each
// streamed entry by one lookup, as k_auto_wl does; a block task whose N+(u)
// exceeds CAP searches it in global memory.
template <int CAP>
__global__ __launch_bounds__(TB) __attribute__((amdgpu_waves_per_eu(8, 8))) void k_batch_wl(const uint32_t* __restrict__ rp,
                                                 const uint32_t* __restrict__ col,
                                                 const uint32_t* __restrict__ wrp,
                                                 const uint2* __restrict__ wse,
                                                 const uint32_t* __restrict__ gid,
                                                 const uint32_t* __restrict__ blk, uint32_t nblk,
                                                 const uint32_t* __restrict__ wav, uint32_t nwav,
                                                 uint32_t wave_chunk, uint32_t* cursor,
                                                 unsigned long long* counts) {
  __shared__ RankLds<CAP> S;
  const uint32_t tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  for (uint32_t i = tid; i < RankLds<CAP>::kRegion; i += TB) S.region[i] = 0;
  __syncthreads();
  for (bool first = true;; first = false) {
    if (tid == 0) S.t = claim(cursor, nblk, 1u, first);
    __syncthreads();
    const uint32_t t = S.t;
    if (t >= nblk) break;
    const uint32_t u = blk[t], wb = wrp[u], wn = wrp[u + 1] - wb;
    unsigned long long c;
    if (uniform(rp[u + 1] - rp[u]) <= (uint32_t)CAP) {
      c = block_rank<CAP>(rp, col, u, S, wse + wb, wn);  // ends with a barrier
    } else {
      c = block_global(rp, col, u, wse + wb, wn);
      __syncthreads();  // every wave has read S.t before it is claimed again
    }
    flush_graph(c, gid[u], counts);
  }
  // the wave slots overlap the block view's table: start them from zero
  for (uint32_t i = tid; i < RankLds<CAP>::kRegion; i += TB) S.region[i] = 0;
  __syncthreads();
  const RankSlot W = rank_wave(S, wv);
  unsigned long long acc = 0;
  uint32_t cur = 0xFFFFFFFFu;
  for (bool first = true;; first = false) {
    uint32_t i0 = 0;
    if (lane == 0) i0 = claim(cursor + 1, nwav, wave_chunk, first);
    i0 = __shfl(i0, 0, 64);
    if (i0 >= nwav) break;
    const uint32_t i1 = min(nwav, i0 + wave_chunk);
    for (uint32_t i = i0; i < i1; ++i) {
      const uint32_t u = wav[i];
      const uint32_t g = gid[u];
      if (g != cur) {
        if (cur != 0xFFFFFFFFu) flush_graph(acc, cur, counts);
        cur = g;
      }
      const uint32_t wb = wrp[u];
      acc += wave_rank<CAP>(rp, col, u, lane, W, wse + wb, wrp[u + 1] - wb);
    }
  }
  if (cur != 0xFFFFFFFFu) flush_graph(acc, cur, counts);
}

// tasks with d+ in [lo, hi] and d+ >= 2 (and, with work lists wrp, a
// non-empty work list)
__global__ void k_bflags(const uint32_t* __restrict__ rp, const uint32_t* __restrict__ wrp,
                         uint32_t n, uint32_t lo, uint32_t hi, uint8_t* __restrict__ f) {
  for (uint64_t u = gtid(); u < n; u += gstride()) {
    const uint32_t d = rp[u + 1] - rp[u];
    f[u] = ((!wrp || wrp[u + 1] > wrp[u]) && d >= 2 && d >= lo && d <= hi) ? 1 : 0;
  }
}

// descending work as the sort key: d+(u) plus the entries the task streams
// (d+(y) over its work list wcol[wrp[u], wrp[u+1]), or over N+(u) in the
// push form, wrp = col and the rows of rp)
__global__ void k_bwork(const uint32_t* __restrict__ rp, const uint32_t* __restrict__ lrp,
                        const uint32_t* __restrict__ lcol, const uint32_t* __restrict__ verts,
                        uint32_t nv, uint64_t* __restrict__ key, uint32_t* __restrict__ val) {
  for (uint64_t i = gtid(); i < nv; i += gstride()) {
    const uint32_t u = verts[i];
    uint64_t w = rp[u + 1] - rp[u];
    for (uint32_t k = lrp[u]; k < lrp[u + 1]; ++k) w += rp[lcol[k] + 1] - rp[lcol[k]];
    key[i] = ~w;
    val[i] = u;
  }
}

// off[g] = first position of list whose vertex is >= voff[g] (list ascending)
__global__ void k_graph_offsets(const uint32_t* __restrict__ list, uint32_t len,
                                const uint32_t* __restrict__ voff, uint32_t ngraphs,
                                uint32_t* __restrict__ off) {
  for (uint64_t g = gtid(); g <= ngraphs; g += gstride()) {
    const uint32_t x = voff[g];
    uint32_t lo = 0, hi = len;
    while (lo < hi) {
      const uint32_t mid = (lo + hi) >> 1;
      if (list[mid] < x) lo = mid + 1;
      else hi = mid;
    }
    off[g] = lo;
  }
}

unsigned grid_for(uint64_t n, int tb = 256) {
  uint64_t b = (n + tb - 1) / tb;
  return (unsigned)std::max<uint64_t>(1, std::min<uint64_t>(b, 1u << 20));
}

struct BuildArgs {
  const uint64_t* eoff;
  const uint32_t* src;
  const uint32_t* dst;
  const uint32_t* nv;
  const uint32_t* voff;
  uint32_t* gs;
  uint32_t* gd;
  uint32_t* gidh;
  std::atomic<int> bad;
};

void build_range(uint64_t b, uint64_t e, void* p) {
  BuildArgs& A = *static_cast<BuildArgs*>(p);
  for (uint64_t g = b; g < e; ++g) {
    const uint32_t base = A.voff[g], n = A.nv[g];
    for (uint64_t k = A.eoff[g]; k < A.eoff[g + 1]; ++k) {
      if (A.src[k] >= n || A.dst[k] >= n) {
        A.bad.store(1, std::memory_order_relaxed);
        return;
      }
      A.gs[k] = A.src[k] + base;
      A.gd[k] = A.dst[k] + base;
    }
    for (uint32_t v = 0; v < n; ++v) A.gidh[base + v] = (uint32_t)g;
  }
}

}  // namespace
}  // namespace tcx

using namespace tcx;

struct tc_batch {
  tc_ctx* ctx = nullptr;
  tc_graph* g = nullptr;
  uint32_t ngraphs = 0;
  std::vector<uint32_t> voff;
  DevBuf d_voff, gid, counts, cursors;
  // schedule for one (wave_max_deg, form)
  uint32_t sched_deg = 0, sched_form = 0;
  DevBuf blk_lpt, blk_id, wav;
  uint64_t nblk = 0, nwav = 0;
  DevBuf blk_off, wav_off;  // per-graph ranges of blk_id / wav
  std::vector<uint32_t> h_blk_off, h_wav_off;
  // captured replay
  bool captured = false;
  uint32_t cap_deg = 0, cap_chunk = 0, cap_grid = 0, cap_fmin = 0, cap_form = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

namespace {

void release_capture(tc_batch* b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
  b->captured = false;
}

tc_batch_sched fill(const tc_batch_sched* s, tc_ctx* c) {
  tc_batch_sched o;
  memset(&o, 0, sizeof(o));
  if (s) o = *s;
  if (!o.wave_max_deg) o.wave_max_deg = 64;
  if (o.wave_max_deg > kWaveMax) o.wave_max_deg = kWaveMax;
  if (!o.wave_chunk) o.wave_chunk = 16;
  if (!o.grid) o.grid = (uint32_t)c->num_cus * 8u;
  o.filter_min = o.filter_min == 1 ? 1u : 16u;
  return o;
}

int check_sched(const tc_batch_sched* s) {
  if (s && s->form > TC_BATCH_PUSH) return fail(TC_ERR_ARG, "unknown batched form %u", s->form);
  return TC_OK;
}

// The batched kernel for the schedule's form (the push form also for its
// wave-variant filter threshold, 1 or 16; the work-list form has none)
template <typename... A>
void launch_batch(const tc_batch_sched& s, dim3 grid, hipStream_t st, A... args) {
  if (s.form == TC_BATCH_WORKLIST)
    hipLaunchKernelGGL((k_batch_wl<2048>), grid, dim3(TB), 0, st, args...);
  else if (s.filter_min == 1)
    hipLaunchKernelGGL((k_batch<2048, 1>), grid, dim3(TB), 0, st, args...);
  else
    hipLaunchKernelGGL((k_batch<2048, 16>), grid, dim3(TB), 0, st, args...);
}

int schedule(tc_ctx* c, tc_batch* b, uint32_t wave_max_deg, uint32_t form) {
  if (b->sched_deg == wave_max_deg && b->sched_form == form) return TC_OK;
  tc_graph* g = b->g;
  const uint32_t n = g->n;
  if (g->m_or == 0 || n == 0) {  // no edges: no work lists and no tasks
    b->nwav = b->nblk = 0;
    b->h_blk_off.assign((size_t)b->ngraphs + 1, 0);
    b->h_wav_off.assign((size_t)b->ngraphs + 1, 0);
    b->sched_deg = wave_max_deg;
    b->sched_form = form;
    release_capture(b);
    return TC_OK;
  }
  const uint32_t* rp = g->rowptr.as<uint32_t>();
  const uint32_t* wrp = nullptr;
  const uint32_t* wcol = nullptr;
  if (form == TC_BATCH_WORKLIST) {
    TCX_TRY(worklists(c, g));
    wrp = g->sc.wl_rp.as<uint32_t>();
    wcol = g->sc.wl_col.as<uint32_t>();
  }
  DevBuf flags, key, ktmp, vtmp;
  TCX_TRY(flags.reserve((size_t)n + 1));
  TCX_TRY(b->wav.reserve(4 * ((size_t)n + 1)));
  TCX_TRY(b->blk_id.reserve(4 * ((size_t)n + 1)));
  TCX_TRY(b->blk_lpt.reserve(4 * ((size_t)n + 1)));
  hipLaunchKernelGGL(k_bflags, dim3(grid_for(n)), dim3(256), 0, c->stream, rp, wrp, n, 2u,
                     wave_max_deg, flags.as<uint8_t>());
  TCX_TRY(select_indices(c, flags.as<uint8_t>(), n, b->wav.as<uint32_t>(), &b->nwav));
  hipLaunchKernelGGL(k_bflags, dim3(grid_for(n)), dim3(256), 0, c->stream, rp, wrp, n,
                     wave_max_deg + 1, 0xFFFFFFFFu, flags.as<uint8_t>());
  TCX_TRY(select_indices(c, flags.as<uint8_t>(), n, b->blk_id.as<uint32_t>(), &b->nblk));
  // block tasks for the batched launch: descending work (longest first)
  const uint64_t nb = b->nblk;
  TCX_TRY(key.reserve(8 * (nb + 1)));
  TCX_TRY(ktmp.reserve(8 * (nb + 1)));
  TCX_TRY(vtmp.reserve(4 * (nb + 1)));
  DevBuf vals;
  TCX_TRY(vals.reserve(4 * (nb + 1)));
  if (nb) {
    hipLaunchKernelGGL(k_bwork, dim3(grid_for(nb)), dim3(256), 0, c->stream, rp, wrp ? wrp : rp,
                       wcol ? wcol : g->col.as<uint32_t>(), b->blk_id.as<uint32_t>(), (uint32_t)nb,
                       key.as<uint64_t>(), vals.as<uint32_t>());
    TCX_HIP(hipGetLastError());
    uint64_t* sk;
    uint32_t* sv;
    TCX_TRY(sort_pairs_u64_u32(c, key.as<uint64_t>(), ktmp.as<uint64_t>(), vals.as<uint32_t>(),
                               vtmp.as<uint32_t>(), nb, 64, &sk, &sv));
    TCX_HIP(hipMemcpyAsync(b->blk_lpt.p, sv, 4 * nb, hipMemcpyDeviceToDevice, c->stream));
  }
  // per-graph ranges of the id-ordered lists
  const uint32_t G = b->ngraphs;
  TCX_TRY(b->blk_off.reserve(4 * ((size_t)G + 1)));
  TCX_TRY(b->wav_off.reserve(4 * ((size_t)G + 1)));
  hipLaunchKernelGGL(k_graph_offsets, dim3(grid_for(G + 1ull)), dim3(256), 0, c->stream,
                     b->blk_id.as<uint32_t>(), (uint32_t)nb, b->d_voff.as<uint32_t>(), G,
                     b->blk_off.as<uint32_t>());
  hipLaunchKernelGGL(k_graph_offsets, dim3(grid_for(G + 1ull)), dim3(256), 0, c->stream,
                     b->wav.as<uint32_t>(), (uint32_t)b->nwav, b->d_voff.as<uint32_t>(), G,
                     b->wav_off.as<uint32_t>());
  TCX_HIP(hipGetLastError());
  b->h_blk_off.resize(G + 1);
  b->h_wav_off.resize(G + 1);
  TCX_HIP(hipMemcpyAsync(b->h_blk_off.data(), b->blk_off.p, 4 * ((size_t)G + 1),
                         hipMemcpyDeviceToHost, c->stream));
  TCX_HIP(hipMemcpyAsync(b->h_wav_off.data(), b->wav_off.p, 4 * ((size_t)G + 1),
                         hipMemcpyDeviceToHost, c->stream));
  TCX_HIP(hipStreamSynchronize(c->stream));
  b->sched_deg = wave_max_deg;
  b->sched_form = form;
  release_capture(b);
  return TC_OK;
}

int enqueue_batched(tc_ctx* c, tc_batch* b, const tc_batch_sched& s) {
  tc_graph* g = b->g;
  hipLaunchKernelGGL(k_batch_reset, dim3(grid_for(2ull + b->ngraphs)), dim3(256), 0, c->stream,
                     b->cursors.as<uint32_t>(), 2u, b->counts.as<unsigned long long>(), b->ngraphs);
  launch_batch(s, dim3(s.grid), c->stream, g->rowptr.as<uint32_t>(), g->col.as<uint32_t>(),
               g->sc.wl_rp.as<uint32_t>(), g->sc.wl_se.as<uint2>(),
               b->gid.as<uint32_t>(), b->blk_lpt.as<uint32_t>(), (uint32_t)b->nblk,
               b->wav.as<uint32_t>(), (uint32_t)b->nwav, s.wave_chunk, b->cursors.as<uint32_t>(),
               b->counts.as<unsigned long long>());
  TCX_HIP(hipGetLastError());
  return TC_OK;
}

int check_batch(tc_ctx* c, tc_batch* b) {
  if (!c || !b) return fail(TC_ERR_ARG, "NULL argument");
  if (b->ctx != c) return fail(TC_ERR_ARG, "batch belongs to another context");
  TCX_HIP(hipSetDevice(c->device));
  return TC_OK;
}

}  // namespace

extern "C" {

int tc_batch_create(tc_ctx* c, uint32_t ngraphs, const uint64_t* eoff, const uint32_t* src,
                    const uint32_t* dst, const uint32_t* nverts, tc_batch** out) {
  if (!c || !out || !eoff || !nverts || ngraphs == 0) return fail(TC_ERR_ARG, "bad arguments");
  *out = nullptr;
  if (eoff[0] != 0) return fail(TC_ERR_ARG, "edge_off[0] must be 0");
  const uint64_t M = eoff[ngraphs];
  if (M && (!src || !dst)) return fail(TC_ERR_ARG, "NULL edge arrays");
  std::vector<uint32_t> voff(ngraphs + 1, 0);
  uint64_t total = 0;
  for (uint32_t g = 0; g < ngraphs; ++g) {
    if (eoff[g + 1] < eoff[g]) return fail(TC_ERR_ARG, "edge offsets must be ascending");
    voff[g] = (uint32_t)total;
    total += nverts[g];
    if (total >= 0x7FFFFFFFull) return fail(TC_ERR_RANGE, "batch exceeds 2^31 vertices");
  }
  voff[ngraphs] = (uint32_t)total;
  std::vector<uint32_t> gs(std::max<uint64_t>(M, 1)), gd(std::max<uint64_t>(M, 1)),
      gidh(std::max<uint64_t>(total, 1));
  BuildArgs A{eoff, src, dst, nverts, voff.data(), gs.data(), gd.data(), gidh.data(), {0}};
  parallel_for(ngraphs, build_range, &A);
  if (A.bad.load()) return fail(TC_ERR_RANGE, "a record's id exceeds its graph's vertex count");
  tc_batch* b = new tc_batch();
  b->ctx = c;
  b->ngraphs = ngraphs;
  b->voff = voff;
  int st = tc_graph_create(c, gs.data(), gd.data(), M, (uint32_t)total, &b->g);
  if (!st) st = tc_graph_orient(c, b->g, TC_ORIENT_DEGREE, 0);
  if (!st) st = b->gid.reserve(4 * std::max<uint64_t>(total, 1));
  if (!st) st = b->d_voff.reserve(4 * ((size_t)ngraphs + 1));
  if (!st) st = b->counts.reserve(8 * (size_t)ngraphs);
  if (!st) st = b->cursors.reserve(8 * ((size_t)ngraphs + 2));
  if (!st && total &&
      hipMemcpyAsync(b->gid.p, gidh.data(), 4 * total, hipMemcpyHostToDevice, c->stream) != hipSuccess)
    st = fail(TC_ERR_HIP, "gid upload failed");
  if (!st && hipMemcpyAsync(b->d_voff.p, voff.data(), 4 * ((size_t)ngraphs + 1),
                            hipMemcpyHostToDevice, c->stream) != hipSuccess)
    st = fail(TC_ERR_HIP, "offset upload failed");
  if (!st && hipStreamSynchronize(c->stream) != hipSuccess) st = fail(TC_ERR_HIP, "sync failed");
  if (st) {
    tc_graph_destroy(b->g);
    delete b;
    return st;
  }
  *out = b;
  return TC_OK;
}

int tc_batch_destroy(tc_batch* b) {
  if (!b) return TC_OK;
  if (b->ctx) (void)hipStreamSynchronize(b->ctx->stream);
  release_capture(b);
  tc_graph_destroy(b->g);
  delete b;
  return TC_OK;
}

int tc_batch_info_get(const tc_batch* b, tc_graph_info* info) {
  if (!b || !info) return fail(TC_ERR_ARG, "NULL argument");
  return tc_graph_info_get(b->g, info);
}

int tc_batch_stream_bytes(tc_ctx* c, tc_batch* b, double* bytes) {
  TCX_TRY(check_batch(c, b));
  if (!bytes) return fail(TC_ERR_ARG, "NULL output");
  return tc_auto_stream_bytes(c, b->g, nullptr, bytes);
}

int tc_batch_count(tc_ctx* c, tc_batch* b, const tc_batch_sched* sched, uint64_t* counts,
                   tc_timing* timing) {
  TCX_TRY(check_batch(c, b));
  if (!counts) return fail(TC_ERR_ARG, "NULL counts");
  const tc_batch_sched s = fill(sched, c);
  tc_timing tm;
  memset(&tm, 0, sizeof(tm));
  TCX_HIP(hipEventRecord(c->ev0, c->stream));
  TCX_TRY(check_sched(sched));
  TCX_TRY(schedule(c, b, s.wave_max_deg, s.form));
  if (!b->captured || b->cap_chunk != s.wave_chunk || b->cap_grid != s.grid ||
      b->cap_deg != s.wave_max_deg || b->cap_fmin != s.filter_min || b->cap_form != s.form) {
    release_capture(b);
    TCX_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    int st = enqueue_batched(c, b, s);
    hipGraph_t gr = nullptr;
    hipError_t e = hipStreamEndCapture(c->stream, &gr);
    if (st) {
      if (gr) (void)hipGraphDestroy(gr);
      return st;
    }
    if (e != hipSuccess) return fail(TC_ERR_HIP, "stream capture failed: %s", hipGetErrorString(e));
    b->graph = gr;
    TCX_HIP(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    b->captured = true;
    b->cap_chunk = s.wave_chunk;
    b->cap_grid = s.grid;
    b->cap_deg = s.wave_max_deg;
    b->cap_fmin = s.filter_min;
    b->cap_form = s.form;
  }
  TCX_HIP(hipEventRecord(c->ev1, c->stream));
  TCX_HIP(hipEventSynchronize(c->ev1));
  TCX_HIP(hipEventElapsedTime(&tm.ms_prep, c->ev0, c->ev1));
  TCX_HIP(hipEventRecord(c->ev0, c->stream));
  TCX_HIP(hipGraphLaunch(b->exec, c->stream));
  TCX_HIP(hipEventRecord(c->ev1, c->stream));
  TCX_HIP(hipEventSynchronize(c->ev1));
  TCX_HIP(hipEventElapsedTime(&tm.ms_count, c->ev0, c->ev1));
  tm.ms_dominant = tm.ms_count;
  tm.launches = 1;
  TCX_HIP(hipMemcpyAsync(counts, b->counts.p, 8 * (size_t)b->ngraphs, hipMemcpyDeviceToHost,
                         c->stream));
  TCX_HIP(hipStreamSynchronize(c->stream));
  if (timing) *timing = tm;
  return TC_OK;
}

int tc_batch_count_pergraph(tc_ctx* c, tc_batch* b, const tc_batch_sched* sched,
                            uint64_t* counts, tc_timing* timing) {
  TCX_TRY(check_batch(c, b));
  if (!counts) return fail(TC_ERR_ARG, "NULL counts");
  const tc_batch_sched s = fill(sched, c);
  tc_timing tm;
  memset(&tm, 0, sizeof(tm));
  TCX_TRY(check_sched(sched));
  TCX_TRY(schedule(c, b, s.wave_max_deg, s.form));
  tc_graph* g = b->g;
  const uint32_t G = b->ngraphs;
  uint32_t* cur = b->cursors.as<uint32_t>();
  hipLaunchKernelGGL(k_batch_reset, dim3(grid_for(2ull * G + G)), dim3(256), 0, c->stream, cur,
                     2u * G, b->counts.as<unsigned long long>(), G);
  TCX_HIP(hipGetLastError());
  TCX_HIP(hipEventRecord(c->ev0, c->stream));
  for (uint32_t q = 0; q < G; ++q) {
    const uint32_t b0 = b->h_blk_off[q], b1 = b->h_blk_off[q + 1];
    const uint32_t w0 = b->h_wav_off[q], w1 = b->h_wav_off[q + 1];
    if (b0 == b1 && w0 == w1) continue;
    const uint64_t need = (b1 - b0) + ((w1 - w0) + 3) / 4;
    const unsigned grid = (unsigned)std::max<uint64_t>(1, std::min<uint64_t>(s.grid, need));
    launch_batch(s, dim3(grid), c->stream, g->rowptr.as<uint32_t>(), g->col.as<uint32_t>(),
                 g->sc.wl_rp.as<uint32_t>(), g->sc.wl_se.as<uint2>(),
                 b->gid.as<uint32_t>(), b->blk_id.as<uint32_t>() + b0, b1 - b0,
                 b->wav.as<uint32_t>() + w0, w1 - w0, s.wave_chunk, cur + 2 * q,
                 b->counts.as<unsigned long long>());
    tm.launches++;
  }
  TCX_HIP(hipGetLastError());
  TCX_HIP(hipEventRecord(c->ev1, c->stream));
  TCX_HIP(hipEventSynchronize(c->ev1));
  TCX_HIP(hipEventElapsedTime(&tm.ms_count, c->ev0, c->ev1));
  tm.ms_dominant = tm.ms_count;
  TCX_HIP(hipMemcpyAsync(counts, b->counts.p, 8 * (size_t)G, hipMemcpyDeviceToHost, c->stream));
  TCX_HIP(hipStreamSynchronize(c->stream));
  if (timing) *timing = tm;
  return TC_OK;
}

int tc_batch_autotune(tc_ctx* c, tc_batch* b, tc_batch_sched* best) {
  TCX_TRY(check_batch(c, b));
  if (!best) return fail(TC_ERR_ARG, "NULL output");
  std::vector<uint64_t> counts(b->ngraphs);
  auto time_of = [&](const tc_batch_sched& s, float* ms) -> int {
    tc_timing t;
    float acc = 0;
    TCX_TRY(tc_batch_count(c, b, &s, counts.data(), &t));  // capture + warm
    for (int r = 0; r < 2; ++r) {
      TCX_TRY(tc_batch_count(c, b, &s, counts.data(), &t));
      acc += t.ms_count;
    }
    *ms = acc / 2;
    return TC_OK;
  };
  // per form: the wave/block degree split, then the filter threshold and the
  // wave-task chunk, each stage keeping the fastest so far; the faster form
  // wins
  float best_ms = 1e30f;
  tc_batch_sched bs = fill(nullptr, c);
  for (uint32_t form : {(uint32_t)TC_BATCH_WORKLIST, (uint32_t)TC_BATCH_PUSH}) {
    tc_batch_sched s = fill(nullptr, c), fb = s;
    float fms = 1e30f;
    auto trial = [&](const tc_batch_sched& t) -> int {
      float ms;
      TCX_TRY(time_of(t, &ms));
      if (ms < fms) { fms = ms; fb = t; }
      return TC_OK;
    };
    s.form = form;
    for (uint32_t d : {16u, 32u, 64u, 128u, 256u, 512u}) {
      s.wave_max_deg = d;
      TCX_TRY(trial(s));
    }
    s = fb;
    s.filter_min = 1;
    TCX_TRY(trial(s));
    s = fb;
    for (uint32_t ch : {4u, 8u, 32u, 64u}) {
      s.wave_chunk = ch;
      TCX_TRY(trial(s));
    }
    if (fms < best_ms) { best_ms = fms; bs = fb; }
  }
  *best = bs;
  return TC_OK;
}

}  // extern "C"
