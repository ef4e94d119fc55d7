// Context, errors, device buffers and edge-file I/O.
#include <cerrno>
#include <cstdarg>
#include <cstring>
#include <fstream>
#include <thread>

#include "internal.h"

namespace tcx {

static thread_local char g_err[512] = "";

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

int fail(int status, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
  return status;
}

int DevBuf::reserve(size_t want) {
  if (want <= bytes && p) return TC_OK;
  release();
  if (want == 0) want = 64;
  hipError_t e = hipMalloc(&p, want);
  if (e != hipSuccess) {
    p = nullptr;
    bytes = 0;
    (void)hipGetLastError();
    return fail(TC_ERR_NOMEM, "hipMalloc(%zu bytes) failed: %s", want, hipGetErrorString(e));
  }
  bytes = want;
  return TC_OK;
}

void DevBuf::release() {
  if (p) (void)hipFree(p);
  p = nullptr;
  bytes = 0;
}

void parallel_for(uint64_t n, void (*fn)(uint64_t, uint64_t, void*), void* arg,
                  unsigned max_threads) {
  unsigned hw = std::thread::hardware_concurrency();
  if (hw == 0) hw = 1;
  if (hw > 16) hw = 16;  // at most 16 host threads per GPU process
  if (max_threads && hw > max_threads) hw = max_threads;
  if (n < 4096 || hw == 1) {
    fn(0, n, arg);
    return;
  }
  std::vector<std::thread> ts;
  uint64_t chunk = (n + hw - 1) / hw;
  for (unsigned t = 0; t < hw; ++t) {
    uint64_t b = t * chunk, e = std::min<uint64_t>(n, b + chunk);
    if (b >= e) break;
    ts.emplace_back(fn, b, e, arg);
  }
  for (auto& t : ts) t.join();
}

}  // namespace tcx

using namespace tcx;

extern "C" {

const char* tc_status_str(int s) {
  switch (s) {
    case TC_OK: return "ok";
    case TC_ERR_ARG: return "invalid argument";
    case TC_ERR_IO: return "i/o error";
    case TC_ERR_FORMAT: return "bad file format";
    case TC_ERR_RANGE: return "value out of range";
    case TC_ERR_NOMEM: return "out of memory";
    case TC_ERR_HIP: return "HIP runtime error";
    case TC_ERR_UNSUPPORTED: return "unsupported input";
    case TC_ERR_STATE: return "invalid call order";
    default: return "unknown status";
  }
}

const char* tc_last_error(void) { return g_err; }

int tc_ctx_create(int device, tc_ctx** out) {
  if (!out) return fail(TC_ERR_ARG, "tc_ctx_create: out is NULL");
  *out = nullptr;
  int ndev = 0;
  hipError_t e = hipGetDeviceCount(&ndev);
  if (e != hipSuccess || ndev == 0) {
    (void)hipGetLastError();
    return fail(TC_ERR_HIP, "no HIP device available (%s)", hipGetErrorString(e));
  }
  if (device < 0 || device >= ndev)
    return fail(TC_ERR_ARG, "device %d out of range (have %d)", device, ndev);
  TCX_HIP(hipSetDevice(device));
  TCX_TRY(tcx::noalloc_setup(device));
  tc_ctx* c = new tc_ctx();
  c->device = device;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) == hipSuccess) {
    c->total_mem = prop.totalGlobalMem;
    c->num_cus = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
  }
  int khz = 0;
  if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, device) == hipSuccess && khz > 0)
    c->wall_clock_khz = khz;
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess) {
    delete c;
    return fail(TC_ERR_HIP, "stream/event creation failed");
  }
  *out = c;
  return TC_OK;
}

int tc_ctx_destroy(tc_ctx* c) {
  if (!c) return TC_OK;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  c->temp.release();
  c->scratch.release();
  c->accum.release();
  c->prim_out.release();
  for (hipEvent_t e : c->lev)
    if (e) (void)hipEventDestroy(e);
  if (c->ev0) (void)hipEventDestroy(c->ev0);
  if (c->ev1) (void)hipEventDestroy(c->ev1);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return TC_OK;
}

int tc_ctx_device_memory(const tc_ctx* c, uint64_t* bytes) {
  if (!c || !bytes) return fail(TC_ERR_ARG, "NULL argument");
  *bytes = c->total_mem;
  return TC_OK;
}

// ---- files ---------------------------------------------------------------

int tc_file_edge_count(const char* path, uint64_t* m) {
  if (!path || !m) return fail(TC_ERR_ARG, "NULL argument");
  FILE* f = fopen(path, "rb");
  if (!f) return fail(TC_ERR_IO, "cannot open %s: %s", path, strerror(errno));
  if (fseeko(f, 0, SEEK_END) != 0) {
    fclose(f);
    return fail(TC_ERR_IO, "cannot seek %s", path);
  }
  off_t size = ftello(f);
  fclose(f);
  if (size < 0) return fail(TC_ERR_IO, "cannot size %s", path);
  if (size % 8 != 0)
    return fail(TC_ERR_FORMAT, "%s: size %lld is not a multiple of 8", path, (long long)size);
  *m = (uint64_t)size / 8;
  return TC_OK;
}

int tc_read_edges(const char* path, uint32_t* src, uint32_t* dst, uint64_t cap,
                  uint64_t* m) {
  if (!path || !m) return fail(TC_ERR_ARG, "NULL argument");
  uint64_t cnt = 0;
  TCX_TRY(tc_file_edge_count(path, &cnt));
  if (cnt > cap) return fail(TC_ERR_RANGE, "%s holds %llu records, capacity %llu", path,
                             (unsigned long long)cnt, (unsigned long long)cap);
  if (cnt && (!src || !dst)) return fail(TC_ERR_ARG, "NULL output arrays");
  FILE* f = fopen(path, "rb");
  if (!f) return fail(TC_ERR_IO, "cannot open %s: %s", path, strerror(errno));
  const uint64_t kChunk = 1u << 20;
  std::vector<uint32_t> buf(2 * kChunk);
  uint64_t done = 0;
  while (done < cnt) {
    uint64_t want = std::min(kChunk, cnt - done);
    if (fread(buf.data(), 8, want, f) != want) {
      fclose(f);
      return fail(TC_ERR_IO, "short read on %s", path);
    }
    for (uint64_t i = 0; i < want; ++i) {
      src[done + i] = buf[2 * i];
      dst[done + i] = buf[2 * i + 1];
    }
    done += want;
  }
  fclose(f);
  *m = cnt;
  return TC_OK;
}

int tc_write_edges(const char* path, const uint32_t* src, const uint32_t* dst, uint64_t m) {
  if (!path || (m && (!src || !dst))) return fail(TC_ERR_ARG, "NULL argument");
  FILE* f = fopen(path, "wb");
  if (!f) return fail(TC_ERR_IO, "cannot create %s: %s", path, strerror(errno));
  const uint64_t kChunk = 1u << 20;
  std::vector<uint32_t> buf(2 * kChunk);
  for (uint64_t done = 0; done < m;) {
    uint64_t want = std::min(kChunk, m - done);
    for (uint64_t i = 0; i < want; ++i) {
      buf[2 * i] = src[done + i];
      buf[2 * i + 1] = dst[done + i];
    }
    if (fwrite(buf.data(), 8, want, f) != want) {
      fclose(f);
      return fail(TC_ERR_IO, "short write on %s", path);
    }
    done += want;
  }
  if (fclose(f) != 0) return fail(TC_ERR_IO, "close failed on %s", path);
  return TC_OK;
}

int tc_list2bin(const char* text_path, const char* bin_path, uint64_t* m) {
  if (!text_path || !bin_path) return fail(TC_ERR_ARG, "NULL argument");
  std::ifstream in(text_path);
  if (!in) return fail(TC_ERR_IO, "cannot open %s", text_path);
  FILE* f = fopen(bin_path, "wb");
  if (!f) return fail(TC_ERR_IO, "cannot create %s", bin_path);
  // Same stream extraction as tools/list2bin.cpp:16 (stops at the first
  // token that is not an integer).
  int s, d;
  uint64_t cnt = 0;
  while (in >> s >> d) {
    int32_t rec[2] = {s, d};
    if (fwrite(rec, 8, 1, f) != 1) {
      fclose(f);
      return fail(TC_ERR_IO, "short write on %s", bin_path);
    }
    ++cnt;
  }
  if (fclose(f) != 0) return fail(TC_ERR_IO, "close failed on %s", bin_path);
  if (m) *m = cnt;
  return TC_OK;
}

// Galois .gr binary CSR (versions 1 and 2): four uint64 header words
// (version, edge-data size, nodes, edges), nodes uint64 row ends, then the
// column ids (uint32 in v1, padded to 8 bytes; uint64 in v2), then edge data.
int tc_read_gr(const char* path, uint32_t* src, uint32_t* dst, uint64_t cap, uint64_t* m,
               uint32_t* n) {
  if (!path || !m) return fail(TC_ERR_ARG, "NULL argument");
  FILE* f = fopen(path, "rb");
  if (!f) return fail(TC_ERR_IO, "cannot open %s: %s", path, strerror(errno));
  auto done = [&](int st) {
    fclose(f);
    return st;
  };
  uint64_t hdr[4];
  if (fread(hdr, 8, 4, f) != 4) return done(fail(TC_ERR_FORMAT, "%s: short .gr header", path));
  const uint64_t ver = hdr[0], nodes = hdr[2], edges = hdr[3];
  if (ver != 1 && ver != 2) return done(fail(TC_ERR_FORMAT, "%s: .gr version %llu", path,
                                              (unsigned long long)ver));
  if (nodes >= 0x7FFFFFFFull) return done(fail(TC_ERR_RANGE, "%s: too many nodes", path));
  if (fseeko(f, 0, SEEK_END) != 0) return done(fail(TC_ERR_IO, "cannot seek %s", path));
  const uint64_t size = (uint64_t)ftello(f);
  const uint64_t colw = ver == 1 ? 4 : 8;
  const uint64_t need = 32 + 8 * nodes + colw * edges;
  if (size < need) return done(fail(TC_ERR_FORMAT, "%s: .gr file truncated", path));
  *m = edges;
  if (n) *n = (uint32_t)nodes;
  if (!src) return done(TC_OK);  // size query
  if (edges > cap) return done(fail(TC_ERR_RANGE, "%s holds %llu edges, capacity %llu", path,
                                    (unsigned long long)edges, (unsigned long long)cap));
  if (edges && !dst) return done(fail(TC_ERR_ARG, "NULL output arrays"));
  std::vector<uint64_t> ends(nodes);
  if (fseeko(f, 32, SEEK_SET) != 0 || fread(ends.data(), 8, nodes, f) != nodes)
    return done(fail(TC_ERR_IO, "short read on %s", path));
  uint64_t prev = 0;
  for (uint64_t u = 0; u < nodes; ++u) {
    if (ends[u] < prev || ends[u] > edges)
      return done(fail(TC_ERR_FORMAT, "%s: row ends not ascending", path));
    for (uint64_t k = prev; k < ends[u]; ++k) src[k] = (uint32_t)u;
    prev = ends[u];
  }
  if (prev != edges) return done(fail(TC_ERR_FORMAT, "%s: row ends do not cover the edges", path));
  const uint64_t kChunk = 1u << 20;
  std::vector<uint64_t> buf(kChunk);
  for (uint64_t k0 = 0; k0 < edges; k0 += kChunk) {
    const uint64_t want = std::min(kChunk, edges - k0);
    if (fread(buf.data(), colw, want, f) != want)
      return done(fail(TC_ERR_IO, "short read on %s", path));
    const uint32_t* b32 = reinterpret_cast<const uint32_t*>(buf.data());
    for (uint64_t i = 0; i < want; ++i) {
      const uint64_t v = ver == 1 ? b32[i] : buf[i];
      if (v >= nodes) return done(fail(TC_ERR_FORMAT, "%s: column id beyond node count", path));
      dst[k0 + i] = (uint32_t)v;
    }
  }
  return done(TC_OK);
}

int tc_bin2list(const char* bin_path, const char* text_path, uint64_t* m) {
  if (!bin_path || !text_path) return fail(TC_ERR_ARG, "NULL argument");
  FILE* f = fopen(bin_path, "rb");
  if (!f) return fail(TC_ERR_IO, "cannot open %s", bin_path);
  FILE* o = fopen(text_path, "w");
  if (!o) {
    fclose(f);
    return fail(TC_ERR_IO, "cannot create %s", text_path);
  }
  int32_t rec[2];
  uint64_t cnt = 0;
  // tools/bin2list.cpp:16-18 reads whole 8-byte records; a trailing partial
  // record is ignored there as well.
  while (fread(rec, 8, 1, f) == 1) {
    fprintf(o, "%d %d\n", rec[0], rec[1]);
    ++cnt;
  }
  fclose(f);
  if (fclose(o) != 0) return fail(TC_ERR_IO, "close failed on %s", text_path);
  if (m) *m = cnt;
  return TC_OK;
}

}  // extern "C"
