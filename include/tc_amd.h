/*
 * tc_amd.h -- C ABI of the MI355X triangle-counting / k-clique engine.
 *
 * The reference (pkumod/accelerating-TC) has no library API: each method is a
 * stand-alone program whose boundary is argv + an edge file + one stdout
 * line.  These entry points are the function-level form of that boundary.
 * Each declaration cites the reference code whose role it takes over.
 *
 * Conventions
 *  - Every int-returning function returns TC_OK (0) or a negative TC_ERR_*;
 *    nothing calls exit().  tc_last_error() gives the message of the last
 *    failure on the calling thread.
 *  - Host pointers passed in are borrowed read-only unless documented as
 *    outputs; outputs are caller-allocated.
 *  - A tc_ctx owns one HIP stream and is not thread-safe; separate contexts
 *    are independent.  A tc_graph / tc_batch is bound to the context that
 *    created it and owns its device memory until destroyed.
 *  - Vertex ids are uint32; the edge-file record is {int32 src, int32 dst},
 *    8 bytes, no header (tools/list2bin.cpp:3-6,16-20).
 */
#ifndef TC_AMD_H
#define TC_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status ------------------------------------------------------------ */
#define TC_OK 0
#define TC_ERR_ARG (-1)         /* bad argument / NULL pointer              */
#define TC_ERR_IO (-2)          /* open/read/write failure                  */
#define TC_ERR_FORMAT (-3)      /* file size not a multiple of the record   */
#define TC_ERR_RANGE (-4)       /* id / size beyond what the engine handles */
#define TC_ERR_NOMEM (-5)       /* device or host allocation failed         */
#define TC_ERR_HIP (-6)         /* HIP runtime error                        */
#define TC_ERR_UNSUPPORTED (-7) /* input outside an algorithm's domain      */
#define TC_ERR_STATE (-8)       /* call order (e.g. count before orient)    */

/* ---- enums ------------------------------------------------------------- */
/* Edge orientation (DAG construction). */
#define TC_ORIENT_ID 0          /* src = larger id: Fox/preTC.cu:92-101, Bisson/preTC.cu:136-144 */
#define TC_ORIENT_DEGREE 1      /* Hu/preTC.cu:125-133 (ties from the larger id) */
#define TC_ORIENT_ADIR 2        /* A-direction peel: edge_directing/A-direction-seri.cpp:35-153,252-273 */
#define TC_ORIENT_KCORE 3       /* edge_directing/kcore_direction.cpp:91-146 */
#define TC_ORIENT_DEGREE_TRICORE 4 /* tricore/src/tricount_gpu.cu:69-80 (ties from the smaller id) */
/* Degeneracy order: the k-core peel of edge_directing/kcore_direction.cpp:91-126
 * run level-synchronously, every edge from the lower (removal step, id).  The
 * reference orients by (core number, id) instead (:137-143); refining the key
 * by the removal step bounds every out-degree by the vertex's core number,
 * hence by the degeneracy (the k-clique path's best orientation). */
#define TC_ORIENT_DEGENERACY 5
/* The OpenMP program's A-direction schedule (edge_directing/A-direction.cpp:
 * 36-161, argv `A-direction <bin> <thre_param>`): set by
 * tc_graph_orient_adir_omp, not by tc_graph_orient. */
#define TC_ORIENT_ADIR_OMP 6

/* Orientation flags. */
#define TC_FLAG_PRUNE_DEG1 1u   /* Hu/preTC.cu:135-169: drop degree<=1 vertices, compact ids */

/* Counting method. */
#define TC_KERNEL_HU 0          /* Hu/tc.cu:69-201      */
#define TC_KERNEL_BISSON 1      /* Bisson/tc.cu:18-104  */
#define TC_KERNEL_FOX 2         /* Fox/tc.cu:64-125     */
#define TC_KERNEL_TRICORE 3     /* tricore/src/tricount_gpu.cu:215-290 */
#define TC_KERNEL_AUTO 4        /* degree-binned engine (fixed or learned bins) */

/* AUTO schedule (tc_sched.auto_thresholds[2]).  Both count every oriented
 * edge u->v as |N+(u) & N+(v)| by a task that holds one of the two lists in
 * LDS and streams the other from HBM.
 *   WORKLIST: the edge goes to the task of whichever endpoint makes the
 *             streamed list the shorter one (u when d+(v) <= d+(u), else v),
 *             from a per-vertex work-list CSR built once per graph: HBM
 *             streams 4 sum_e min(d+(u), d+(v)) bytes of lists.
 *   PUSH:     task u streams N+(v) for every v in N+(u): 4 sum_v d-(v) d+(v). */
#define TC_AUTO_WORKLIST 0
#define TC_AUTO_PUSH 1

/* k-clique variant. */
#define KCL_VARIANT_AUTO 0      /* k = 3: triangle engine; else pivot (d+ <= 1024) or merge */
#define KCL_VARIANT_BITMAP 1    /* LDS bitmap depth-first path (d+ <= 512, k <= 9)  */
#define KCL_VARIANT_MERGE 2     /* sorted-list intersection path                   */
#define KCL_VARIANT_PIVOT 3     /* pivoting over the clique tree, LDS bit matrix   */

typedef struct tc_ctx tc_ctx;
typedef struct tc_graph tc_graph;
typedef struct tc_batch tc_batch;

typedef struct tc_graph_info {
  uint32_t n;               /* vertex count (max id + 1, or after pruning)      */
  uint32_t max_out_degree;  /* max d+ of the oriented CSR                       */
  uint64_t m_raw;           /* records currently held as raw edges              */
  uint64_t m_oriented;      /* oriented unique edges (= undirected simple edges) */
  int32_t orient_mode;      /* TC_ORIENT_*, -1 before tc_graph_orient           */
  uint32_t n_raw;           /* every raw record id is below this (n differs
                               after pruning); A-order's perm has n_raw entries */
  uint64_t sum_out_sq;      /* sum_u d+(u)^2                                    */
  uint64_t sum_in_out;      /* sum_v d-(v) d+(v)  (= wedge count)               */
  double bytes_alg;         /* B_alg = 4(n+1) + 8m+ + 4 sum d+^2 + 4 sum d- d+  */
  double bytes_min;         /* 4(n+1) + 4m+                                     */
} tc_graph_info;

/* Per-call schedule.  Zero fields take the defaults listed. */
typedef struct tc_sched {
  int32_t kernel;           /* TC_KERNEL_*                                       */
  uint32_t block_size;      /* reserved (workgroups are 256 threads)             */
  uint32_t grid_size;       /* Hu: work-balanced blocks (default 30000);
                               Bisson/TriCore/AUTO: resident workgroups          */
  uint32_t lds_slice;       /* reserved                                          */
  uint32_t bitmap_min_deg;  /* Bisson: d+ above which the bitmap path runs (64)  */
  uint32_t fox_max_lanes;   /* Fox: widest lanes-per-edge bin, 1..64 (64)        */
  uint32_t tricore_parts;   /* TriCore: node-range partitions (1)                */
  uint32_t auto_bins;       /* reserved                                          */
  uint32_t auto_thresholds[4]; /* AUTO: [0] largest d+ run as a wave task (64,
                               at most 512); [1] largest d+ run by the 2048-entry
                               block kernel, above it the 8192-entry one (2048);
                               [2] schedule: TC_AUTO_WORKLIST (0, default) or
                               TC_AUTO_PUSH; [3] WORKLIST: largest d+ of a task
                               run by one lane (N+(x) in registers) when its
                               work list has at most 16 entries (8; at most 8;
                               1 turns lane tasks off)                         */
  uint32_t fox_rearrange;   /* Fox: 1 = rearrangeEdge (Fox/tc.cu:127-183,345):
                               each bin's edges sorted by max(d+(u), d+(v)) and
                               dealt in groups of up to 32 round-robin over the
                               launch's workgroups, so every workgroup gets
                               light and heavy edges; counts are unchanged     */
  uint32_t reserved[3];
} tc_sched;

typedef struct tc_timing {
  float ms_count;           /* hipEvent time of the count kernels only           */
  float ms_prep;            /* schedule building (work model, bins) on device    */
  uint32_t launches;        /* count-kernel launches                             */
  uint32_t dominant_kernel; /* index of the kernel with the most time            */
  float ms_dominant;        /* its hipEvent time                                 */
  float reserved;
} tc_timing;

/* ---- errors / context -------------------------------------------------- */
const char* tc_status_str(int status);
const char* tc_last_error(void);
/* Picks the device and creates the stream; replaces the per-program
 * cudaSetDevice + warm-up (Hu/tc.cu:209-219, tricore/src/tricount_gpu.cu:29-35). */
int tc_ctx_create(int device, tc_ctx** out);
int tc_ctx_destroy(tc_ctx* ctx);
/* Total device memory in bytes (tricore init_gpu(), tricount_gpu.cu:29-35). */
int tc_ctx_device_memory(const tc_ctx* ctx, uint64_t* bytes);

/* ---- L0: edge files (tools/list2bin.cpp, tools/bin2list.cpp) ---------- */
/* Record count = size/8; a size that is not a multiple of 8 is TC_ERR_FORMAT
 * (tricore/src/main.cpp:38-41). */
int tc_file_edge_count(const char* path, uint64_t* m);
/* Whole-file read into caller arrays of capacity cap (Hu/preTC.cu:68-86). */
int tc_read_edges(const char* path, uint32_t* src, uint32_t* dst, uint64_t cap,
                  uint64_t* m);
int tc_write_edges(const char* path, const uint32_t* src, const uint32_t* dst,
                   uint64_t m);
/* Text "src dst" lines <-> binary records (tools/list2bin.cpp:8-24,
 * tools/bin2list.cpp:8-22). */
int tc_list2bin(const char* text_path, const char* bin_path, uint64_t* m);
int tc_bin2list(const char* bin_path, const char* text_path, uint64_t* m);
/* Galois .gr binary CSR, versions 1 and 2 (the k-clique input,
 * k-clique/kcl.cu:149 reads it through Galois' CSRGraph::read): one record
 * per stored arc (u, v).  src = NULL queries *m (arcs) and *n (nodes). */
int tc_read_gr(const char* path, uint32_t* src, uint32_t* dst, uint64_t cap, uint64_t* m,
               uint32_t* n);

/* ---- seeded synthetic generators (deterministic for any thread count) -- */
/* Erdos-Renyi: exactly m unique undirected pairs, no self-loops. */
int tc_gen_er(uint32_t n, uint64_t m, uint64_t seed, uint32_t* src, uint32_t* dst);
/* R-MAT: nedges raw draws over 2^scale ids, quadrant probs (a,b,c,1-a-b-c). */
int tc_gen_rmat(uint32_t scale, uint64_t nedges, double a, double b, double c,
                uint64_t seed, uint32_t* src, uint32_t* dst);
/* Barabasi-Albert: n vertices, k attachments each; m_out = edges written
 * (capacity needed: n*k). */
int tc_gen_ba(uint32_t n, uint32_t k, uint64_t seed, uint32_t* src, uint32_t* dst,
              uint64_t* m_out);
/* Random geometric graph in the unit square, radius r.  Call with src=NULL
 * to get the edge count in *m_out, then again with arrays of that size. */
int tc_gen_rgg(uint32_t n, double r, uint64_t seed, uint32_t* src, uint32_t* dst,
               uint64_t* m_out);

/* ---- L1/L2: device graph ---------------------------------------------- */
/* Uploads raw records (host arrays) to HBM.  n = max id + 1 when n_hint is
 * 0 (Hu/preTC.cu:90-113); ids must be < n_hint otherwise. */
int tc_graph_create(tc_ctx* ctx, const uint32_t* src, const uint32_t* dst,
                    uint64_t m, uint32_t n_hint, tc_graph** out);
int tc_graph_destroy(tc_graph* g);
/* Replaces the raw records by the simple undirected edge set: self-loops
 * dropped, duplicates merged, each edge once as (min,max), sorted. */
int tc_graph_simplify(tc_ctx* ctx, tc_graph* g);
/* A-order relabelling of the raw records (vertex_ordering/A-order/A-order.cpp:62-230);
 * the records are rewritten with new ids.  perm (optional, n_raw entries as
 * tc_graph_info reports before the call) gets old->new.  The graph's n becomes the largest new id
 * + 1 over the records, as reading A-order's "reordered" file back gives
 * (edge_directing/A-direction-seri.cpp:187-193), so an A-direction peel after
 * it uses the reference chain's threshold.  TC_ERR_UNSUPPORTED outside the
 * reference's defined range. */
int tc_graph_relabel_aorder(tc_ctx* ctx, tc_graph* g, uint32_t* perm);
/* Bandwidth model behind A-order's bucket scores (A-order.cpp:20-36): the
 * bandwidth of reading a list of len entries is table[len] for len < 10 and
 * a * len^b (computed in double, rounded to float) from 10 on; the
 * "equal bandwidth" term is lambda[len] * sqrt(bw) / sqrt(len) for len < 10
 * and lambda_tail * sqrt(bw) / sqrt(len) from 10 on (float). */
typedef struct tc_bw_model {
  float table[10];
  double a, b;
  float lambda[10];
  float lambda_tail;
} tc_bw_model;
/* The reference's constants (table {.., 271, 310, ..}, 384.1 len^0.1491,
 * lambda {.., 145.51, ..}, 12), measured on the GPU the reference targets. */
int tc_bw_model_reference(tc_bw_model* out);
/* Achieved bandwidth (GB/s of list entries) of reading lists of each of the
 * count lengths lens[i] at random 4-byte offsets of a 1 GiB array, one wave
 * per list, the whole GPU busy (the access the counting kernels make to
 * N+(v)).  No counterpart in the reference, which ships measured numbers. */
int tc_measure_list_bandwidth(tc_ctx* ctx, const uint32_t* lens, uint32_t count, float* gbps);
/* Vector-ALU issue-rate probe (no counterpart in the reference; it prices
 * the issue-bound kernels): one timed launch of `iters` rounds of eight
 * independent integer-operation chains per lane, eight waves on every SIMD;
 * *ms receives its hipEvent time.  The wave instructions it issued come from
 * a profiled run (SQ_INSTS_VALU); their count over *ms is the GPU's
 * attainable issue rate for such code. */
int tc_measure_valu_rate(tc_ctx* ctx, uint32_t iters, float* ms);
/* A model fitted to such measurements: table[len] = the measured value for
 * 2 <= len <= 9 (all of them must be present), (a, b) the least-squares fit
 * of log gbps on log len over the lengths >= 10 (at least two), lambda and
 * lambda_tail the reference's (the reference gives no way to re-derive
 * them). */
int tc_bw_model_fit(const uint32_t* lens, const float* gbps, uint32_t count, tc_bw_model* out);
/* A-order with another bandwidth model; tc_graph_relabel_aorder uses the
 * reference's. */
int tc_graph_relabel_aorder_model(tc_ctx* ctx, tc_graph* g, const tc_bw_model* model,
                                  uint32_t* perm);
/* Id relabellings of the reference's format tools, applied to the raw
 * records on the GPU (records keep their order).  perm (optional) gets
 * old -> new for ids 0 .. max id (max id + 1 entries, max over the records
 * before the call).  The graph's n becomes the largest new id + 1.
 *   TC_RELABEL_COMPACT   tools/serialize.cpp:17-47: ids that occur keep their
 *                        order, ids that never occur are squeezed out
 *   TC_RELABEL_FIRST_SRC tools/adjust2tc.cpp:13-35: numbered by first
 *                        appearance as a source, then the rest by id
 *   TC_RELABEL_DEGREE    tools/directionByReorder.cpp:38-70 (and Bisson's
 *                        chooseIndex != 0, Bisson/preTC.cu:146-161): sorted
 *                        by (degree, id), the highest degree gets id 0 */
#define TC_RELABEL_COMPACT 0
#define TC_RELABEL_FIRST_SRC 1
#define TC_RELABEL_DEGREE 2
/*   TC_RELABEL_ADIR      Bisson/A-direction.cpp:287-303: the A-direction peel
 *                        of A-direction-seri.cpp (threshold 2m/n + 1 over
 *                        n = max id + 1), then ids sorted by (priority,
 *                        runtime degree when peeled, id), the last gets id 0,
 *                        so that the ID orientation (src = larger id, Bisson's
 *                        chooseIndex 0) points every edge up the peel order.
 *                        The degree recorded is the peeled vertex's own
 *                        (the reference stores it at the frontier index
 *                        instead, :118, which leaves cascade-peeled entries
 *                        unset). */
#define TC_RELABEL_ADIR 3
int tc_graph_relabel(tc_ctx* ctx, tc_graph* g, int mode, uint32_t* perm);
/* Orients the raw records and builds the oriented CSR in HBM
 * (Hu/preTC.cu:116-291, Fox/preTC.cu:88-203, tricore/src/tricount_gpu.cu:37-203,
 * edge_directing/A-direction-seri.cpp:35-273, edge_directing/kcore_direction.cpp:25-146). */
int tc_graph_orient(tc_ctx* ctx, tc_graph* g, int mode, uint32_t flags);
/* The OpenMP A-direction program (edge_directing/A-direction.cpp:36-161,
 * 197-198) restated deterministically: threshold0 = 2 totalEdge / n (float);
 * each round peels every unpeeled vertex whose runtime degree is within the
 * threshold, and cascades (as the serial peel does) only while the frontier
 * holds more than n / 5000 and more than 200 vertices, for at most six
 * levels; the threshold then grows by (int)threshold0 * thre_param / 5.
 * Edges point up (priority, id) (:259-268).  The reference runs this with
 * unsynchronised OpenMP threads, so its labels vary from run to run; this
 * is the one-thread order of its loops with a frontier of exactly the
 * vertices peeled at that step.  thre_param must be positive. */
int tc_graph_orient_adir_omp(tc_ctx* ctx, tc_graph* g, double thre_param);
/* Orientation tie rules (no counterpart in the reference, which breaks
 * every tie by id).  For TC_ORIENT_ADIR (ties: equal priority) and
 * TC_ORIENT_DEGREE (ties: equal raw degree) a tie key t[v] is compared
 * before the id:
 *   TC_TIE_ID           no key: the reference's rule, = tc_graph_orient;
 *   TC_TIE_FORCED_ASC   t = forced out-degree h(v), the neighbours strictly
 *                       above v in the primary order (ADIR: over the simple
 *                       graph; DEGREE: over the raw records), so vertices
 *                       with few forced out-edges take the tied edges;
 *   TC_TIE_FORCED_DESC  t = ~h(v), the reverse.
 * Every rule is a strict total order, so the orientation is acyclic and the
 * count is the same. */
#define TC_TIE_ID 0
#define TC_TIE_FORCED_ASC 1
#define TC_TIE_FORCED_DESC 2
#define TC_TIE_CANDIDATES 3
int tc_graph_orient_ties(tc_ctx* ctx, tc_graph* g, int mode, int tie);
/* The A-direction cost function of the current orientation: sum over the
 * CSR's rows of |d+(v) - m+/n| (tools/verification.cpp:43-45's f_lp, in
 * double; deterministic). */
int tc_graph_adir_cost(tc_ctx* ctx, tc_graph* g, double* cost);
int tc_graph_info_get(const tc_graph* g, tc_graph_info* info);
/* Copies of device state (host arrays of the documented sizes). */
int tc_graph_get_csr(const tc_graph* g, uint32_t* rowptr /*n+1*/, uint32_t* col /*m+*/);
int tc_graph_get_edges(const tc_graph* g, uint32_t* src, uint32_t* dst /*m_raw*/);
/* ADIR: peel priorities (A-direction-seri.cpp numbering); KCORE: core numbers;
 * DEGENERACY: removal steps. */
int tc_graph_get_vertex_key(const tc_graph* g, int32_t* key /*n*/);
/* Stage file of the oriented CSR: the preprocessing result (n, m+, the
 * orientation, rowptr, col and, after ADIR / KCORE / DEGENERACY, the vertex
 * keys), so a later run loads it instead of repeating the upload, sort and
 * peel that every reference program re-runs (Hu/preTC.cu:186-291).  A loaded
 * graph holds no raw records: every count works on it; simplify, relabel and
 * orient return TC_ERR_STATE.  load checks the file's structure (rows
 * ascending, loop-free, ids below n) and returns TC_ERR_FORMAT otherwise.
 * No counterpart in the reference. */
int tc_graph_save_csr(const tc_graph* g, const char* path);
int tc_graph_load_csr(tc_ctx* ctx, const char* path, tc_graph** out);
/* Device addresses of the CSR (for callers sharing HBM buffers). */
int tc_graph_device_csr(const tc_graph* g, const uint32_t** rowptr,
                        const uint32_t** col);

/* ---- L3: work models / schedules -------------------------------------- */
/* Hu per-vertex work W(u), Hu/tc.cu:22-35 (n entries). */
int tc_hu_workload(tc_ctx* ctx, tc_graph* g, int64_t* w);
/* Hu block start vertices, Hu/tc.cu:286-301 (nblocks+1 entries). */
int tc_hu_block_starts(tc_ctx* ctx, tc_graph* g, uint32_t nblocks, int32_t* start);
/* Fox per-edge work in CSR edge order, Fox/tc.cu:18-33 (m+ entries). */
int tc_fox_workload(tc_ctx* ctx, tc_graph* g, int32_t* w);
/* Fox bin sizes: count[j] = edges in the bin of 2^j lanes, j = 0..log2(max_lanes),
 * bins per Fox/tc.cu:316-331. */
int tc_fox_bin_sizes(tc_ctx* ctx, tc_graph* g, uint32_t max_lanes, uint64_t* count);

/* ---- orientation cost functions of the raw records -------------------
 * The values the reference's analysis tools print, from the raw records
 * (ids as read, nodes = max id + 1):
 *   out[0..2] = idPartCost, degreePartCost, lpPartCost of
 *               tools/costFuncCal.cpp:40-88 with threshold factor k;
 *   out[3..4] = f_lp, f_degree of tools/verification.cpp:36-60.
 * Counts run on the GPU; the float sums follow the tools' order exactly. */
int tc_cost_functions(tc_ctx* ctx, tc_graph* g, int k, float* out /*5*/);
/* Approximation ratio bound of tools/approximateRatio.cpp:37-94. */
int tc_approx_ratio(tc_ctx* ctx, tc_graph* g, double* ratio);

/* ---- L4/L5: counting --------------------------------------------------- */
/* Exact triangle count of the oriented CSR (uint64; the reference prints it as
 * "There are N triangles in the input graph.", Hu/tc.cu:337).  sched may be
 * NULL (AUTO with defaults).  timing may be NULL. */
int tc_count(tc_ctx* ctx, tc_graph* g, const tc_sched* sched, uint64_t* count,
             tc_timing* timing);
/* Bytes one AUTO count streams under sched (NULL: defaults): for
 * TC_AUTO_WORKLIST the sum over tasks x of 20 + 4 d+(x) (task id, row bounds
 * in both CSRs, N+(x)) and over the work list of x of 8 + 4 d+(y) (the row
 * bounds of y, stored with the entry, and N+(y)); for TC_AUTO_PUSH 4(n+1) + 12 m+ + 4 sum_v d-(v) d+(v).
 * Builds the schedule when needed.  No counterpart in the reference. */
int tc_auto_stream_bytes(tc_ctx* ctx, tc_graph* g, const tc_sched* sched, double* bytes);
/* Per-vertex forward counts tri[u] (n entries): the triangles whose first
 * vertex (the tail of the two other arcs) is u, from the AUTO kernel's block
 * and wave tasks with default thresholds.  They sum to tc_count's total. */
int tc_count_per_vertex(tc_ctx* ctx, tc_graph* g, uint64_t* tri);

/* ---- counting without device allocation --------------------------------
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): the exact triangle count of an oriented CSR the caller
 * already holds in device memory (rowptr[n+1], col[rowptr[n]], every row
 * ascending and without repeats, rowptr[n] below 2^32 - 512, as
 * tc_graph_device_csr gives), written to
 * the caller's device uint64 *count.  The work is enqueued on ctx's stream
 * and not waited for.  The kernels are tc_count's AUTO push form without
 * its task lists.  Their scratch memory (two claim cursors and one partial
 * count per workgroup) is a static device array of the library,
 * TC_NOALLOC_SCRATCH_BYTES on each device, held until the process ends; its
 * address is looked up when a context is created on the device.  grid =
 * workgroups (0: 8 per CU, at most TC_NOALLOC_MAX_GRID); a larger grid is
 * TC_ERR_RANGE.  Every call on one device uses that device's array, and the
 * cursors are cleared at the start of each call: calls that overlap on one
 * device (other streams, other contexts) race and must be serialised. */
#define TC_NOALLOC_SCRATCH_BYTES 65536
#define TC_NOALLOC_MAX_GRID 4096
int tc_count_csr_noalloc(tc_ctx* ctx, const uint32_t* rowptr, const uint32_t* col, uint32_t n,
                         uint32_t grid, unsigned long long* count);

/* ---- learned scheduling of the AUTO engine ------------------------------
 * No counterpart in the reference.  The reference's schedules come from
 * fixed work formulas (Hu/tc.cu:22-35, Fox/tc.cu:18-33,316-331) and its
 * orientation quality from fixed cost functions (tools/costFuncCal.cpp,
 * tools/verification.cpp).  Here a small MLP (8 features -> 32 -> 32 -> 2,
 * ReLU, BF16 weights, FP32 accumulate) chooses the per-bin variant of AUTO's
 * work-list schedule (task x holds N+(x) and streams N+(y) for y in W(x)).
 * The bins ("cells") are d+(x) = 2..8 (cells 2..8) and d+ in (2^(b-1), 2^b]
 * for b = 4..9 (cells 9..14, up to d+ 512).  A cell's tasks run either their
 * narrow variant (a lane for d+ <= 8, a wave above) or their wide one (a wave
 * for d+ <= 8, a workgroup above); the fixed rule runs cells 2..11
 * (d+ <= 64) narrow, which sets AUTO's bin thresholds.  Whatever the cell
 * runs, a lane task keeps at most 16 streamed lists and a wave task at most
 * 2^16 entries of work (else it runs one size up), and tasks above d+ 512
 * run as workgroups.  The model predicts, from 8 features of a cell's tasks,
 * the log2 microseconds of the cell's count launch under each variant; it is
 * trained on launches the engine times on itself (tc_graph_measure_cells).
 * Cell 15 (the tasks above d+ 512, always workgroups) is timed and predicted
 * wide only, so the model predicts a whole count; learned orientation ties
 * compare candidate orientations by that prediction.
 * The schedule only decides HOW each task runs; every task runs exactly
 * once, so the count is exact whatever the model predicts. */
#define TC_MODEL_FEATURES 8
#define TC_MODEL_OUTPUTS 2
#define TC_VARIANT_BLOCK 0   /* a workgroup per task                          */
#define TC_VARIANT_WAVE 1    /* a wave per task, d+ <= 512                     */
#define TC_VARIANT_LANE 2    /* a lane per task, d+ <= 8 (N+(x) in registers)  */
#define TC_CELLS 16          /* cell arrays are indexed by cell, 0 and 1 unused;
                                cell 15 holds the tasks above d+ 512 (no choice) */
#define TC_CELL_MAX_LEVEL 16 /* sampling levels of the cell calls               */

typedef struct tc_model tc_model;

typedef struct tc_train_opts {
  uint32_t steps;         /* Adam steps (default 3000)          */
  uint32_t batch;         /* samples per step (default 4096)    */
  float lr;               /* learning rate (default 3e-3)       */
  uint32_t seed;          /* sample order seed                  */
  uint32_t reserved[4];
} tc_train_opts;

uint32_t tc_model_num_params(void);
int tc_model_create(tc_ctx* ctx, uint64_t seed, tc_model** out);
int tc_model_destroy(tc_model* m);
/* fp32 master parameters: W1[8][32], b1[32], W2[32][32], b2[32], W3[32][2], b3[2]
 * (W[i][j] connects input i to output j). */
int tc_model_set_params(tc_model* m, const float* params, uint32_t n);
int tc_model_get_params(const tc_model* m, float* params, uint32_t n);
/* Feature / target standardisation: x' = (x - x_mean) / x_std, y = y' * y_std + y_mean. */
int tc_model_set_norm(tc_model* m, const float* x_mean, const float* x_std,
                      const float* y_mean, const float* y_std);
int tc_model_get_norm(const tc_model* m, float* x_mean, float* x_std, float* y_mean,
                      float* y_std);
/* Forward pass on host arrays: X[N][8] raw features -> Y[N][2] log2 cost (us). */
int tc_model_forward(tc_ctx* ctx, tc_model* m, const float* X, uint32_t N, float* Y);
/* Masked mean-squared error in standardised target units and its gradient
 * with respect to the fp32 master parameters (the forward pass uses their
 * BF16 rounding).  T[N][2] raw log2 costs, M[N][2] in {0,1}. */
int tc_model_loss_grad(tc_ctx* ctx, tc_model* m, const float* X, const float* T,
                       const float* M, uint32_t N, float* loss, float* grad);
/* Sets the standardisation from the data, then trains with Adam on the GPU. */
int tc_model_train(tc_ctx* ctx, tc_model* m, const float* X, const float* T, const float* M,
                   uint32_t N, const tc_train_opts* opts, float* final_loss);
/* The list-read bandwidth model behind feature 4 (default: the reference's,
 * tc_bw_model_reference; the MI355X one is measured by
 * tc_measure_list_bandwidth / tc_bw_model_fit). */
int tc_model_set_bw_model(tc_model* m, const tc_bw_model* bw);
/* Per-task figures (diagnostics; the schedule uses the cell calls below).
 * Per-task features X[n][8] of the work-list schedule (0 for vertices with
 * no task: d+ < 2 or an empty work list): log2(1+d+(x)), log2(1+|W(x)|),
 * log2(1+sum d+(y)), log2(1+max d+(y)), log2(1+t) with t the nanoseconds
 * the lists N+(y) take to read under the model's bandwidth model (the
 * reference's when m is NULL), floor(log2(1+d+(x))), the fraction of y with
 * d+(y) < 64, log2(1+sum ceil(d+(y)/64)). */
int tc_graph_features(tc_ctx* ctx, tc_graph* g, const tc_model* m, float* X);
/* Measured cost of every task under one variant (TC_VARIANT_*), in
 * wave-microseconds: a block task's time x 4 (it holds four waves), a wave
 * task's time, a lane task's round / 64; from a timed run of that variant
 * over every task it can run (wave: d+ <= 512, lane: d+ <= 8); 0 where not
 * measured.  count (optional) receives the run's exact count. */
int tc_graph_measure_costs(tc_ctx* ctx, tc_graph* g, int variant, float* cost_us,
                           uint64_t* count);
/* Cell features X[TC_CELLS][8] over the tasks of sampling level `level`
 * (level 0: all tasks; level s: those whose id hash h(x) = (x * 0x9E3779B1)
 * xor'ed with its upper 16 bits has s low zero bits): log2(1+tasks),
 * log2(1+sum |W(x)|), log2(1+sum of streamed entries d+(y)), log2(1+sum
 * d+(x)), log2(1+t) with t the nanoseconds the streamed lists take under
 * the model's bandwidth model (the reference's when m is NULL; per task
 * rounded to picoseconds, then summed), the fraction of streamed lists with
 * d+(y) < 64, log2(1+largest task work d+(x) + sum d+(y)), the fraction of
 * tasks within their narrow variant's limit.  Empty cells are all zero. */
int tc_graph_cell_features(tc_ctx* ctx, tc_graph* g, const tc_model* m, uint32_t level,
                           float* X);
/* Times each non-empty cell's count launch alone, over the tasks of
 * `level`, with the cell run narrow (narrow != 0; not cell 15) or wide: us[cell] = the
 * fastest of `reps` (1..64) timed launches; counts[cell] (optional) = the
 * triangles those tasks count (the same in every repetition, else
 * TC_ERR_STATE).  Entries of empty cells are 0. */
int tc_graph_measure_cells(tc_ctx* ctx, tc_graph* g, int narrow, uint32_t level, uint32_t reps,
                           float* us, uint64_t* counts);
/* The model's choice: *narrow_cells = bit mask of the cells run narrow, Y
 * (optional, TC_CELLS x 2) = its predictions [wide, narrow] from the cells'
 * level-0 features (rows of empty cells are not used).  The
 * fixed rule (cells 2..11) is the prior: a non-empty cell switches only
 * when the other variant is predicted more than 10 % faster.  With m NULL,
 * the fixed rule's mask and Y = 0. */
int tc_learned_cells(tc_ctx* ctx, tc_graph* g, tc_model* m, uint32_t* narrow_cells, float* Y);
/* Learned orientation ties (TC_ORIENT_ADIR or TC_ORIENT_DEGREE): orients g
 * under every tie rule TC_TIE_* in turn and scores each: predicted_us[t] =
 * the model's predicted count time (tc_learned_cells' total over all cells)
 * and adir_cost[t] = the A-direction cost function (tc_graph_adir_cost).
 * With a model the lowest prediction wins, but a rule other than
 * TC_TIE_ID only for a predicted saving above 2 %; with m NULL the lowest
 * A-direction cost wins (the fallback).  g is left oriented under the
 * chosen rule (*chosen).  Both score arrays (optional) have
 * TC_TIE_CANDIDATES entries; predicted_us is 0 without a model. */
int tc_graph_orient_learned(tc_ctx* ctx, tc_graph* g, int mode, tc_model* m, int* chosen,
                            double* predicted_us, double* adir_cost);
/* Variant chosen for each vertex's task (TC_VARIANT_*, 255 = no task): the
 * cell's choice within the task's limits (above).  model NULL gives the
 * fixed rule's variants (tc_count's AUTO work-list schedule). */
int tc_learned_assignment(tc_ctx* ctx, tc_graph* g, tc_model* m, uint8_t* variant);
/* Exact count with the learned schedule (block tasks by descending work, as
 * the fixed rule orders them; wave and lane tasks in id order); model NULL
 * counts with the fixed rule (tc_count, TC_KERNEL_AUTO). */
int tc_count_learned(tc_ctx* ctx, tc_graph* g, tc_model* m, uint64_t* count,
                     tc_timing* timing);

/* ---- batched counting of many small graphs -----------------------------
 * No counterpart in the reference (one graph per process run).  A batch is
 * the disjoint union of its graphs in one degree-oriented CSR (orientation
 * rules are local, so this equals orienting each graph on its own); one
 * persistent kernel counts every graph over AUTO's work lists (TC_AUTO_WORKLIST:
 * an edge never leaves its graph, so neither does a work list) and
 * accumulates per-graph uint64 counts.  The count sequence is captured once
 * in a hipGraph and replayed. */
typedef struct tc_batch_sched {
  uint32_t wave_max_deg;  /* vertices with 2 <= d+ <= this run as wave tasks (64) */
  uint32_t wave_chunk;    /* wave tasks fetched per cursor bump (16)              */
  uint32_t grid;          /* persistent workgroups (0 = 8 per CU)                  */
  uint32_t filter_min;    /* push form: wave tasks hash N+(u) into a bitmap filter
                             above this d+: 1 or 16 (0 = 16); the work-list form
                             decides every entry by a rank-indexed table of N+(u) */
  uint32_t form;          /* TC_BATCH_WORKLIST (0): tasks stream AUTO's work
                             lists; TC_BATCH_PUSH (1): task u streams N+(v) for
                             v in N+(u)                                           */
  uint32_t reserved[3];
} tc_batch_sched;
#define TC_BATCH_WORKLIST 0
#define TC_BATCH_PUSH 1

/* Graph g holds records [edge_off[g], edge_off[g+1]) of src/dst with local
 * ids < nverts[g]; edge_off has ngraphs + 1 ascending entries, edge_off[0] = 0.
 * Orientation: TC_ORIENT_DEGREE per graph. */
int tc_batch_create(tc_ctx* ctx, uint32_t ngraphs, const uint64_t* edge_off,
                    const uint32_t* src, const uint32_t* dst, const uint32_t* nverts,
                    tc_batch** out);
int tc_batch_destroy(tc_batch* b);
/* Sizes of the batch's union CSR (n, m+, algorithmic bytes of one count). */
int tc_batch_info_get(const tc_batch* b, tc_graph_info* info);
/* Bytes one batched count streams in the work-list form: the work-list
 * bytes of the union CSR (tc_auto_stream_bytes), whatever the task split.
 * The push form streams 4(n+1) + 12 m+ + 4 sum d- d+ (tc_batch_info_get). */
int tc_batch_stream_bytes(tc_ctx* ctx, tc_batch* b, double* bytes);
/* For each form, tries the wave_max_deg candidates, then filter_min, then
 * the wave_chunk candidates (each stage from the fastest so far) on the
 * batch, and returns the fastest schedule of either form (timed with
 * hipEvents). */
int tc_batch_autotune(tc_ctx* ctx, tc_batch* b, tc_batch_sched* best);
/* Exact per-graph counts (ngraphs entries) from one replay of the captured
 * hipGraph (capture happens on the first call per schedule). */
int tc_batch_count(tc_ctx* ctx, tc_batch* b, const tc_batch_sched* sched, uint64_t* counts,
                   tc_timing* timing);
/* Same counts with one kernel launch per graph (the per-graph baseline). */
int tc_batch_count_pergraph(tc_ctx* ctx, tc_batch* b, const tc_batch_sched* sched,
                            uint64_t* counts, tc_timing* timing);

/* ---- k-clique (k-clique/kcl.cu:16-215) -------------------------------- */
/* Exact k-clique count over the graph's oriented CSR.  KCL_VARIANT_AUTO on
 * an orientation other than TC_ORIENT_DEGENERACY that leaves roots with
 * d+ > 1024 counts on a degeneracy-oriented copy of the same records when
 * that copy is narrower (the count does not depend on the acyclic
 * orientation); the copy is built once, kept with the graph, and its build
 * time is added to timing->ms_prep. */
int kcl_count(tc_ctx* ctx, tc_graph* g, uint32_t k, int variant, uint64_t* count,
              tc_timing* timing);
/* k-cliques whose source in the DAG (the lowest vertex in the orientation
 * order) is one of roots[0 .. nroots), each listed root counted once, on
 * the graph's own orientation (3 <= k <= 12).  Disjoint root sets add up to
 * kcl_count's total: the unit of work to split one count over several calls
 * or devices.  No counterpart in the reference. */
int kcl_count_roots(tc_ctx* ctx, tc_graph* g, uint32_t k, const uint32_t* roots,
                    uint64_t nroots, uint64_t* count, tc_timing* timing);
/* File-level form of kcl_gpu_solver (k-clique/kcl.h:6-7): reads a Galois
 * .gr file (path ending in ".gr", the input of k-clique/kcl_gpu.cpp:22-27)
 * or a list2bin file, degeneracy orientation (TC_ORIENT_DEGENERACY),
 * counts k-cliques.  n_chunk > 1 counts the DAG in n_chunk parts one after
 * the other, as N_CHUNK splits its edges (k-clique/kcl.cu:163-209): the roots
 * whose out-edges start in each of n_chunk equal ranges of the CSR's edges
 * form one part (a root's edges are never split); 0 is TC_ERR_ARG. */
int kcl_gpu_solver(const char* path, uint32_t k, uint64_t* total, size_t n_chunk);

/* ---- learned k-clique plan --------------------------------------------
 * No counterpart in the reference (k-clique/kcl.cu runs one path for every
 * root).  The cost model of tc_model_* (8 features -> 2 log2 costs) is
 * trained on per-root costs measured by the kernels themselves and picks,
 * per root, the LDS bit-matrix pivot path or the merge path.  The count is
 * exact whatever the model predicts. */
/* Per-root features X[n][8] for k (zero for roots with d+ < k-1):
 * log2(1+d+), log2(1+m) with m the edges of the sub-graph of N+(u),
 * density m / C(d+,2), k, log2(1+sum_v d+(v)), log2(C(d+,k-1) p^C(k-1,2))
 * (the (k-1)-cliques of a random graph of density p), the bit-matrix width
 * class (0: d+ <= 64, 1: <= 128, 2: <= 256, 3: <= 512, 4: <= 1024, 5: above),
 * log2(1+max_v d+(v)). */
int kcl_root_features(tc_ctx* ctx, tc_graph* g, uint32_t k, float* X);
/* Measured cost of every root on one path, in CU-microseconds (the root's
 * time times the share of a CU its workgroup holds); 0 where not measured.
 * KCL_VARIANT_PIVOT measures the roots with k-1 <= d+ <= 1024,
 * KCL_VARIANT_MERGE every root with d+ >= k-1; count (optional) receives
 * the k-cliques of the measured roots. */
int kcl_measure_costs(tc_ctx* ctx, tc_graph* g, uint32_t k, int variant, float* cost_us,
                      uint64_t* count);
/* Path per root (0 pivot, 1 merge, 255 none: d+ < k-1).  The fixed rule
 * (pivot iff d+ <= 1024) is the prior: a root moves to the merge path only
 * for a predicted saving above 25 %. */
int kcl_learned_plan(tc_ctx* ctx, tc_graph* g, uint32_t k, tc_model* m, uint8_t* path);
/* Exact k-clique count (3 <= k <= 12) with the learned plan; each launch
 * takes its roots widest first, as kcl_count.  The plan is built once per
 * (graph, k, model parameters) and kept with the graph.  m NULL runs
 * kcl_count(KCL_VARIANT_AUTO).  Like kcl_count, it counts on the kept
 * degeneracy-oriented copy when the graph's orientation leaves roots with
 * d+ > 1024.  timing->ms_prep holds the build time of the plan (and of the
 * copy), 0 when both were kept. */
int kcl_count_learned(tc_ctx* ctx, tc_graph* g, uint32_t k, tc_model* m, uint64_t* count,
                      tc_timing* timing);

/* ---- TriCore function API (tricore/include/tricount.h:7-17) ----------- */
/* Same signatures and meaning as the reference; edges are the uint64 view of
 * the record, FIRST = high word (tricore/include/util.h:14-16).  Failures are
 * reported through tc_last_error() and a 0 / no-op result. */
uint64_t tc_tricore_init_gpu(void);
void tc_tricore_cal_degree(const uint64_t* edges, uint64_t edge_count, uint32_t* deg,
                           uint32_t node_count);
void tc_tricore_redirect_edges(uint64_t* edges, uint64_t edge_count, const uint32_t* deg,
                               uint32_t node_count);
uint32_t tc_tricore_cal_part_num(const uint64_t* edges, uint64_t edge_count,
                                 uint32_t node_count);
void tc_tricore_split(uint64_t* edges, uint64_t edge_count, uint64_t part_num,
                      uint32_t** node_index, const uint32_t* node_split,
                      uint64_t* adj_count);
uint64_t tc_tricore_tricount_gpu(uint64_t part_num, uint32_t* adj,
                                 const uint64_t* adj_count, const uint32_t* node_split,
                                 uint32_t** node_index);

#ifdef __cplusplus
}
#endif
#endif /* TC_AMD_H */
