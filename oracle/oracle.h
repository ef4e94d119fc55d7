/*
 * CPU oracle for the triangle-counting / k-clique path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is a plain, serial restatement of
 * the reference's CPU-side algorithms.  It is linked or loaded by tests/,
 * by __graft_entry__.smoke() and by bench.py's cpu_baseline leg, and only as
 * the checker.  The product library (libtcamd.so) never links, loads or
 * calls it.
 *
 * Every function names the reference file:line it restates.  Pinning status
 * (what independent evidence checks each function) is listed next to it:
 *   pinned: brute force / closed form / networkx / oracle/_ref build;
 *   invariants: only structural invariants and hand-traced cases pin it.
 *
 * Vertex ids are uint32; counts are uint64.  Edge lists are two parallel
 * arrays (src[i], dst[i]), the record layout of tools/list2bin.cpp:3-6.
 */
#pragma once
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* n = max id + 1 (0 for an empty list).  Hu/preTC.cu:90-113.  pinned: numpy. */
uint32_t orc_num_nodes(const uint32_t* src, const uint32_t* dst, uint64_t m);

/* Degree over the raw list, duplicates and self-loops counted.
 * Hu/preTC.cu:116-123.  pinned: numpy bincount. */
void orc_degree_raw(const uint32_t* src, const uint32_t* dst, uint64_t m,
                    uint32_t n, uint32_t* deg);

/* Orientation rules, applied record by record (out may alias in).
 *  mode 0  ID       : src = larger id            Fox/preTC.cu:92-101
 *  mode 1  DEGREE   : Hu / A-order rule           Hu/preTC.cu:125-133,
 *                     swap if deg[s]>deg[d] || (deg[s]==deg[d] && s<d)
 *  mode 4  DEG_TRI  : TriCore rule                tricore/src/tricount_gpu.cu:69-80,
 *                     swap if deg[s]>deg[d] || (deg[s]==deg[d] && s>d)
 * pinned: numpy restatement + count invariance. */
void orc_orient_simple(const uint32_t* src, const uint32_t* dst, uint64_t m,
                       const uint32_t* deg, int mode,
                       uint32_t* osrc, uint32_t* odst);

/* Hu's degree<=1 pruning and id compaction, Hu/preTC.cu:135-169.
 * deg: raw degree.  Writes new ids in place; an edge touching a degree<=1
 * vertex gets src (or dst) = 0xFFFFFFFF ("wiped").  Returns the new n.
 * pinned: hand cases + count invariance. */
uint32_t orc_prune_deg1(uint32_t* src, uint32_t* dst, uint64_t m, uint32_t n,
                        const uint32_t* deg);

/* Oriented CSR: sort by (src,dst), drop self-loops and wiped records, dedupe.
 * Hu/preTC.cu:225-291, tricore/src/tricount_gpu.cu:171-174.
 * rowptr has n+1 entries, col has room for m.  Returns m+ (oriented edges).
 * pinned: numpy unique. */
uint64_t orc_build_csr(const uint32_t* src, const uint32_t* dst, uint64_t m,
                       uint32_t n, uint32_t* rowptr, uint32_t* col);

/* Simple symmetric adjacency (both directions, sorted, unique, no loops),
 * edge_directing/A-direction-seri.cpp:214-229.  col needs room for 2m.
 * Returns 2*m_simple.  pinned: numpy unique. */
uint64_t orc_build_sym(const uint32_t* src, const uint32_t* dst, uint64_t m,
                       uint32_t n, uint32_t* rowptr, uint32_t* col);

/* A-direction peel priorities, edge_directing/A-direction-seri.cpp:35-153,
 * run on the simple symmetric adjacency (the reference's originEdges holds
 * both directions, :242-252).  prio[v] >= 0.
 * invariants only (the reference needs Boost and cannot be built here). */
void orc_adir_priority(uint32_t n, const uint32_t* sym_rowptr,
                       const uint32_t* sym_col, int32_t* prio);
/* Bisson/A-direction.cpp:180-304: A-direction peel, then ids ranked by
 * (priority, runtime degree at peel, id), the last rank gets id 0. */
/* edge_directing/A-direction.cpp:36-161 (OpenMP program, one-thread order,
 * each frontier exactly the vertices peeled at its step); thre_param > 0. */
void orc_adir_omp_priority(uint32_t n, const uint32_t* sym_rowptr, const uint32_t* sym_col,
                           double thre_param, int32_t* prio);
void orc_relabel_adir(const uint32_t* src, const uint32_t* dst, uint64_t m, uint32_t* perm);

/* A-direction emit rule, edge_directing/A-direction-seri.cpp:254-273:
 * edge i->j iff prio[j] > prio[i] || (prio[j]==prio[i] && i<j).
 * Writes (osrc,odst) pairs in the reference's output order.  Returns m+. */
uint64_t orc_adir_orient(uint32_t n, const uint32_t* sym_rowptr,
                         const uint32_t* sym_col, const int32_t* prio,
                         uint32_t* osrc, uint32_t* odst);

/* Core numbers (Batagelj-Zaversnik), edge_directing/kcore_direction.cpp:91-126
 * with the degree loop over edges as intended (:35 loops over nodes).
 * pinned: networkx.core_number. */
void orc_kcore(uint32_t n, const uint32_t* sym_rowptr, const uint32_t* sym_col,
               uint32_t* core);

/* k-core orientation, edge_directing/kcore_direction.cpp:132-146:
 * src is the lower (core, id).  Emits each undirected simple edge once. */
uint64_t orc_kcore_orient(uint32_t n, const uint32_t* sym_rowptr,
                          const uint32_t* sym_col, const uint32_t* core,
                          uint32_t* osrc, uint32_t* odst);

/* Removal steps of the level-synchronous k-core peel (the engine's
 * degeneracy order; edge_directing/kcore_direction.cpp:91-126 peels the same
 * levels one vertex at a time).  Level k starts at 0 and each level then
 * begins at max(k + 1, least remaining degree).  Within a level, step t
 * removes every remaining vertex whose degree among the remaining vertices
 * is <= k; steps repeat until none is left at that degree.  Steps are
 * numbered 0, 1, ... over the whole peel; core[v] gets the level.  Orient
 * with orc_kcore_orient(step) (lower (step, id) is the source).
 * pinned: core == orc_kcore (networkx.core_number), out-degree <= core on
 * every vertex, hand-traced paths / stars / cliques; the step numbering is
 * this engine's own (no reference counterpart). */
void orc_degeneracy(uint32_t n, const uint32_t* sym_rowptr, const uint32_t* sym_col,
                    uint32_t* step, uint32_t* core);

/* A-order relabelling, vertex_ordering/A-order/A-order.cpp:62-230.
 * Inputs: raw edge list.  Outputs: perm[old] = new id (the badLookUpTable of
 * :206-210) and the oriented edges (the "normal" file, :219-224).
 * Returns 0, or -1 where the reference's bucket loops leave their defined
 * range: it needs 1 <= (#out-degree 2..4 vertices) <= 40 * (#buckets).
 * pinned: oracle/_ref/A-order on small inputs. */
int orc_aorder(const uint32_t* src, const uint32_t* dst, uint64_t m, uint32_t n,
               uint32_t* perm, uint32_t* osrc, uint32_t* odst);

/* Forward (edge-iterator) merge counter over an oriented CSR:
 * sum over u, v in N+(u) of |N+(u) & N+(v)|.  This is the count every
 * reference kernel computes (Hu/tc.cu:143-187, Bisson/tc.cu:42-86,
 * Fox/tc.cu:84-105, tricore/src/tricount_gpu.cu:248-286).
 * pinned: brute force, closed forms, networkx. */
uint64_t orc_count(uint32_t n, const uint32_t* rowptr, const uint32_t* col);

/* Same counter with OpenMP over vertices (labelled "omp" wherever timed). */
uint64_t orc_count_omp(uint32_t n, const uint32_t* rowptr, const uint32_t* col);

/* Forward counter restricted to rows u in [u_begin, u_end) (timed samples). */
uint64_t orc_count_rows(uint32_t n, const uint32_t* rowptr, const uint32_t* col,
                        uint32_t u_begin, uint32_t u_end);

/* Forward counter over rows start, start+stride, ... (timed CPU samples). */
uint64_t orc_count_rows_strided(uint32_t n, const uint32_t* rowptr, const uint32_t* col,
                                uint32_t start, uint32_t stride);

/* orc_count_rows_strided with OpenMP over the sampled rows, `threads`
 * threads (labelled as parallel wherever timed). */
uint64_t orc_count_rows_strided_omp(uint32_t n, const uint32_t* rowptr, const uint32_t* col,
                                    uint32_t start, uint32_t stride, int threads);

/* Per-vertex triangle counts of the forward counter: tri[u] = sum over
 * v in N+(u) of |N+(u) & N+(v)|. */
void orc_count_per_vertex(uint32_t n, const uint32_t* rowptr,
                          const uint32_t* col, uint64_t* tri);

/* k-clique count over a DAG (any acyclic orientation), k >= 2.
 * Semantics of k-clique/kcl.cu:16-50,147-215 (every k-clique counted once).
 * pinned: C(n,k) on K_n, Turan graphs, networkx clique enumeration. */
uint64_t orc_kclique(uint32_t n, const uint32_t* rowptr, const uint32_t* col,
                     uint32_t k);
/* The same count restricted to the roots start, start + stride, ... (a
 * bounded sample for timing; the rows partition the cliques by root). */
uint64_t orc_kclique_rows_strided(uint32_t n, const uint32_t* rowptr, const uint32_t* col,
                                  uint32_t k, uint32_t start, uint32_t stride);
/* k-cliques rooted at each listed row (lowest vertex = rows[i]); counts[i]
 * per row, ids >= n give 0.  OpenMP over the list when threads > 1. */
void orc_kclique_row_list(uint32_t n, const uint32_t* rowptr, const uint32_t* col, uint32_t k,
                          const uint32_t* rows, uint64_t nrows, uint64_t* counts, int threads);

/* Hu work model, Hu/tc.cu:22-35: W(u) = (sum_{v in N+(u)} d+(v)) *
 * (unsigned)log2(d+(u)); 0 when d+(u) <= 1.  pinned: closed form. */
void orc_hu_workload(uint32_t n, const uint32_t* rowptr, const uint32_t* col,
                     int64_t* w);

/* Hu block partition, Hu/tc.cu:36-67,286-301: inclusive prefix of W, then
 * start[i] = binarySearchValue(prefix, i*step+1, n, -1) for 0<i<nblocks,
 * start[0]=0, start[nblocks]=n.  start has nblocks+1 entries. */
void orc_hu_block_starts(uint32_t n, const int64_t* w, uint32_t nblocks,
                         int32_t* start);

/* Fox per-edge work, Fox/tc.cu:18-33: int(min(a,b) * log2(max(a,b)+2)). */
void orc_fox_workload(uint64_t m, const uint32_t* esrc, const uint32_t* edst,
                      const uint32_t* adjlen, int32_t* w);

/* Fox bins, Fox/tc.cu:316-331, over ascending work w[0..size).
 * max_lanes = 32 in the reference.  bin_start gets max_lanes+2 entries with
 * the reference's indexing (bin of i lanes = [bin_start[i], bin_start[i+1])). */
void orc_fox_bins(const int32_t* w, uint32_t size, uint32_t max_lanes,
                  uint32_t* bin_start);

/* Orientation cost functions (float, reference summation order):
 * tools/costFuncCal.cpp:15-88 -> out[0..2] = idPartCost, degreePartCost,
 * lpPartCost (k = argv[2]). */
void orc_cost_func(const uint32_t* src, const uint32_t* dst, uint64_t m,
                   int k, float* out);
/* tools/verification.cpp:11-60 -> out[0] = f_lp, out[1] = f_degree. */
void orc_verification(const uint32_t* src, const uint32_t* dst, uint64_t m,
                      float* out);
/* Id relabellings of the format tools; perm (max id + 1 entries) gets
 * old -> new.  pinned: the reference tools built in oracle/_ref.
 *   tools/serialize.cpp:17-47           present ids compacted in order
 *   tools/adjust2tc.cpp:13-35           first appearance as a source first
 *   tools/directionByReorder.cpp:38-70  by (degree, id), highest degree -> 0 */
void orc_relabel_compact(const uint32_t* src, const uint32_t* dst, uint64_t m, uint32_t* perm);
void orc_relabel_first_src(const uint32_t* src, const uint32_t* dst, uint64_t m,
                           uint32_t* perm);
void orc_relabel_degree(const uint32_t* src, const uint32_t* dst, uint64_t m, uint32_t* perm);
/* tools/approximateRatio.cpp:14-94 -> approximation ratio. */
double orc_approx_ratio(const uint32_t* src, const uint32_t* dst, uint64_t m);

#ifdef __cplusplus
}
#endif
