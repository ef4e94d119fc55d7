"""ctypes binding of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, always as the checker.  The engine never uses it.
See oracle.h for what each function restates and how it is pinned.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# ORACLE_LIB selects another build of the same source (the sanitizer build
# of the host sanitizer test)
LIB = os.environ.get("ORACLE_LIB") or os.path.join(HERE, "liboracle.so")
REF_DIR = os.path.join(HERE, "_ref")


def build(quiet: bool = True) -> None:
    subprocess.run(["make", "-C", HERE, "oracle"], check=True,
                   stdout=subprocess.DEVNULL if quiet else None)


if not os.path.exists(LIB):
    build()

_lib = C.CDLL(LIB)
_u32p = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
_u64p = np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS")
_f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")


def _sig(name, res, args):
    f = getattr(_lib, name)
    f.restype = res
    f.argtypes = args


_sig("orc_num_nodes", C.c_uint32, [_u32p, _u32p, C.c_uint64])
_sig("orc_degree_raw", None, [_u32p, _u32p, C.c_uint64, C.c_uint32, _u32p])
_sig("orc_orient_simple", None, [_u32p, _u32p, C.c_uint64, _u32p, C.c_int, _u32p, _u32p])
_sig("orc_prune_deg1", C.c_uint32, [_u32p, _u32p, C.c_uint64, C.c_uint32, _u32p])
_sig("orc_build_csr", C.c_uint64, [_u32p, _u32p, C.c_uint64, C.c_uint32, _u32p, _u32p])
_sig("orc_build_sym", C.c_uint64, [_u32p, _u32p, C.c_uint64, C.c_uint32, _u32p, _u32p])
_sig("orc_adir_priority", None, [C.c_uint32, _u32p, _u32p, _i32p])
_sig("orc_adir_omp_priority", None, [C.c_uint32, _u32p, _u32p, C.c_double, _i32p])
_sig("orc_adir_orient", C.c_uint64, [C.c_uint32, _u32p, _u32p, _i32p, _u32p, _u32p])
_sig("orc_kcore", None, [C.c_uint32, _u32p, _u32p, _u32p])
_sig("orc_kcore_orient", C.c_uint64, [C.c_uint32, _u32p, _u32p, _u32p, _u32p, _u32p])
_sig("orc_degeneracy", None, [C.c_uint32, _u32p, _u32p, _u32p, _u32p])
_sig("orc_aorder", C.c_int, [_u32p, _u32p, C.c_uint64, C.c_uint32, _u32p, _u32p, _u32p])
_sig("orc_count", C.c_uint64, [C.c_uint32, _u32p, _u32p])
_sig("orc_count_omp", C.c_uint64, [C.c_uint32, _u32p, _u32p])
_sig("orc_count_rows", C.c_uint64, [C.c_uint32, _u32p, _u32p, C.c_uint32, C.c_uint32])
_sig("orc_count_rows_strided", C.c_uint64, [C.c_uint32, _u32p, _u32p, C.c_uint32, C.c_uint32])
_sig("orc_count_rows_strided_omp", C.c_uint64, [C.c_uint32, _u32p, _u32p, C.c_uint32, C.c_uint32,
                                                C.c_int])
_sig("orc_count_row_list", C.c_uint64, [C.c_uint32, _u32p, _u32p, _u32p, C.c_uint64, C.c_int])
_sig("orc_count_per_vertex", None, [C.c_uint32, _u32p, _u32p, _u64p])
_sig("orc_kclique", C.c_uint64, [C.c_uint32, _u32p, _u32p, C.c_uint32])
_sig("orc_kclique_rows_strided", C.c_uint64, [C.c_uint32, _u32p, _u32p, C.c_uint32, C.c_uint32,
                                               C.c_uint32])
_sig("orc_kclique_row_list", None, [C.c_uint32, _u32p, _u32p, C.c_uint32, _u32p, C.c_uint64,
                                    _u64p, C.c_int])
_sig("orc_hu_workload", None, [C.c_uint32, _u32p, _u32p, _i64p])
for _name in ("orc_relabel_compact", "orc_relabel_first_src", "orc_relabel_degree",
              "orc_relabel_adir"):
    _sig(_name, None, [_u32p, _u32p, C.c_uint64, _u32p])
_sig("orc_hu_block_starts", None, [C.c_uint32, _i64p, C.c_uint32, _i32p])
_sig("orc_fox_workload", None, [C.c_uint64, _u32p, _u32p, _u32p, _i32p])
_sig("orc_fox_bins", None, [_i32p, C.c_uint32, C.c_uint32, _u32p])
_sig("orc_cost_func", None, [_u32p, _u32p, C.c_uint64, C.c_int, _f32p])
_sig("orc_verification", None, [_u32p, _u32p, C.c_uint64, _f32p])
_sig("orc_approx_ratio", C.c_double, [_u32p, _u32p, C.c_uint64])


def _u32(a):
    return np.ascontiguousarray(a, dtype=np.uint32)


def _nz(a, dtype):
    """ctypes ndpointer refuses empty arrays' NULL data; pad to one element."""
    return a if len(a) else np.zeros(1, dtype)


def num_nodes(src, dst) -> int:
    src, dst = _u32(src), _u32(dst)
    return int(_lib.orc_num_nodes(_nz(src, np.uint32), _nz(dst, np.uint32), len(src)))


def degree_raw(src, dst, n):
    src, dst = _u32(src), _u32(dst)
    deg = np.zeros(max(n, 1), np.uint32)
    _lib.orc_degree_raw(_nz(src, np.uint32), _nz(dst, np.uint32), len(src), n, deg)
    return deg[:n]


ORIENT_ID, ORIENT_DEGREE, ORIENT_DEG_TRICORE = 0, 1, 4


def orient_simple(src, dst, deg, mode):
    src, dst = _u32(src), _u32(dst)
    m = len(src)
    os_, od = np.empty(max(m, 1), np.uint32), np.empty(max(m, 1), np.uint32)
    _lib.orc_orient_simple(_nz(src, np.uint32), _nz(dst, np.uint32), m,
                           _nz(_u32(deg), np.uint32), mode, os_, od)
    return os_[:m], od[:m]


def prune_deg1(src, dst, n, deg):
    src, dst = _u32(src).copy(), _u32(dst).copy()
    m = len(src)
    s, d = _nz(src, np.uint32), _nz(dst, np.uint32)
    n2 = _lib.orc_prune_deg1(s, d, m, n, _nz(_u32(deg), np.uint32))
    return s[:m], d[:m], int(n2)


def build_csr(src, dst, n):
    src, dst = _u32(src), _u32(dst)
    m = len(src)
    rp = np.zeros(n + 1, np.uint32)
    col = np.zeros(max(m, 1), np.uint32)
    mo = _lib.orc_build_csr(_nz(src, np.uint32), _nz(dst, np.uint32), m, n, rp, col)
    return rp, col[:mo]


def build_sym(src, dst, n):
    src, dst = _u32(src), _u32(dst)
    m = len(src)
    rp = np.zeros(n + 1, np.uint32)
    col = np.zeros(max(2 * m, 1), np.uint32)
    m2 = _lib.orc_build_sym(_nz(src, np.uint32), _nz(dst, np.uint32), m, n, rp, col)
    return rp, col[:m2]


def adir_priority(n, srp, scol):
    prio = np.zeros(max(n, 1), np.int32)
    _lib.orc_adir_priority(n, _u32(srp), _nz(_u32(scol), np.uint32), prio)
    return prio[:n]


def adir_omp_priority(n, srp, scol, thre_param):
    """edge_directing/A-direction.cpp (OpenMP program) in one-thread order."""
    prio = np.zeros(max(n, 1), np.int32)
    _lib.orc_adir_omp_priority(n, _u32(srp), _nz(_u32(scol), np.uint32), float(thre_param), prio)
    return prio[:n]


def adir_orient(n, srp, scol, prio):
    m2 = len(scol)
    os_, od = np.empty(max(m2, 1), np.uint32), np.empty(max(m2, 1), np.uint32)
    k = _lib.orc_adir_orient(n, _u32(srp), _nz(_u32(scol), np.uint32),
                             _nz(np.ascontiguousarray(prio, np.int32), np.int32), os_, od)
    return os_[:k], od[:k]


def kcore(n, srp, scol):
    core = np.zeros(max(n, 1), np.uint32)
    _lib.orc_kcore(n, _u32(srp), _nz(_u32(scol), np.uint32), core)
    return core[:n]


def kcore_orient(n, srp, scol, core):
    m2 = len(scol)
    os_, od = np.empty(max(m2, 1), np.uint32), np.empty(max(m2, 1), np.uint32)
    k = _lib.orc_kcore_orient(n, _u32(srp), _nz(_u32(scol), np.uint32),
                              _nz(_u32(core), np.uint32), os_, od)
    return os_[:k], od[:k]


def degeneracy(n, srp, scol):
    """(removal step, core number) per vertex of the level-synchronous peel."""
    step = np.zeros(max(n, 1), np.uint32)
    core = np.zeros(max(n, 1), np.uint32)
    _lib.orc_degeneracy(n, _u32(srp), _nz(_u32(scol), np.uint32), step, core)
    return step[:n], core[:n]


def aorder(src, dst, n):
    """Returns (perm old->new, oriented src, oriented dst) or None outside the
    reference's defined range."""
    src, dst = _u32(src), _u32(dst)
    m = len(src)
    perm = np.zeros(max(n, 1), np.uint32)
    os_, od = np.empty(max(m, 1), np.uint32), np.empty(max(m, 1), np.uint32)
    r = _lib.orc_aorder(_nz(src, np.uint32), _nz(dst, np.uint32), m, n, perm, os_, od)
    if r != 0:
        return None
    return perm[:n], os_[:m], od[:m]


def count(rp, col) -> int:
    rp = _u32(rp)
    return int(_lib.orc_count(len(rp) - 1, rp, _nz(_u32(col), np.uint32)))


def count_omp(rp, col) -> int:
    rp = _u32(rp)
    return int(_lib.orc_count_omp(len(rp) - 1, rp, _nz(_u32(col), np.uint32)))


def count_rows(rp, col, u_begin, u_end) -> int:
    rp = _u32(rp)
    return int(_lib.orc_count_rows(len(rp) - 1, rp, _nz(_u32(col), np.uint32), u_begin, u_end))


def count_rows_strided(rp, col, start, stride) -> int:
    rp = _u32(rp)
    return int(_lib.orc_count_rows_strided(len(rp) - 1, rp, _nz(_u32(col), np.uint32), start,
                                           stride))


def count_rows_strided_omp(rp, col, start, stride, threads) -> int:
    rp = _u32(rp)
    return int(_lib.orc_count_rows_strided_omp(len(rp) - 1, rp, _nz(_u32(col), np.uint32), start,
                                               stride, threads))


def count_row_list(rp, col, rows, threads=1) -> int:
    """Triangles at the listed rows (oracle/oracle.cpp orc_count_row_list)."""
    rp = _u32(rp)
    rows = _u32(rows)
    return int(_lib.orc_count_row_list(len(rp) - 1, rp, _nz(_u32(col), np.uint32),
                                       _nz(rows, np.uint32), len(rows), threads))


def count_per_vertex(rp, col):
    rp = _u32(rp)
    n = len(rp) - 1
    tri = np.zeros(max(n, 1), np.uint64)
    _lib.orc_count_per_vertex(n, rp, _nz(_u32(col), np.uint32), tri)
    return tri[:n]


def kclique(rp, col, k) -> int:
    rp = _u32(rp)
    return int(_lib.orc_kclique(len(rp) - 1, rp, _nz(_u32(col), np.uint32), k))


def kclique_rows_strided(rp, col, k, start, stride) -> int:
    rp = _u32(rp)
    return int(_lib.orc_kclique_rows_strided(len(rp) - 1, rp, _nz(_u32(col), np.uint32), k,
                                             start, stride))


def kclique_row_list(rp, col, k, rows, threads=1):
    """Per-row counts of the k-cliques rooted at each listed row (oracle/oracle.cpp
    orc_kclique_row_list)."""
    rp = _u32(rp)
    rows = _u32(rows)
    out = np.zeros(max(len(rows), 1), np.uint64)
    _lib.orc_kclique_row_list(len(rp) - 1, rp, _nz(_u32(col), np.uint32), k,
                              _nz(rows, np.uint32), len(rows), out, threads)
    return out[:len(rows)]


def hu_workload(rp, col):
    rp = _u32(rp)
    n = len(rp) - 1
    w = np.zeros(max(n, 1), np.int64)
    _lib.orc_hu_workload(n, rp, _nz(_u32(col), np.uint32), w)
    return w[:n]


def hu_block_starts(w, nblocks):
    w = np.ascontiguousarray(w, np.int64)
    s = np.zeros(nblocks + 1, np.int32)
    _lib.orc_hu_block_starts(len(w), _nz(w, np.int64), nblocks, s)
    return s


def fox_workload(esrc, edst, adjlen):
    esrc, edst = _u32(esrc), _u32(edst)
    w = np.zeros(max(len(esrc), 1), np.int32)
    _lib.orc_fox_workload(len(esrc), _nz(esrc, np.uint32), _nz(edst, np.uint32),
                          _nz(_u32(adjlen), np.uint32), w)
    return w[:len(esrc)]


def fox_bins(sorted_w, max_lanes=32):
    w = np.ascontiguousarray(sorted_w, np.int32)
    out = np.zeros(max_lanes + 2, np.uint32)
    _lib.orc_fox_bins(_nz(w, np.int32), len(w), max_lanes, out)
    return out


def cost_func(src, dst, k):
    out = np.zeros(3, np.float32)
    _lib.orc_cost_func(_u32(src), _u32(dst), len(src), k, out)
    return out


def verification(src, dst):
    out = np.zeros(2, np.float32)
    _lib.orc_verification(_u32(src), _u32(dst), len(src), out)
    return out


def relabel(src, dst, mode):
    """old -> new id map of tools/serialize.cpp ("compact"),
    tools/adjust2tc.cpp ("first_src"), tools/directionByReorder.cpp ("degree")
    or Bisson/A-direction.cpp ("adir")."""
    src, dst = _u32(src), _u32(dst)
    n = num_nodes(src, dst)
    perm = np.zeros(max(n, 1), np.uint32)
    fn = {"compact": _lib.orc_relabel_compact, "first_src": _lib.orc_relabel_first_src,
          "degree": _lib.orc_relabel_degree, "adir": _lib.orc_relabel_adir}[mode]
    fn(_nz(src, np.uint32), _nz(dst, np.uint32), len(src), perm)
    return perm[:n]


def approx_ratio(src, dst) -> float:
    return float(_lib.orc_approx_ratio(_u32(src), _u32(dst), len(src)))


def pipeline_count(src, dst, mode="degree"):
    """Oracle end to end: orientation + CSR + serial forward count."""
    n = num_nodes(src, dst)
    if mode == "adir":
        srp, scol = build_sym(src, dst, n)
        prio = adir_priority(n, srp, scol)
        os_, od = adir_orient(n, srp, scol, prio)
    else:
        deg = degree_raw(src, dst, n)
        os_, od = orient_simple(src, dst, deg,
                                {"id": ORIENT_ID, "degree": ORIENT_DEGREE,
                                 "tricore": ORIENT_DEG_TRICORE}[mode])
    rp, col = build_csr(os_, od, n)
    return count(rp, col), rp, col
