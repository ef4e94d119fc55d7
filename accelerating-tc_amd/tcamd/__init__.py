"""Python binding of libtcamd.so, the MI355X triangle-counting / k-clique engine.

The engine's boundary is the C ABI in include/tc_amd.h; this module binds it
with ctypes and mirrors the reference's pipeline one step per call:

    edges = read_edges("g.bin")                    # tools/list2bin format
    with Context() as ctx:
        g = ctx.graph(*edges)                      # upload raw records to HBM
        g.orient(ORIENT_DEGREE)                    # Hu/preTC.cu orientation + CSR
        count, timing = g.count(KERNEL_HU)         # Hu/tc.cu triangleCountKernel

Every call goes to the HIP library; there is no CPU fallback.  A missing or
unbuilt library raises at import time.
"""
from __future__ import annotations

import atexit
import ctypes as C
import json
import os
import weakref
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "libtcamd.so")

if not os.path.exists(LIB_PATH):
    raise ImportError(
        f"libtcamd.so not found at {LIB_PATH}; build it with "
        "`make -C accelerating-tc_amd lib` (or __graft_entry__.build())")

_lib = C.CDLL(LIB_PATH)

OK = 0
ERR_ARG, ERR_IO, ERR_FORMAT, ERR_RANGE, ERR_NOMEM, ERR_HIP, ERR_UNSUPPORTED, ERR_STATE = (
    -1, -2, -3, -4, -5, -6, -7, -8)

ORIENT_ID, ORIENT_DEGREE, ORIENT_ADIR, ORIENT_KCORE, ORIENT_DEGREE_TRICORE = 0, 1, 2, 3, 4
ORIENT_DEGENERACY = 5
ORIENT_ADIR_OMP = 6  # set by Graph.orient_adir_omp
TIE_ID, TIE_FORCED_ASC, TIE_FORCED_DESC = 0, 1, 2
TIE_CANDIDATES = 3
FLAG_PRUNE_DEG1 = 1
KERNEL_HU, KERNEL_BISSON, KERNEL_FOX, KERNEL_TRICORE, KERNEL_AUTO = 0, 1, 2, 3, 4
KERNEL_NAMES = {"hu": KERNEL_HU, "bisson": KERNEL_BISSON, "fox": KERNEL_FOX,
                "tricore": KERNEL_TRICORE, "auto": KERNEL_AUTO}
AUTO_WORKLIST, AUTO_PUSH = 0, 1
BATCH_WORKLIST, BATCH_PUSH = 0, 1
KCL_AUTO, KCL_BITMAP, KCL_MERGE, KCL_PIVOT = 0, 1, 2, 3
RELABEL_COMPACT, RELABEL_FIRST_SRC, RELABEL_DEGREE, RELABEL_ADIR = 0, 1, 2, 3

_u32p = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
_u64p = np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS")
_vp = C.c_void_p


class GraphInfo(C.Structure):
    _fields_ = [("n", C.c_uint32), ("max_out_degree", C.c_uint32), ("m_raw", C.c_uint64),
                ("m_oriented", C.c_uint64), ("orient_mode", C.c_int32), ("n_raw", C.c_uint32),
                ("sum_out_sq", C.c_uint64), ("sum_in_out", C.c_uint64),
                ("bytes_alg", C.c_double), ("bytes_min", C.c_double)]


class Sched(C.Structure):
    _fields_ = [("kernel", C.c_int32), ("block_size", C.c_uint32), ("grid_size", C.c_uint32),
                ("lds_slice", C.c_uint32), ("bitmap_min_deg", C.c_uint32),
                ("fox_max_lanes", C.c_uint32), ("tricore_parts", C.c_uint32),
                ("auto_bins", C.c_uint32), ("auto_thresholds", C.c_uint32 * 4),
                ("fox_rearrange", C.c_uint32), ("reserved", C.c_uint32 * 3)]


class Timing(C.Structure):
    _fields_ = [("ms_count", C.c_float), ("ms_prep", C.c_float), ("launches", C.c_uint32),
                ("dominant_kernel", C.c_uint32), ("ms_dominant", C.c_float),
                ("reserved", C.c_float)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "reserved"}


def _sig(name, res, args):
    f = getattr(_lib, name)
    f.restype = res
    f.argtypes = args
    return f


_sig("tc_status_str", C.c_char_p, [C.c_int])
_sig("tc_last_error", C.c_char_p, [])
_sig("tc_ctx_create", C.c_int, [C.c_int, C.POINTER(_vp)])
_sig("tc_ctx_destroy", C.c_int, [_vp])
_sig("tc_ctx_device_memory", C.c_int, [_vp, C.POINTER(C.c_uint64)])
_sig("tc_file_edge_count", C.c_int, [C.c_char_p, C.POINTER(C.c_uint64)])
_sig("tc_read_edges", C.c_int, [C.c_char_p, _u32p, _u32p, C.c_uint64, C.POINTER(C.c_uint64)])
_sig("tc_write_edges", C.c_int, [C.c_char_p, _u32p, _u32p, C.c_uint64])
_sig("tc_list2bin", C.c_int, [C.c_char_p, C.c_char_p, C.POINTER(C.c_uint64)])
_sig("tc_bin2list", C.c_int, [C.c_char_p, C.c_char_p, C.POINTER(C.c_uint64)])
_sig("tc_read_gr", C.c_int, [C.c_char_p, _vp, _vp, C.c_uint64, C.POINTER(C.c_uint64),
                             C.POINTER(C.c_uint32)])
_sig("tc_gen_er", C.c_int, [C.c_uint32, C.c_uint64, C.c_uint64, _u32p, _u32p])
_sig("tc_gen_rmat", C.c_int, [C.c_uint32, C.c_uint64, C.c_double, C.c_double, C.c_double,
                              C.c_uint64, _u32p, _u32p])
_sig("tc_gen_ba", C.c_int, [C.c_uint32, C.c_uint32, C.c_uint64, _u32p, _u32p,
                            C.POINTER(C.c_uint64)])
_sig("tc_gen_rgg", C.c_int, [C.c_uint32, C.c_double, C.c_uint64, _vp, _vp,
                             C.POINTER(C.c_uint64)])
_sig("tc_graph_create", C.c_int, [_vp, _vp, _vp, C.c_uint64, C.c_uint32, C.POINTER(_vp)])
_sig("tc_graph_destroy", C.c_int, [_vp])
_sig("tc_graph_simplify", C.c_int, [_vp, _vp])
_sig("tc_graph_relabel_aorder", C.c_int, [_vp, _vp, _vp])


class BwModel(C.Structure):
    """A-order's list-read bandwidth model (tc_bw_model)."""
    _fields_ = [("table", C.c_float * 10), ("a", C.c_double), ("b", C.c_double),
                ("lambda_", C.c_float * 10), ("lambda_tail", C.c_float)]

    def as_dict(self):
        return {"table": list(self.table), "a": self.a, "b": self.b,
                "lambda": list(self.lambda_), "lambda_tail": self.lambda_tail}

    @classmethod
    def from_dict(cls, d):
        m = cls()
        m.table[:] = [float(x) for x in d["table"]]
        m.a, m.b = float(d["a"]), float(d["b"])
        m.lambda_[:] = [float(x) for x in d["lambda"]]
        m.lambda_tail = float(d["lambda_tail"])
        return m


_sig("tc_graph_relabel_aorder_model", C.c_int, [_vp, _vp, C.POINTER(BwModel), _vp])
_sig("tc_bw_model_reference", C.c_int, [C.POINTER(BwModel)])
_sig("tc_measure_list_bandwidth", C.c_int, [_vp, _vp, C.c_uint32, _vp])
_sig("tc_measure_valu_rate", C.c_int, [_vp, C.c_uint32, C.POINTER(C.c_float)])
_sig("tc_bw_model_fit", C.c_int, [_vp, _vp, C.c_uint32, C.POINTER(BwModel)])
_sig("tc_graph_orient", C.c_int, [_vp, _vp, C.c_int, C.c_uint32])
_sig("tc_graph_orient_adir_omp", C.c_int, [_vp, _vp, C.c_double])
_sig("tc_graph_orient_ties", C.c_int, [_vp, _vp, C.c_int, C.c_int])
_sig("tc_graph_adir_cost", C.c_int, [_vp, _vp, C.POINTER(C.c_double)])
_sig("tc_graph_relabel", C.c_int, [_vp, _vp, C.c_int, _vp])
_sig("tc_graph_info_get", C.c_int, [_vp, C.POINTER(GraphInfo)])
_sig("tc_graph_get_csr", C.c_int, [_vp, _vp, _vp])
_sig("tc_graph_get_edges", C.c_int, [_vp, _vp, _vp])
_sig("tc_graph_get_vertex_key", C.c_int, [_vp, _i32p])
_sig("tc_hu_workload", C.c_int, [_vp, _vp, _i64p])
_sig("tc_hu_block_starts", C.c_int, [_vp, _vp, C.c_uint32, _i32p])
_sig("tc_fox_workload", C.c_int, [_vp, _vp, _i32p])
_sig("tc_fox_bin_sizes", C.c_int, [_vp, _vp, C.c_uint32, _u64p])
_sig("tc_count", C.c_int, [_vp, _vp, C.POINTER(Sched), C.POINTER(C.c_uint64),
                           C.POINTER(Timing)])
_sig("tc_count_per_vertex", C.c_int, [_vp, _vp, _u64p])
_sig("tc_batch_stream_bytes", C.c_int, [_vp, _vp, C.POINTER(C.c_double)])
_sig("tc_graph_save_csr", C.c_int, [_vp, C.c_char_p])
_sig("tc_graph_load_csr", C.c_int, [_vp, C.c_char_p, C.POINTER(_vp)])
_sig("tc_auto_stream_bytes", C.c_int, [_vp, _vp, C.c_void_p, C.POINTER(C.c_double)])
_sig("tc_count_csr_noalloc", C.c_int, [_vp, C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32,
                                       C.c_void_p])
_f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")


class TrainOpts(C.Structure):
    _fields_ = [("steps", C.c_uint32), ("batch", C.c_uint32), ("lr", C.c_float),
                ("seed", C.c_uint32), ("reserved", C.c_uint32 * 4)]


_sig("tc_cost_functions", C.c_int, [_vp, _vp, C.c_int, _f32p])
_sig("tc_approx_ratio", C.c_int, [_vp, _vp, C.POINTER(C.c_double)])
_sig("tc_model_num_params", C.c_uint32, [])
_sig("tc_model_create", C.c_int, [_vp, C.c_uint64, C.POINTER(_vp)])
_sig("tc_model_destroy", C.c_int, [_vp])
_sig("tc_model_set_params", C.c_int, [_vp, _f32p, C.c_uint32])
_sig("tc_model_get_params", C.c_int, [_vp, _f32p, C.c_uint32])
_sig("tc_model_set_norm", C.c_int, [_vp, _f32p, _f32p, _f32p, _f32p])
_sig("tc_model_get_norm", C.c_int, [_vp, _f32p, _f32p, _f32p, _f32p])
_sig("tc_model_forward", C.c_int, [_vp, _vp, _f32p, C.c_uint32, _f32p])
_sig("tc_model_loss_grad", C.c_int, [_vp, _vp, _f32p, _f32p, _f32p, C.c_uint32,
                                     C.POINTER(C.c_float), _f32p])
_sig("tc_model_train", C.c_int, [_vp, _vp, _f32p, _f32p, _f32p, C.c_uint32,
                                 C.POINTER(TrainOpts), C.POINTER(C.c_float)])
_sig("tc_graph_features", C.c_int, [_vp, _vp, _vp, _f32p])
_sig("tc_model_set_bw_model", C.c_int, [_vp, C.c_void_p])
_sig("tc_graph_measure_costs", C.c_int, [_vp, _vp, C.c_int, _f32p, C.POINTER(C.c_uint64)])
_sig("tc_graph_cell_features", C.c_int, [_vp, _vp, _vp, C.c_uint32, _f32p])
_sig("tc_graph_measure_cells", C.c_int, [_vp, _vp, C.c_int, C.c_uint32, C.c_uint32, _f32p,
                                         C.c_void_p])
_sig("tc_learned_cells", C.c_int, [_vp, _vp, _vp, C.POINTER(C.c_uint32), C.c_void_p])
_sig("tc_graph_orient_learned", C.c_int, [_vp, _vp, C.c_int, _vp, C.POINTER(C.c_int),
                                          C.c_void_p, C.c_void_p])
_sig("tc_learned_assignment", C.c_int, [_vp, _vp, _vp, _u8p])
_sig("tc_count_learned", C.c_int, [_vp, _vp, _vp, C.POINTER(C.c_uint64), C.POINTER(Timing)])
_sig("kcl_count", C.c_int, [_vp, _vp, C.c_uint32, C.c_int, C.POINTER(C.c_uint64),
                            C.POINTER(Timing)])
_sig("kcl_gpu_solver", C.c_int, [C.c_char_p, C.c_uint32, C.POINTER(C.c_uint64), C.c_size_t])
_sig("kcl_root_features", C.c_int, [_vp, _vp, C.c_uint32, _f32p])
_sig("kcl_count_roots", C.c_int, [_vp, _vp, C.c_uint32, _vp, C.c_uint64, C.POINTER(C.c_uint64),
                                  C.POINTER(Timing)])
_sig("kcl_measure_costs", C.c_int, [_vp, _vp, C.c_uint32, C.c_int, _f32p, C.POINTER(C.c_uint64)])
_sig("kcl_learned_plan", C.c_int, [_vp, _vp, C.c_uint32, _vp, _u8p])
_sig("kcl_count_learned", C.c_int, [_vp, _vp, C.c_uint32, _vp, C.POINTER(C.c_uint64),
                                    C.POINTER(Timing)])
_sig("tc_tricore_init_gpu", C.c_uint64, [])
_sig("tc_tricore_cal_degree", None, [_u64p, C.c_uint64, _u32p, C.c_uint32])
_sig("tc_tricore_redirect_edges", None, [_u64p, C.c_uint64, _u32p, C.c_uint32])
_sig("tc_tricore_cal_part_num", C.c_uint32, [_u64p, C.c_uint64, C.c_uint32])
_sig("tc_tricore_split", None, [_u64p, C.c_uint64, C.c_uint64, C.POINTER(C.POINTER(C.c_uint32)),
                                _u32p, _u64p])
_sig("tc_tricore_tricount_gpu", C.c_uint64, [C.c_uint64, _vp, _u64p, _u32p,
                                             C.POINTER(C.POINTER(C.c_uint32))])


class BatchSched(C.Structure):
    _fields_ = [("wave_max_deg", C.c_uint32), ("wave_chunk", C.c_uint32), ("grid", C.c_uint32),
                ("filter_min", C.c_uint32), ("form", C.c_uint32), ("reserved", C.c_uint32 * 3)]

    def as_dict(self):
        return {"wave_max_deg": self.wave_max_deg, "wave_chunk": self.wave_chunk,
                "grid": self.grid, "filter_min": self.filter_min, "form": self.form}


_sig("tc_batch_create", C.c_int, [_vp, C.c_uint32, _u64p, _vp, _vp, _u32p, C.POINTER(_vp)])
_sig("tc_batch_destroy", C.c_int, [_vp])
_sig("tc_batch_info_get", C.c_int, [_vp, C.POINTER(GraphInfo)])
_sig("tc_batch_autotune", C.c_int, [_vp, _vp, C.POINTER(BatchSched)])
_sig("tc_batch_count", C.c_int, [_vp, _vp, C.POINTER(BatchSched), _u64p, C.POINTER(Timing)])
_sig("tc_batch_count_pergraph", C.c_int, [_vp, _vp, C.POINTER(BatchSched), _u64p,
                                          C.POINTER(Timing)])

EXPORTED = [
    "tc_status_str", "tc_last_error", "tc_ctx_create", "tc_ctx_destroy", "tc_ctx_device_memory",
    "tc_file_edge_count", "tc_read_edges", "tc_write_edges", "tc_list2bin", "tc_bin2list", "tc_read_gr",
    "tc_gen_er", "tc_gen_rmat", "tc_gen_ba", "tc_gen_rgg", "tc_graph_create", "tc_graph_destroy",
    "tc_graph_simplify", "tc_graph_relabel_aorder", "tc_graph_relabel", "tc_graph_orient", "tc_graph_info_get",
    "tc_graph_relabel_aorder_model", "tc_bw_model_reference", "tc_measure_list_bandwidth",
    "tc_measure_valu_rate", "tc_bw_model_fit",
    "tc_graph_get_csr", "tc_graph_get_edges", "tc_graph_get_vertex_key", "tc_graph_device_csr",
    "tc_hu_workload", "tc_hu_block_starts", "tc_fox_workload", "tc_fox_bin_sizes", "tc_count",
    "tc_count_per_vertex", "tc_auto_stream_bytes", "tc_cost_functions", "tc_approx_ratio", "kcl_count", "kcl_gpu_solver",
    "kcl_root_features", "kcl_measure_costs", "kcl_learned_plan", "kcl_count_learned",
    "kcl_count_roots",
    "tc_tricore_init_gpu",
    "tc_tricore_cal_degree", "tc_tricore_redirect_edges", "tc_tricore_cal_part_num",
    "tc_tricore_split", "tc_tricore_tricount_gpu", "tc_model_num_params", "tc_model_create",
    "tc_model_destroy", "tc_model_set_params", "tc_model_get_params", "tc_model_set_norm",
    "tc_model_get_norm", "tc_model_forward", "tc_model_loss_grad", "tc_model_train",
    "tc_graph_orient_adir_omp", "tc_graph_orient_ties", "tc_graph_adir_cost",
    "tc_graph_orient_learned",
    "tc_graph_features", "tc_model_set_bw_model", "tc_graph_measure_costs",
    "tc_graph_cell_features", "tc_graph_measure_cells", "tc_learned_cells",
    "tc_learned_assignment", "tc_count_learned",
    "tc_batch_create", "tc_batch_destroy", "tc_batch_info_get", "tc_batch_autotune", "tc_batch_count",
    "tc_batch_count_pergraph", "tc_batch_stream_bytes", "tc_count_csr_noalloc", "tc_graph_save_csr", "tc_graph_load_csr",
]
VARIANT_BLOCK, VARIANT_WAVE, VARIANT_LANE = 0, 1, 2
CELLS, CELL_MAX_LEVEL = 16, 16
BLOCK_CELL = 15  # tasks above d+ 512: timed and predicted wide only
FIXED_NARROW = 0xFFC  # the fixed rule runs cells 2..11 (d+ <= 64) narrow
N_FEATURES, N_OUTPUTS = 8, 2


class TCError(RuntimeError):
    def __init__(self, status: int, where: str):
        msg = _lib.tc_last_error().decode(errors="replace")
        self.status = status
        super().__init__(f"{where}: {_lib.tc_status_str(status).decode()} ({status}): {msg}")


def _check(status: int, where: str):
    if status != OK:
        raise TCError(status, where)


def last_error() -> str:
    return _lib.tc_last_error().decode(errors="replace")


def _u32(a):
    return np.ascontiguousarray(a, dtype=np.uint32)


# ---- files and generators ----------------------------------------------------

def file_edge_count(path: str) -> int:
    m = C.c_uint64()
    _check(_lib.tc_file_edge_count(path.encode(), C.byref(m)), "tc_file_edge_count")
    return m.value


def read_edges(path: str):
    m = file_edge_count(path)
    src = np.empty(max(m, 1), np.uint32)
    dst = np.empty(max(m, 1), np.uint32)
    got = C.c_uint64()
    _check(_lib.tc_read_edges(path.encode(), src, dst, m, C.byref(got)), "tc_read_edges")
    return src[:m], dst[:m]


def write_edges(path: str, src, dst):
    src, dst = _u32(src), _u32(dst)
    if len(src) == 0:
        src = np.zeros(1, np.uint32)
        dst = np.zeros(1, np.uint32)
        m = 0
    else:
        m = len(src)
    _check(_lib.tc_write_edges(path.encode(), src, dst, m), "tc_write_edges")


def list2bin(text_path: str, bin_path: str) -> int:
    m = C.c_uint64()
    _check(_lib.tc_list2bin(text_path.encode(), bin_path.encode(), C.byref(m)), "tc_list2bin")
    return m.value


def bin2list(bin_path: str, text_path: str) -> int:
    m = C.c_uint64()
    _check(_lib.tc_bin2list(bin_path.encode(), text_path.encode(), C.byref(m)), "tc_bin2list")
    return m.value


def read_gr(path: str):
    """(src, dst, n) of a Galois .gr file (one record per stored arc)."""
    m, n = C.c_uint64(), C.c_uint32()
    _check(_lib.tc_read_gr(path.encode(), None, None, 0, C.byref(m), C.byref(n)), "tc_read_gr")
    s = np.empty(max(m.value, 1), np.uint32)
    d = np.empty(max(m.value, 1), np.uint32)
    _check(_lib.tc_read_gr(path.encode(), s.ctypes.data, d.ctypes.data, m.value, C.byref(m),
                           C.byref(n)), "tc_read_gr")
    return s[:m.value], d[:m.value], n.value


def write_gr(path: str, src, dst, n: int):
    """Writes a symmetric Galois .gr v1 file (both arcs of every record,
    rows sorted, no edge data) - the k-clique CLI's input format."""
    s, d = _u32(src).astype(np.int64), _u32(dst).astype(np.int64)
    a = np.concatenate([s, d])
    b = np.concatenate([d, s])
    keep = a != b
    key = np.unique(a[keep] * (1 << 32) + b[keep])
    a, b = key >> 32, key & 0xFFFFFFFF
    ends = np.cumsum(np.bincount(a, minlength=n)).astype(np.uint64)
    with open(path, "wb") as f:
        np.array([1, 0, n, len(b)], np.uint64).tofile(f)
        ends.tofile(f)
        b.astype(np.uint32).tofile(f)
        if len(b) % 2:
            np.zeros(1, np.uint32).tofile(f)


def gen_er(n: int, m: int, seed: int):
    s = np.empty(max(m, 1), np.uint32)
    d = np.empty(max(m, 1), np.uint32)
    _check(_lib.tc_gen_er(n, m, seed, s, d), "tc_gen_er")
    return s[:m], d[:m]


def gen_rmat(scale: int, nedges: int, a=0.57, b=0.19, c=0.19, seed: int = 0):
    s = np.empty(max(nedges, 1), np.uint32)
    d = np.empty(max(nedges, 1), np.uint32)
    _check(_lib.tc_gen_rmat(scale, nedges, a, b, c, seed, s, d), "tc_gen_rmat")
    return s[:nedges], d[:nedges]


def gen_ba(n: int, k: int, seed: int):
    m = (n - k) * k
    s = np.empty(max(m, 1), np.uint32)
    d = np.empty(max(m, 1), np.uint32)
    got = C.c_uint64()
    _check(_lib.tc_gen_ba(n, k, seed, s, d, C.byref(got)), "tc_gen_ba")
    return s[:got.value], d[:got.value]


def gen_rgg(n: int, r: float, seed: int):
    m = C.c_uint64()
    _check(_lib.tc_gen_rgg(n, r, seed, None, None, C.byref(m)), "tc_gen_rgg")
    s = np.empty(max(m.value, 1), np.uint32)
    d = np.empty(max(m.value, 1), np.uint32)
    _check(_lib.tc_gen_rgg(n, r, seed, s.ctypes.data, d.ctypes.data, C.byref(m)), "tc_gen_rgg")
    return s[:m.value], d[:m.value]


def bw_model_reference() -> BwModel:
    m = BwModel()
    _check(_lib.tc_bw_model_reference(C.byref(m)), "tc_bw_model_reference")
    return m


def measure_list_bandwidth(ctx: "Context", lens) -> np.ndarray:
    """GB/s of list entries read, one wave per list of each length."""
    lens = np.ascontiguousarray(lens, np.uint32)
    out = np.zeros(max(len(lens), 1), np.float32)
    _check(_lib.tc_measure_list_bandwidth(ctx.handle, lens.ctypes.data, len(lens),
                                          out.ctypes.data), "tc_measure_list_bandwidth")
    return out[:len(lens)]


def measure_valu_rate(ctx: "Context", iters: int) -> float:
    """hipEvent milliseconds of one issue-rate probe launch of `iters` rounds
    (its wave instructions come from a profiled run's SQ_INSTS_VALU)."""
    ms = C.c_float()
    _check(_lib.tc_measure_valu_rate(ctx.handle, iters, C.byref(ms)), "tc_measure_valu_rate")
    return ms.value


def fit_bw_model(lens, gbps) -> BwModel:
    lens = np.ascontiguousarray(lens, np.uint32)
    gbps = np.ascontiguousarray(gbps, np.float32)
    m = BwModel()
    _check(_lib.tc_bw_model_fit(lens.ctypes.data, gbps.ctypes.data, len(lens), C.byref(m)),
           "tc_bw_model_fit")
    return m


BW_MODEL_MI355X = os.path.join(_HERE, "aorder_bw_mi355x.json")


def load_bw_model(path: str = BW_MODEL_MI355X) -> BwModel:
    with open(path) as f:
        return BwModel.from_dict(json.load(f)["model"])


def rgg_radius(n: int, avg_degree: float) -> float:
    """Radius giving the target average degree in the unit square."""
    return float(np.sqrt(avg_degree / (np.pi * (n - 1))))


# ---- context / graph ---------------------------------------------------------

@dataclass
class CountResult:
    count: int
    timing: dict


# Contexts still open when the interpreter exits are closed (children first)
# by an atexit hook, while the library and the HIP runtime are intact:
# __del__ during interpreter teardown runs in no defined order, if at all.
_live_contexts = weakref.WeakSet()


@atexit.register
def _close_live_contexts():
    for c in list(_live_contexts):
        try:
            c.close()
        except Exception:
            pass


class Context:
    def __init__(self, device: int = 0):
        h = _vp()
        _check(_lib.tc_ctx_create(device, C.byref(h)), "tc_ctx_create")
        self._h = h
        self.device = device
        # graphs and models live on this context; closing it closes them first
        self._children = weakref.WeakSet()
        _live_contexts.add(self)

    @property
    def handle(self):
        return self._h

    def device_memory(self) -> int:
        b = C.c_uint64()
        _check(_lib.tc_ctx_device_memory(self._h, C.byref(b)), "tc_ctx_device_memory")
        return b.value

    def close(self):
        if getattr(self, "_h", None):
            for child in list(self._children):
                child.close()
            _lib.tc_ctx_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def graph(self, src, dst, n: int = 0) -> "Graph":
        return Graph(self, src, dst, n)

    def load_csr(self, path: str) -> "Graph":
        """A graph from an oriented-CSR stage file (tc_graph_load_csr)."""
        h = _vp()
        _check(_lib.tc_graph_load_csr(self.handle, path.encode(), C.byref(h)), "tc_graph_load_csr")
        g = Graph.__new__(Graph)
        g.ctx = self
        g._h = h
        self._children.add(g)
        return g

    def batch(self, graphs) -> "Batch":
        return Batch(self, graphs)


class Graph:
    def __init__(self, ctx: Context, src, dst, n: int = 0):
        src, dst = _u32(src), _u32(dst)
        if len(src) != len(dst):
            raise ValueError("src and dst differ in length")
        self.ctx = ctx
        h = _vp()
        m = len(src)
        _check(_lib.tc_graph_create(ctx.handle, src.ctypes.data if m else None,
                                    dst.ctypes.data if m else None, m, n, C.byref(h)),
               "tc_graph_create")
        self._h = h
        ctx._children.add(self)

    def close(self):
        if getattr(self, "_h", None):
            _lib.tc_graph_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- preprocessing --
    def save_csr(self, path: str):
        """Writes the oriented CSR stage file (tc_graph_save_csr)."""
        _check(_lib.tc_graph_save_csr(self._h, path.encode()), "tc_graph_save_csr")
        return self

    def simplify(self):
        _check(_lib.tc_graph_simplify(self.ctx.handle, self._h), "tc_graph_simplify")
        return self

    def relabel(self, mode: int):
        """Id relabelling of tools/serialize.cpp (RELABEL_COMPACT),
        tools/adjust2tc.cpp (RELABEL_FIRST_SRC),
        tools/directionByReorder.cpp (RELABEL_DEGREE) or
        Bisson/A-direction.cpp (RELABEL_ADIR); returns old -> new."""
        s, d = self.edges()
        nids = int(max(s.max(), d.max())) + 1 if len(s) else 1
        perm = np.empty(nids, np.uint32)
        _check(_lib.tc_graph_relabel(self.ctx.handle, self._h, mode, perm.ctypes.data),
               "tc_graph_relabel")
        return perm

    def relabel_aorder(self, model: "BwModel | None" = None):
        """A-order; model (optional) replaces the reference's bandwidth model.
        Returns old -> new for every raw id (n_raw entries; the CSR's n is
        smaller after a pruned orientation)."""
        n = self.info().n_raw
        perm = np.empty(max(n, 1), np.uint32)
        if model is None:
            _check(_lib.tc_graph_relabel_aorder(self.ctx.handle, self._h, perm.ctypes.data),
                   "tc_graph_relabel_aorder")
        else:
            _check(_lib.tc_graph_relabel_aorder_model(self.ctx.handle, self._h, C.byref(model),
                                                      perm.ctypes.data),
                   "tc_graph_relabel_aorder_model")
        return perm[:n]

    def orient(self, mode: int, prune_deg1: bool = False):
        _check(_lib.tc_graph_orient(self.ctx.handle, self._h, mode,
                                    FLAG_PRUNE_DEG1 if prune_deg1 else 0), "tc_graph_orient")
        return self

    def orient_adir_omp(self, thre_param: float):
        """The OpenMP A-direction program's schedule
        (edge_directing/A-direction.cpp, argv thre_param)."""
        _check(_lib.tc_graph_orient_adir_omp(self.ctx.handle, self._h, float(thre_param)),
               "tc_graph_orient_adir_omp")
        return self

    def orient_ties(self, mode: int, tie: int):
        """Orientation with a tie rule TIE_* (tc_graph_orient_ties)."""
        _check(_lib.tc_graph_orient_ties(self.ctx.handle, self._h, mode, tie),
               "tc_graph_orient_ties")
        return self

    def adir_cost(self) -> float:
        """sum |d+(v) - m+/n| of the current orientation."""
        out = C.c_double()
        _check(_lib.tc_graph_adir_cost(self.ctx.handle, self._h, C.byref(out)),
               "tc_graph_adir_cost")
        return out.value

    def orient_learned(self, mode: int, model=None):
        """Learned orientation ties: (chosen TIE_*, predicted us per rule,
        A-direction cost per rule); the graph keeps the chosen rule."""
        ch = C.c_int()
        pred = np.zeros(TIE_CANDIDATES, np.float64)
        cost = np.zeros(TIE_CANDIDATES, np.float64)
        _check(_lib.tc_graph_orient_learned(self.ctx.handle, self._h, mode,
                                            model.handle if model else None, C.byref(ch),
                                            pred.ctypes.data, cost.ctypes.data),
               "tc_graph_orient_learned")
        return ch.value, pred, cost

    def info(self) -> GraphInfo:
        i = GraphInfo()
        _check(_lib.tc_graph_info_get(self._h, C.byref(i)), "tc_graph_info_get")
        return i

    def csr(self):
        i = self.info()
        rp = np.empty(i.n + 1, np.uint32)
        col = np.empty(max(i.m_oriented, 1), np.uint32)
        _check(_lib.tc_graph_get_csr(self._h, rp.ctypes.data, col.ctypes.data), "tc_graph_get_csr")
        return rp, col[:i.m_oriented]

    def edges(self):
        i = self.info()
        s = np.empty(max(i.m_raw, 1), np.uint32)
        d = np.empty(max(i.m_raw, 1), np.uint32)
        _check(_lib.tc_graph_get_edges(self._h, s.ctypes.data, d.ctypes.data), "tc_graph_get_edges")
        return s[:i.m_raw], d[:i.m_raw]

    def vertex_key(self):
        n = self.info().n
        k = np.empty(max(n, 1), np.int32)
        _check(_lib.tc_graph_get_vertex_key(self._h, k), "tc_graph_get_vertex_key")
        return k[:n]

    # -- work models --
    def hu_workload(self):
        n = self.info().n
        w = np.empty(max(n, 1), np.int64)
        _check(_lib.tc_hu_workload(self.ctx.handle, self._h, w), "tc_hu_workload")
        return w[:n]

    def hu_block_starts(self, nblocks: int):
        s = np.empty(nblocks + 1, np.int32)
        _check(_lib.tc_hu_block_starts(self.ctx.handle, self._h, nblocks, s), "tc_hu_block_starts")
        return s

    def fox_workload(self):
        m = self.info().m_oriented
        w = np.empty(max(m, 1), np.int32)
        _check(_lib.tc_fox_workload(self.ctx.handle, self._h, w), "tc_fox_workload")
        return w[:m]

    def fox_bin_sizes(self, max_lanes: int = 32):
        nb = int(np.log2(max_lanes)) + 1
        out = np.zeros(nb, np.uint64)
        _check(_lib.tc_fox_bin_sizes(self.ctx.handle, self._h, max_lanes, out), "tc_fox_bin_sizes")
        return out

    # -- counting --
    @staticmethod
    def _sched(kernel, sched):
        if isinstance(kernel, str):
            kernel = KERNEL_NAMES[kernel]
        s = Sched()
        s.kernel = kernel
        for k, v in sched.items():
            if k == "auto_thresholds":
                for i, x in enumerate(v):
                    s.auto_thresholds[i] = x
            elif k == "auto_schedule":
                s.auto_thresholds[2] = {"worklist": AUTO_WORKLIST, "push": AUTO_PUSH}.get(v, v)
            else:
                setattr(s, k, v)
        return s

    def auto_stream_bytes(self, **sched) -> float:
        """Bytes one AUTO count streams under the schedule (tc_auto_stream_bytes)."""
        s = self._sched(KERNEL_AUTO, sched)
        out = C.c_double()
        _check(_lib.tc_auto_stream_bytes(self.ctx.handle, self._h, C.byref(s), C.byref(out)),
               "tc_auto_stream_bytes")
        return out.value

    def count(self, kernel=KERNEL_AUTO, **sched) -> CountResult:
        s = self._sched(kernel, sched)
        out = C.c_uint64()
        t = Timing()
        _check(_lib.tc_count(self.ctx.handle, self._h, C.byref(s), C.byref(out), C.byref(t)),
               "tc_count")
        return CountResult(out.value, t.as_dict())

    def cost_functions(self, k: int = 1):
        """(idPartCost, degreePartCost, lpPartCost, f_lp, f_degree) of the raw
        records: tools/costFuncCal.cpp and tools/verification.cpp."""
        out = np.zeros(5, np.float32)
        _check(_lib.tc_cost_functions(self.ctx.handle, self._h, k, out), "tc_cost_functions")
        return out

    def approx_ratio(self) -> float:
        r = C.c_double()
        _check(_lib.tc_approx_ratio(self.ctx.handle, self._h, C.byref(r)), "tc_approx_ratio")
        return r.value

    def count_per_vertex(self):
        n = self.info().n
        tri = np.zeros(max(n, 1), np.uint64)
        _check(_lib.tc_count_per_vertex(self.ctx.handle, self._h, tri), "tc_count_per_vertex")
        return tri[:n]

    def features(self, model=None):
        """Per-task features of the work-list schedule (tc_graph_features);
        feature 4 uses the model's bandwidth model (the reference's without
        one)."""
        n = self.info().n
        X = np.zeros((max(n, 1), N_FEATURES), np.float32)
        _check(_lib.tc_graph_features(self.ctx.handle, self._h, model.handle if model else None,
                                      X.reshape(-1)), "tc_graph_features")
        return X[:n]

    def measure_costs(self, variant: int):
        """(wave-microseconds per task under the variant, exact count of the
        timed run)."""
        n = self.info().n
        c = np.zeros(max(n, 1), np.float32)
        cnt = C.c_uint64()
        _check(_lib.tc_graph_measure_costs(self.ctx.handle, self._h, variant, c, C.byref(cnt)),
               "tc_graph_measure_costs")
        return c[:n], cnt.value

    def cell_features(self, model=None, level: int = 0):
        """Features (CELLS x 8) of the decision cells over the tasks of a
        sampling level (tc_graph_cell_features)."""
        X = np.zeros((CELLS, N_FEATURES), np.float32)
        _check(_lib.tc_graph_cell_features(self.ctx.handle, self._h,
                                           model.handle if model else None, level, X.reshape(-1)),
               "tc_graph_cell_features")
        return X

    def measure_cells(self, narrow: bool, level: int = 0, reps: int = 3):
        """(microseconds of each cell's count launch alone, its exact count),
        the cell run narrow or wide (tc_graph_measure_cells)."""
        us = np.zeros(CELLS, np.float32)
        cnt = np.zeros(CELLS, np.uint64)
        _check(_lib.tc_graph_measure_cells(self.ctx.handle, self._h, int(bool(narrow)), level,
                                           reps, us, cnt.ctypes.data), "tc_graph_measure_cells")
        return us, cnt

    def learned_cells(self, model=None):
        """(narrow-cell mask, predictions CELLS x 2) of the model's choice."""
        mask = C.c_uint32()
        Y = np.zeros((CELLS, 2), np.float32)
        _check(_lib.tc_learned_cells(self.ctx.handle, self._h, model.handle if model else None,
                                     C.byref(mask), Y.ctypes.data), "tc_learned_cells")
        return mask.value, Y

    def learned_assignment(self, model=None):
        n = self.info().n
        v = np.zeros(max(n, 1), np.uint8)
        _check(_lib.tc_learned_assignment(self.ctx.handle, self._h,
                                          model.handle if model else None, v),
               "tc_learned_assignment")
        return v[:n]

    def count_learned(self, model=None) -> CountResult:
        out = C.c_uint64()
        t = Timing()
        _check(_lib.tc_count_learned(self.ctx.handle, self._h, model.handle if model else None,
                                     C.byref(out), C.byref(t)), "tc_count_learned")
        return CountResult(out.value, t.as_dict())

    def kclique(self, k: int, variant: int = KCL_AUTO) -> CountResult:
        out = C.c_uint64()
        t = Timing()
        _check(_lib.kcl_count(self.ctx.handle, self._h, k, variant, C.byref(out), C.byref(t)),
               "kcl_count")
        return CountResult(out.value, t.as_dict())

    def kclique_roots(self, k: int, roots) -> CountResult:
        """k-cliques whose DAG source is one of `roots` (each counted once)."""
        r = np.ascontiguousarray(roots, np.uint32)
        out = C.c_uint64()
        t = Timing()
        _check(_lib.kcl_count_roots(self.ctx.handle, self._h, k, r.ctypes.data if len(r) else None,
                                    len(r), C.byref(out), C.byref(t)), "kcl_count_roots")
        return CountResult(out.value, t.as_dict())

    def kclique_features(self, k: int):
        """Per-root features (n x 8) of the learned k-clique plan."""
        n = self.info().n
        X = np.zeros((max(n, 1), N_FEATURES), np.float32)
        _check(_lib.kcl_root_features(self.ctx.handle, self._h, k, X.reshape(-1)),
               "kcl_root_features")
        return X[:n]

    def kclique_costs(self, k: int, variant: int):
        """(CU-microseconds per root on KCL_PIVOT or KCL_MERGE, k-cliques of
        the measured roots)."""
        n = self.info().n
        c = np.zeros(max(n, 1), np.float32)
        cnt = C.c_uint64()
        _check(_lib.kcl_measure_costs(self.ctx.handle, self._h, k, variant, c, C.byref(cnt)),
               "kcl_measure_costs")
        return c[:n], cnt.value

    def kclique_plan(self, k: int, model):
        """Path per root: 0 pivot, 1 merge, 255 none."""
        n = self.info().n
        out = np.zeros(max(n, 1), np.uint8)
        _check(_lib.kcl_learned_plan(self.ctx.handle, self._h, k, model.handle, out),
               "kcl_learned_plan")
        return out[:n]

    def kclique_learned(self, k: int, model=None) -> CountResult:
        out = C.c_uint64()
        t = Timing()
        _check(_lib.kcl_count_learned(self.ctx.handle, self._h, k, model.handle if model else None,
                                      C.byref(out), C.byref(t)), "kcl_count_learned")
        return CountResult(out.value, t.as_dict())


class Batch:
    """Many small graphs counted together: graphs is a sequence of
    (src, dst, n) with local ids < n.  Each graph is degree-oriented."""

    def __init__(self, ctx: Context, graphs):
        graphs = list(graphs)
        if not graphs:
            raise ValueError("empty batch")
        off = np.zeros(len(graphs) + 1, np.uint64)
        nv = np.zeros(len(graphs), np.uint32)
        for i, (s, d, n) in enumerate(graphs):
            if len(s) != len(d):
                raise ValueError(f"graph {i}: src and dst differ in length")
            off[i + 1] = off[i] + len(s)
            nv[i] = n
        m = int(off[-1])
        src = _u32(np.concatenate([g[0] for g in graphs])) if m else np.zeros(0, np.uint32)
        dst = _u32(np.concatenate([g[1] for g in graphs])) if m else np.zeros(0, np.uint32)
        h = _vp()
        _check(_lib.tc_batch_create(ctx.handle, len(graphs), off,
                                    src.ctypes.data if m else None,
                                    dst.ctypes.data if m else None, nv, C.byref(h)),
               "tc_batch_create")
        self._h = h
        self.ctx = ctx
        self.ngraphs = len(graphs)
        self.nedges = m
        self.nverts = int(nv.sum(dtype=np.uint64))
        ctx._children.add(self)

    def close(self):
        if getattr(self, "_h", None):
            _lib.tc_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @staticmethod
    def _sched(sched):
        if sched is None:
            return None
        if isinstance(sched, BatchSched):
            return C.byref(sched)
        s = BatchSched()
        for k, v in dict(sched).items():
            setattr(s, k, int(v))
        return C.byref(s)

    def info(self) -> GraphInfo:
        i = GraphInfo()
        _check(_lib.tc_batch_info_get(self._h, C.byref(i)), "tc_batch_info_get")
        return i

    def stream_bytes(self, form: int = BATCH_WORKLIST) -> float:
        """Bytes one batched count streams: the work-list bytes
        (tc_batch_stream_bytes), or for the push form 4(n+1) + 12 m+ +
        4 sum d- d+."""
        if form == BATCH_PUSH:
            i = self.info()
            return 4.0 * (i.n + 1) + 12.0 * i.m_oriented + 4.0 * i.sum_in_out
        out = C.c_double()
        _check(_lib.tc_batch_stream_bytes(self.ctx.handle, self._h, C.byref(out)),
               "tc_batch_stream_bytes")
        return out.value

    def autotune(self) -> BatchSched:
        s = BatchSched()
        _check(_lib.tc_batch_autotune(self.ctx.handle, self._h, C.byref(s)), "tc_batch_autotune")
        return s

    def count(self, sched=None):
        """Per-graph counts (uint64 array) and timing of one hipGraph replay."""
        out = np.zeros(self.ngraphs, np.uint64)
        t = Timing()
        _check(_lib.tc_batch_count(self.ctx.handle, self._h, self._sched(sched), out,
                                   C.byref(t)), "tc_batch_count")
        return out, t.as_dict()

    def count_pergraph(self, sched=None):
        """Same counts with one kernel launch per graph."""
        out = np.zeros(self.ngraphs, np.uint64)
        t = Timing()
        _check(_lib.tc_batch_count_pergraph(self.ctx.handle, self._h, self._sched(sched), out,
                                            C.byref(t)), "tc_batch_count_pergraph")
        return out, t.as_dict()


class Model:
    """The AUTO engine's cost model (8 -> 32 -> 32 -> 2 MLP)."""
    n_params = int(_lib.tc_model_num_params())

    def __init__(self, ctx: Context, seed: int = 0):
        h = _vp()
        _check(_lib.tc_model_create(ctx.handle, seed, C.byref(h)), "tc_model_create")
        self._h = h
        self.ctx = ctx
        ctx._children.add(self)

    @property
    def handle(self):
        return self._h

    def close(self):
        if getattr(self, "_h", None):
            _lib.tc_model_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def params(self):
        p = np.zeros(self.n_params, np.float32)
        _check(_lib.tc_model_get_params(self._h, p, self.n_params), "tc_model_get_params")
        return p

    @params.setter
    def params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        _check(_lib.tc_model_set_params(self._h, p, len(p)), "tc_model_set_params")

    @property
    def norm(self):
        xm, xs = np.zeros(N_FEATURES, np.float32), np.zeros(N_FEATURES, np.float32)
        ym, ys = np.zeros(N_OUTPUTS, np.float32), np.zeros(N_OUTPUTS, np.float32)
        _check(_lib.tc_model_get_norm(self._h, xm, xs, ym, ys), "tc_model_get_norm")
        return xm, xs, ym, ys

    @norm.setter
    def norm(self, v):
        xm, xs, ym, ys = (np.ascontiguousarray(a, np.float32) for a in v)
        _check(_lib.tc_model_set_norm(self._h, xm, xs, ym, ys), "tc_model_set_norm")

    def set_bw_model(self, bw: "BwModel"):
        """The list-read bandwidth model behind feature 4 (tc_model_set_bw_model)."""
        _check(_lib.tc_model_set_bw_model(self._h, C.byref(bw)), "tc_model_set_bw_model")
        return self

    def forward(self, X):
        X = np.ascontiguousarray(X, np.float32)
        Y = np.zeros((max(len(X), 1), N_OUTPUTS), np.float32)
        _check(_lib.tc_model_forward(self.ctx.handle, self._h, X.reshape(-1), len(X),
                                     Y.reshape(-1)), "tc_model_forward")
        return Y[:len(X)]

    def loss_grad(self, X, T, M):
        X = np.ascontiguousarray(X, np.float32)
        T = np.ascontiguousarray(T, np.float32)
        M = np.ascontiguousarray(M, np.float32)
        g = np.zeros(self.n_params, np.float32)
        loss = C.c_float()
        _check(_lib.tc_model_loss_grad(self.ctx.handle, self._h, X.reshape(-1), T.reshape(-1),
                                       M.reshape(-1), len(X), C.byref(loss), g),
               "tc_model_loss_grad")
        return float(loss.value), g

    def train(self, X, T, M, steps=3000, batch=4096, lr=3e-3, seed=0) -> float:
        X = np.ascontiguousarray(X, np.float32)
        T = np.ascontiguousarray(T, np.float32)
        M = np.ascontiguousarray(M, np.float32)
        o = TrainOpts()
        o.steps, o.batch, o.lr, o.seed = steps, batch, lr, seed
        loss = C.c_float()
        _check(_lib.tc_model_train(self.ctx.handle, self._h, X.reshape(-1), T.reshape(-1),
                                   M.reshape(-1), len(X), C.byref(o), C.byref(loss)),
               "tc_model_train")
        return float(loss.value)


def kcl_gpu_solver(path: str, k: int, n_chunk: int = 1) -> int:
    out = C.c_uint64()
    _check(_lib.kcl_gpu_solver(path.encode(), k, C.byref(out), n_chunk), "kcl_gpu_solver")
    return out.value


# ---- TriCore function API (tricore/include/tricount.h:7-17) ------------------

class tricore:
    @staticmethod
    def init_gpu() -> int:
        return int(_lib.tc_tricore_init_gpu())

    @staticmethod
    def pack(src, dst):
        """uint64 view of the record: FIRST = high word (tricore/include/util.h:14-16)."""
        return (_u32(src).astype(np.uint64) << np.uint64(32)) | _u32(dst).astype(np.uint64)

    @staticmethod
    def cal_degree(edges, node_count: int):
        edges = np.ascontiguousarray(edges, np.uint64)
        deg = np.zeros(max(node_count, 1), np.uint32)
        _lib.tc_tricore_cal_degree(edges, len(edges), deg, node_count)
        return deg[:node_count]

    @staticmethod
    def redirect_edges(edges, deg):
        deg = _u32(deg)
        _lib.tc_tricore_redirect_edges(edges, len(edges), deg, len(deg))
        return edges

    @staticmethod
    def cal_part_num(edges, node_count: int) -> int:
        return int(_lib.tc_tricore_cal_part_num(edges, len(edges), node_count))

    @staticmethod
    def count(edges, node_count: int, part_num: int) -> int:
        """split + tricount_gpu as tricore/src/main.cpp:75-95 chains them."""
        edges = np.ascontiguousarray(edges, np.uint64).copy()
        split = np.array([node_count if i == part_num else (node_count // part_num) * i
                          for i in range(part_num + 1)], np.uint32)
        node_index = (C.POINTER(C.c_uint32) * part_num)()
        adj_count = np.zeros(part_num + 1, np.uint64)
        _lib.tc_tricore_split(edges, len(edges), part_num, node_index, split, adj_count)
        if any(not bool(node_index[i]) for i in range(part_num)):
            raise RuntimeError("tricore split failed: " + last_error())
        total = _lib.tc_tricore_tricount_gpu(part_num, edges.ctypes.data, adj_count, split,
                                             node_index)
        libc = C.CDLL(None)
        libc.free.argtypes = [C.c_void_p]
        for i in range(part_num):
            libc.free(C.cast(node_index[i], C.c_void_p))
        return int(total)
