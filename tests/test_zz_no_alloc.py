"""tc_count_csr_noalloc (counting with the library's static scratch only)
against tc_count's AUTO count of the same oriented CSR, through the C ABI:
two calls in a row, the second on a smaller input, then a third on the
first input again (the cursors are cleared by every call)."""
import ctypes as C

import numpy as np
import pytest

import tcamd
from graphs import hub, rmat

pytestmark = pytest.mark.gpu


# The output buffer comes from the HIP runtime the engine itself is linked
# against (the same soname resolves to the already loaded library), so the
# test shares its device context.
_hip = C.CDLL("libamdhip64.so.7")
_hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
_hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
_hip.hipFree.argtypes = [C.c_void_p]
HIP_D2H, HIP_H2D = 2, 1


def noalloc_count(ctx, g, grid=0):
    rp = C.c_void_p()
    col = C.c_void_p()
    assert tcamd._lib.tc_graph_device_csr(g._h, C.byref(rp), C.byref(col)) == 0
    out = C.c_void_p()
    assert _hip.hipMalloc(C.byref(out), 8) == 0
    try:
        poison = C.c_uint64(0xFFFFFFFFFFFFFFFF)
        assert _hip.hipMemcpy(out, C.byref(poison), 8, HIP_H2D) == 0
        st = tcamd._lib.tc_count_csr_noalloc(ctx.handle, rp, col, g.info().n, grid, out)
        if st != 0:
            raise tcamd.TCError(st, "tc_count_csr_noalloc")
        # the call is asynchronous on the context's stream: wait for the device
        assert _hip.hipDeviceSynchronize() == 0
        got = C.c_uint64()
        assert _hip.hipMemcpy(C.byref(got), out, 8, HIP_D2H) == 0
        return got.value
    finally:
        _hip.hipFree(out)


def test_noalloc_matches_tc_count(ctx):
    big = rmat(15, 16, 3)
    small = rmat(10, 8, 5)
    results = []
    for (s, d), orient in ((big, tcamd.ORIENT_ADIR), (small, tcamd.ORIENT_DEGREE),
                           (big, tcamd.ORIENT_ADIR)):
        g = ctx.graph(s, d).orient(orient)
        want = g.count(tcamd.KERNEL_AUTO).count
        got = noalloc_count(ctx, g)
        print(f"n={g.info().n} m={g.info().m_oriented}: noalloc {got}, tc_count {want}", flush=True)
        assert got == want
        assert noalloc_count(ctx, g, grid=3) == want
        results.append(got)
        g.close()
    assert results[0] == results[2] and results[1] < results[0]


def test_noalloc_rows_beyond_lds_and_empty(ctx):
    # ID orientation puts the hub's 6 000 out-neighbours in one row (beyond the
    # 2048-entry LDS slot: searched in global memory)
    s, d = hub()
    g = ctx.graph(s, d).orient(tcamd.ORIENT_ID)
    assert g.info().max_out_degree > 2048
    assert noalloc_count(ctx, g) == g.count(tcamd.KERNEL_AUTO).count
    g.close()
    g = ctx.graph(np.array([0, 1], np.uint32), np.array([1, 2], np.uint32))
    g.orient(tcamd.ORIENT_DEGREE)
    assert noalloc_count(ctx, g) == 0
    g.close()


def test_noalloc_rejects_oversized_grid(ctx):
    s, d = rmat(8, 8, 1)
    g = ctx.graph(s, d).orient(tcamd.ORIENT_DEGREE)
    with pytest.raises(tcamd.TCError) as e:
        noalloc_count(ctx, g, grid=4097)
    assert e.value.status == tcamd.ERR_RANGE
    g.close()
