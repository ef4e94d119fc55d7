"""GPU parity: every kernel, orientation and preprocessing step of libtcamd.so
against the CPU oracle on the same seeded inputs (bit-exact)."""
import numpy as np
import pytest

import oracle as orc
import tcamd
from graphs import CASES, clique, er, hub, rmat

pytestmark = pytest.mark.gpu

ORIENTS = {
    "id": tcamd.ORIENT_ID,
    "degree": tcamd.ORIENT_DEGREE,
    "tricore": tcamd.ORIENT_DEGREE_TRICORE,
    "adir": tcamd.ORIENT_ADIR,
    "kcore": tcamd.ORIENT_KCORE,
    "degeneracy": tcamd.ORIENT_DEGENERACY,
}
KERNELS = ["hu", "bisson", "fox", "tricore", "auto"]

_truth = {}


def truth(name):
    if name not in _truth:
        s, d = CASES[name]()
        _truth[name] = orc.pipeline_count(s, d, "degree")[0]
    return _truth[name]


def oracle_csr(s, d, mode):
    n = orc.num_nodes(s, d)
    if mode in ("id", "degree", "tricore"):
        deg = orc.degree_raw(s, d, n)
        m = {"id": orc.ORIENT_ID, "degree": orc.ORIENT_DEGREE, "tricore": orc.ORIENT_DEG_TRICORE}[mode]
        os_, od = orc.orient_simple(s, d, deg, m)
        return orc.build_csr(os_, od, n)
    srp, scol = orc.build_sym(s, d, n)
    if mode == "adir":
        key = orc.adir_priority(n, srp, scol)
        os_, od = orc.adir_orient(n, srp, scol, key)
    elif mode == "kcore":
        key = orc.kcore(n, srp, scol)
        os_, od = orc.kcore_orient(n, srp, scol, key)
    else:
        key, _ = orc.degeneracy(n, srp, scol)
        os_, od = orc.kcore_orient(n, srp, scol, key)
    return orc.build_csr(os_, od, n)


@pytest.mark.parametrize("case", list(CASES))
@pytest.mark.parametrize("orient", list(ORIENTS))
def test_count_every_kernel(ctx, case, orient):
    s, d = CASES[case]()
    expect = truth(case)
    print(f"case={case} orient={orient} m={len(s)} expect={expect}", flush=True)
    g = ctx.graph(s, d).orient(ORIENTS[orient])
    for k in KERNELS:
        got = g.count(k).count
        assert got == expect, f"{k} on {case}/{orient}: {got} != {expect}"
    g.close()


@pytest.mark.parametrize("case", list(CASES))
@pytest.mark.parametrize("orient", list(ORIENTS))
def test_oriented_csr_matches_oracle(ctx, case, orient):
    s, d = CASES[case]()
    rp_o, col_o = oracle_csr(s, d, orient)
    g = ctx.graph(s, d).orient(ORIENTS[orient])
    rp, col = g.csr()
    np.testing.assert_array_equal(rp, rp_o)
    np.testing.assert_array_equal(col, col_o)
    info = g.info()
    dplus = np.diff(rp_o.astype(np.int64))
    indeg = np.bincount(col_o, minlength=len(dplus)).astype(np.int64)
    assert info.sum_out_sq == int((dplus * dplus).sum())
    assert info.sum_in_out == int((dplus * indeg).sum())
    assert info.max_out_degree == (int(dplus.max()) if len(dplus) else 0)
    g.close()


@pytest.mark.parametrize("case", ["er", "rmat12", "ba", "messy", "hub"])
def test_adir_priorities_exact(ctx, case):
    s, d = CASES[case]()
    n = orc.num_nodes(s, d)
    srp, scol = orc.build_sym(s, d, n)
    expect = orc.adir_priority(n, srp, scol)
    g = ctx.graph(s, d).orient(tcamd.ORIENT_ADIR)
    np.testing.assert_array_equal(g.vertex_key(), expect)
    g.close()


@pytest.mark.parametrize("case", ["er", "rmat12", "ba", "rgg", "messy", "hub"])
@pytest.mark.parametrize("thre_param", [1, 5, 20])
def test_adir_omp_priorities_exact(ctx, case, thre_param):
    """edge_directing/A-direction.cpp's schedule: priorities and oriented CSR
    equal the oracle's; the count is unchanged."""
    s, d = CASES[case]()
    n = orc.num_nodes(s, d)
    srp, scol = orc.build_sym(s, d, n)
    expect = orc.adir_omp_priority(n, srp, scol, thre_param)
    g = ctx.graph(s, d).orient_adir_omp(thre_param)
    np.testing.assert_array_equal(g.vertex_key(), expect)
    es, ed = orc.adir_orient(n, srp, scol, expect)
    rp, col = g.csr()
    np.testing.assert_array_equal(np.repeat(np.arange(len(rp) - 1), np.diff(rp)), es)
    np.testing.assert_array_equal(col, ed)
    assert g.info().orient_mode == tcamd.ORIENT_ADIR_OMP
    assert g.count("auto").count == truth(case)
    g.close()


def test_adir_omp_known_answers(ctx):
    """The oracle's hand-traced graphs: a 150-leaf star (no cascade below 201
    vertices), a 300-leaf star (cascade) and 16 layers of 210 vertices (the
    six-level cascade cap)."""
    g = ctx.graph(np.zeros(150, np.uint32), np.arange(1, 151, dtype=np.uint32))
    p = g.orient_adir_omp(10).vertex_key()
    assert p[0] == 50 and set(p[1:].tolist()) == {0}
    g.close()
    g = ctx.graph(np.zeros(300, np.uint32), np.arange(1, 301, dtype=np.uint32))
    p = g.orient_adir_omp(5).vertex_key()
    assert p[0] == 1 and set(p[1:].tolist()) == {0}
    g.close()
    src, dst = [], []
    for i in range(15):
        for j in range(210):
            for t in range(3 + (i == 7)):
                src.append(i * 210 + j)
                dst.append((i + 1) * 210 + (j + t) % 210)
    g = ctx.graph(np.array(src, np.uint32), np.array(dst, np.uint32))
    p = g.orient_adir_omp(1).vertex_key().reshape(16, 210)
    want = [0, 1, 2, 3, 4, 5, 6, 8, 8, 6, 5, 4, 3, 2, 1, 0]
    assert [sorted(set(r.tolist())) for r in p] == [[w] for w in want]
    with pytest.raises(tcamd.TCError):
        g.orient_adir_omp(0)
    g.close()


@pytest.mark.parametrize("case", ["er", "rmat12", "ba", "rgg", "messy"])
def test_kcore_numbers_exact(ctx, case):
    s, d = CASES[case]()
    n = orc.num_nodes(s, d)
    srp, scol = orc.build_sym(s, d, n)
    expect = orc.kcore(n, srp, scol)
    g = ctx.graph(s, d).orient(tcamd.ORIENT_KCORE)
    np.testing.assert_array_equal(g.vertex_key(), expect.astype(np.int32))
    g.close()


@pytest.mark.parametrize("case", ["er", "rmat12", "ba", "rgg", "messy"])
def test_degeneracy_steps_exact(ctx, case):
    s, d = CASES[case]()
    n = orc.num_nodes(s, d)
    srp, scol = orc.build_sym(s, d, n)
    step, core = orc.degeneracy(n, srp, scol)
    g = ctx.graph(s, d).orient(tcamd.ORIENT_DEGENERACY)
    np.testing.assert_array_equal(g.vertex_key(), step.astype(np.int32))
    rp, _ = g.csr()
    assert np.all(np.diff(rp.astype(np.int64)) <= core)
    g.close()


@pytest.mark.parametrize("case", ["er", "rmat12", "ba", "messy"])
def test_prune_deg1_matches_oracle(ctx, case):
    s, d = CASES[case]()
    n = orc.num_nodes(s, d)
    deg = orc.degree_raw(s, d, n)
    os_, od = orc.orient_simple(s, d, deg, orc.ORIENT_DEGREE)
    ps, pd, n2 = orc.prune_deg1(os_, od, n, deg)
    rp_o, col_o = orc.build_csr(ps, pd, n2)
    g = ctx.graph(s, d).orient(tcamd.ORIENT_DEGREE, prune_deg1=True)
    rp, col = g.csr()
    assert g.info().n == n2
    np.testing.assert_array_equal(rp, rp_o)
    np.testing.assert_array_equal(col, col_o)
    assert g.count("hu").count == truth(case)
    g.close()


@pytest.mark.parametrize("case", ["er", "rmat12", "ba", "rgg"])
def test_aorder_matches_oracle(ctx, case):
    s, d = CASES[case]()
    n = orc.num_nodes(s, d)
    res = orc.aorder(s, d, n)
    assert res is not None
    perm_o, os_, od = res
    g = ctx.graph(s, d)
    perm = g.relabel_aorder()
    np.testing.assert_array_equal(perm, perm_o)
    rs, rd = g.edges()
    np.testing.assert_array_equal(rs, perm_o[os_])
    np.testing.assert_array_equal(rd, perm_o[od])
    # the reference chain: A-direction-seri reads A-order's "reordered"
    # records and takes nodeNum = largest id + 1 (A-direction-seri.cpp:187-193)
    n_file = orc.num_nodes(rs, rd)
    assert g.info().n == n_file
    srp, scol = orc.build_sym(rs, rd, n_file)
    prio_o = orc.adir_priority(n_file, srp, scol)
    os2, od2 = orc.adir_orient(n_file, srp, scol, prio_o)
    rp_o, col_o = orc.build_csr(os2, od2, n_file)
    g.orient(tcamd.ORIENT_ADIR)
    np.testing.assert_array_equal(g.vertex_key(), prio_o)
    rp, col = g.csr()
    np.testing.assert_array_equal(rp, rp_o)
    np.testing.assert_array_equal(col, col_o)
    for k in KERNELS:
        assert g.count(k).count == truth(case)
    g.close()


@pytest.mark.parametrize("case", ["er", "rmat12", "ba", "messy", "hub"])
def test_reorient_after_prune(ctx, case):
    """Pruning shrinks the CSR's rows, not the raw records' id range: every
    later orientation of the same graph (and simplify, A-order) sees all ids."""
    s, d = CASES[case]()
    expect = truth(case)
    g = ctx.graph(s, d).orient(tcamd.ORIENT_DEGREE, prune_deg1=True)
    assert g.count("auto").count == expect
    for mode in ("id", "adir", "kcore", "degeneracy", "tricore", "degree"):
        g.orient(ORIENTS[mode])
        rp_o, col_o = oracle_csr(s, d, mode)
        rp, col = g.csr()
        np.testing.assert_array_equal(rp, rp_o)
        np.testing.assert_array_equal(col, col_o)
        for k in KERNELS:
            assert g.count(k).count == expect, (mode, k)
        g.orient(tcamd.ORIENT_DEGREE, prune_deg1=True)
    g.simplify()
    g.orient(tcamd.ORIENT_ADIR)
    assert g.count("auto").count == expect
    if case in ("er", "rmat12", "ba"):  # inside A-order's defined range
        g.orient(tcamd.ORIENT_DEGREE, prune_deg1=True)
        n_raw = g.info().n_raw
        assert n_raw == orc.num_nodes(s, d) and g.info().n <= n_raw
        perm = g.relabel_aorder()  # old -> new for every raw id, not just the CSR's rows
        assert len(perm) == n_raw
        np.testing.assert_array_equal(np.sort(perm), np.arange(n_raw))
        g.orient(tcamd.ORIENT_ADIR)
        assert g.count("auto").count == expect
    g.close()


@pytest.mark.parametrize("case", ["er", "rmat12", "hub"])
def test_work_models_match_oracle(ctx, case):
    s, d = CASES[case]()
    rp_o, col_o = oracle_csr(s, d, "degree")
    g = ctx.graph(s, d).orient(tcamd.ORIENT_DEGREE)
    w = g.hu_workload()
    w_o = orc.hu_workload(rp_o, col_o)
    np.testing.assert_array_equal(w, w_o)
    for nb in (1, 7, 300, 30000):
        np.testing.assert_array_equal(g.hu_block_starts(nb), orc.hu_block_starts(w_o, nb))
    fw = g.fox_workload()
    esrc = np.repeat(np.arange(len(rp_o) - 1, dtype=np.uint32), np.diff(rp_o.astype(np.int64)))
    fw_o = orc.fox_workload(esrc, col_o, np.diff(rp_o.astype(np.int64)).astype(np.uint32))
    np.testing.assert_array_equal(fw, fw_o)
    for lanes in (32, 64):
        bins = g.fox_bin_sizes(lanes)
        ref = orc.fox_bins(np.sort(fw_o), lanes)
        starts = [int(ref[1 << j]) for j in range(len(bins))] + [len(fw_o)]
        assert [int(x) for x in bins] == [starts[j + 1] - starts[j] for j in range(len(bins))]
    g.close()


@pytest.mark.parametrize("orient", ["degree", "id"])
@pytest.mark.parametrize("case", ["er", "rmat12", "hub", "messy", "k6", "ba"])
def test_per_vertex_counts(ctx, case, orient):
    # under ID orientation the hub's row (6 000 out-neighbours) is a huge task
    s, d = CASES[case]()
    rp_o, col_o = oracle_csr(s, d, orient)
    g = ctx.graph(s, d).orient(ORIENTS[orient])
    per = g.count_per_vertex()
    want = orc.count_per_vertex(rp_o, col_o)
    np.testing.assert_array_equal(per, want)
    assert int(per.sum()) == g.count("auto").count
    g.close()


@pytest.mark.parametrize("sched", [
    dict(kernel="bisson", bitmap_min_deg=2),
    dict(kernel="bisson", bitmap_min_deg=16, grid_size=7),
    dict(kernel="fox", fox_max_lanes=32),
    dict(kernel="fox", fox_max_lanes=1),
    dict(kernel="fox", fox_rearrange=1),
    dict(kernel="fox", fox_max_lanes=4, fox_rearrange=1),
    dict(kernel="tricore", tricore_parts=3),
    dict(kernel="tricore", tricore_parts=8, grid_size=5),
    dict(kernel="hu", grid_size=1),
    dict(kernel="hu", grid_size=977),
    dict(kernel="auto", grid_size=3),
    dict(kernel="auto", auto_thresholds=[2, 2]),
    dict(kernel="auto", auto_thresholds=[512, 512], grid_size=5),
    dict(kernel="auto", auto_thresholds=[16, 40]),
    dict(kernel="auto", auto_schedule="push"),
    dict(kernel="auto", auto_schedule="push", auto_thresholds=[2, 2]),
    dict(kernel="auto", auto_schedule="push", auto_thresholds=[512, 512], grid_size=5),
    dict(kernel="auto", auto_schedule="worklist", auto_thresholds=[16, 40], grid_size=3),
    dict(kernel="auto", auto_thresholds=[64, 2048, 0, 1]),          # no lane tasks
    dict(kernel="auto", auto_thresholds=[2, 2, 0, 4], grid_size=2),  # lane tasks up to d+ 4
    dict(kernel="auto", auto_thresholds=[8, 8, 0, 8], grid_size=1),
])
@pytest.mark.parametrize("case", ["rmat12", "hub", "k6", "messy"])
def test_schedule_variants(ctx, case, sched):
    s, d = CASES[case]()
    sched = dict(sched)
    k = sched.pop("kernel")
    for orient in (tcamd.ORIENT_ID, tcamd.ORIENT_DEGREE):
        g = ctx.graph(s, d).orient(orient)
        assert g.count(k, **sched).count == truth(case)
        g.close()


@pytest.mark.parametrize("width", [3, 8, 9])
def test_lane_rows_full_and_partial(ctx, width):
    # circulant graph C_n(1..width) under the ID orientation: almost every
    # vertex has d+ = width and a work list of width entries, so with width
    # 8 the lane tasks read full 32-byte rows (no padding), with 3 mostly
    # padding, and with 9 no task runs as a lane
    n = 3000
    i = np.arange(n, dtype=np.uint32)
    s = np.concatenate([i] * width)
    d = np.concatenate([(i + k) % n for k in range(1, width + 1)]).astype(np.uint32)
    rp_o, col_o = oracle_csr(s, d, "id")
    expect = orc.count(rp_o, col_o)
    print(f"width={width} expect={expect}", flush=True)
    assert np.median(np.diff(rp_o)) == width  # the wrap-around vertices hold up to 2 width
    g = ctx.graph(s, d).orient(tcamd.ORIENT_ID)
    for sched in (dict(), dict(auto_thresholds=[8, 8, 0, 8], grid_size=1),
                  dict(auto_thresholds=[2, 2, 0, 4]), dict(auto_thresholds=[64, 2048, 0, 1])):
        assert g.count("auto", **sched).count == expect, sched
    g.close()


def test_auto_hub_beyond_lds(ctx):
    """A row longer than the 8192-entry LDS slot is searched in global memory."""
    n = 12000
    hub = np.full(n - 1, n - 1, np.uint32)
    spokes = np.arange(n - 1, dtype=np.uint32)
    ring_s = np.arange(n - 1, dtype=np.uint32)
    ring_d = np.roll(ring_s, -1)
    s = np.concatenate([hub, ring_s])
    d = np.concatenate([spokes, ring_d])
    expect = orc.pipeline_count(s, d, "id")[0]
    assert expect == n - 1
    g = ctx.graph(s, d).orient(tcamd.ORIENT_ID)
    assert g.info().max_out_degree == n - 1
    for th in ([64, 2048], [512, 600], [2, 2]):
        assert g.count("auto", auto_thresholds=th).count == expect
    g.close()


def test_auto_rejects_bad_thresholds(ctx):
    s, d = CASES["k6"]()
    g = ctx.graph(s, d).orient(tcamd.ORIENT_DEGREE)
    with pytest.raises(tcamd.TCError):
        g.count("auto", auto_thresholds=[1024, 2048])
    with pytest.raises(tcamd.TCError):
        g.count("auto", auto_thresholds=[64, 32])
    g.close()


def test_empty_and_tiny(ctx):
    e = np.zeros(0, np.uint32)
    g = ctx.graph(e, e)
    g.orient(tcamd.ORIENT_DEGREE)
    for k in KERNELS:
        assert g.count(k).count == 0
    assert g.kclique(4).count == 0
    g.close()
    for (s, d, t) in [([0], [1], 0), ([0, 1, 2], [1, 2, 0], 1), ([3, 3], [3, 3], 0)]:
        g = ctx.graph(np.array(s, np.uint32), np.array(d, np.uint32))
        for o in ORIENTS.values():
            g.orient(o)
            for k in KERNELS:
                assert g.count(k).count == t


@pytest.mark.parametrize("case", ["er", "rmat12", "rgg", "k6", "hub"])
@pytest.mark.parametrize("variant", [tcamd.KCL_AUTO, tcamd.KCL_MERGE, tcamd.KCL_BITMAP,
                                     tcamd.KCL_PIVOT])
def test_kclique_matches_oracle(ctx, case, variant):
    s, d = CASES[case]()
    rp_o, col_o = oracle_csr(s, d, "kcore")
    g = ctx.graph(s, d).orient(tcamd.ORIENT_KCORE)
    for k in range(3, 8):
        expect = orc.kclique(rp_o, col_o, k)
        got = g.kclique(k, variant).count
        assert got == expect, f"k={k} variant={variant}: {got} != {expect}"
    g.close()


def test_kclique_dense_bitmap_widths(ctx):
    # K_n with n around each LDS bitmap width boundary: C(n, k) exactly
    from math import comb
    for n in (40, 70, 130, 200, 300):
        s, d = clique(n)
        g = ctx.graph(s, d).orient(tcamd.ORIENT_KCORE)
        for k in (3, 4, 5):
            assert g.kclique(k).count == comb(n, k), (n, k)
            assert g.kclique(k, tcamd.KCL_BITMAP).count == comb(n, k), (n, k)
        g.close()


def test_kclique_pivot_deep_and_wide(ctx):
    """Pivot path: cliques of up to 1024 vertices (every bit-matrix width, the
    deepest clique trees) and k up to 12; K_n holds C(n, k) k-cliques, and a
    disjoint union adds."""
    from math import comb
    for n in (12, 64, 65, 129, 300, 513, 1000):
        s, d = clique(n)
        g = ctx.graph(s, d).orient(tcamd.ORIENT_DEGREE)
        for k in (3, 4, 7, 12):
            if comb(n, k) < 2 ** 63:
                assert g.kclique(k, tcamd.KCL_PIVOT).count == comb(n, k), (n, k)
        g.close()
    # K_40 + K_25 + a sparse random part sharing no vertex
    a, b = clique(40)
    c, e = clique(25)
    rs, rd = er(2000, 12000, 9)
    s = np.concatenate([a, c + 40, rs + 65]).astype(np.uint32)
    d = np.concatenate([b, e + 40, rd + 65]).astype(np.uint32)
    rp_o, col_o = oracle_csr(rs, rd, "degree")
    g = ctx.graph(s, d).orient(tcamd.ORIENT_ADIR)
    for k in (3, 5, 8, 11):
        expect = comb(40, k) + comb(25, k) + orc.kclique(rp_o, col_o, k)
        assert g.kclique(k, tcamd.KCL_PIVOT).count == expect, k
        assert g.kclique(k).count == expect, k
    g.close()


def test_tricore_function_api(ctx):
    s, d = rmat(11, 8, 5)
    n = orc.num_nodes(s, d)
    edges = tcamd.tricore.pack(s, d)
    deg = tcamd.tricore.cal_degree(edges, n)
    np.testing.assert_array_equal(deg, orc.degree_raw(s, d, n))
    red = tcamd.tricore.redirect_edges(edges.copy(), deg)
    os_, od = orc.orient_simple(s, d, deg, orc.ORIENT_DEG_TRICORE)
    np.testing.assert_array_equal(red, tcamd.tricore.pack(os_, od))
    assert tcamd.tricore.cal_part_num(red, n) == 1
    expect = orc.pipeline_count(s, d, "degree")[0]
    for parts in (1, 2, 3, 7):
        assert tcamd.tricore.count(red, n, parts) == expect


def test_kcl_gpu_solver_file(ctx, tmp_path):
    s, d = er(500, 6000, 9)
    p = str(tmp_path / "g.bin")
    tcamd.write_edges(p, s, d)
    rp_o, col_o = oracle_csr(s, d, "kcore")
    for k in (3, 4, 5):
        want = orc.kclique(rp_o, col_o, k)
        for n_chunk in (1, 2, 7, 64):
            assert tcamd.kcl_gpu_solver(p, k, n_chunk) == want, (k, n_chunk)


def test_larger_rmat_all_kernels(ctx):
    s, d = rmat(16, 16, 22)
    expect, rp_o, col_o = orc.pipeline_count(s, d, "degree")
    g = ctx.graph(s, d)
    g.simplify()
    g.relabel_aorder()
    g.orient(tcamd.ORIENT_ADIR)
    for k in KERNELS:
        assert g.count(k).count == expect, k
    g.close()


@pytest.mark.parametrize("case", ["er", "rmat12", "ba", "rgg", "hub", "messy", "k6"])
def test_cost_functions_match_oracle(ctx, case):
    """tools/costFuncCal.cpp, tools/verification.cpp, tools/approximateRatio.cpp:
    the same float / double values, bit for bit."""
    s, d = CASES[case]()
    g = ctx.graph(s, d)
    for k in (0, 1, 3):
        got = g.cost_functions(k)
        ref = np.concatenate([np.asarray(orc.cost_func(s, d, k), np.float32),
                              np.asarray(orc.verification(s, d), np.float32)])
        np.testing.assert_array_equal(got, ref)
    assert g.approx_ratio() == orc.approx_ratio(s, d) or (
        np.isnan(g.approx_ratio()) and np.isnan(orc.approx_ratio(s, d)))
    g.close()


@pytest.mark.parametrize("case", ["er", "rmat12", "messy", "hub", "k6"])
def test_relabel_tools_match_oracle(ctx, case):
    """tools/serialize.cpp, adjust2tc.cpp, directionByReorder.cpp relabellings
    (the oracle is pinned against those tools' own builds) and Bisson's
    A-direction relabel (pinned by hand traces)."""
    s, d = CASES[case]()
    s = (s.astype(np.int64) * 3 + 2).astype(np.uint32)   # holes in the id range
    d = (d.astype(np.int64) * 3 + 2).astype(np.uint32)
    for mode, name in ((tcamd.RELABEL_COMPACT, "compact"), (tcamd.RELABEL_FIRST_SRC, "first_src"),
                       (tcamd.RELABEL_DEGREE, "degree"), (tcamd.RELABEL_ADIR, "adir")):
        g = ctx.graph(s, d)
        perm = g.relabel(mode)
        ref = orc.relabel(s, d, name)
        np.testing.assert_array_equal(perm, ref)
        rs, rd = g.edges()
        np.testing.assert_array_equal(rs, ref[s])
        np.testing.assert_array_equal(rd, ref[d])
        assert g.info().n == int(ref.max()) + 1
        g.orient(tcamd.ORIENT_DEGREE)
        assert g.count("auto").count == truth(case)
        g.close()
    e = np.zeros(0, np.uint32)
    g = ctx.graph(e, e)
    assert g.relabel(tcamd.RELABEL_DEGREE).tolist() == [0]
    g.close()


def test_bisson_relabel_hand_traces(ctx):
    """Bisson/A-direction.cpp on the two hand-traced record sets of the
    oracle's known-answer test (one needs the record slot of the frontier
    index), and on an R-MAT graph whose cascades run many levels."""
    for s, d, want in (([4, 3, 4, 2], [4, 1, 2, 3], [4, 3, 0, 1, 2]),
                       ([0, 0, 0, 1, 1, 2, 4, 5, 7], [1, 2, 3, 2, 3, 3, 5, 6, 7],
                        [3, 2, 1, 0, 6, 4, 5, 7])):
        g = ctx.graph(np.array(s, np.uint32), np.array(d, np.uint32))
        assert g.relabel(tcamd.RELABEL_ADIR).tolist() == want
        g.close()
    s, d = rmat(14, 16, 21)
    ref = orc.relabel(s, d, "adir")
    g = ctx.graph(s, d)
    np.testing.assert_array_equal(g.relabel(tcamd.RELABEL_ADIR), ref)
    g.close()


def test_aorder_bandwidth_model(ctx):
    # the reference model through the model entry point gives the same
    # permutation; the measured MI355X model gives a bijection and the count
    s, d = rmat(12, 16, 5)
    g1 = ctx.graph(s, d)
    g1.simplify()
    p_default = g1.relabel_aorder()
    g2 = ctx.graph(s, d)
    g2.simplify()
    p_ref = g2.relabel_aorder(tcamd.bw_model_reference())
    np.testing.assert_array_equal(p_default, p_ref)
    g3 = ctx.graph(s, d)
    g3.simplify()
    p_gpu = g3.relabel_aorder(tcamd.load_bw_model())
    assert np.array_equal(np.sort(p_gpu), np.arange(len(p_gpu)))
    expect = orc.pipeline_count(s, d, "degree")[0]
    for g in (g1, g2, g3):
        assert g.orient(tcamd.ORIENT_ADIR).count(tcamd.KERNEL_AUTO).count == expect
        g.close()
    # models the bucket heaps cannot order are rejected
    g4 = ctx.graph(s, d)
    for field, value in (("a", float("nan")), ("a", -1.0), ("b", float("inf"))):
        bad = tcamd.bw_model_reference()
        setattr(bad, field, value)
        with pytest.raises(tcamd.TCError):
            g4.relabel_aorder(bad)
    bad = tcamd.bw_model_reference()
    bad.table[3] = 0.0
    with pytest.raises(tcamd.TCError):
        g4.relabel_aorder(bad)
    g4.close()


def test_list_bandwidth_measurement(ctx):
    lens = [2, 9, 64, 512]
    bw = tcamd.measure_list_bandwidth(ctx, lens)
    assert np.all(bw > 0)
    assert bw[2] > bw[0]  # a wave reading 64 entries moves more than one reading 2
    with pytest.raises(tcamd.TCError):
        tcamd.measure_list_bandwidth(ctx, [0])


def test_valu_rate_probe(ctx):
    # the probe's work is a fixed number of rounds per lane: its time grows
    # with the rounds (eight times the rounds, four to twelve times the time)
    t1 = min(tcamd.measure_valu_rate(ctx, 2000) for _ in range(2))
    t8 = min(tcamd.measure_valu_rate(ctx, 16000) for _ in range(2))
    assert t1 > 0 and 4.0 < t8 / t1 < 12.0
    with pytest.raises(tcamd.TCError):
        tcamd.measure_valu_rate(ctx, 0)


@pytest.mark.parametrize("case", list(CASES))
def test_auto_strided_and_claimed_wave_tasks(ctx, case):
    # AUTO strides over the wave tasks when they are few (at most 16 per
    # wave) and claims them from a counter otherwise; a small grid makes the
    # same graph take the claimed path
    s, d = CASES[case]()
    expect = truth(case)
    g = ctx.graph(s, d).orient(tcamd.ORIENT_ADIR)
    for grid in (0, 1, 2, 7, 64):
        for sched in ("worklist", "push"):
            assert g.count(tcamd.KERNEL_AUTO, grid_size=grid, auto_schedule=sched).count == expect, (
                grid, sched)
    g.close()


def worklist_bytes(rp, col):
    """Bytes of the work-list schedule, recounted from a CSR: each edge u->v
    goes to the task of u when d+(v) <= d+(u), else of v; it is dropped when
    v has no out-edge or the owner u has d+(u) < 2.  A task x (d+(x) >= 2,
    non-empty list) reads 20 + 4 d+(x) bytes, each list entry y 8 + 4 d+(y)
    (the row bounds of y, stored with the entry, and N+(y))."""
    n = len(rp) - 1
    dp = np.diff(rp.astype(np.int64))
    src = np.repeat(np.arange(n), dp)
    du, dv = dp[src], dp[col]
    keep = ~((dv == 0) | ((dv <= du) & (du < 2)))
    push = dv <= du
    owner = np.where(push, src, col)[keep]
    streamed = np.where(push, col, src)[keep]
    nw = np.bincount(owner, minlength=n)
    tasks = (nw > 0) & (dp >= 2)
    assert nw[~tasks].sum() == 0  # every kept entry belongs to a runnable task
    return int((20 + 4 * dp[tasks]).sum() + (8 + 4 * dp[streamed]).sum())


@pytest.mark.parametrize("case", list(CASES))
@pytest.mark.parametrize("orient", ["degree", "adir", "id"])
def test_auto_stream_bytes(ctx, case, orient):
    s, d = CASES[case]()
    rp_o, col_o = oracle_csr(s, d, orient)
    g = ctx.graph(s, d).orient(ORIENTS[orient])
    assert g.auto_stream_bytes() == worklist_bytes(rp_o, col_o)
    info = g.info()
    assert g.auto_stream_bytes(auto_schedule="push") == (
        4 * (info.n + 1) + 12 * info.m_oriented + 4 * info.sum_in_out)
    g.close()


def test_kclique_auto_recounts_wide_orientations(ctx):
    # a hub joined to 1 500 vertices carrying a 40-clique: under the ID
    # orientation the hub's out-degree exceeds the LDS bit matrix, so AUTO
    # counts on a degeneracy-oriented copy; every variant and the oracle agree
    from math import comb
    a, b = clique(40)
    hub = np.full(1500, 6000, np.uint32)
    leaves = np.arange(1500, dtype=np.uint32)
    rs, rd = er(1500, 6000, 11)
    s = np.concatenate([a, hub, rs])
    d = np.concatenate([b, leaves, rd])
    rp_o, col_o = oracle_csr(s, d, "id")
    g = ctx.graph(s, d).orient(tcamd.ORIENT_ID)
    assert g.info().max_out_degree > 1024
    for k in (3, 4, 5, 6):
        expect = orc.kclique(rp_o, col_o, k)
        r1 = g.kclique(k)
        r2 = g.kclique(k)
        assert r1.count == r2.count == expect, k
        assert g.kclique(k, tcamd.KCL_MERGE).count == expect
    assert expect >= comb(40, 6) + comb(40, 5)
    g.close()


@pytest.mark.parametrize("case", ["rmat12", "rgg", "hub", "k6", "messy"])
def test_kclique_root_subsets_add_up(ctx, case):
    s, d = CASES[case]()
    g = ctx.graph(s, d).orient(tcamd.ORIENT_DEGENERACY)
    n = g.info().n
    rng = np.random.default_rng(7)
    part = rng.integers(0, 3, n)
    for k in (3, 4, 5):
        full = g.kclique(k).count
        parts = [g.kclique_roots(k, np.nonzero(part == j)[0]).count for j in range(3)]
        assert sum(parts) == full, (k, parts, full)
        assert g.kclique_roots(k, np.arange(n)).count == full
        # a listed root counts once however often it is listed
        twice = np.concatenate([np.nonzero(part == 0)[0]] * 2)
        assert g.kclique_roots(k, twice).count == parts[0]
    assert g.kclique_roots(4, np.zeros(0, np.uint32)).count == 0
    with pytest.raises(tcamd.TCError):
        g.kclique_roots(4, np.array([n], np.uint32))
    g.close()


@pytest.mark.parametrize("case", ["er", "rmat12", "hub", "messy"])
def test_csr_stage_file_round_trip(ctx, case, tmp_path):
    """An oriented CSR saved and loaded back counts like the original under
    every method (k-cliques too, through the degeneracy copy that wide
    orientations take), and refuses the passes that need raw records."""
    s, d = CASES[case]()
    expect = truth(case)
    for orient in ("id", "adir", "degeneracy"):
        g = ctx.graph(s, d).orient(ORIENTS[orient])
        p = str(tmp_path / f"{orient}.csr")
        g.save_csr(p)
        h = ctx.load_csr(p)
        a, b = g.info(), h.info()
        assert (a.n, a.m_oriented, a.orient_mode, a.sum_out_sq, a.sum_in_out, a.max_out_degree) == (
            b.n, b.m_oriented, b.orient_mode, b.sum_out_sq, b.sum_in_out, b.max_out_degree)
        for x, y in zip(g.csr(), h.csr()):
            np.testing.assert_array_equal(x, y)
        if orient != "id":
            np.testing.assert_array_equal(g.vertex_key(), h.vertex_key())
        for k in KERNELS:
            assert h.count(k).count == expect, (orient, k)
        for k in (4, 5):
            assert h.kclique(k).count == g.kclique(k).count, (orient, k)
        with pytest.raises(tcamd.TCError) as e:
            h.orient(tcamd.ORIENT_DEGREE)
        assert e.value.status == tcamd.ERR_STATE
        g.close()
        h.close()


def test_csr_stage_file_rejects_bad_files(ctx, tmp_path):
    s, d = CASES["k6"]()
    g = ctx.graph(s, d).orient(tcamd.ORIENT_DEGREE)
    p = tmp_path / "g.csr"
    g.save_csr(str(p))
    rp, _ = g.csr()
    g.close()
    raw = bytearray(p.read_bytes())
    n = len(rp) - 1
    c0 = 32 + 4 * (n + 1)  # 32-byte header, then rowptr[n + 1], then col
    row = int(np.nonzero(np.diff(rp.astype(np.int64)) >= 2)[0][0])  # a row of 2+ entries
    e = c0 + 4 * int(rp[row])
    swapped = raw[:e] + raw[e + 4:e + 8] + raw[e:e + 4] + raw[e + 8:]
    bad = {
        "magic": b"XXXXXXXX" + raw[8:],
        "truncated": raw[:-4],
        "trailing": raw + b"\0",
        "col_range": raw[:c0] + (n + 93).to_bytes(4, "little") + raw[c0 + 4:],
        "col_order": swapped,
    }
    for name, data in bad.items():
        q = tmp_path / f"{name}.csr"
        q.write_bytes(bytes(data))
        with pytest.raises(tcamd.TCError) as e:
            ctx.load_csr(str(q))
        assert e.value.status == tcamd.ERR_FORMAT, name


@pytest.mark.parametrize("case", list(CASES))
def test_fox_rearrange_keeps_bins_and_count(ctx, case):
    """rearrangeEdge permutes each bin's edges: the bin sizes and the count
    stay those of the plain Fox schedule, and switching back and forth
    rebuilds the schedule each time."""
    s, d = CASES[case]()
    g = ctx.graph(s, d).orient(tcamd.ORIENT_ID)
    expect = truth(case)
    for flag in (0, 1, 0, 1):
        assert g.count("fox", fox_rearrange=flag).count == expect, flag
    plain = g.fox_bin_sizes(64)
    g.count("fox", fox_rearrange=1)
    np.testing.assert_array_equal(g.fox_bin_sizes(64), plain)
    with pytest.raises(tcamd.TCError):
        g.count("fox", fox_rearrange=2)
    g.close()
