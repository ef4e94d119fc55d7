"""CPU tests that pin the oracle: brute force, closed forms, networkx, the
reference's own buildable tools (oracle/_ref), and hand-traced cases."""
import itertools
import os
import subprocess
from math import comb

import networkx as nx
import numpy as np
import pytest

import oracle as orc
import tcamd
from graphs import clique, er, messy, rmat

REF = orc.REF_DIR


def brute_triangles(src, dst):
    """O(n^3) count on the simple graph of the raw records."""
    n = int(max(src.max(), dst.max())) + 1 if len(src) else 0
    A = np.zeros((n, n), bool)
    for s, d in zip(src, dst):
        if s != d:
            A[s, d] = A[d, s] = True
    t = 0
    for i, j, k in itertools.combinations(range(n), 3):
        t += A[i, j] and A[j, k] and A[i, k]
    return t


def nx_graph(src, dst):
    G = nx.Graph()
    G.add_edges_from(zip(src.tolist(), dst.tolist()))
    G.remove_edges_from(list(nx.selfloop_edges(G)))
    return G


def nx_triangles(src, dst):
    return sum(nx.triangles(nx_graph(src, dst)).values()) // 3


MODES = ["id", "degree", "tricore", "adir"]


@pytest.mark.parametrize("seed", range(6))
def test_count_vs_brute_force(seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(5, 40))
    m = int(rng.integers(0, 3 * n))
    src = rng.integers(0, n, m).astype(np.uint32)
    dst = rng.integers(0, n, m).astype(np.uint32)  # duplicates and self-loops on purpose
    if m == 0:
        return
    expect = brute_triangles(src, dst)
    for mode in MODES:
        assert orc.pipeline_count(src, dst, mode)[0] == expect, mode


def wheel(n):  # hub n + cycle 0..n-1
    s = list(range(n)) + [n] * n
    d = [(i + 1) % n for i in range(n)] + list(range(n))
    return np.array(s, np.uint32), np.array(d, np.uint32)


def friendship(k):  # k triangles sharing vertex 0
    s, d = [], []
    for i in range(k):
        a, b = 2 * i + 1, 2 * i + 2
        s += [0, 0, a]
        d += [a, b, b]
    return np.array(s, np.uint32), np.array(d, np.uint32)


def biclique(a, b):
    s, d = zip(*[(i, a + j) for i in range(a) for j in range(b)])
    return np.array(s, np.uint32), np.array(d, np.uint32)


def grid(r, c):
    s, d = [], []
    for i in range(r):
        for j in range(c):
            v = i * c + j
            if j + 1 < c:
                s.append(v); d.append(v + 1)
            if i + 1 < r:
                s.append(v); d.append(v + c)
    return np.array(s, np.uint32), np.array(d, np.uint32)


@pytest.mark.parametrize("graph,expect", [
    (clique(12), comb(12, 3)),
    (clique(3), 1),
    (wheel(9), 9),
    (wheel(4), 4),  # hub + C_4: one triangle per cycle edge
    (friendship(7), 7),
    (biclique(6, 9), 0),
    (grid(7, 9), 0),
])
def test_count_closed_forms(graph, expect):
    s, d = graph
    for mode in MODES:
        assert orc.pipeline_count(s, d, mode)[0] == expect


def test_disjoint_union_adds():
    s1, d1 = clique(7)
    s2, d2 = friendship(5)
    off = np.uint32(7)
    s = np.concatenate([s1, s2 + off])
    d = np.concatenate([d1, d2 + off])
    assert orc.pipeline_count(s, d, "degree")[0] == comb(7, 3) + 5


@pytest.mark.parametrize("make", [er, lambda: rmat(10, 8, 3), messy])
def test_count_vs_networkx(make):
    s, d = make()
    expect = nx_triangles(s, d)
    for mode in MODES:
        assert orc.pipeline_count(s, d, mode)[0] == expect


def test_count_invariant_under_relabel():
    s, d = rmat(11, 8, 9)
    n = orc.num_nodes(s, d)
    perm = np.random.default_rng(0).permutation(n).astype(np.uint32)
    a = orc.pipeline_count(s, d, "degree")[0]
    b = orc.pipeline_count(perm[s], perm[d], "degree")[0]
    assert a == b


def test_oracle_internal_consistency():
    s, d = rmat(11, 8, 4)
    c, rp, col = orc.pipeline_count(s, d, "degree")
    assert orc.count_omp(rp, col) == c
    per = orc.count_per_vertex(rp, col)
    assert int(per.sum()) == c
    n = len(rp) - 1
    assert orc.count_rows(rp, col, 0, n // 3) + orc.count_rows(rp, col, n // 3, n) == c
    assert sum(orc.count_rows_strided(rp, col, k, 3) for k in range(3)) == c


def _is_dag(rp, col):
    G = nx.DiGraph()
    n = len(rp) - 1
    G.add_nodes_from(range(n))
    for u in range(n):
        for e in range(rp[u], rp[u + 1]):
            G.add_edge(u, int(col[e]))
    return nx.is_directed_acyclic_graph(G)


@pytest.mark.parametrize("mode", MODES + ["kcore", "degeneracy"])
def test_orientations_are_dags_covering_each_edge_once(mode):
    s, d = rmat(9, 8, 11)
    n = orc.num_nodes(s, d)
    if mode in ("kcore", "degeneracy"):
        srp, scol = orc.build_sym(s, d, n)
        key = orc.kcore(n, srp, scol) if mode == "kcore" else orc.degeneracy(n, srp, scol)[0]
        os_, od = orc.kcore_orient(n, srp, scol, key)
        rp, col = orc.build_csr(os_, od, n)
    else:
        _, rp, col = orc.pipeline_count(s, d, mode)
    assert _is_dag(rp, col)
    und = {(min(a, b), max(a, b)) for a, b in zip(s.tolist(), d.tolist()) if a != b}
    got = []
    for u in range(n):
        for e in range(rp[u], rp[u + 1]):
            got.append((min(u, int(col[e])), max(u, int(col[e]))))
    assert len(got) == len(set(got)) == len(und) and set(got) == und
    assert all(np.all(np.diff(col[rp[u]:rp[u + 1]].astype(np.int64)) > 0) for u in range(n))


# ---- A-direction: hand-traced priorities (A-direction-seri.cpp:35-153) ------

def _prio(s, d):
    s, d = np.array(s, np.uint32), np.array(d, np.uint32)
    n = orc.num_nodes(s, d)
    srp, scol = orc.build_sym(s, d, n)
    return orc.adir_priority(n, srp, scol).tolist()


def test_adir_hand_traces():
    # path 0-1-2: threshold 4/3+1; both ends peel in round 0 (label 0), the
    # middle drops to 2 after its first peeled neighbour -> label 1.
    assert _prio([0, 1], [1, 2]) == [0, 1, 0]
    # star with 4 leaves: threshold 8/5+1 = 2.6; leaves label 0, the centre
    # reaches 2 during the leaves' level -> label 1.
    assert _prio([0, 0, 0, 0], [1, 2, 3, 4]) == [1, 0, 0, 0, 0]
    # K4: doubled degree 6 > threshold 4; round 0 peels nothing (label 0
    # unused), round 1 (threshold 8) peels all with label 1.
    s, d = clique(4)
    assert _prio(s, d) == [1, 1, 1, 1]
    # triangle with a pendant vertex 3 on vertex 0: threshold 8/4+1 = 3;
    # vertex 3 (2) and vertices 1, 2 (4 > 3)...: round 0 peels 3 (label 0),
    # vertex 0 drops 6 -> 4 -> still > 3; nothing else.  Round 1 (thr 6)
    # peels 0, 1, 2 together (label 2).
    assert _prio([0, 1, 2, 0], [1, 2, 0, 3]) == [2, 2, 2, 0]


def test_adir_emit_rule_orders_by_priority_then_id():
    s, d = rmat(9, 8, 5)
    n = orc.num_nodes(s, d)
    srp, scol = orc.build_sym(s, d, n)
    p = orc.adir_priority(n, srp, scol)
    os_, od = orc.adir_orient(n, srp, scol, p)
    key = lambda v: (int(p[v]), int(v))
    assert all(key(a) < key(b) for a, b in zip(os_.tolist(), od.tolist()))
    assert len(os_) * 2 == len(scol)


def test_kcore_vs_networkx():
    for s, d in [rmat(10, 8, 2), er(400, 3000, 8)]:
        n = orc.num_nodes(s, d)
        srp, scol = orc.build_sym(s, d, n)
        core = orc.kcore(n, srp, scol)
        G = nx_graph(s, d)
        G.add_nodes_from(range(n))
        ref = nx.core_number(G)
        assert [int(core[v]) for v in range(n)] == [ref[v] for v in range(n)]


# ---- degeneracy order (level-synchronous k-core peel) -----------------------

def _degen(s, d):
    s, d = np.array(s, np.uint32), np.array(d, np.uint32)
    n = orc.num_nodes(s, d)
    srp, scol = orc.build_sym(s, d, n)
    step, core = orc.degeneracy(n, srp, scol)
    return step.tolist(), core.tolist()


def test_degeneracy_hand_traces():
    # path 0-1-2-3: level 1 removes both ends (step 0), then the middle pair,
    # whose degrees fall to 1 (step 1)
    assert _degen([0, 1, 2], [1, 2, 3]) == ([0, 1, 1, 0], [1, 1, 1, 1])
    # path of five: the removal runs inwards one step at a time
    assert _degen([0, 1, 2, 3], [1, 2, 3, 4]) == ([0, 1, 2, 1, 0], [1] * 5)
    # star, centre 0: leaves at step 0, the centre drops to 1 at step 1
    assert _degen([0, 0, 0, 0], [1, 2, 3, 4]) == ([1, 0, 0, 0, 0], [1] * 5)
    # K4: nothing below level 3, then all four at once
    s, d = clique(4)
    assert _degen(s, d) == ([0] * 4, [3] * 4)
    # triangle with pendant 3 on vertex 0: the pendant goes at level 1; vertex
    # 0 falls to 2, which is not <= 1; level 2 takes the triangle in one step
    assert _degen([0, 1, 2, 0], [1, 2, 0, 3]) == ([1, 1, 1, 0], [2, 2, 2, 1])


@pytest.mark.parametrize("graph", ["rmat", "er", "messy"])
def test_degeneracy_order_bounds_out_degree_by_core(graph):
    s, d = {"rmat": lambda: rmat(11, 16, 3), "er": lambda: er(2000, 30000, 4),
            "messy": messy}[graph]()
    n = orc.num_nodes(s, d)
    srp, scol = orc.build_sym(s, d, n)
    step, core = orc.degeneracy(n, srp, scol)
    # the levels are the core numbers (networkx-pinned orc_kcore)
    np.testing.assert_array_equal(core, orc.kcore(n, srp, scol))
    os_, od = orc.kcore_orient(n, srp, scol, step)
    rp, col = orc.build_csr(os_, od, n)
    dp = np.diff(rp.astype(np.int64))
    assert np.all(dp <= core) and dp.max() <= core.max()
    # with the (core, id) key of kcore_direction.cpp the bound fails on R-MAT
    rk, _ = orc.build_csr(*orc.kcore_orient(n, srp, scol, core), n)
    if graph == "rmat":
        assert np.any(np.diff(rk.astype(np.int64)) > core)
    # every orientation keeps the triangle count
    assert orc.count_rows(rp, col, 0, n) == orc.count_rows(rk, _, 0, n)


# ---- k-clique ---------------------------------------------------------------

def _kcore_csr(s, d):
    n = orc.num_nodes(s, d)
    srp, scol = orc.build_sym(s, d, n)
    os_, od = orc.kcore_orient(n, srp, scol, orc.kcore(n, srp, scol))
    return orc.build_csr(os_, od, n)


def turan(n, r):
    part = [i * r // n for i in range(n)]
    pairs = [(i, j) for i in range(n) for j in range(i + 1, n) if part[i] != part[j]]
    s, d = zip(*pairs)
    return np.array(s, np.uint32), np.array(d, np.uint32)


def test_kclique_closed_forms():
    rp, col = _kcore_csr(*clique(14))
    for k in range(1, 9):
        assert orc.kclique(rp, col, k) == comb(14, k)
    rp, col = _kcore_csr(*turan(20, 5))  # 5 parts of 4
    for k in range(2, 7):
        assert orc.kclique(rp, col, k) == comb(5, k) * 4 ** k


def test_kclique_vs_networkx():
    for s, d in [er(120, 1500, 3), rmat(8, 8, 7)]:
        rp, col = _kcore_csr(s, d)
        G = nx_graph(s, d)
        sizes = {}
        for c in nx.enumerate_all_cliques(G):
            sizes[len(c)] = sizes.get(len(c), 0) + 1
        for k in range(3, 7):
            assert orc.kclique(rp, col, k) == sizes.get(k, 0)
        assert orc.kclique(rp, col, 3) == nx_triangles(s, d)


# ---- work models: structural properties --------------------------------------

def test_hu_workload_and_block_starts_properties():
    s, d = rmat(11, 8, 6)
    _, rp, col = orc.pipeline_count(s, d, "degree")
    w = orc.hu_workload(rp, col)
    dp = np.diff(rp.astype(np.int64))
    # W(u) = 0 whenever d+(u) <= 1 and is a multiple of floor(log2 d+(u))
    assert np.all(w[dp <= 1] == 0)
    lg = np.floor(np.log2(np.maximum(dp, 1))).astype(np.int64)
    ok = lg > 0
    assert np.all(w[ok] % lg[ok] == 0)
    wedges = np.array([dp[col[rp[u]:rp[u + 1]]].sum() for u in range(len(dp))])
    assert np.array_equal(w[ok] // lg[ok], wedges[ok])
    pre = np.cumsum(w)
    for nb in (1, 5, 64, 1000):
        st = orc.hu_block_starts(w, nb)
        step = (int(pre[-1]) + nb - 1) // nb
        assert st[0] == 0 and st[nb] == len(w) and np.all(np.diff(st) >= 0)
        for i in range(1, nb):  # lower_bound of i*step+1 in the prefix
            assert st[i] == np.searchsorted(pre, i * step + 1, side="left")


def test_work_models_known_answers():
    # oriented CSR by hand: 0->{1,2,3,4}, 1->{2}, 2->{3,4}, 3->{4,5,6}, 4->{}, ...
    rp = np.array([0, 4, 5, 7, 10, 10, 10, 10], np.uint32)
    col = np.array([1, 2, 3, 4, 2, 3, 4, 4, 5, 6], np.uint32)
    w = orc.hu_workload(rp, col)
    # W(0) = (1+2+3+0) * floor(log2 4) = 12; W(1) = 0 (d+ = 1);
    # W(2) = (3+0) * 1 = 3; W(3) = 0 * 1 = 0
    assert w.tolist() == [12, 0, 3, 0, 0, 0, 0]
    esrc = np.repeat(np.arange(7, dtype=np.uint32), np.diff(rp.astype(np.int64)))
    fw = orc.fox_workload(esrc, col, np.diff(rp.astype(np.int64)).astype(np.uint32))
    # edge 0->3: min(4,3)=3, log2(4+2)=2.585 -> 7; edge 0->1: 1*log2(6) -> 2;
    # edge 2->3: min(2,3)=2 * log2(5) = 4.64 -> 4; edge 3->5: 0 -> 0
    assert fw[2] == 7 and fw[0] == 2 and fw[5] == 4 and fw[8] == 0


def test_fox_bins_partition_sorted_work():
    s, d = rmat(11, 8, 6)
    _, rp, col = orc.pipeline_count(s, d, "id")
    dp = np.diff(rp.astype(np.int64))
    esrc = np.repeat(np.arange(len(dp), dtype=np.uint32), dp)
    ws = np.sort(orc.fox_workload(esrc, col, dp.astype(np.uint32)))
    bins = orc.fox_bins(ws, 32)
    starts = [int(bins[1 << j]) for j in range(6)] + [len(ws)]
    assert starts[0] == 0 and all(x <= y for x, y in zip(starts, starts[1:]))
    # every bin boundary is a lower_bound position in the sorted work array
    for j in range(1, 6):
        p = starts[j]
        assert p == len(ws) or p == 0 or ws[p - 1] < ws[p]


# ---- reference tools built from /root/reference (oracle/_ref) ---------------

needs_ref = pytest.mark.skipif(not os.path.exists(os.path.join(REF, "A-order")),
                               reason="oracle/_ref not built (reference tree absent)")


def _run_ref(tool, args, cwd):
    return subprocess.run([os.path.join(REF, tool)] + args, cwd=cwd, check=True,
                          capture_output=True, text=True).stdout


@needs_ref
@pytest.mark.parametrize("make", [er, lambda: rmat(12, 8, 21), lambda: tcamd.gen_ba(3000, 6, 2)])
def test_aorder_matches_reference_binary(tmp_path, make):
    s, d = make()
    p = str(tmp_path / "g.bin")
    tcamd.write_edges(p, s, d)
    _run_ref("A-order", [p], str(tmp_path))
    ns, nd = tcamd.read_edges(str(tmp_path / "normal"))
    rs, rd = tcamd.read_edges(str(tmp_path / "reordered"))
    n = orc.num_nodes(s, d)
    res = orc.aorder(s, d, n)
    assert res is not None
    perm, os_, od = res
    assert np.array_equal(ns, os_) and np.array_equal(nd, od)
    assert np.array_equal(rs, perm[os_]) and np.array_equal(rd, perm[od])


def _ref_floats(out):
    vals = {}
    for line in out.splitlines():
        if ":" in line:
            k, v = line.split(":", 1)
            vals[k.strip()] = v.strip()
        elif " is " in line:
            k, v = line.rsplit(" is ", 1)
            vals[k.strip()] = v.strip()
    return vals


@needs_ref
@pytest.mark.parametrize("k", [0, 1, 3])
def test_cost_functions_match_reference(tmp_path, k):
    s, d = rmat(12, 8, 13)
    p = str(tmp_path / "g.bin")
    tcamd.write_edges(p, s, d)
    got = orc.cost_func(s, d, k)
    ref = _ref_floats(_run_ref("costFuncCal", [p, str(k)], str(tmp_path)))
    for i, key in enumerate(["idPartCost", "degreePartCost", "lpPartCost"]):
        assert ref[key] == f"{float(got[i]):g}"


@needs_ref
def test_verification_and_approx_ratio_match_reference(tmp_path):
    s, d = rmat(12, 8, 17)
    txt = str(tmp_path / "g.txt")
    with open(txt, "w") as f:
        for a, b in zip(s.tolist(), d.tolist()):
            f.write(f"{a} {b}\n")
    out = _run_ref("verification", [txt], str(tmp_path))
    v = orc.verification(s, d)
    lines = out.splitlines()
    assert lines[0].endswith(f"{float(v[0]):g}") and lines[1].endswith(f"{float(v[1]):g}")
    p = str(tmp_path / "g.bin")
    tcamd.write_edges(p, s, d)
    out = _run_ref("approximateRatio", [p], str(tmp_path))
    assert out.strip().endswith(f"{orc.approx_ratio(s, d):g}")


@needs_ref
def test_list2bin_bin2list_match_reference(tmp_path):
    s, d = er(300, 2000, 4)
    txt = str(tmp_path / "g.txt")
    with open(txt, "w") as f:
        for a, b in zip(s.tolist(), d.tolist()):
            f.write(f"{a} {b}\n")
    _run_ref("list2bin", [txt, str(tmp_path / "ref.bin")], str(tmp_path))
    assert tcamd.list2bin(txt, str(tmp_path / "ours.bin")) == len(s)
    assert open(tmp_path / "ref.bin", "rb").read() == open(tmp_path / "ours.bin", "rb").read()
    _run_ref("bin2list", [str(tmp_path / "ref.bin"), str(tmp_path / "ref.txt")], str(tmp_path))
    tcamd.bin2list(str(tmp_path / "ours.bin"), str(tmp_path / "ours.txt"))
    assert open(tmp_path / "ref.txt").read() == open(tmp_path / "ours.txt").read()


def test_kclique_rows_strided_partitions_the_count():
    s, d = rmat(9, 8, 3)
    n = orc.num_nodes(s, d)
    deg = orc.degree_raw(s, d, n)
    os_, od = orc.orient_simple(s, d, deg, orc.ORIENT_DEGREE)
    rp, col = orc.build_csr(os_, od, n)
    for k in (3, 4):
        total = orc.kclique(rp, col, k)
        parts = sum(orc.kclique_rows_strided(rp, col, k, r, 3) for r in range(3))
        assert parts == total


@needs_ref
@pytest.mark.parametrize("make", [lambda: er(300, 2000, 4), messy, lambda: rmat(9, 8, 2)])
def test_relabel_tools_match_reference(tmp_path, make):
    """tools/serialize.cpp, tools/adjust2tc.cpp, tools/directionByReorder.cpp
    pin the oracle's relabellings (records rewritten through perm)."""
    s, d = make()
    # leave holes in the id range so that serialize has ids to squeeze out
    s = (s.astype(np.int64) * 3 + 1).astype(np.uint32)
    d = (d.astype(np.int64) * 3 + 1).astype(np.uint32)
    p = str(tmp_path / "g.bin")
    tcamd.write_edges(p, s, d)
    for tool, mode in (("serialize", "compact"), ("directionByReorder", "degree")):
        out = str(tmp_path / f"{tool}.bin")
        _run_ref(tool, [p, out], str(tmp_path))
        rs, rd = tcamd.read_edges(out)
        perm = orc.relabel(s, d, mode)
        np.testing.assert_array_equal(rs, perm[s])
        np.testing.assert_array_equal(rd, perm[d])
    txt = str(tmp_path / "g.txt")
    with open(txt, "w") as f:
        for a, b in zip(s.tolist(), d.tolist()):
            f.write(f"{a} {b}\n")
    _run_ref("adjust2tc", [txt, str(tmp_path / "adj.txt")], str(tmp_path))
    got = np.loadtxt(str(tmp_path / "adj.txt"), dtype=np.int64).reshape(-1, 2)
    perm = orc.relabel(s, d, "first_src")
    np.testing.assert_array_equal(got[:, 0], perm[s])
    np.testing.assert_array_equal(got[:, 1], perm[d])


def test_relabel_oracle_properties():
    s, d = rmat(8, 8, 5)
    n = orc.num_nodes(s, d)
    for mode in ("compact", "first_src", "degree"):
        perm = orc.relabel(s, d, mode)
        assert len(perm) == n
        if mode != "compact":
            assert sorted(perm.tolist()) == list(range(n))  # a bijection on [0, n)
    deg = np.bincount(np.concatenate([s, d]), minlength=n)
    perm = orc.relabel(s, d, "degree")
    assert perm[np.argmax(deg)] == 0 or deg[perm == 0][0] == deg.max()


def test_parallel_sample_counter_equals_serial():
    s, d = rmat(11, 16, 9)
    c, rp, col = orc.pipeline_count(s, d, "degree")
    n = len(rp) - 1
    assert orc.count_rows_strided_omp(rp, col, 0, 1, 4) == c
    for start, stride in ((0, 3), (5, 7), (n - 1, 2), (n + 3, 1)):
        assert (orc.count_rows_strided_omp(rp, col, start, stride, 3)
                == orc.count_rows_strided(rp, col, start, stride))


def test_row_list_counter_partitions_the_count():
    # the CPU baseline times systematic samples of rows in out-degree order:
    # the S samples of one stride partition the rows, so their counts add up
    s, d = rmat(11, 16, 9)
    c, rp, col = orc.pipeline_count(s, d, "degree")
    n = len(rp) - 1
    order = np.argsort(np.diff(rp.astype(np.int64)), kind="stable").astype(np.uint32)
    for stride in (1, 3, 8):
        parts = [orc.count_row_list(rp, col, order[o::stride]) for o in range(stride)]
        assert sum(parts) == c
        assert [orc.count_row_list(rp, col, order[o::stride], 4) for o in range(stride)] == parts
    per = orc.count_per_vertex(rp, col)
    rows = np.array([3, 3, n - 1, 0], np.uint32)
    assert orc.count_row_list(rp, col, rows) == int(per[3] * 2 + per[n - 1] + per[0])
    assert orc.count_row_list(rp, col, np.array([n, n + 7], np.uint32)) == 0  # ids >= n skipped
    assert orc.count_row_list(rp, col, np.zeros(0, np.uint32)) == 0
