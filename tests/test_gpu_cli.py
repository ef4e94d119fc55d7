"""GPU tests of the command-line drivers: argv forms and result lines of the
reference programs, counts and written files against the oracle."""
import os
import re
import subprocess

import numpy as np
import pytest

import oracle as orc
import tcamd
from graphs import clique, er, rmat

pytestmark = pytest.mark.gpu

BIN = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                   "accelerating-tc_amd", "bin")


def run(name, *args, cwd=None):
    exe = os.path.join(BIN, name)
    if not os.path.exists(exe):
        pytest.fail(f"{exe} is not built (make -C accelerating-tc_amd cli)")
    r = subprocess.run([exe, *map(str, args)], cwd=cwd, capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    return r.stdout


def triangles(out):
    m = re.search(r"There are (\d+) triangles in the input graph\.", out)
    assert m, out
    return int(m.group(1))


@pytest.fixture(scope="module")
def graph_file(tmp_path_factory):
    s, d = rmat(12, 16, 22)
    p = str(tmp_path_factory.mktemp("cli") / "g.bin")
    tcamd.write_edges(p, s, d)
    return p, s, d, orc.pipeline_count(s, d, "degree")[0]


def test_method_drivers(graph_file):
    p, _, _, t = graph_file
    assert triangles(run("tc_hu", "-f", p)) == t
    assert triangles(run("tc_bisson", "-f", p, 0)) == t
    assert triangles(run("tc_bisson", "-f", p, 1)) == t
    assert triangles(run("tc_fox", "-f", p, 0)) == t
    assert triangles(run("tc_fox", "-f", p, 1)) == t  # rearrangeEdge
    assert triangles(run("tric_gpu", "-f", p)) == t
    for k in ("auto", "hu", "tricore"):
        for o in ("adir", "kcore", "degree", "degeneracy", "id", "tricore"):
            assert triangles(run("tc", "-f", p, "--kernel", k, "--orient", o)) == t
    assert triangles(run("tc", "-f", p, "--aorder", "--simplify")) == t


def test_usage_lines():
    assert "Usage: ./TC -f inputFileName" in run("tc_hu")
    assert "Usage: ./Peel inputFileName outFileName" in run("peel")


def test_a_direction_writes_oriented_file(graph_file, tmp_path):
    p, s, d, t = graph_file
    run("a_direction_seri", p, cwd=tmp_path)
    os_, od = tcamd.read_edges(str(tmp_path / "lpBasedGraph.bin"))
    n = orc.num_nodes(s, d)
    srp, scol = orc.build_sym(s, d, n)
    prio = orc.adir_priority(n, srp, scol)
    es, ed = orc.adir_orient(n, srp, scol, prio)
    np.testing.assert_array_equal(os_, es)
    np.testing.assert_array_equal(od, ed)
    rp, col = orc.build_csr(os_, od, n)
    assert orc.count(rp, col) == t


def test_openmp_a_direction_writes_oriented_file(graph_file, tmp_path):
    """edge_directing/A-direction.cpp's argv form <bin> <thre_param>."""
    p, s, d, t = graph_file
    run("a_direction", p, 3, cwd=tmp_path)
    os_, od = tcamd.read_edges(str(tmp_path / "lpBasedGraph.bin"))
    n = orc.num_nodes(s, d)
    srp, scol = orc.build_sym(s, d, n)
    es, ed = orc.adir_orient(n, srp, scol, orc.adir_omp_priority(n, srp, scol, 3))
    np.testing.assert_array_equal(os_, es)
    np.testing.assert_array_equal(od, ed)
    assert "Usage: ./A-direction inputFileName thre_param" in run("a_direction")


def test_bisson_a_direction_relabels(graph_file, tmp_path):
    """Bisson/A-direction.cpp: both directions of every simple edge, old id
    ascending then neighbour ascending, through the A-direction lookUpTable."""
    p, s, d, t = graph_file
    run("bisson_a_direction", p, cwd=tmp_path)
    rs, rd = tcamd.read_edges(str(tmp_path / "lpBasedGraph.bin"))
    keep = s != d
    pairs = np.unique(np.concatenate([s[keep].astype(np.uint64) << np.uint64(32) | d[keep],
                                      d[keep].astype(np.uint64) << np.uint64(32) | s[keep]]))
    perm = orc.relabel(s, d, "adir")
    np.testing.assert_array_equal(rs, perm[(pairs >> np.uint64(32)).astype(np.int64)])
    np.testing.assert_array_equal(rd, perm[(pairs & np.uint64(0xFFFFFFFF)).astype(np.int64)])
    assert triangles(run("tc", "-f", str(tmp_path / "lpBasedGraph.bin"))) == t


def test_peel_orients_every_record(graph_file, tmp_path):
    p, s, d, _ = graph_file
    out = str(tmp_path / "peeled.bin")
    run("peel", p, out)
    ps, pd = tcamd.read_edges(out)
    n = orc.num_nodes(s, d)
    srp, scol = orc.build_sym(s, d, n)
    prio = orc.adir_priority(n, srp, scol).astype(np.int64)
    assert len(ps) == len(s)
    # same multiset of undirected records, each with the higher-priority end first
    np.testing.assert_array_equal(np.sort(np.minimum(ps, pd).astype(np.int64) << 32 | np.maximum(ps, pd)),
                                  np.sort(np.minimum(s, d).astype(np.int64) << 32 | np.maximum(s, d)))
    hi = (prio[ps] > prio[pd]) | ((prio[ps] == prio[pd]) & (ps >= pd))
    assert hi.all()


def test_a_order_writes_normal_and_reordered(graph_file, tmp_path):
    p, s, d, _ = graph_file
    run("a_order", p, cwd=tmp_path)
    n = orc.num_nodes(s, d)
    perm, os_, od = orc.aorder(s, d, n)
    ns, nd = tcamd.read_edges(str(tmp_path / "normal"))
    rs, rd = tcamd.read_edges(str(tmp_path / "reordered"))
    np.testing.assert_array_equal(ns, os_)
    np.testing.assert_array_equal(nd, od)
    np.testing.assert_array_equal(rs, perm[os_])
    np.testing.assert_array_equal(rd, perm[od])


@pytest.mark.parametrize("k", [3, 4, 5])
def test_kclique_driver_reads_gr(tmp_path, k):
    s, d = er(300, 6000, 5)
    n = orc.num_nodes(s, d)
    gr = str(tmp_path / "g.gr")
    tcamd.write_gr(gr, s, d, n)
    out = run("kcl", "-symmetricGraph", "-simpleGraph", gr, f"-k={k}", "-ft=gr")
    got = int(re.search(r"total_num_cliques = (\d+)", out).group(1))
    srp, scol = orc.build_sym(s, d, n)
    core = orc.kcore(n, srp, scol)
    ks, kd = orc.kcore_orient(n, srp, scol, core)
    rp, col = orc.build_csr(ks, kd, n)
    assert got == orc.kclique(rp, col, k)
    cs, cd = clique(8)
    tcamd.write_gr(gr, cs, cd, 8)
    out = run("kcl", "-symmetricGraph", "-simpleGraph", gr, f"-k={k}")
    assert int(re.search(r"total_num_cliques = (\d+)", out).group(1)) == \
        [56, 70, 56][k - 3]


def test_relabel_drivers(graph_file, tmp_path):
    _, s, d, t = graph_file
    s = (s.astype(np.int64) * 2 + 1).astype(np.uint32)
    d = (d.astype(np.int64) * 2 + 1).astype(np.uint32)
    p = str(tmp_path / "holes.bin")
    tcamd.write_edges(p, s, d)
    for exe, mode in (("serialize", "compact"), ("direction_by_reorder", "degree")):
        out = str(tmp_path / f"{exe}.bin")
        run(exe, p, out)
        rs, rd = tcamd.read_edges(out)
        perm = orc.relabel(s, d, mode)
        np.testing.assert_array_equal(rs, perm[s])
        np.testing.assert_array_equal(rd, perm[d])
        assert triangles(run("tc", "-f", out)) == t
    txt = tmp_path / "g.txt"
    txt.write_text("".join(f"{a} {b}\n" for a, b in zip(s.tolist(), d.tolist())))
    run("adjust2tc", str(txt), str(tmp_path / "adj.txt"))
    got = np.loadtxt(str(tmp_path / "adj.txt"), dtype=np.int64).reshape(-1, 2)
    perm = orc.relabel(s, d, "first_src")
    np.testing.assert_array_equal(got[:, 0], perm[s])
    np.testing.assert_array_equal(got[:, 1], perm[d])


def test_kcore_direction_driver(graph_file, tmp_path):
    p, s, d, _ = graph_file
    out = str(tmp_path / "kcore.bin")
    run("kcore_direction", p, out)
    got = np.fromfile(out, dtype=np.uint64)
    # the file holds records {src, dst}: as uint64 FIRST = dst, SECOND = src
    a = d.astype(np.int64)
    b = s.astype(np.int64)
    n = orc.num_nodes(s, d)
    srp, scol = orc.build_sym(s, d, n)
    core = orc.kcore(n, srp, scol).astype(np.int64)
    swap = (core[a] > core[b]) | ((core[a] == core[b]) & (a > b))
    lo = np.where(swap, b, a).astype(np.uint64)
    hi = np.where(swap, a, b).astype(np.uint64)
    np.testing.assert_array_equal(got, lo << np.uint64(32) | hi)
