"""Full-size parity of the five benchmark configurations, through the C ABI.

Every input is built at its stated size from the engine's seeded generators,
preprocessed on the GPU exactly as bench.py does, and compared with the
oracle's results for the same input in tests/golden/configs.json (written by
tests/golden/make_config_fixtures.py, which runs the CPU oracle here):

* the oriented CSR, the A-direction priorities / degeneracy steps and the
  A-order permutation, byte for byte through their SHA-256 digests;
* the triangle count of every method (Hu, Bisson, Fox, TriCore, AUTO and the
  learned schedule), the per-graph counts of the 4,096-graph batch, and the
  4-clique counts of the k-clique graphs, exactly;
* for k = 5, 6, 7, the k-cliques rooted at seeded root samples of every
  bit-matrix width class.

Live oracle checks run beside them on the GPU's own CSR: the per-vertex
triangle counts of a systematic row sample, and k-clique roots.
"""
import hashlib
import json
import os
import time

import numpy as np
import pytest

import oracle as orc
import tcamd
from tcamd import batch as tb

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
with open(os.path.join(HERE, "golden", "configs.json")) as _f:
    FIX = json.load(_f)

KERNELS = ["hu", "bisson", "fox", "tricore", "auto"]
THREADS = 16


def sha(*arrays):
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def say(*a):
    print(*a, flush=True)


def tri_input(name):
    if name == "er":
        return tcamd.gen_er(65536, 1 << 20, 1), False
    if name in ("rmat22", "rmat24"):
        scale = int(name[4:])
        return tcamd.gen_rmat(scale, 16 << scale, 0.57, 0.19, 0.19, scale), True
    if name == "ba24":
        return tcamd.gen_ba(1 << 24, 8, 3), False
    raise ValueError(name)


def row_sample(rp, target_rows):
    """Every S-th row in out-degree order: each degree class in proportion."""
    deg = np.diff(rp.astype(np.int64))
    order = np.argsort(deg, kind="stable").astype(np.uint32)
    stride = max(1, len(order) // target_rows)
    return order[stride // 2::stride]


@pytest.mark.parametrize("name", ["er", "rmat22", "ba24", "rmat24"])
def test_config_triangles_full_size(ctx, name):
    fx = FIX[name]
    t0 = time.time()
    (src, dst), aorder = tri_input(name)
    say(f"[{name}] {len(src)} records generated")
    g = ctx.graph(src, dst)
    del src, dst
    g.simplify()
    if aorder:
        perm = g.relabel_aorder()
        assert sha(perm) == fx["perm_sha256"]
    g.orient(tcamd.ORIENT_ADIR)
    info = g.info()
    assert (info.n, info.m_oriented) == (fx["n"], fx["m_or"])
    rp, col = g.csr()
    assert sha(rp, col) == fx["csr_sha256"], "oriented CSR differs from the oracle's"
    assert sha(g.vertex_key()) == fx["key_sha256"], "A-direction priorities differ"
    say(f"[{name}] CSR and priorities equal the oracle's ({time.time() - t0:.1f}s)")
    from tcamd import learned
    model = learned.load(ctx)
    for k in KERNELS:
        got = g.count(k).count
        say(f"[{name}] {k}: {got}")
        assert got == fx["count"], k
    assert g.count_learned(model).count == fx["count"]
    # live oracle on the GPU's CSR: per-vertex counts of a row sample
    per = g.count_per_vertex()
    assert int(per.sum()) == fx["count"]
    rows = row_sample(rp, 400000 if name != "rmat24" else 200000)
    assert orc.count_row_list(rp, col, rows, THREADS) == int(per[rows].sum())
    say(f"[{name}] done ({time.time() - t0:.1f}s)")
    g.close()


def test_config_batched_full_size(ctx):
    fx = FIX["batched"]
    graphs = tb.make_batch(4096, 5)
    assert len(graphs) == fx["graphs"]
    assert sum(len(s) for s, _, _ in graphs) == fx["records"]
    b = ctx.batch(graphs)
    want = np.array(fx["counts"], np.uint64)
    for sched in (None, dict(wave_max_deg=128, wave_chunk=32, filter_min=1)):
        for _ in range(3):  # hipGraph replays after the first
            c, _ = b.count(sched)
            np.testing.assert_array_equal(c, want)
        say(f"[batched] hipGraph counts equal the oracle's for {sched}")
    c, _ = b.count_pergraph(None)
    np.testing.assert_array_equal(c, want)
    # one graph of each kind recounted live by the oracle
    for i in (0, 1, 2):
        s, d, _ = graphs[i]
        assert orc.pipeline_count(s, d, "degree")[0] == int(want[i])
    b.close()


@pytest.mark.parametrize("name", ["kclique_rgg", "kclique_rmat20"])
def test_config_kclique_full_size(ctx, name):
    fx = FIX[name]
    t0 = time.time()
    if name == "kclique_rgg":
        src, dst = tcamd.gen_rgg(1000000, tcamd.rgg_radius(1000000, 48), 4)
    else:
        src, dst = tcamd.gen_rmat(20, 32 << 20, 0.57, 0.19, 0.19, 20)
    g = ctx.graph(src, dst)
    del src, dst
    g.simplify()
    g.orient(tcamd.ORIENT_DEGENERACY)
    info = g.info()
    assert (info.n, info.m_oriented, info.max_out_degree) == (
        fx["n"], fx["m_or"], fx["max_out_degree"])
    rp, col = g.csr()
    assert sha(rp, col) == fx["csr_sha256"], "degeneracy-oriented CSR differs"
    assert sha(g.vertex_key()) == fx["key_sha256"], "degeneracy steps differ"
    assert g.count(tcamd.KERNEL_AUTO).count == fx["triangles"]
    assert g.kclique(3).count == fx["triangles"]
    k4 = g.kclique(4).count
    say(f"[{name}] k=4: {k4} ({time.time() - t0:.1f}s)")
    assert k4 == fx["k4"]
    for k in ("5", "6", "7"):
        for cls in fx["root_samples"][k]:
            roots = np.array(cls["roots"], np.uint32)
            if len(roots) == 0:
                continue
            got = g.kclique_roots(int(k), roots).count
            assert got == sum(cls["counts"]), (k, cls["max_deg"])
        # a few roots one by one, and the same roots recounted by the oracle live
        cls = max(fx["root_samples"][k], key=lambda c: len(c["roots"]))
        for r, c in list(zip(cls["roots"], cls["counts"]))[:4]:
            assert g.kclique_roots(int(k), np.array([r], np.uint32)).count == c
            assert int(orc.kclique_row_list(rp, col, int(k), np.array([r], np.uint32))[0]) == c
        say(f"[{name}] k={k}: root samples equal the oracle's ({time.time() - t0:.1f}s)")
    g.close()
