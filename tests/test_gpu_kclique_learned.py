"""GPU tests of the learned k-clique plan: per-root features against a numpy
restatement, measured costs on every root of each path, and exact counts
under untrained, trained and deliberately extreme models (all roots on one
path)."""
import os
from math import lgamma, log, log2

import numpy as np
import pytest

import oracle as orc
import tcamd
from graphs import CASES, clique
from tcamd import kclique_learned as kl

pytestmark = pytest.mark.gpu


def oracle_degen_csr(s, d):
    n = orc.num_nodes(s, d)
    srp, scol = orc.build_sym(s, d, n)
    step, _ = orc.degeneracy(n, srp, scol)
    return orc.build_csr(*orc.kcore_orient(n, srp, scol, step), n)


def constant_model(ctx, y_pivot, y_merge):
    """All-zero weights: every root is predicted (y_pivot, y_merge)."""
    m = tcamd.Model(ctx, 0)
    m.params = np.zeros(tcamd.Model.n_params, np.float32)
    m.norm = ([0.0] * 8, [1.0] * 8, [y_pivot, y_merge], [1.0, 1.0])
    return m


@pytest.mark.parametrize("case", ["rmat12", "rgg", "hub", "messy"])
def test_kclique_features_match_restatement(ctx, case):
    s, d = CASES[case]()
    rp, col = oracle_degen_csr(s, d)
    n = len(rp) - 1
    g = ctx.graph(s, d).orient(tcamd.ORIENT_DEGENERACY)
    k = 5
    X = g.kclique_features(k)
    tri = g.count_per_vertex() if hasattr(g, "count_per_vertex") else None
    deg = np.diff(rp.astype(np.int64))
    for u in range(n):
        du = int(deg[u])
        if du + 1 < k:
            assert not X[u].any()
            continue
        nb = col[rp[u]:rp[u + 1]].astype(np.int64)
        # edges of the sub-graph of N+(u): arcs v -> w inside N+(u)
        m = sum(int(np.isin(col[rp[v]:rp[v + 1]], nb).sum()) for v in nb)
        if tri is not None:
            assert int(tri[u]) == m
        dv = deg[nb]
        pairs = du * (du - 1) / 2
        dens = m / pairs
        r = k - 1
        est = -30.0
        if dens > 0:
            lc = (lgamma(du + 1) - lgamma(r + 1) - lgamma(du - r + 1)) / log(2)
            est = max(-30.0, lc + 0.5 * r * (r - 1) * log2(dens))
        cls = 0 if du <= 64 else 1 if du <= 128 else 2 if du <= 256 else 3 if du <= 512 else \
            4 if du <= 1024 else 5
        want = [log2(1 + du), log2(1 + m), dens, k, log2(1 + dv.sum()), est, cls,
                log2(1 + dv.max())]
        np.testing.assert_allclose(X[u], want, rtol=1e-5, atol=1e-5, err_msg=f"root {u}")
    g.close()


@pytest.mark.parametrize("case", ["rmat12", "rgg", "hub"])
def test_kclique_costs_cover_every_root(ctx, case):
    s, d = CASES[case]()
    rp, col = oracle_degen_csr(s, d)
    deg = np.diff(rp.astype(np.int64))
    g = ctx.graph(s, d).orient(tcamd.ORIENT_DEGENERACY)
    for k in (4, 5):
        expect = orc.kclique(rp, col, k)
        cp, n_p = g.kclique_costs(k, tcamd.KCL_PIVOT)
        cm, n_m = g.kclique_costs(k, tcamd.KCL_MERGE)
        roots = deg >= k - 1
        assert n_m == expect
        assert np.all(cm[roots] > 0) and not cm[~roots].any()
        piv = roots & (deg <= 1024)
        assert np.all(cp[piv] > 0) and not cp[~piv].any()
        if piv.sum() == roots.sum():
            assert n_p == expect
    g.close()


@pytest.mark.parametrize("case", list(CASES))
def test_kclique_learned_exact_under_any_model(ctx, case):
    s, d = CASES[case]()
    rp, col = oracle_degen_csr(s, d)
    deg = np.diff(rp.astype(np.int64))
    g = ctx.graph(s, d).orient(tcamd.ORIENT_DEGENERACY)
    models = {"untrained": tcamd.Model(ctx, 3),
              "all_merge": constant_model(ctx, 10.0, 0.0),
              "all_pivot": constant_model(ctx, 0.0, 10.0),
              "tie": constant_model(ctx, 5.0, 5.0)}
    if os.path.exists(kl.DEFAULT_MODEL):
        models["trained"] = kl.load(ctx)
    for k in (3, 4, 5, 6):
        expect = orc.kclique(rp, col, k)
        for name, m in models.items():
            got = g.kclique_learned(k, m).count
            assert got == expect, f"{case} k={k} model={name}: {got} != {expect}"
        roots = deg >= k - 1
        plan = g.kclique_plan(k, models["all_merge"])
        assert np.all(plan[roots] == 1) and np.all(plan[~roots] == 255)
        plan = g.kclique_plan(k, models["all_pivot"])
        assert np.all(plan[roots & (deg <= 1024)] == 0)
        assert np.all(plan[roots & (deg > 1024)] == 1)  # beyond the LDS matrix
        # within the 25 % margin the fixed rule (pivot) stays
        plan = g.kclique_plan(k, models["tie"])
        assert np.all(plan[roots & (deg <= 1024)] == 0)
    assert g.kclique_learned(5, None).count == orc.kclique(rp, col, 5)
    g.close()


def test_kclique_learned_hub_beyond_lds(ctx):
    # a vertex of out-degree 1 200 above a clique: only the merge path holds it
    from math import comb
    a, b = clique(30)
    hub = np.full(1200, 5000, np.uint32)
    leaves = np.arange(1200, dtype=np.uint32)
    s = np.concatenate([a, hub])
    d = np.concatenate([b, leaves])
    g = ctx.graph(s, d).orient(tcamd.ORIENT_ID)
    for k in (3, 4, 5):
        expect = comb(30, k) + (comb(30, k - 1) if k - 1 <= 30 else 0)
        for m in (tcamd.Model(ctx, 1), constant_model(ctx, 0.0, 10.0)):
            assert g.kclique_learned(k, m).count == expect
    g.close()


def test_kclique_plan_kept_until_model_or_k_changes(ctx):
    s, d = CASES["rmat12"]()
    rp, col = oracle_degen_csr(s, d)
    g = ctx.graph(s, d).orient(tcamd.ORIENT_DEGENERACY)
    m = constant_model(ctx, 0.0, 10.0)
    expect4, expect5 = orc.kclique(rp, col, 4), orc.kclique(rp, col, 5)
    r1 = g.kclique_learned(4, m)
    r2 = g.kclique_learned(4, m)
    assert r1.count == r2.count == expect4
    assert r1.timing["ms_prep"] > 0.0 and r2.timing["ms_prep"] == 0.0  # kept with the graph
    assert set(np.unique(g.kclique_plan(4, m))) <= {0, 255}
    # new parameters: the plan is rebuilt and follows them (all to merge)
    m.norm = ([0.0] * 8, [1.0] * 8, [10.0, 0.0], [1.0, 1.0])
    r3 = g.kclique_learned(4, m)
    assert r3.count == expect4 and r3.timing["ms_prep"] > 0.0
    plan = g.kclique_plan(4, m)
    assert set(np.unique(plan)) <= {1, 255} and (plan == 1).any()
    r4 = g.kclique_learned(5, m)  # another k: rebuilt
    assert r4.count == expect5 and r4.timing["ms_prep"] > 0.0
    g.close()
