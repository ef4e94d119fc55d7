"""GPU tests of the learned scheduling: MLP kernels against the numpy
oracle, training, cost measurement, and exactness of learned schedules."""
import numpy as np
import pytest

import mlp_ref
import oracle as orc
import tcamd
from graphs import ba, er, hub, rmat, rgg
from tcamd import learned

pytestmark = pytest.mark.gpu


def rand_data(n, seed):
    rng = np.random.default_rng(seed)
    X = rng.normal(0, 1, (n, 8)).astype(np.float32) * 2 + 3
    T = rng.normal(0, 1, (n, 2)).astype(np.float32)
    M = (rng.random((n, 2)) > 0.25).astype(np.float32)
    return X, T, M


def test_forward_matches_oracle(ctx):
    m = tcamd.Model(ctx, seed=3)
    X, _, _ = rand_data(1000, 1)
    norm = (np.full(8, 3, np.float32), np.full(8, 2, np.float32),
            np.array([1.5, -0.5], np.float32), np.array([2.0, 0.5], np.float32))
    m.norm = norm
    got = m.forward(X)
    ref = mlp_ref.predict(m.params, X, norm)
    np.testing.assert_allclose(got, ref, rtol=1e-4, atol=1e-4)


@pytest.mark.parametrize("n", [1, 77, 128, 1000, 5000])
def test_loss_and_gradient_match_oracle(ctx, n):
    m = tcamd.Model(ctx, seed=5)
    X, T, M = rand_data(n, n)
    if M.sum() == 0:
        M[0, 0] = 1
    loss, g = m.loss_grad(X, T, M)
    loss_ref, g_ref = mlp_ref.loss_grad(m.params, X, T, M)
    assert abs(loss - loss_ref) <= 1e-4 * max(1.0, abs(loss_ref))
    # float32 sums in different orders: compare against the gradient scale
    scale = np.abs(g_ref).max()
    np.testing.assert_allclose(g, g_ref, rtol=1e-3, atol=1e-5 * scale)


def test_gradient_is_deterministic(ctx):
    m = tcamd.Model(ctx, seed=7)
    X, T, M = rand_data(3000, 9)
    _, g1 = m.loss_grad(X, T, M)
    _, g2 = m.loss_grad(X, T, M)
    assert np.array_equal(g1, g2)


def test_training_fits_a_known_function(ctx):
    rng = np.random.default_rng(2)
    X = rng.uniform(0, 10, (20000, 8)).astype(np.float32)
    T = np.stack([X[:, 0] + 0.5 * X[:, 2], np.maximum(X[:, 1], X[:, 3])], 1).astype(np.float32)
    M = np.ones_like(T)
    m = tcamd.Model(ctx, seed=1)
    before = m.loss_grad(X, T, M)[0]
    loss = m.train(X, T, M, steps=1500, batch=2048, lr=3e-3)
    assert loss < 0.05 * before and loss < 0.05
    Y = m.forward(X[:500])
    assert np.abs(Y - T[:500]).mean() < 0.35


def test_measure_costs_counts_exactly(ctx):
    s, d = rmat(13, 16, 3)
    expect = orc.pipeline_count(s, d, "degree")[0]
    g = ctx.graph(s, d).orient(tcamd.ORIENT_ADIR)
    X = g.features()
    dplus = np.rint(np.exp2(X[:, 0]) - 1)
    rp, col = g.csr()
    assert np.array_equal(dplus, np.diff(rp.astype(np.int64)))
    cb, nb = g.measure_costs(tcamd.VARIANT_BLOCK)
    cw, nw = g.measure_costs(tcamd.VARIANT_WAVE)
    assert nb == expect and nw == expect
    assert np.all(cb[dplus >= 2] > 0) and np.all(cb[dplus < 2] == 0)
    assert np.all(cw[(dplus >= 2) & (dplus <= 512)] > 0) and np.all(cw[dplus < 2] == 0)
    g.close()


@pytest.mark.parametrize("make", [er, rmat, ba, rgg, hub])
def test_learned_schedules_are_exact(ctx, make):
    s, d = make()
    expect = orc.pipeline_count(s, d, "degree")[0]
    g = ctx.graph(s, d).orient(tcamd.ORIENT_ADIR)
    assert g.count_learned(None).count == expect            # analytic fallback
    for seed in (0, 1):                                     # untrained models: any schedule
        m = tcamd.Model(ctx, seed)
        assert g.count_learned(m).count == expect
        v = g.learned_assignment(m)
        rp, _ = g.csr()
        dp = np.diff(rp.astype(np.int64))
        assert np.all(v[dp < 2] == 255) and np.all(v[dp >= 2] <= 1)
        assert np.all(v[dp > 512] == 0)
    g.close()


def test_learned_pipeline_small(ctx):
    """collect -> train -> evaluate -> count on a held-out graph (small sizes)."""
    graphs = [("rmat", dict(scale=14, ef=16, abc=(0.57, 0.19, 0.19), seed=11)),
              ("ba", dict(n=1 << 14, k=8, seed=12))]
    X, T, M = learned.collect(ctx, graphs, per_graph=5000, log=lambda *a: None)
    model, loss = learned.train(ctx, X, T, M, steps=800)
    assert np.isfinite(loss)
    j = learned.evaluate(model, X, T, M)
    assert j["block"]["mae_model"] <= j["block"]["mae_analytic"] * 1.05
    s, d = rmat(15, 16, 99)
    expect = orc.pipeline_count(s, d, "degree")[0]
    g = ctx.graph(s, d).orient(tcamd.ORIENT_ADIR)
    assert g.count_learned(model).count == expect
    g.close()
