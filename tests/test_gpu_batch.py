"""GPU tests of batched counting: per-graph counts of one hipGraph replay and
of the one-launch-per-graph baseline against the oracle, under every
schedule the autotuner can pick, with replays interleaved with other work."""
import numpy as np
import pytest

import oracle as orc
import tcamd
from graphs import clique
from tcamd import batch as tb

pytestmark = pytest.mark.gpu


def oracle_counts(graphs):
    out = []
    for s, d, _ in graphs:
        out.append(orc.pipeline_count(s, d, "degree")[0] if len(s) else 0)
    return np.array(out, np.uint64)


@pytest.fixture(scope="module")
def small_batch():
    graphs = tb.make_batch(count=60, seed=5, n_max=4096)
    return graphs, oracle_counts(graphs)


def test_batched_and_pergraph_match_oracle(ctx, small_batch):
    graphs, expect = small_batch
    b = ctx.batch(graphs)
    got, t = b.count()
    assert np.array_equal(got, expect)
    assert t["launches"] == 1
    per, tp = b.count_pergraph()
    assert np.array_equal(per, expect)
    assert tp["launches"] >= 1
    b.close()


def test_replays_interleaved_with_other_work(ctx, small_batch):
    graphs, expect = small_batch
    b = ctx.batch(graphs)
    s, d = clique(7)
    g = ctx.graph(s, d).orient(tcamd.ORIENT_DEGREE)
    for _ in range(4):
        got, _ = b.count()
        assert np.array_equal(got, expect)
        assert g.count(tcamd.KERNEL_AUTO).count == 35
    g.close()
    b.close()


@pytest.mark.parametrize("sched", [
    dict(wave_max_deg=1), dict(wave_max_deg=16, wave_chunk=1), dict(wave_max_deg=512, grid=1),
    dict(wave_max_deg=128, wave_chunk=64, grid=7), dict(wave_max_deg=32, wave_chunk=8),
    dict(wave_max_deg=64, filter_min=1), dict(wave_max_deg=512, filter_min=1, grid=5),
    dict(form=1), dict(form=1, wave_max_deg=1), dict(form=1, wave_max_deg=512, filter_min=1),
    dict(form=1, wave_max_deg=16, wave_chunk=1, grid=3)])
def test_every_schedule_is_exact(ctx, small_batch, sched):
    graphs, expect = small_batch
    b = ctx.batch(graphs)
    assert np.array_equal(b.count(sched)[0], expect)
    assert np.array_equal(b.count_pergraph(sched)[0], expect)
    b.close()


def test_autotune_picks_a_candidate_and_stays_exact(ctx, small_batch):
    graphs, expect = small_batch
    b = ctx.batch(graphs)
    s = b.autotune()
    assert s.wave_max_deg in (16, 32, 64, 128, 256, 512)
    assert s.wave_chunk in (4, 8, 16, 32, 64)
    assert s.filter_min in (1, 16) and s.form in (tcamd.BATCH_WORKLIST, tcamd.BATCH_PUSH)
    assert s.grid > 0
    assert np.array_equal(b.count(s)[0], expect)
    b.close()


def test_edge_cases(ctx):
    e = np.zeros(0, np.uint32)
    hub_n = 5000
    hs = np.zeros(hub_n - 1, np.uint32)
    hd = np.arange(1, hub_n, dtype=np.uint32)
    ring_s = np.arange(1, hub_n, dtype=np.uint32)
    ring_d = np.roll(ring_s, -1)
    hub = (np.concatenate([hs, ring_s]), np.concatenate([hd, ring_d]), hub_n)
    ks, kd = clique(6)
    graphs = [
        (e, e, 1),                                                    # empty
        (np.array([0, 1, 2], np.uint32), np.array([1, 2, 0], np.uint32), 3),  # one triangle
        (e, e, 10),                                                   # isolated vertices
        (np.array([3, 3, 0, 1, 1], np.uint32), np.array([3, 3, 1, 0, 0], np.uint32), 4),
        (ks, kd, 6),
        hub,                                                          # one vertex of degree 4999
        (ks, kd, 100),                                                # clique + isolated tail
    ]
    expect = oracle_counts(graphs)
    assert expect.tolist()[:5] == [0, 1, 0, 0, 20]
    b = ctx.batch(graphs)
    for sched in (None, dict(wave_max_deg=1), dict(wave_max_deg=512, grid=3),
                  dict(wave_max_deg=512, filter_min=1), dict(form=1),
                  dict(form=1, wave_max_deg=1)):
        assert np.array_equal(b.count(sched)[0], expect)
        assert np.array_equal(b.count_pergraph(sched)[0], expect)
    with pytest.raises(tcamd.TCError):
        b.count(dict(form=2))
    b.close()
    only_empty = ctx.batch([(e, e, 5), (e, e, 1)])
    assert only_empty.count()[0].tolist() == [0, 0]
    only_empty.close()


def test_rejects_ids_beyond_graph(ctx):
    s = np.array([0, 1], np.uint32)
    d = np.array([1, 5], np.uint32)
    with pytest.raises(tcamd.TCError):
        ctx.batch([(s, d, 5)])
    # offsets must start at 0 and ascend
    import ctypes as C
    h = C.c_void_p()
    nv = np.array([6], np.uint32)
    for off in ([1, 2], [2, 1]):
        st = tcamd._lib.tc_batch_create(ctx.handle, 1, np.array(off, np.uint64), s.ctypes.data,
                                        d.ctypes.data, nv, C.byref(h))
        assert st == tcamd.ERR_ARG
