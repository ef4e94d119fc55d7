"""bench.py's host-side helpers and argument checks (no GPU)."""
import importlib.util
import os
import subprocess
import sys

import numpy as np
import pytest

import oracle as orc
from graphs import rmat

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def bench():
    spec = importlib.util.spec_from_file_location("bench", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_vs_plain_ratios(bench):
    res = {"hu": {"edges_per_s": 2.0e8, "ms_kernels": 200.0},
           "fox": {"edges_per_s": 6.0e8, "ms_kernels": 70.0},
           "tricore": {"edges_per_s": 8.0e8, "ms_kernels": 50.0},
           "auto": {"edges_per_s": 3.0e9, "ms_kernels": 20.0},
           "learned": {"edges_per_s": 3.3e9, "ms_kernels": 16.0}}
    r = bench.vs_plain(res, "learned", 64e9)
    assert r["fastest_plain"] == "tricore"
    assert r["vs_fastest_plain"] == pytest.approx(3.3e9 / 8.0e8)
    assert r["vs_auto_fixed_rule"] == pytest.approx(1.1)
    assert r["vs_auto_fixed_rule_kernel"] == pytest.approx(20.0 / 16.0)
    assert r["hbm_gbps_measured"] == pytest.approx(4000.0)       # 64 GB in 16 ms
    assert r["hbm_frac_measured"] == pytest.approx(0.5)
    r = bench.vs_plain(res, "auto", None)
    assert "vs_auto_fixed_rule" not in r and "hbm_frac_measured" not in r
    assert bench.vs_plain({"auto": res["auto"]}, "auto", None) is None


def test_cpu_baseline_sample_is_exact_and_proportional(bench):
    s, d = rmat(11, 16, 9)
    c, rp, col = orc.pipeline_count(s, d, "degree")
    cpu, (rows, edges, c1) = bench.cpu_baseline(rp, col, target_s=0.01)
    deg = np.diff(rp.astype(np.int64))
    assert edges == int(deg[rows].sum()) and cpu["value"] > 0 and cpu["cores"] == 1
    assert c1 == orc.count_row_list(rp, col, rows)
    # a systematic sample in degree order: the sampled rows' degrees span
    # the whole range and no row is taken twice
    assert len(set(rows.tolist())) == len(rows)
    assert deg[rows].max() >= np.sort(deg)[-len(deg) // len(rows) - 1]
    par = bench.cpu_baseline_parallel(rp, col, (rows, edges, c1))
    assert par["kind"] == "port-parallel"


def test_bench_refuses_several_gpus():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 2
    assert "one GPU" in r.stderr


def test_cpu_baseline_checks_the_gpu_per_vertex_counts(bench):
    s, d = rmat(10, 16, 4)
    _, rp, col = orc.pipeline_count(s, d, "degree")
    per = orc.count_per_vertex(rp, col)
    cpu, (rows, _, c1) = bench.cpu_baseline(rp, col, per, target_s=0.01)
    assert c1 == int(per[rows].sum())
    wrong = per.copy()
    wrong[rows[len(rows) // 2]] += 1
    with pytest.raises(SystemExit):
        bench.cpu_baseline(rp, col, wrong, target_s=0.01)


def test_kernel_bytes_rule(bench):
    # the bytes the AUTO kernel streams: row offsets, each out-list once, and
    # per edge u->v the two bounds of row v and N+(v)
    s, d = rmat(9, 8, 2)
    _, rp, col = orc.pipeline_count(s, d, "degree")
    n, m = len(rp) - 1, len(col)
    dp = np.diff(rp.astype(np.int64))
    want = 4 * (n + 1) + 4 * m
    for u in range(n):
        for v in col[rp[u]:rp[u + 1]]:
            want += 8 + 4 * int(dp[v])

    class Info:
        pass
    info = Info()
    info.n, info.m_oriented = n, m
    info.sum_in_out = int((dp * np.bincount(col, minlength=n)).sum())
    assert bench.kernel_bytes(info) == want


def test_kclique_roofline_from_profile(bench, tmp_path, monkeypatch):
    import json
    (tmp_path / "profiles").mkdir()
    (tmp_path / "profiles" / "valu_kx.json").write_text(json.dumps({"valu_insts": {"5": 6.144e10}}))
    monkeypatch.setattr(bench, "ROOT", str(tmp_path))
    r = bench.kclique_roofline("kx", {4: {"ms_kernels": 10.0}, 5: {"ms_kernels": 200.0}})
    assert r["bound"] == "valu" and r["k"] == 5
    assert r["achieved"] == pytest.approx(6.144e10 / 0.2)
    assert r["frac"] == pytest.approx(0.5)                 # peak 6.144e11 wave instructions/s
    assert r["peak_source"].startswith("nominal")
    assert bench.kclique_roofline("missing", {5: {"ms_kernels": 1.0}}) is None
    # a measured issue rate replaces the nominal peak
    (tmp_path / "profiles" / "valu_calibration.json").write_text(
        json.dumps({"wave_insts_per_s": 1.2288e12}))
    r = bench.kclique_roofline("kx", {5: {"ms_kernels": 200.0}})
    assert r["peak"] == 1.2288e12 and r["frac"] == pytest.approx(0.25)
    assert r["peak_source"].startswith("measured")
