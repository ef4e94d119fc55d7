"""Host sanitizers on the oracle: the same C++ source built with
AddressSanitizer + UBSan (oracle/Makefile target `san`) runs the oracle's
functions on seeded graphs in a subprocess that halts on the first report;
its results must equal the normal build's."""
import json
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle")

SCRIPT = r"""
import json, sys
import numpy as np
sys.path.insert(0, sys.argv[1])
import oracle as orc

def graphs():
    rng = np.random.default_rng(3)
    out = []
    for n, m in ((40, 300), (300, 2500), (900, 9000)):
        s = rng.integers(0, n, m).astype(np.uint32)
        d = (n * rng.random(m) ** 2).astype(np.uint32)   # skewed ends, loops, duplicates
        out.append((s, d))
    iu = np.triu_indices(24, 1)                          # a clique
    out.append((iu[0].astype(np.uint32), iu[1].astype(np.uint32)))
    s = np.concatenate([np.zeros(500, np.uint32), np.arange(1, 501, dtype=np.uint32)])
    d = np.concatenate([np.arange(1, 501, dtype=np.uint32), np.roll(np.arange(1, 501), 1).astype(np.uint32)])
    out.append((s, d))                                   # a wheel: hub + cycle
    return out

res = []
for s, d in graphs():
    n = orc.num_nodes(s, d)
    deg = orc.degree_raw(s, d, n)
    r = {}
    for name, mode in (("id", orc.ORIENT_ID), ("degree", orc.ORIENT_DEGREE),
                       ("tricore", orc.ORIENT_DEG_TRICORE)):
        rp, col = orc.build_csr(*orc.orient_simple(s, d, deg, mode), n)
        r[name] = orc.count(rp, col)
    r["omp"] = orc.count_omp(rp, col)
    r["strided"] = orc.count_rows_strided(rp, col, 1, 3) + orc.count_rows_strided_omp(rp, col, 0, 2, 4)
    rows = np.arange(n + 2, dtype=np.uint32)[::-3]                # includes ids >= n
    r["row_list"] = [orc.count_row_list(rp, col, rows), orc.count_row_list(rp, col, rows, 4)]
    r["per_vertex"] = int(orc.count_per_vertex(rp, col).sum())
    r["hu"] = int(orc.hu_workload(rp, col).sum())
    r["hu_blocks"] = orc.hu_block_starts(orc.hu_workload(rp, col), 7).tolist()
    ps, pd, n2 = orc.prune_deg1(*orc.orient_simple(s, d, deg, orc.ORIENT_DEGREE), n, deg)
    r["pruned"] = orc.count(*orc.build_csr(ps, pd, n2))
    srp, scol = orc.build_sym(s, d, n)
    prio = orc.adir_priority(n, srp, scol)
    rp_a, col_a = orc.build_csr(*orc.adir_orient(n, srp, scol, prio), n)
    r["adir"] = [orc.count(rp_a, col_a), int(prio.sum())]
    core = orc.kcore(n, srp, scol)
    step, _ = orc.degeneracy(n, srp, scol)
    rp_k, col_k = orc.build_csr(*orc.kcore_orient(n, srp, scol, step), n)
    r["kcore"] = [int(core.sum()), int(step.sum()), orc.count(rp_k, col_k)]
    r["kclique"] = [orc.kclique(rp_k, col_k, k) for k in (3, 4, 5)]
    r["kclique_rows"] = orc.kclique_rows_strided(rp_k, col_k, 4, 1, 2)
    a = orc.aorder(s, d, n)
    r["aorder"] = None if a is None else int(a[0].sum())
    r["relabel"] = [int(orc.relabel(s, d, mode)[0].sum()) for mode in ("compact", "first_src", "degree")]
    r["costs"] = [float(x) for x in orc.cost_func(s, d, 2)] + [float(x) for x in orc.verification(s, d)]
    r["approx"] = float(orc.approx_ratio(s, d))
    res.append(r)
print(json.dumps(res))
print("oracle library:", orc.LIB, file=sys.stderr)
"""


def run(env_extra):
    env = dict(os.environ)
    env.update(env_extra)
    return subprocess.run([sys.executable, "-c", SCRIPT, ORACLE], capture_output=True, text=True,
                          env=env, timeout=600)


def test_oracle_under_asan_and_ubsan():
    asan = subprocess.run(["gcc", "-print-file-name=libasan.so"], capture_output=True,
                          text=True).stdout.strip()
    if not os.path.isabs(asan) or not os.path.exists(asan):
        pytest.skip("no AddressSanitizer runtime for gcc in this image")
    b = subprocess.run(["make", "-C", ORACLE, "san"], capture_output=True, text=True)
    assert b.returncode == 0, b.stderr[-2000:]
    plain = run({})
    assert plain.returncode == 0, plain.stderr[-3000:]
    san = run({"ORACLE_LIB": os.path.join(ORACLE, "_san", "liboracle_san.so"),
               "LD_PRELOAD": asan,
               "ASAN_OPTIONS": "detect_leaks=0:halt_on_error=1:abort_on_error=0",
               "UBSAN_OPTIONS": "halt_on_error=1:print_stacktrace=1"})
    assert san.returncode == 0, san.stderr[-4000:]
    assert "runtime error" not in san.stderr and "AddressSanitizer" not in san.stderr
    assert "_san/liboracle_san.so" in san.stderr  # the sanitizer build was the one loaded
    assert json.loads(san.stdout) == json.loads(plain.stdout)
