#!/usr/bin/env python3
"""Writes tests/golden/configs.json: the oracle's results for the five
benchmark configurations at their full sizes, for the GPU parity tests in
tests/test_gpu_configs.py.

Everything here is the CPU oracle (oracle/oracle.cpp) run on the seeded
inputs of the engine's own generators (host code of libtcamd.so, no GPU):

* er, rmat22, rmat24, ba24: simplify (self-loops dropped, duplicates merged),
  A-order where the configuration has it, A-direction; the full triangle
  count; SHA-256 digests of the oriented CSR (rowptr then col, uint32 LE),
  of the A-direction priorities (int32 LE) and of the A-order permutation.
* batched: the triangle count of every one of the 4,096 graphs (degree
  orientation per graph).
* kclique_rgg, kclique_rmat20: simplify + degeneracy orientation; digests
  of the CSR and of the removal steps; the triangle count and the full
  4-clique count; for k = 5, 6, 7, per-root counts on seeded root samples
  in each bit-matrix width class (d+ <= 64, 128, 256, 512, 1024).  A root
  is sampled only when the expected number of its k-cliques in a random
  graph of its neighbourhood's density, C(d+, k-1) p^C(k-1, 2), is at most
  2*10^6, which bounds the depth-first oracle's time per root.

Run from the repository root (about 15-25 minutes on 8 cores):
    python tests/golden/make_config_fixtures.py [config ...]
"""
import hashlib
import json
import math
import os
import sys
import time
from multiprocessing import Pool

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path[:0] = [os.path.join(ROOT, "accelerating-tc_amd"), os.path.join(ROOT, "oracle")]
import oracle as orc  # noqa: E402
import tcamd  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "configs.json")
THREADS = min(8, os.cpu_count() or 1)
CLASSES = (64, 128, 256, 512, 1024)
SAMPLE_PER_CLASS = 24
MAX_EXPECTED = 2e6


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def sha(*arrays):
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def simplify(s, d):
    lo = np.minimum(s, d).astype(np.uint64)
    hi = np.maximum(s, d).astype(np.uint64)
    keep = lo != hi
    k = np.unique((lo[keep] << np.uint64(32)) | hi[keep])
    return (k >> np.uint64(32)).astype(np.uint32), (k & np.uint64(0xFFFFFFFF)).astype(np.uint32)


def adir_chain(s, d, n, aorder):
    """Oracle form of simplify -> [A-order] -> A-direction -> CSR."""
    t0 = time.time()
    ss, dd = simplify(s, d)
    out = {}
    if aorder:
        perm, os_, od = orc.aorder(ss, dd, n)
        out["perm_sha256"] = sha(perm)
        ss, dd = perm[os_], perm[od]
        n = orc.num_nodes(ss, dd)  # A-direction-seri.cpp:187-193 on the relabelled file
        log(f"  A-order done {time.time() - t0:.1f}s, n -> {n}")
    srp, scol = orc.build_sym(ss, dd, n)
    prio = orc.adir_priority(n, srp, scol)
    os_, od = orc.adir_orient(n, srp, scol, prio)
    del srp, scol
    rp, col = orc.build_csr(os_, od, n)
    log(f"  CSR done {time.time() - t0:.1f}s: n={n} m={len(col)}")
    out.update(n=int(n), m_or=int(len(col)), csr_sha256=sha(rp, col), key_sha256=sha(prio))
    out["count"] = orc.count_omp(rp, col)
    log(f"  count {out['count']} done {time.time() - t0:.1f}s")
    return out


def tri_config(name):
    if name == "er":
        s, d = tcamd.gen_er(65536, 1 << 20, 1)
        return adir_chain(s, d, orc.num_nodes(s, d), False)
    if name in ("rmat22", "rmat24"):
        scale = int(name[4:])
        s, d = tcamd.gen_rmat(scale, 16 << scale, 0.57, 0.19, 0.19, scale)
        return adir_chain(s, d, orc.num_nodes(s, d), True)
    if name == "ba24":
        s, d = tcamd.gen_ba(1 << 24, 8, 3)
        return adir_chain(s, d, orc.num_nodes(s, d), False)
    raise ValueError(name)


def _one_graph(args):
    s, d = args
    if len(s) == 0:
        return 0
    return orc.pipeline_count(s, d, "degree")[0]


def batched_config():
    from tcamd import batch as tb
    graphs = tb.make_batch(4096, 5)
    with Pool(THREADS) as p:
        counts = p.map(_one_graph, [(s, d) for s, d, _ in graphs], chunksize=16)
    return {"graphs": len(graphs), "counts": [int(c) for c in counts],
            "total": int(sum(counts)), "records": int(sum(len(g[0]) for g in graphs))}


def kclique_config(name):
    t0 = time.time()
    if name == "kclique_rgg":
        s, d = tcamd.gen_rgg(1000000, tcamd.rgg_radius(1000000, 48), 4)
    else:
        s, d = tcamd.gen_rmat(20, 32 << 20, 0.57, 0.19, 0.19, 20)
    n = orc.num_nodes(s, d)
    ss, dd = simplify(s, d)
    srp, scol = orc.build_sym(ss, dd, n)
    step, core = orc.degeneracy(n, srp, scol)
    os_, od = orc.kcore_orient(n, srp, scol, step)
    del srp, scol
    rp, col = orc.build_csr(os_, od, n)
    dp = np.diff(rp.astype(np.int64))
    log(f"  {name}: n={n} m={len(col)} max d+ {dp.max()} ({time.time() - t0:.1f}s)")
    tri = orc.count_per_vertex(rp, col)
    out = {"n": int(n), "m_or": int(len(col)), "csr_sha256": sha(rp, col),
           "key_sha256": sha(step.astype(np.int32)), "max_out_degree": int(dp.max()),
           "triangles": int(tri.sum())}
    all_rows = np.arange(n, dtype=np.uint32)
    out["k4"] = int(orc.kclique_row_list(rp, col, 4, all_rows, THREADS).sum())
    log(f"  k=4 full count {out['k4']} ({time.time() - t0:.1f}s)")
    rng = np.random.default_rng(12345)
    samples = {}
    for k in (5, 6, 7):
        per = []
        lo = k - 1
        for hi in CLASSES:
            cand = np.nonzero((dp >= lo) & (dp <= hi))[0]
            rng.shuffle(cand)
            pick = []
            for u in cand:
                dd_ = int(dp[u])
                pairs = dd_ * (dd_ - 1) / 2
                p = float(tri[u]) / pairs if pairs else 0.0
                if p == 0.0:
                    continue
                est = math.comb(dd_, k - 1) * p ** math.comb(k - 1, 2)
                if est <= MAX_EXPECTED:
                    pick.append(int(u))
                if len(pick) == SAMPLE_PER_CLASS:
                    break
            lo = hi + 1
            roots = np.array(sorted(pick), np.uint32)
            cnt = orc.kclique_row_list(rp, col, k, roots, THREADS) if len(roots) else []
            per.append({"max_deg": hi, "roots": roots.tolist(), "counts": [int(c) for c in cnt]})
        samples[str(k)] = per
        log(f"  k={k}: sampled {[len(c['roots']) for c in per]} roots ({time.time() - t0:.1f}s)")
    out["root_samples"] = samples
    return out


def main():
    names = sys.argv[1:] or ["er", "rmat22", "ba24", "kclique_rgg", "kclique_rmat20",
                             "batched", "rmat24"]
    data = {}
    if os.path.exists(OUT):
        with open(OUT) as f:
            data = json.load(f)
    for name in names:
        log(f"[fixtures] {name}")
        t0 = time.time()
        if name == "batched":
            data[name] = batched_config()
        elif name.startswith("kclique"):
            data[name] = kclique_config(name)
        else:
            data[name] = tri_config(name)
        data[name]["oracle_seconds"] = round(time.time() - t0, 1)
        with open(OUT, "w") as f:
            json.dump(data, f, indent=1, sort_keys=True)
        log(f"[fixtures] {name} written ({time.time() - t0:.1f}s)")


if __name__ == "__main__":
    main()
