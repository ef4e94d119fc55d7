#!/usr/bin/env python3
"""Benchmark of the triangle-counting hot path on one MI355X.

One step = one exact triangle count (tc_count) over the oriented CSR that is
already resident in HBM.  Default workload: R-MAT scale 24, edge factor 16,
(a,b,c,d) = (0.57,0.19,0.19,0.05), seed 24, symmetrised and deduplicated,
A-order relabelling + A-direction orientation; every method (Hu, Bisson,
Fox, TriCore), the degree-binned AUTO engine and the learned schedule are
timed, and their counts must agree.  The default run adds two legs of the
same configuration set on the same line: Barabasi-Albert n = 2^24, m = 8
(A-direction; learned vs the fastest plain method) and the 4,096-graph batch
(one hipGraph replay vs one launch per graph).

    python bench.py [--steps K] [--warmup W] [--legs ba24,batched]
                    [--workload rmat24|rmat22|er|ba24|batched|kclique_rgg|kclique_rmat20]

Prints ONE JSON line.  `value` = TC edges/s of the head engine (the learned
schedule when a trained cost model is present, else AUTO): oriented edges /
count time, wall clock over K steps (the plain methods, 5-20x slower, are
timed over min(K, 5) steps after one warm-up).  `roofline.achieved` = the
head kernel's own streamed bytes per count, 4(n+1) + 12m+ + 4 sum_v d-(v) d+(v)
(row offsets, each out-list once, and for every edge u->v the bounds and the
list of v), over the hipEvent time of the count launch.  The `batched` workload counts 4,096
small graphs (vertex counts log-uniform in [64, 65536], average degree
uniform in [4, 64], 1/3 each Erdos-Renyi / R-MAT / random geometric, seed 5)
with one hipGraph replay per step and reports the one-launch-per-graph
baseline beside it.  The k-clique workloads count k-cliques (k = 4..7 on a
random geometric graph, n = 10^6, average degree 48; k = 4, 5 on R-MAT scale
20, edge factor 32) after checking that 3-cliques equal the triangle count.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "accelerating-tc_amd"))

HBM_PEAK_GBPS = 8000.0
METRIC = "TC edges/s + HBM GB/s % peak; learned/batched vs plain HIP, exact counts"
KERNELS = ["hu", "bisson", "fox", "tricore", "auto", "learned"]
KERNEL_SYMBOL = {
    "hu": "k_hu<2048, 1024>", "bisson": "k_bisson", "fox": "k_fox<L>",
    "tricore": "k_tricore", "auto": "k_auto<2048>", "learned": "k_auto<2048>",
}

WORKLOADS = {
    "rmat22": dict(kind="rmat", scale=22, ef=16, seed=22, aorder=True, orient="adir"),
    "rmat24": dict(kind="rmat", scale=24, ef=16, seed=24, aorder=True, orient="adir"),
    "er": dict(kind="er", n=65536, m=1048576, seed=1, aorder=False, orient="adir"),
    "ba24": dict(kind="ba", n=1 << 24, k=8, seed=3, aorder=False, orient="adir"),
    "batched": dict(kind="batched", count=4096, seed=5),
    "kclique_rgg": dict(kind="kclique", graph="rgg", n=1000000, deg=48, seed=4, ks=(4, 5, 6, 7),
                        orient="degeneracy"),
    "kclique_rmat20": dict(kind="kclique", graph="rmat", scale=20, ef=32, seed=20, ks=(4, 5),
                           orient="degeneracy"),
}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_input(tcamd, w):
    if w["kind"] == "rmat":
        return tcamd.gen_rmat(w["scale"], w["ef"] << w["scale"], 0.57, 0.19, 0.19, w["seed"])
    if w["kind"] == "ba":
        return tcamd.gen_ba(w["n"], w["k"], w["seed"])
    return tcamd.gen_er(w["n"], w["m"], w["seed"])


def describe(w):
    if w["kind"] == "rmat":
        d = f"R-MAT scale {w['scale']} ef{w['ef']} (0.57,0.19,0.19,0.05) seed {w['seed']}"
    elif w["kind"] == "ba":
        d = f"Barabasi-Albert n={w['n']} m={w['k']} seed {w['seed']}"
    else:
        d = f"Erdos-Renyi n={w['n']} m={w['m']} seed {w['seed']}"
    return d + (", A-order + A-direction" if w["aorder"] else ", A-direction")


def kernel_bytes(info):
    """Bytes the push form of the AUTO kernel streams per count: the row
    offsets, every out-list once (staged in LDS per task), and for every edge
    u->v the two bounds of row v and the list N+(v):
    4(n+1) + 4m+ + 8m+ + 4 sum_v d-(v) d+(v)."""
    return 4.0 * (info.n + 1) + 12.0 * info.m_oriented + 4.0 * info.sum_in_out


def cpu_baseline(rp, col, per_vertex=None, target_s=15.0):
    """Serial oracle counter on a systematic sample of the same oriented CSR:
    every S-th row in out-degree order, so each degree class is represented
    in proportion (a stride over ids can take or miss the few heavy rows).
    With the GPU's per-vertex counts, the sample's count must equal their sum
    over the same rows."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    import oracle as orc

    n = len(rp) - 1
    deg = np.diff(rp.astype(np.int64))
    order = np.argsort(deg, kind="stable").astype(np.uint32)
    # calibrate on a sparse sample, then size the sample to ~target_s seconds
    s0 = max(1, n // 2000)
    t0 = time.perf_counter()
    orc.count_row_list(rp, col, order[s0 // 2::s0])
    t1 = time.perf_counter() - t0
    est_full = max(t1, 1e-4) * s0
    stride = max(1, int(round(est_full / target_s)))
    rows = order[stride // 2::stride]
    t0 = time.perf_counter()
    c1 = orc.count_row_list(rp, col, rows)
    t = time.perf_counter() - t0
    edges = int(deg[rows].sum())
    if per_vertex is not None and c1 != int(per_vertex[rows].sum()):
        raise SystemExit(f"oracle sample count {c1} != GPU per-vertex counts "
                         f"{int(per_vertex[rows].sum())} over the same {len(rows)} rows")
    return {
        "value": edges / t,
        "unit": "edges/s",
        "cores": 1,
        "kind": "port",
        "sample": (f"serial forward merge counter (oracle) over every {stride}-th row in "
                   f"out-degree order of the same oriented CSR: {edges} of {int(deg.sum())} "
                   f"edges, {t:.1f} s"),
        "host": host_info(),
    }, (rows, edges, c1)


def cpu_baseline_parallel(rp, col, sample):
    """The same counter and sample with OpenMP over the host's cores (the
    optional second CPU column; labelled parallel)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as orc

    rows, edges, c1 = sample
    threads = max(1, min(16, os.cpu_count() or 1))
    t0 = time.perf_counter()
    c = orc.count_row_list(rp, col, rows, threads)
    t = time.perf_counter() - t0
    if c != c1:
        raise SystemExit(f"parallel oracle sample {c} != serial {c1}")
    return {"value": edges / t, "unit": "edges/s", "cores": threads, "kind": "port-parallel",
            "sample": f"the serial sample above with OpenMP over {threads} threads, {t:.2f} s"}


def vs_plain(results, head, traffic):
    """The headline engine against the four reference methods (Hu, Bisson,
    Fox, TriCore) and against AUTO's fixed rule, by edges/s; with the HBM
    fraction from measured traffic when a traffic profile exists."""
    plain = {k: results[k]["edges_per_s"] for k in ("hu", "bisson", "fox", "tricore")
             if k in results}
    if not plain:
        return None
    best = max(plain, key=plain.get)
    out = {"fastest_plain": best,
           "vs_fastest_plain": results[head]["edges_per_s"] / plain[best]}
    if "auto" in results and head != "auto":
        out["vs_auto_fixed_rule"] = results[head]["edges_per_s"] / results["auto"]["edges_per_s"]
        out["vs_auto_fixed_rule_kernel"] = (results["auto"]["ms_kernels"]
                                            / results[head]["ms_kernels"])
    if traffic:
        gbps = traffic / (results[head]["ms_kernels"] * 1e-3) / 1e9
        out["hbm_gbps_measured"] = gbps
        out["hbm_frac_measured"] = gbps / HBM_PEAK_GBPS
    return out


def host_info():
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "cpu_model": model}


def cpu_baseline_batch(graphs, gpu_counts, target_s=15.0):
    """Oracle pipeline (degree orientation + serial count) over every k-th
    graph; each graph's count must equal the GPU's."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as orc

    def run(step):
        edges, t0 = 0, time.perf_counter()
        for i in range(0, len(graphs), step):
            s, d, _ = graphs[i]
            if len(s):
                c, rp, _ = orc.pipeline_count(s, d, "degree")
                edges += int(rp[-1])
                if c != int(gpu_counts[i]):
                    raise SystemExit(f"graph {i}: oracle {c} != GPU {int(gpu_counts[i])}")
        return edges, time.perf_counter() - t0

    step = max(1, len(graphs) // 32)
    e, t = run(step)
    step2 = max(1, int(round(t * step / target_s)))
    e, t = run(step2)
    return {
        "value": e / t, "unit": "edges/s", "cores": 1, "kind": "port",
        "sample": (f"oracle pipeline (degree orientation, CSR build, serial merge count) over "
                   f"every {step2}-th graph of the batch: {e} oriented edges, {t:.1f} s"),
        "host": host_info(),
    }


def run_batched(args, w, steps, warmup, with_cpu):
    """The batch counted by one hipGraph replay per step, and by one launch
    per graph; both must give the same per-graph counts."""
    import numpy as np
    import tcamd
    from tcamd import batch as tb

    t0 = time.perf_counter()
    graphs = tb.make_batch(w["count"], w["seed"])
    t_gen = time.perf_counter() - t0
    log(f"[bench] generated {len(graphs)} graphs, {sum(len(g[0]) for g in graphs)} records "
        f"in {t_gen:.1f}s")
    ctx = tcamd.Context(0)
    t0 = time.perf_counter()
    b = ctx.batch(graphs)
    t_prep = time.perf_counter() - t0
    info = b.info()
    m = info.m_oriented
    stream = b.stream_bytes()
    log(f"[bench] union n={info.n} m={m} max_out={info.max_out_degree} "
        f"B_alg={info.bytes_alg/1e9:.2f} GB work-list stream {stream/1e9:.2f} GB prep {t_prep:.1f}s")
    t0 = time.perf_counter()
    sched = b.autotune()
    t_tune = time.perf_counter() - t0
    stream = b.stream_bytes(sched.form)  # the bytes of the chosen form
    log(f"[bench] autotune {sched.as_dict()} in {t_tune:.1f}s; the chosen form streams "
        f"{stream/1e9:.2f} GB")

    def timed(fn, nsteps, nwarm):
        ref, _ = fn(sched)
        for _ in range(nwarm):
            fn(sched)
        ev = []
        t0 = time.perf_counter()
        for _ in range(nsteps):
            c, t = fn(sched)
            ev.append(t["ms_count"])
            if not np.array_equal(c, ref):
                raise SystemExit("per-graph counts changed between steps")
        wall = (time.perf_counter() - t0) / nsteps
        return ref, wall, float(np.mean(ev)), t["launches"]

    cb, wall_b, ev_b, _ = timed(b.count, steps, warmup)
    cp, wall_p, ev_p, launches_p = timed(b.count_pergraph, min(steps, 3), min(warmup, 1))
    if not np.array_equal(cb, cp):
        raise SystemExit("batched and per-graph counts disagree")
    log(f"[bench] batched  wall {wall_b*1e3:.3f} ms graph {ev_b:.3f} ms; per-graph wall "
        f"{wall_p*1e3:.3f} ms kernels {ev_p:.3f} ms ({launches_p} launches); "
        f"speedup {wall_p / wall_b:.2f}x")
    cpu = cpu_baseline_batch(graphs, cb) if with_cpu else None
    if cpu:
        log(f"[bench] cpu baseline {cpu['value']:.3e} edges/s ({cpu['sample']})")
    b.close()
    ctx.close()
    return dict(info=info, m=m, graphs=len(graphs), counts=cb, sched=sched.as_dict(), stream=stream,
                wall_b=wall_b, ev_b=ev_b, wall_p=wall_p, ev_p=ev_p, launches_p=launches_p,
                cpu=cpu, t_gen=t_gen, t_prep=t_prep, t_tune=t_tune)


def batched_desc(w):
    return (f"batched: {w['count']} graphs, n log-uniform [64, 65536], average degree "
            f"U[4, 64], 1/3 ER / R-MAT (0.57,0.19,0.19,0.05) / random geometric, seed "
            f"{w['seed']}, degree orientation per graph")


def main_batched(args, w):
    r = run_batched(args, w, args.steps, args.warmup, not args.no_cpu)
    info, m = r["info"], r["m"]
    kb = r["stream"]  # the bytes k_batch streams under the chosen form
    achieved = kb / (r["ev_b"] * 1e-3) / 1e9
    line = {
        "metric": METRIC,
        "value": m / r["wall_b"],
        "unit": "edges/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": r["wall_b"] * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "u32", "data": "synthetic",
        "config": {
            "workload": batched_desc(w),
            "n": info.n, "m": m, "graphs": r["graphs"], "total_count": int(r["counts"].sum()),
            "engine": "batched", "sched": r["sched"], "parallelism": "single GPU",
        },
        "roofline": {
            "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBPS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBPS, "traffic": load_traffic("batched", "batched"),
            "kernel": "k_batch<2048>", "bytes_per_count": kb,
            "bytes_rule": "work-list form: tc_batch_stream_bytes (per task 20 + 4 d+(x), per "
                          "work-list entry 8 + 4 d+(y)); push form: 4(n+1) + 12 m+ + "
                          "4 sum d- d+",
            "bytes_push_form": kernel_bytes(info),
            "bytes_alg_edge_model": info.bytes_alg,
        },
        "cpu_baseline": r["cpu"],
        "batched_vs_pergraph": r["wall_p"] / r["wall_b"],
        "pergraph": {"ms_wall": r["wall_p"] * 1e3, "ms_kernels": r["ev_p"],
                     "launches": r["launches_p"], "edges_per_s": m / r["wall_p"]},
        "batched": {"ms_wall": r["wall_b"] * 1e3, "ms_graph": r["ev_b"],
                    "edges_per_s": m / r["wall_b"]},
        "end_to_end_s": {"generate": r["t_gen"], "preprocess": r["t_prep"],
                         "autotune": r["t_tune"]},
    }
    print(json.dumps(line), flush=True)


def main_kclique(args, w):
    import numpy as np
    import tcamd

    KCL_ORIENTS = {"degeneracy": (tcamd.ORIENT_DEGENERACY, "degeneracy (k-core peel order)"),
                   "adir": (tcamd.ORIENT_ADIR, "A-direction"),
                   "degree": (tcamd.ORIENT_DEGREE, "degree"),
                   "kcore": (tcamd.ORIENT_KCORE, "k-core (core number, id)")}

    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    t0 = time.perf_counter()
    if w["graph"] == "rgg":
        src, dst = tcamd.gen_rgg(w["n"], tcamd.rgg_radius(w["n"], w["deg"]), w["seed"])
        desc = f"random geometric n={w['n']} average degree {w['deg']} seed {w['seed']}"
    else:
        src, dst = tcamd.gen_rmat(w["scale"], w["ef"] << w["scale"], 0.57, 0.19, 0.19, w["seed"])
        desc = (f"R-MAT scale {w['scale']} ef{w['ef']} (0.57,0.19,0.19,0.05) seed {w['seed']}")
    t_gen = time.perf_counter() - t0
    ctx = tcamd.Context(0)
    t0 = time.perf_counter()
    g = ctx.graph(src, dst)
    g.simplify()
    orient = args.orient or w["orient"]
    g.orient(KCL_ORIENTS[orient][0])
    t_prep = time.perf_counter() - t0
    info = g.info()
    m = info.m_oriented
    log(f"[bench] {desc}: {orient} orientation n={info.n} m={m} max_out={info.max_out_degree} "
        f"prep {t_prep:.2f}s")
    tri = g.count(tcamd.KERNEL_AUTO).count
    k3 = g.kclique(3, tcamd.KCL_PIVOT).count
    if k3 != tri:
        raise SystemExit(f"3-cliques {k3} != triangles {tri}")
    ks = tuple(int(x) for x in args.ks.split(",") if x) or w["ks"]
    chunks = max(1, args.kcl_chunks)
    if chunks > 1:
        # roots dealt round-robin in widest-first order: every chunk gets the
        # same mix of wide and narrow roots, and each call prints when done
        rp_h, _ = g.csr()
        dp = np.diff(rp_h.astype(np.int64))
        order = np.argsort(-dp, kind="stable")
        order = order[dp[order] >= 2]
        root_chunks = [np.sort(order[j::chunks]).astype(np.uint32) for j in range(chunks)]

    def kcount(k):
        if chunks == 1:
            return g.kclique(k)
        total, ms = 0, 0.0
        for j, rc in enumerate(root_chunks):
            r = g.kclique_roots(k, rc)
            total += r.count
            ms += r.timing["ms_count"]
            log(f"[bench] k={k}: chunk {j + 1}/{chunks} ({len(rc)} roots) {r.count} cliques, "
                f"{r.timing['ms_count'] / 1e3:.2f} s")
        return tcamd.CountResult(total, {"ms_count": ms})
    kmodel = None
    from tcamd import kclique_learned as kl
    if os.path.exists(kl.DEFAULT_MODEL):
        kmodel = kl.load(ctx)
    per_k = {}
    for k in ks:
        t0 = time.perf_counter()
        r = kcount(k)  # the reference count, also the warm-up
        ref = r.count
        log(f"[bench] k={k}: reference count {ref} in {time.perf_counter() - t0:.2f} s")
        if args.steps == 0:  # a count too long to repeat: the reference run is the timed one
            wall = time.perf_counter() - t0
            ms = [r.timing["ms_count"]]
        else:
            ms = []
            t0 = time.perf_counter()
            for _ in range(args.steps):
                r = kcount(k)
                if r.count != ref:
                    raise SystemExit(f"k={k}: count changed between steps")
                ms.append(r.timing["ms_count"])
            wall = (time.perf_counter() - t0) / args.steps
        per_k[k] = dict(count=r.count, ms_wall=wall * 1e3, ms_kernels=float(np.mean(ms)),
                        edges_per_s=m / wall)
        log(f"[bench] k={k}: {r.count} cliques, {wall*1e3:.2f} ms ({m / wall:.3e} edges/s)")
        if kmodel is not None and args.steps > 0:
            rl = g.kclique_learned(k, kmodel)  # builds and keeps the plan
            prep = rl.timing["ms_prep"]
            lm = []
            for _ in range(max(args.steps, 1)):
                rl = g.kclique_learned(k, kmodel)
                if rl.count != ref:
                    raise SystemExit(f"k={k}: learned plan count {rl.count} != {ref}")
                lm.append(rl.timing["ms_count"])
            merge = int((g.kclique_plan(k, kmodel) == 1).sum())
            per_k[k]["learned"] = dict(ms_kernels=float(np.mean(lm)), ms_plan=prep,
                                       merge_roots=merge)
            log(f"[bench] k={k}: learned plan {np.mean(lm):.2f} ms (plan {prep:.1f} ms once, "
                f"{merge} roots on the merge path)")
    cpu = None
    if not args.no_cpu:
        import oracle as orc
        rp, col = g.csr()
        k = min(ks)
        deg = np.diff(rp.astype(np.int64))
        # roots in a seeded random order, one oracle call each, until about
        # 15 s have passed (a root's cost grows steeply with its out-degree,
        # so a fixed stride can land on the densest roots)
        order = np.random.default_rng(0).permutation(info.n)
        edges = roots = 0
        t0 = time.perf_counter()
        for r in order:
            orc.kclique_rows_strided(rp, col, k, int(r), info.n)
            edges += int(deg[r])
            roots += 1
            if time.perf_counter() - t0 > 15.0:
                break
        t = time.perf_counter() - t0
        cpu = {"value": edges / t, "unit": "edges/s", "cores": 1, "kind": "port",
               "sample": (f"oracle depth-first {k}-clique counter over {roots} roots drawn at "
                          f"random (seed 0) from the same oriented CSR: {edges} of {m} edges, "
                          f"{t:.1f} s"),
               "host": host_info()}
        log(f"[bench] cpu baseline {cpu['value']:.3e} edges/s ({cpu['sample']})")
    total_wall = sum(v["ms_wall"] for v in per_k.values()) * 1e-3
    line = {
        "metric": METRIC,
        "value": m * len(per_k) / total_wall,
        "unit": "edges/s",
        "n_gpus": 1, "steps": args.steps, "warmup": 1,
        "ms_per_step": total_wall * 1e3 / len(per_k),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "u32", "data": "synthetic",
        "config": {"workload": f"k-clique k={','.join(map(str, ks))} on {desc}, "
                               f"{KCL_ORIENTS[orient][1]} orientation", "n": info.n, "m": m,
                   "max_out_degree": info.max_out_degree,
                   "triangles": tri, "engine": "k-clique pivot", "root_chunks": chunks,
                   "parallelism": "single GPU"},
        "roofline": kclique_roofline(args.workload, per_k),
        "cpu_baseline": cpu,
        "kclique": {str(k): v for k, v in per_k.items()},
        "end_to_end_s": {"generate": t_gen, "preprocess": t_prep},
    }
    print(json.dumps(line), flush=True)
    g.close()
    ctx.close()


# Nominal VALU issue peak of MI355X: 1,024 SIMDs at 2.4 GHz, a wave64
# instruction every 4 cycles on a 16-lane SIMD.  The measured peak (the
# issue-rate probe, profiles/valu_calibration.json) replaces it when present.
VALU_NOMINAL_WAVE_INSTS = 1024 * 2.4e9 / 4


def valu_peak():
    """(wave instructions per second, source) of the VALU issue roofline."""
    prof = os.path.join(ROOT, "profiles", "valu_calibration.json")
    if os.path.exists(prof):
        with open(prof) as f:
            return json.load(f)["wave_insts_per_s"], "measured (valu_calibration.json)"
    return VALU_NOMINAL_WAVE_INSTS, "nominal (4 cycles per wave64 instruction)"


def kclique_roofline(workload, per_k):
    """The pivot kernels are bound by VALU issue, not HBM: the wave
    instructions a count issues (SQ_INSTS_VALU summed over the launch's
    kernels, profiles/valu_<workload>.json) over its hipEvent time, against
    the issue peak.  The largest k with a profile is reported."""
    prof = os.path.join(ROOT, "profiles", f"valu_{workload}.json")
    if not os.path.exists(prof):
        return None
    with open(prof) as f:
        insts = json.load(f)["valu_insts"]
    ks = [k for k in per_k if str(k) in insts]
    if not ks:
        return None
    k = max(ks)
    peak, source = valu_peak()
    achieved = insts[str(k)] / (per_k[k]["ms_kernels"] * 1e-3)
    return {"bound": "valu", "achieved": achieved, "peak": peak,
            "unit": "wave-instructions/s", "frac": achieved / peak, "peak_source": source,
            "traffic": None, "kernel": "k_kcl_pivot (all widths)", "k": k,
            "valu_insts_per_count": insts[str(k)]}


def run_triangles(args, w, kernels, with_cpu):
    """One triangle workload: every method timed on the same resident CSR;
    the counts must agree with each other (and the CPU sample with the
    GPU's per-vertex counts)."""
    import numpy as np
    import tcamd

    t_all = time.perf_counter()
    src, dst = make_input(tcamd, w)
    t_gen = time.perf_counter() - t_all
    log(f"[bench] {describe(w)}: generated {len(src)} records in {t_gen:.2f}s")
    ctx = tcamd.Context(0)
    t0 = time.perf_counter()
    g = ctx.graph(src, dst)
    del src, dst
    g.simplify()
    t_simpl = time.perf_counter() - t0
    if w["aorder"]:
        g.relabel_aorder()
    t_aorder = time.perf_counter() - t0 - t_simpl
    g.orient({"adir": tcamd.ORIENT_ADIR, "degree": tcamd.ORIENT_DEGREE}[w["orient"]])
    t_prep = time.perf_counter() - t0
    info = g.info()
    m = info.m_oriented
    stream = g.auto_stream_bytes()  # the work-list schedule's bytes per count
    log(f"[bench] n={info.n} m={m} max_out={info.max_out_degree} B_alg={info.bytes_alg/1e9:.2f} GB "
        f"work-list stream {stream/1e9:.2f} GB (push form {kernel_bytes(info)/1e9:.2f} GB) "
        f"prep {t_prep:.2f}s (simplify {t_simpl:.2f}s, aorder {t_aorder:.2f}s)")

    model = None
    if "learned" in kernels:
        from tcamd import learned
        model = learned.load(ctx) if os.path.exists(learned.DEFAULT_MODEL) else None
        if model is None:
            log("[bench] no trained cost model; 'learned' uses the analytic fallback")

    def run(k):
        return g.count_learned(model) if k == "learned" else g.count(k)

    head = "learned" if "learned" in kernels else ("auto" if "auto" in kernels else kernels[0])
    results = {}
    counts = set()
    for k in kernels:
        plain = k in ("hu", "bisson", "fox", "tricore")
        nsteps = min(args.steps, 5) if plain and k != head else args.steps
        nwarm = min(args.warmup, 1) if plain and k != head else args.warmup
        r = run(k)  # builds and caches the schedule
        counts.add(r.count)
        for _ in range(nwarm):
            run(k)
        ev_ms, dom_ms = [], []
        t0 = time.perf_counter()
        for _ in range(nsteps):
            r = run(k)
            ev_ms.append(r.timing["ms_count"])
            dom_ms.append(r.timing["ms_dominant"])
            counts.add(r.count)
        wall = (time.perf_counter() - t0) / nsteps
        ev = float(np.mean(ev_ms))
        results[k] = dict(count=r.count, steps=nsteps, warmup=nwarm, ms_wall=wall * 1e3,
                          ms_kernels=ev, ms_dominant=float(np.mean(dom_ms)),
                          launches=r.timing["launches"], ms_prep=r.timing["ms_prep"],
                          edges_per_s=m / wall, edges_per_s_kernel=m / (ev * 1e-3),
                          gbps_alg=info.bytes_alg / (ev * 1e-3) / 1e9)
        log(f"[bench] {k:8s} count={r.count} wall {wall*1e3:.3f} ms kernels {ev:.3f} ms "
            f"({results[k]['gbps_alg']:.0f} GB/s alg, {nsteps} steps)")
    if len(counts) != 1:
        raise SystemExit(f"kernels disagree on the count: {sorted(counts)}")

    cpu = cpu_par = None
    if with_cpu:
        rp, col = g.csr()
        per = g.count_per_vertex()
        if int(per.sum()) != results[head]["count"]:
            raise SystemExit("per-vertex counts do not add up to the count")
        cpu, sample = cpu_baseline(rp, col, per)
        log(f"[bench] cpu baseline {cpu['value']:.3e} edges/s ({cpu['sample']}); the sample's "
            f"count equals the GPU's per-vertex counts over the same rows")
        cpu_par = cpu_baseline_parallel(rp, col, sample)
        log(f"[bench] cpu baseline, parallel: {cpu_par['value']:.3e} edges/s ({cpu_par['sample']})")
        del rp, col, per
    g.close()
    ctx.close()
    return dict(info=info, m=m, head=head, results=results, cpu=cpu, cpu_par=cpu_par,
                t_gen=t_gen, t_prep=t_prep, stream=stream)


def triangle_roofline(info, hr, traffic, stream):
    """The head kernel (k_auto_wl, AUTO and the learned schedule) streams the
    work-list bytes of tc_auto_stream_bytes whatever variant runs each task."""
    achieved = stream / (hr["ms_kernels"] * 1e-3) / 1e9
    return {
        "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBPS, "unit": "GB/s",
        "frac": achieved / HBM_PEAK_GBPS, "traffic": traffic, "kernel": "k_auto_wl<2048>",
        "bytes_per_count": stream,
        "bytes_rule": "tc_auto_stream_bytes: per task x, 20 + 4 d+(x) (task id, row bounds in "
                      "both CSRs, N+(x)); per work-list entry y, 8 + 4 d+(y) (its row bounds, "
                      "stored with the entry, and N+(y))",
        "bytes_push_form": kernel_bytes(info),
        "bytes_alg_edge_model": info.bytes_alg,
    }


def leg_summary(name, w, r, traffic=None):
    """The compact record of an extra leg on the headline line."""
    if w["kind"] == "batched":
        info, m = r["info"], r["m"]
        return {"workload": batched_desc(w), "graphs": r["graphs"],
                "total_count": int(r["counts"].sum()), "m": m,
                "batched_edges_per_s": m / r["wall_b"], "pergraph_edges_per_s": m / r["wall_p"],
                "batched_vs_pergraph": r["wall_p"] / r["wall_b"],
                "ms_batched": r["wall_b"] * 1e3, "ms_pergraph": r["wall_p"] * 1e3,
                "sched": r["sched"],
                "gbps_stream": r["stream"] / (r["ev_b"] * 1e-3) / 1e9}
    hr = r["results"][r["head"]]
    return {"workload": describe(w), "n": r["info"].n, "m": r["m"], "count": hr["count"],
            "engine": r["head"], "edges_per_s": hr["edges_per_s"], "ms_per_step": hr["ms_wall"],
            "roofline": triangle_roofline(r["info"], hr, traffic, r["stream"]),
            "learned_vs_plain": vs_plain(r["results"], r["head"], traffic),
            "kernels": {k: {kk: (round(v, 4) if isinstance(v, float) else v)
                            for kk, v in x.items()} for k, x in r["results"].items()}}


def load_traffic(workload, head):
    prof = os.path.join(ROOT, "profiles", f"traffic_{workload}.json")
    if os.path.exists(prof):
        with open(prof) as f:
            return json.load(f).get(head)
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="rmat24", choices=sorted(WORKLOADS))
    ap.add_argument("--kernels", default=",".join(KERNELS))
    ap.add_argument("--legs", default=None,
                    help="extra legs on the same line (default for rmat24: ba24,er,batched; "
                         "'' for none)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--ks", default="", help="k-clique workloads: comma-separated k values")
    ap.add_argument("--kcl-chunks", type=int, default=1,
                    help="k-clique workloads: count each k over this many root chunks "
                         "(kcl_count_roots), printing after each")
    ap.add_argument("--orient", default="", choices=["", "degeneracy", "adir", "degree", "kcore"],
                    help="k-clique workloads: orientation (default: degeneracy)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a 1 GiB device-to-device copy (HBM counter calibration)")
    ap.add_argument("--valu-calibration", action="store_true",
                    help="only run the vector-ALU issue-rate probe (profile it for SQ_INSTS_VALU)")
    args = ap.parse_args()
    if args.valu_calibration:
        import tcamd
        ctx = tcamd.Context(0)
        iters = 100000
        ms = [tcamd.measure_valu_rate(ctx, iters) for _ in range(3)]
        import torch
        waves = torch.cuda.get_device_properties(0).multi_processor_count * 8 * 4
        print(json.dumps({"probe_iters": iters, "waves": waves, "ms": ms,
                          "valu_insts_per_round": 48}))
        return
    if args.copy_calibration:
        import torch
        x = torch.empty(1 << 30, dtype=torch.uint8, device="cuda")
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": 1 << 30, "copy_bytes_written": 1 << 30, "launches": 3}))
        return
    if args.gpus != 1:
        print("bench.py: this engine measures one GPU; multi-GPU runs are not part of this "
              "build (see the sharding design for how the path would split).", file=sys.stderr)
        sys.exit(2)

    w = WORKLOADS[args.workload]
    if w["kind"] == "batched":
        return main_batched(args, w)
    if w["kind"] == "kclique":
        return main_kclique(args, w)

    kernels = [k for k in args.kernels.split(",") if k]
    legs = args.legs if args.legs is not None else ("ba24,er,batched"
                                                    if args.workload == "rmat24" else "")
    legs = [x for x in legs.split(",") if x and x != args.workload]
    r = run_triangles(args, w, kernels, not args.no_cpu)
    info, m, head = r["info"], r["m"], r["head"]
    hr = r["results"][head]
    traffic = load_traffic(args.workload, head)
    extra = {}
    for name in legs:
        lw = WORKLOADS[name]
        if lw["kind"] == "batched":
            lr = run_batched(args, lw, min(args.steps, 10), min(args.warmup, 2), False)
        elif lw["kind"] in ("rmat", "er", "ba"):
            lk = [k for k in ("tricore", "auto", "learned") if k in kernels]
            lr = run_triangles(args, lw, lk, False)
        else:
            raise SystemExit(f"leg {name}: only triangle and batched workloads")
        extra[name] = leg_summary(name, lw, lr, load_traffic(name, lr.get("head", "")))
    line = {
        "metric": METRIC,
        "value": hr["edges_per_s"],
        "unit": "edges/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": hr["ms_wall"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic",
        "config": {
            "workload": describe(w),
            "n": info.n, "m": m, "count": hr["count"], "engine": head,
            "parallelism": "single GPU",
        },
        "roofline": triangle_roofline(info, hr, traffic, r["stream"]),
        "cpu_baseline": r["cpu"],
        "cpu_baseline_parallel": r["cpu_par"],
        "learned_vs_plain": vs_plain(r["results"], head, traffic),
        "kernels": {k: {kk: (round(v, 4) if isinstance(v, float) else v) for kk, v in x.items()}
                    for k, x in r["results"].items()},
        "legs": extra,
        "end_to_end_s": {"generate": r["t_gen"], "preprocess": r["t_prep"]},
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
