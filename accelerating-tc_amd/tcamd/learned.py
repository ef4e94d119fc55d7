"""Learned scheduling workflow for the AUTO engine.

1. collect(): on training graphs, the engine measures the per-vertex cost of
   its two variants (block / wave) on itself and computes 8 features per
   vertex (tc_graph_measure_costs, tc_graph_features);
2. train(): the MLP is trained on the GPU by the engine's own forward /
   backward / Adam kernels (tc_model_train);
3. evaluate(): the model's prediction error is compared with the analytic
   A-direction work estimate log2(1 + sum_{v in N+(u)} d+(v)) fitted to the
   same data - the fallback the engine uses without a model;
4. Graph.count_learned(model) counts with the learned schedule.

The count is exact under any schedule: the model only picks which variant
counts each vertex and in what order.
"""
from __future__ import annotations

import json
import os

import numpy as np

import tcamd

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_MODEL = os.path.join(HERE, "cost_model.json")

# Training graphs: small and mid-size members of the benchmark families
# (R-MAT with several skews, Barabasi-Albert, Erdos-Renyi, random geometric),
# prepared as the benchmarks prepare theirs (R-MAT: A-order relabelling).
TRAIN_GRAPHS = [
    ("rmat", dict(scale=18, ef=16, abc=(0.57, 0.19, 0.19), seed=101, aorder=True)),
    ("rmat", dict(scale=19, ef=16, abc=(0.57, 0.19, 0.19), seed=102, aorder=True)),
    ("rmat", dict(scale=20, ef=16, abc=(0.57, 0.19, 0.19), seed=108, aorder=True)),
    ("rmat", dict(scale=21, ef=16, abc=(0.57, 0.19, 0.19), seed=109, aorder=True)),
    ("rmat", dict(scale=22, ef=16, abc=(0.57, 0.19, 0.19), seed=122, aorder=True)),
    ("rmat", dict(scale=23, ef=16, abc=(0.57, 0.19, 0.19), seed=123, aorder=True)),
    ("rmat", dict(scale=18, ef=32, abc=(0.45, 0.22, 0.22), seed=103, aorder=True)),
    ("rmat", dict(scale=17, ef=8, abc=(0.65, 0.15, 0.15), seed=104, aorder=True)),
    ("rmat", dict(scale=19, ef=16, abc=(0.57, 0.19, 0.19), seed=110, aorder=False)),
    ("ba", dict(n=1 << 18, k=8, seed=105)),
    ("ba", dict(n=1 << 21, k=8, seed=111)),
    ("er", dict(n=1 << 17, m=1 << 21, seed=106)),
    ("er", dict(n=1 << 20, m=1 << 24, seed=112)),
    ("rgg", dict(n=1 << 17, deg=32, seed=107)),
    ("rgg", dict(n=1 << 20, deg=48, seed=113)),
]


def make_graph(kind, p):
    if kind == "rmat":
        a, b, c = p["abc"]
        return tcamd.gen_rmat(p["scale"], p["ef"] << p["scale"], a, b, c, p["seed"])
    if kind == "ba":
        return tcamd.gen_ba(p["n"], p["k"], p["seed"])
    if kind == "er":
        return tcamd.gen_er(p["n"], p["m"], p["seed"])
    if kind == "rgg":
        return tcamd.gen_rgg(p["n"], tcamd.rgg_radius(p["n"], p["deg"]), p["seed"])
    raise ValueError(kind)


def samples_of(g, per_graph=60000, seed=0, repeats=2):
    """Features X, targets T = log2(cost us) [block, wave], mask M."""
    X = g.features()
    tb = np.zeros(len(X), np.float64)
    tw = np.zeros(len(X), np.float64)
    for _ in range(repeats):  # median-of-runs smoothing would need more runs; mean of 2
        cb, _ = g.measure_costs(tcamd.VARIANT_BLOCK)
        cw, _ = g.measure_costs(tcamd.VARIANT_WAVE)
        tb += cb
        tw += cw
    tb /= repeats
    tw /= repeats
    dplus = np.rint(np.exp2(X[:, 0]) - 1)
    keep = np.nonzero(dplus >= 2)[0]
    rng = np.random.default_rng(seed)
    # stratified by degree bin so that rare large vertices are represented
    bins = X[keep, 5].astype(int)
    chosen = []
    per_bin = max(1, per_graph // max(1, len(np.unique(bins))))
    for b in np.unique(bins):
        idx = keep[bins == b]
        if len(idx) > per_bin:
            idx = rng.choice(idx, per_bin, replace=False)
        chosen.append(idx)
    sel = np.sort(np.concatenate(chosen))
    T = np.zeros((len(sel), 2), np.float32)
    M = np.zeros((len(sel), 2), np.float32)
    okb = tb[sel] > 0
    okw = (tw[sel] > 0) & (dplus[sel] <= 512)
    T[okb, 0] = np.log2(tb[sel][okb])
    T[okw, 1] = np.log2(tw[sel][okw])
    M[okb, 0] = 1
    M[okw, 1] = 1
    return X[sel], T, M


def collect(ctx, graphs=TRAIN_GRAPHS, per_graph=60000, log=print):
    Xs, Ts, Ms = [], [], []
    for i, (kind, p) in enumerate(graphs):
        s, d = make_graph(kind, p)
        g = ctx.graph(s, d)
        g.simplify()
        if p.get("aorder"):
            g.relabel_aorder()
        g.orient(tcamd.ORIENT_ADIR)
        X, T, M = samples_of(g, per_graph, seed=i)
        log(f"[learned] {kind} {p}: {len(X)} samples")
        Xs.append(X)
        Ts.append(T)
        Ms.append(M)
        g.close()
    return np.concatenate(Xs), np.concatenate(Ts), np.concatenate(Ms)


def train(ctx, X, T, M, steps=3000, seed=0):
    m = tcamd.Model(ctx, seed)
    loss = m.train(X, T, M, steps=steps, batch=4096, lr=3e-3, seed=seed)
    return m, loss


def evaluate(model, X, T, M):
    """Mean absolute error (log2 us) of the model and of the analytic
    A-direction work estimate (least-squares fit of log2 cost on
    log2(1 + sum d+(v)), per variant) on the same samples."""
    Y = model.forward(X)
    out = {}
    for j, name in enumerate(["block", "wave"]):
        m = M[:, j] > 0
        if not m.any():
            continue
        a = X[m, 2].astype(np.float64)
        A = np.stack([np.ones_like(a), a], 1)
        coef, *_ = np.linalg.lstsq(A, T[m, j], rcond=None)
        base = A @ coef
        out[name] = {
            "mae_model": float(np.abs(Y[m, j] - T[m, j]).mean()),
            "mae_analytic": float(np.abs(base - T[m, j]).mean()),
            "samples": int(m.sum()),
        }
    return out


def save(model, path=DEFAULT_MODEL, meta=None):
    xm, xs, ym, ys = model.norm
    with open(path, "w") as f:
        json.dump({"params": model.params.tolist(), "x_mean": xm.tolist(), "x_std": xs.tolist(),
                   "y_mean": ym.tolist(), "y_std": ys.tolist(), "meta": meta or {}}, f)


def load(ctx, path=DEFAULT_MODEL):
    with open(path) as f:
        d = json.load(f)
    m = tcamd.Model(ctx)
    m.params = np.array(d["params"], np.float32)
    m.norm = (d["x_mean"], d["x_std"], d["y_mean"], d["y_std"])
    return m
