"""Learned plan for k-clique counting (kcl_count_learned).

1. collect(): on training graphs, the kernels time every root on the LDS
   bit-matrix pivot path and on the merge path (kcl_measure_costs, in
   CU-microseconds) and compute 8 features per root (kcl_root_features);
2. train(): the same 8-32-32-2 MLP as the triangle scheduler, trained by the
   engine's own forward / backward / Adam kernels (tc_model_train), predicts
   log2 of both costs;
3. evaluate(): its error next to an analytic estimate (least-squares fit of the
   log2 cost on log2 d+ and on the log2 clique estimate of a random graph of
   the root's density), on held-out roots;
4. Graph.kclique_learned(k, model) counts with the learned plan: a path per
   root, each launch widest first (ordering launches by predicted cost was
   measured slower on R-MAT-20 ef32 at k = 5: 808 against 785 ms).

The count is exact whatever the model predicts: every root is counted once,
on one of two exact paths.
"""
from __future__ import annotations

import json
import os

import numpy as np

import tcamd
from tcamd.learned import make_graph

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_MODEL = os.path.join(HERE, "kcl_cost_model.json")

# (family, parameters, k values): small members of the k-clique benchmark
# families (random geometric, R-MAT ef16/ef32) and two others, oriented by
# the degeneracy order as the benchmarks are.  Sizes keep the merge path's
# timing run of every root to seconds.
TRAIN_GRAPHS = [
    ("rgg", dict(n=1 << 16, deg=32, seed=201), (4, 5, 6, 7)),
    ("rgg", dict(n=1 << 17, deg=48, seed=202), (4, 5, 6, 7)),
    ("rmat", dict(scale=14, ef=16, abc=(0.57, 0.19, 0.19), seed=203), (4, 5, 6)),
    ("rmat", dict(scale=15, ef=32, abc=(0.57, 0.19, 0.19), seed=204), (4, 5)),
    ("rmat", dict(scale=15, ef=16, abc=(0.45, 0.22, 0.22), seed=206), (4, 5)),
    ("er", dict(n=1 << 15, m=1 << 19, seed=207), (4, 5)),
    ("ba", dict(n=1 << 15, k=8, seed=208), (4, 5)),
]
PATHS = ("pivot", "merge")


def prepare(ctx, kind, p):
    s, d = make_graph(kind, p)
    g = ctx.graph(s, d)
    g.simplify()
    g.orient(tcamd.ORIENT_DEGENERACY)
    return g


def samples_of(g, k, per_graph=20000, seed=0):
    """Features X, targets T = log2(CU-us) [pivot, merge], mask M."""
    X = g.kclique_features(k)
    cp, n_p = g.kclique_costs(k, tcamd.KCL_PIVOT)
    cm, n_m = g.kclique_costs(k, tcamd.KCL_MERGE)
    d = np.rint(np.exp2(X[:, 0]) - 1)
    roots = np.nonzero(d >= k - 1)[0]
    # with every root on the pivot path both timed runs counted all roots
    if not (d > 1024).any() and n_p != n_m:
        raise RuntimeError(f"k={k}: pivot and merge counts differ ({n_p} != {n_m})")
    rng = np.random.default_rng(seed)
    # stratified by bit-matrix width class so that the rare wide roots count
    cls = X[roots, 6].astype(int)
    per = max(1, per_graph // max(1, len(np.unique(cls))))
    chosen = []
    for c in np.unique(cls):
        idx = roots[cls == c]
        if len(idx) > per:
            idx = rng.choice(idx, per, replace=False)
        chosen.append(idx)
    sel = np.sort(np.concatenate(chosen)) if chosen else np.zeros(0, np.int64)
    T = np.zeros((len(sel), 2), np.float32)
    M = np.zeros((len(sel), 2), np.float32)
    okp = cp[sel] > 0
    okm = cm[sel] > 0
    T[okp, 0] = np.log2(cp[sel][okp])
    T[okm, 1] = np.log2(cm[sel][okm])
    M[okp, 0] = 1
    M[okm, 1] = 1
    return X[sel], T, M, (n_p, n_m)


def collect(ctx, graphs=TRAIN_GRAPHS, per_graph=20000, log=print):
    Xs, Ts, Ms = [], [], []
    for i, (kind, p, ks) in enumerate(graphs):
        g = prepare(ctx, kind, p)
        for k in ks:
            log(f"[kcl-learned] {kind} {p} k={k}: timing both paths")
            X, T, M, counts = samples_of(g, k, per_graph, seed=100 * i + k)
            log(f"[kcl-learned] {kind} {p} k={k}: {len(X)} roots, counts {counts}")
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
    """Mean absolute error (log2 CU-us) of the model and of the analytic
    estimate (least-squares fit on [1, log2(1+d+), clique estimate]) on the
    same roots."""
    Y = model.forward(X)
    out = {}
    for j, name in enumerate(PATHS):
        m = M[:, j] > 0
        if not m.any():
            continue
        A = np.stack([np.ones(int(m.sum())), X[m, 0], X[m, 5]], 1).astype(np.float64)
        coef, *_ = np.linalg.lstsq(A, T[m, j], rcond=None)
        base = A @ coef
        out[name] = {
            "mae_model": float(np.abs(Y[m, j] - T[m, j]).mean()),
            "mae_analytic": float(np.abs(base - T[m, j]).mean()),
            "roots": int(m.sum()),
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
