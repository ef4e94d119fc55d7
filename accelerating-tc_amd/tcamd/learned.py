"""Learned scheduling workflow for the AUTO engine.

AUTO counts with a work-list schedule: task x holds N+(x) and streams N+(y)
for y in W(x).  Tasks fall in decision cells by d+(x) (2..8, then
(2^(b-1), 2^b] up to 512); a cell's tasks run narrow (a lane for d+ <= 8,
a wave above) or wide (a wave, a workgroup).  The fixed rule runs d+ <= 64
narrow.

1. collect(): on training graphs and at sampling levels 0..LEVELS-1 (level s
   keeps about 2^-s of each cell's tasks, so cells of every size occur), the
   engine times each cell's count launch alone under both variants
   (tc_graph_measure_cells) and computes the cell's 8 features
   (tc_graph_cell_features), one of them the lists' read time under the
   list-read bandwidth model measured on MI355X (tcamd/aorder_bw_mi355x.json);
2. train(): the MLP learns the log2 microseconds of [wide, narrow] on the
   GPU with the engine's own forward / backward / Adam kernels
   (tc_model_train);
3. evaluate(): the model's prediction error next to the analytic estimate
   (log2 time fitted on the bandwidth-model time), and what its per-cell
   choices cost in measured launch time next to the fixed rule's choices and
   the best ones;
4. Graph.count_learned(model) counts with the learned schedule.

The count is exact under any schedule: the model only picks how each task
runs.
"""
from __future__ import annotations

import json
import os

import numpy as np

import tcamd

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_MODEL = os.path.join(HERE, "cost_model.json")

# Training graphs: members of the benchmark families at several sizes
# (R-MAT with several skews, Barabasi-Albert, Erdos-Renyi, random geometric),
# prepared as the benchmarks prepare theirs (R-MAT: A-order relabelling),
# with seeds other than the benchmarks' own.
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
    ("ba", dict(n=1 << 22, k=8, seed=114)),
    ("er", dict(n=1 << 15, m=1 << 19, seed=116)),
    ("er", dict(n=1 << 16, m=1 << 20, seed=117)),
    ("er", dict(n=1 << 17, m=1 << 21, seed=106)),
    ("er", dict(n=1 << 18, m=1 << 21, seed=118)),
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


LEVELS = 7
SWITCH_MARGIN = 0.10


def feature_model(ctx, bw=None):
    """A Model used only to carry the bandwidth model into the features."""
    m = tcamd.Model(ctx)
    m.set_bw_model(bw if bw is not None else tcamd.load_bw_model())
    return m


def cell_samples(g, fm, levels=LEVELS, reps=3):
    """Per (level, non-empty cell): features X, targets T = log2 us of
    [wide, narrow] launches, mask M (the block cell has no narrow variant),
    and (level, cell)."""
    Xs, Ts, Ms, grp = [], [], [], []
    for lv in range(levels):
        X = g.cell_features(fm, lv)
        tw, cw = g.measure_cells(False, lv, reps)
        tn, cn = g.measure_cells(True, lv, reps)
        if not np.array_equal(cw[:tcamd.BLOCK_CELL], cn[:tcamd.BLOCK_CELL]):
            raise RuntimeError(f"level {lv}: the two variants count {cw} and {cn}")
        for c in range(2, tcamd.CELLS):
            block = c == tcamd.BLOCK_CELL
            if X[c, 0] > 0 and tw[c] > 0 and (block or tn[c] > 0):
                Xs.append(X[c])
                Ts.append([np.log2(tw[c]), 0.0 if block else np.log2(tn[c])])
                Ms.append([1.0, 0.0 if block else 1.0])
                grp.append((lv, c))
    X = np.array(Xs, np.float32).reshape(-1, tcamd.N_FEATURES)
    T = np.array(Ts, np.float32).reshape(-1, 2)
    M = np.array(Ms, np.float32).reshape(-1, 2)
    return X, T, M, grp


def collect(ctx, graphs=TRAIN_GRAPHS, levels=LEVELS, log=print, bw=None):
    """Cell samples of every training graph; G holds (graph, level, cell)."""
    fm = feature_model(ctx, bw)
    Xs, Ts, Ms, G = [], [], [], []
    for i, (kind, p) in enumerate(graphs):
        s, d = make_graph(kind, p)
        g = ctx.graph(s, d)
        del s, d
        g.simplify()
        if p.get("aorder"):
            g.relabel_aorder()
        g.orient(tcamd.ORIENT_ADIR)
        X, T, M, grp = cell_samples(g, fm, levels)
        log(f"[learned] {kind} {p}: {len(X)} cell samples")
        Xs.append(X)
        Ts.append(T)
        Ms.append(M)
        G += [(i, lv, c) for lv, c in grp]
        g.close()
    return np.concatenate(Xs), np.concatenate(Ts), np.concatenate(Ms), np.array(G, np.int64)


def choose(Y, cells):
    """The engine's per-cell choice (tc_learned_cells) from predictions Y:
    True = narrow.  The fixed rule is the prior; a cell switches when the
    other variant is predicted faster by more than SWITCH_MARGIN."""
    prior = (tcamd.FIXED_NARROW >> np.asarray(cells)) & 1 == 1
    wide, narrow = np.exp2(Y[:, 0].astype(np.float64)), np.exp2(Y[:, 1].astype(np.float64))
    keep = np.where(prior, narrow, wide)
    other = np.where(prior, wide, narrow)
    return np.where(other < (1 - SWITCH_MARGIN) * keep, ~prior, prior)


def train(ctx, X, T, M, steps=3000, seed=0):
    m = tcamd.Model(ctx, seed)
    loss = m.train(X, T, M, steps=steps, batch=min(512, len(X)), lr=3e-3, seed=seed)
    return m, loss


def evaluate(model, X, T, M, G=None):
    """Mean absolute error (log2 us) of the model and of the analytic
    estimate (least-squares fit of log2 time on the log2 bandwidth-model
    time of the cell's lists, feature 4, per variant) on the same samples.
    With the groups G, also the measured launch time of the cells under the
    model's choices, the fixed rule's and the best ones (sums over samples)."""
    Y = model.forward(X)
    out = {}
    for j, name in enumerate(["wide", "narrow"]):
        m = M[:, j] > 0
        if not m.any():
            continue
        a = X[m, 4].astype(np.float64)
        A = np.stack([np.ones_like(a), a], 1)
        coef, *_ = np.linalg.lstsq(A, T[m, j], rcond=None)
        base = A @ coef
        out[name] = {
            "mae_model": float(np.abs(Y[m, j] - T[m, j]).mean()),
            "mae_analytic": float(np.abs(base - T[m, j]).mean()),
            "samples": int(m.sum()),
        }
    choice = G[:, 2] < tcamd.BLOCK_CELL if G is not None and len(G) else None
    if choice is not None and choice.any():
        cells = G[choice, 2]
        t = np.exp2(T[choice].astype(np.float64))
        Y = Y[choice]
        fixed = (tcamd.FIXED_NARROW >> cells) & 1 == 1
        pick = choose(Y, cells)
        out["choice_us"] = {
            "model": float(np.where(pick, t[:, 1], t[:, 0]).sum()),
            "fixed_rule": float(np.where(fixed, t[:, 1], t[:, 0]).sum()),
            "best": float(t.min(1).sum()),
            "switched": int((pick != fixed).sum()),
            "agree_with_best": float((pick == (t[:, 1] < t[:, 0])).mean()),
        }
    return out


def save(model, path=DEFAULT_MODEL, meta=None, bw=None):
    xm, xs, ym, ys = model.norm
    bw = bw if bw is not None else tcamd.load_bw_model()
    with open(path, "w") as f:
        json.dump({"params": model.params.tolist(), "x_mean": xm.tolist(), "x_std": xs.tolist(),
                   "y_mean": ym.tolist(), "y_std": ys.tolist(), "bw_model": bw.as_dict(),
                   "meta": meta or {}}, f)


def load(ctx, path=DEFAULT_MODEL):
    with open(path) as f:
        d = json.load(f)
    m = tcamd.Model(ctx)
    m.params = np.array(d["params"], np.float32)
    m.norm = (d["x_mean"], d["x_std"], d["y_mean"], d["y_std"])
    m.set_bw_model(tcamd.BwModel.from_dict(d["bw_model"]) if "bw_model" in d
                   else tcamd.load_bw_model())
    return m
