#!/usr/bin/env python3
"""Train the k-clique plan's cost model on a GPU and write its weights.

    python accelerating-tc_amd/tcamd/train_kcl_model.py [--out PATH] [--steps N]

Times every root of the graphs in tcamd.kclique_learned.TRAIN_GRAPHS on the
pivot and merge paths, trains the MLP with the engine's own kernels, reports
its held-out error next to the analytic estimate, and saves the model as JSON
(default: tcamd/kcl_cost_model.json).
"""
import argparse
import json
import os
import sys
import time

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))

import numpy as np  # noqa: E402

import tcamd  # noqa: E402
from tcamd import kclique_learned as kl  # noqa: E402


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=kl.DEFAULT_MODEL)
    ap.add_argument("--steps", type=int, default=3000)
    ap.add_argument("--per-graph", type=int, default=20000)
    args = ap.parse_args()
    t0 = time.time()
    with tcamd.Context(0) as ctx:
        X, T, M = kl.collect(ctx, per_graph=args.per_graph, log=lambda *a: print(*a, flush=True))
        print(f"[train] {len(X)} roots collected in {time.time() - t0:.1f}s", flush=True)
        perm = np.random.default_rng(0).permutation(len(X))
        cut = int(0.85 * len(X))
        tr, te = perm[:cut], perm[cut:]
        t1 = time.time()
        model, loss = kl.train(ctx, X[tr], T[tr], M[tr], steps=args.steps)
        print(f"[train] {args.steps} Adam steps in {time.time() - t1:.1f}s, loss {loss:.4f}",
              flush=True)
        j_tr = kl.evaluate(model, X[tr], T[tr], M[tr])
        j_te = kl.evaluate(model, X[te], T[te], M[te])
        print("[train] held-out error (log2 CU-us):", json.dumps(j_te), flush=True)
        kl.save(model, args.out, meta={"train": j_tr, "heldout": j_te, "roots": len(X),
                                       "steps": args.steps, "loss": loss})
        print(f"[train] wrote {args.out}", flush=True)


if __name__ == "__main__":
    main()
