#!/usr/bin/env python3
"""Train the AUTO engine's cost model on a GPU and write its weights.

    python accelerating-tc_amd/tcamd/train_cost_model.py [--out PATH] [--steps N]

Collects per-vertex costs that the engine measures on the training graphs of
tcamd.learned.TRAIN_GRAPHS, trains the MLP with the engine's own kernels,
reports its error next to the analytic A-direction work estimate, and saves
the model as JSON (default: tcamd/cost_model.json).
"""
import argparse
import json
import os
import sys
import time

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))

import numpy as np  # noqa: E402

import tcamd  # noqa: E402
from tcamd import learned  # noqa: E402


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=learned.DEFAULT_MODEL)
    ap.add_argument("--steps", type=int, default=4000)
    ap.add_argument("--per-graph", type=int, default=60000)
    args = ap.parse_args()
    t0 = time.time()
    with tcamd.Context(0) as ctx:
        X, T, M = learned.collect(ctx, per_graph=args.per_graph,
                                  log=lambda *a: print(*a, flush=True))
        print(f"[train] {len(X)} samples collected in {time.time() - t0:.1f}s", flush=True)
        rng = np.random.default_rng(0)
        perm = rng.permutation(len(X))
        cut = int(0.85 * len(X))
        tr, te = perm[:cut], perm[cut:]
        t1 = time.time()
        model, loss = learned.train(ctx, X[tr], T[tr], M[tr], steps=args.steps)
        print(f"[train] {args.steps} Adam steps in {time.time() - t1:.1f}s, loss {loss:.4f}",
              flush=True)
        j_tr = learned.evaluate(model, X[tr], T[tr], M[tr])
        j_te = learned.evaluate(model, X[te], T[te], M[te])
        print("[train] held-out error (log2 us):", json.dumps(j_te), flush=True)
        learned.save(model, args.out, meta={"train": j_tr, "heldout": j_te, "samples": len(X),
                                            "steps": args.steps, "loss": loss})
        print(f"[train] wrote {args.out}", flush=True)


if __name__ == "__main__":
    main()
