#!/usr/bin/env python3
"""Train the AUTO engine's cost model on a GPU and write its weights.

    python accelerating-tc_amd/tcamd/train_cost_model.py [--out PATH] [--steps N]

Times every decision cell's count launch under both variants on the training
graphs of tcamd.learned.TRAIN_GRAPHS (at several sampling levels), trains the
MLP with the engine's own kernels on all graphs but every fifth, reports its
error and its choices' measured cost on the held-out graphs next to the
analytic estimate and the fixed rule, and saves the model as JSON (default:
tcamd/cost_model.json).
"""
import argparse
import json
import os
import sys
import time

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))

import tcamd  # noqa: E402
from tcamd import learned  # noqa: E402


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=learned.DEFAULT_MODEL)
    ap.add_argument("--steps", type=int, default=3000)
    ap.add_argument("--levels", type=int, default=learned.LEVELS)
    args = ap.parse_args()
    t0 = time.time()
    with tcamd.Context(0) as ctx:
        X, T, M, G = learned.collect(ctx, levels=args.levels,
                                     log=lambda *a: print(*a, flush=True))
        print(f"[train] {len(X)} cell samples collected in {time.time() - t0:.1f}s", flush=True)
        held = G[:, 0] % 5 == 4
        tr, te = ~held, held
        t1 = time.time()
        model, loss = learned.train(ctx, X[tr], T[tr], M[tr], steps=args.steps)
        print(f"[train] {args.steps} Adam steps in {time.time() - t1:.1f}s, loss {loss:.4f}",
              flush=True)
        j_tr = learned.evaluate(model, X[tr], T[tr], M[tr], G[tr])
        j_te = learned.evaluate(model, X[te], T[te], M[te], G[te])
        print("[train] training graphs:", json.dumps(j_tr), flush=True)
        print("[train] held-out graphs:", json.dumps(j_te), flush=True)
        learned.save(model, args.out, meta={"train": j_tr, "heldout": j_te, "samples": len(X),
                                            "steps": args.steps, "loss": loss,
                                            "levels": args.levels})
        print(f"[train] wrote {args.out}", flush=True)


if __name__ == "__main__":
    main()
