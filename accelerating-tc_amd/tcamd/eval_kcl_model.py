#!/usr/bin/env python3
"""Compare the learned k-clique plan with the fixed rule on a GPU.

    python accelerating-tc_amd/tcamd/eval_kcl_model.py [--model PATH] [--graphs rgg,rmat20,ba20,er20]

For each graph (degeneracy orientation) and k, prints the median of three
counts with the fixed rule (kcl_count) and with the model's plan
(kcl_count_learned), the plan's one-off build time and how many roots the
plan sends to the merge path.  Every count must agree.
"""
import argparse
import os
import sys

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))

import numpy as np  # noqa: E402

import tcamd  # noqa: E402
from tcamd import kclique_learned as kl  # noqa: E402

GRAPHS = {
    "rgg": (lambda: tcamd.gen_rgg(1000000, tcamd.rgg_radius(1000000, 48), 4),
            "random geometric n=10^6 average degree 48 seed 4", (4, 5, 6, 7)),
    "rmat20": (lambda: tcamd.gen_rmat(20, 32 << 20, 0.57, 0.19, 0.19, 20),
               "R-MAT scale 20 ef32 (0.57,0.19,0.19,0.05) seed 20", (4, 5)),
    "ba20": (lambda: tcamd.gen_ba(1 << 20, 8, 3), "Barabasi-Albert n=2^20 m=8 seed 3", (4, 5)),
    "er20": (lambda: tcamd.gen_er(1 << 20, 1 << 24, 1), "Erdos-Renyi n=2^20 m=2^24 seed 1", (4, 5)),
}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--model", default=kl.DEFAULT_MODEL)
    ap.add_argument("--graphs", default=",".join(GRAPHS))
    args = ap.parse_args()
    with tcamd.Context(0) as ctx:
        model = kl.load(ctx, args.model)
        for name in args.graphs.split(","):
            make, what, ks = GRAPHS[name]
            s, d = make()
            g = ctx.graph(s, d)
            g.simplify()
            g.orient(tcamd.ORIENT_DEGENERACY)
            for k in ks:
                learned = [g.kclique_learned(k, model) for _ in range(3)]  # first builds the plan
                fixed = [g.kclique(k) for _ in range(3)]
                plan = g.kclique_plan(k, model)
                counts = {r.count for r in fixed + learned}
                if len(counts) != 1:
                    raise SystemExit(f"{name} k={k}: counts differ: {sorted(counts)}")
                ms_f = np.median([r.timing["ms_count"] for r in fixed])
                ms_l = np.median([r.timing["ms_count"] for r in learned])
                print(f"[eval] {what} k={k}: fixed rule {ms_f:.2f} ms, learned {ms_l:.2f} ms "
                      f"(plan {learned[0].timing['ms_prep']:.2f} ms once), "
                      f"merge roots {int((plan == 1).sum())}, count {fixed[0].count}", flush=True)
            g.close()


if __name__ == "__main__":
    main()
