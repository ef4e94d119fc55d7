#!/usr/bin/env python3
"""Measure this GPU's list-read bandwidth and fit A-order's bandwidth model.

    python accelerating-tc_amd/tcamd/measure_bw_model.py [--out PATH]

The reference's A-order (vertex_ordering/A-order/A-order.cpp:20-36) scores
its buckets with bandwidths measured on the GPU it targets: a table for list
lengths 2..9 and a power law from 10 on.  This script measures the same
curve here (tc_measure_list_bandwidth), fits the model (tc_bw_model_fit) and
writes lengths, measurements and model as JSON (default:
tcamd/aorder_bw_mi355x.json), which Graph.relabel_aorder(tcamd.load_bw_model())
uses.
"""
import argparse
import json
import os
import sys

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))

import tcamd  # noqa: E402

LENS = list(range(2, 10)) + [10, 12, 16, 20, 24, 32, 48, 64, 96, 128, 192, 256, 384, 512, 768,
                             1024]


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=tcamd.BW_MODEL_MI355X)
    args = ap.parse_args()
    with tcamd.Context(0) as ctx:
        gbps = tcamd.measure_list_bandwidth(ctx, LENS)
    for l, b in zip(LENS, gbps):
        print(f"[bw] len {l:5d}: {b:8.1f} GB/s", flush=True)
    model = tcamd.fit_bw_model(LENS, gbps)
    ref = tcamd.bw_model_reference()
    print(f"[bw] fit: {model.a:.1f} * len^{model.b:.4f} (reference {ref.a} * len^{ref.b})",
          flush=True)
    with open(args.out, "w") as f:
        json.dump({"lens": LENS, "gbps": [float(x) for x in gbps], "model": model.as_dict(),
                   "command": "python accelerating-tc_amd/tcamd/measure_bw_model.py"}, f, indent=1)
    print(f"[bw] wrote {args.out}", flush=True)


if __name__ == "__main__":
    main()
