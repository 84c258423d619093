"""Drift report: the GPU's binary32 step against the oracle in binary32
(bit-identical) and binary64 (per-step relative drift of fields and
positions), at full size, one step from each checkpoint.

  python tests/drift_report.py --out profiles/drift_report.json [--stress]

Checkpoints: the 1M-cell workload (256^3, O2 + drug) at steps 0 and 20;
with --stress also the stress workload (16M cells, 512^3, three fields,
division age 1 for steps [100, 110)) at steps 0 and 105, inside the burst.
This is test infrastructure: it runs the oracle as the checker (see
tests/_drift.py for what each figure means).
"""
import argparse
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
sys.path.insert(0, os.path.dirname(HERE))

from _drift import checkpoint  # noqa: E402
from oracle import oracle as O  # noqa: E402
from oncolytica_amd import workloads as W  # noqa: E402
from oncolytica_amd.hip import _abi  # noqa: E402


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--out", default=None, help="write the JSON record here (default: stdout only)")
    ap.add_argument("--stress", action="store_true", help="also the 16M-cell stress workload")
    args = ap.parse_args()
    log = lambda m: print(m, file=sys.stderr, flush=True)  # noqa: E731
    ctx = _abi.Context(0)
    info = ctx.info()
    t0 = time.perf_counter()
    rec = {"device": info.name.decode(), "cus": info.compute_units, "oracle_threads": O.threads(),
           "tolerance_rel_per_step": 1e-5, "checkpoints": []}
    w = W.large_c2()
    cols = W.initial_cells(w)
    for k in (0, 20):
        rec["checkpoints"].append(checkpoint(ctx, w, cols, k, log=log))
    if args.stress:
        w = W.stress_c5()
        base = list(w.params)
        hot = [1.0] + base[1:]
        cols = W.initial_cells(w)
        for k in (0, 105):
            rec["checkpoints"].append(checkpoint(ctx, w, cols, k, lambda s: hot if 100 <= s < 110 else base, log=log))
    rec["wall_s"] = time.perf_counter() - t0
    ctx.close()
    text = json.dumps(rec, indent=1)
    if args.out:
        with open(args.out, "w") as f:
            f.write(text + "\n")
    print(text)


if __name__ == "__main__":
    main()
