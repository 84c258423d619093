#!/usr/bin/env python3
"""Benchmark of the per-timestep hot path on one MI355X.

Workload (default): large_1m_256 — 1,000,000 cells uniform in a radius-80
sphere (+N(0, 0.3^2) jitter) in a 256^3 grid with two binary32 fields (O2,
and a drug field with decay and a point source on face z=0), repulsion +
adhesion forces, the built-in tumour rule with division, hypoxic death and
drug-induced death. One step = Morton keys, stable radix sort, permute + bins,
forces + uptake, one 7-point sweep per field, the rule, compaction.

Prints ONE JSON line with:
  value      cell-updates/s (sum of live cells at the start of each timed step
             / wall seconds), state resident in HBM, eager launches with HIP
             events between phases (the same timed region gives `roofline`);
  roofline   the dominant kernel's algorithmic bytes per launch / its mean
             event-timed duration, against 8 TB/s;
  cpu_baseline  the CPU oracle (binary32 C restatement, one thread) on a
             bounded sample of the same workload.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "cell-updates/s at 1M cells + 256³ field; ensemble p50/p99 job latency"
HBM_PEAK_GBS = 8000.0


def _log(*a):
    print(*a, file=sys.stderr, flush=True)


def algorithmic_bytes(phase: str, n_mean: float, nvox: int, n_fields: int, cell_bytes: int, occ_bins: float):
    """Algorithmic HBM bytes of ONE launch of a phase (units: cells/voxels of
    the timed steps). Per-unit figures: Morton 20 B/cell; radix pass 16 B/cell;
    permute 2S + key 4 B/cell + 8 B per occupied bin; forces 12 B read + 12 B
    written + 4 B key per cell; stencil 8 B/voxel/field; rule 2S + 4F + 4 B
    flags per cell; compaction 4 B flag + 8 B scan + 2S per cell."""
    S = cell_bytes
    return {
        "morton": 20.0 * n_mean,
        "sort": 16.0 * n_mean,
        "permute_bins": (2 * S + 4) * n_mean + 8.0 * occ_bins,
        "forces": 28.0 * n_mean,
        "stencil": 8.0 * nvox * n_fields,
        "rule": (2 * S + 4 * n_fields + 4) * n_mean,
        "compact": (4 + 8 + 2 * S) * n_mean,
    }.get(phase)


def cpu_baseline(w, steps: int):
    """The oracle (binary32 C, one thread) on `steps` steps of the workload
    from its seeded initial state."""
    from oracle import oracle as O
    from oncolytica_amd import workloads as W
    cols = W.initial_cells(w)
    st = O.TumorState(w.desc(), cols)
    cu = 0
    t0 = time.perf_counter()
    for _ in range(steps):
        cu += st.n.value
        st.run(1)
    dt = time.perf_counter() - t0
    return {"value": cu / dt, "unit": "cell-updates/s", "cores": 1, "kind": "port",
            "sample": f"{steps} steps of {w.name} from its seeded initial state "
                      f"({w.n_cells:,} cells, {w.dims[0]}^3, {len(w.fields)} fields), "
                      f"oracle/onco_oracle_real.c binary32, one thread, {dt:.1f} s"}


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--cells", type=int, default=1_000_000)
    ap.add_argument("--cpu-steps", type=int, default=5, help="oracle steps in the CPU baseline sample")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--calibrate", action="store_true",
                    help="run a plain 256 MiB copy kernel 5 times first (HBM byte-counter calibration)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(sims and ensembles shard by job across GPUs; see DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    from oncolytica_amd.hip import _abi
    from oncolytica_amd.hip._sim import Batch, Sim
    from oncolytica_amd import workloads as W

    if not os.path.exists(_abi.LIB_PATH):
        from oncolytica_amd import _build
        _build.build()

    ctx = _abi.Context(0)
    info = ctx.info()
    _log(f"device {info.name.decode()} {info.arch.decode()} CUs={info.compute_units}")

    if args.calibrate:
        import numpy as np
        nw = 64 << 20
        a, b = ctx.alloc(4 * nw), ctx.alloc(4 * nw)
        for _ in range(5):
            _abi.call("onc_copy_u32", ctx.h, a.p, b.p, nw)
        a.free()
        b.free()

    w = W.large_c2(n_cells=args.cells)
    desc = w.desc()
    cols0 = W.initial_cells(w)

    import ctypes as C

    def fresh():
        sim = Sim(ctx, desc)
        sim.upload_cells(cols0)
        b = Batch(ctx)
        b.add(sim)
        return sim, b

    def cell_updates(sim):
        v = C.c_uint64()
        _abi.call("onc_sim_cell_updates", sim.h, C.byref(v))
        return int(v.value)

    # warmup on a throwaway sim (first launches, hipGraph capture path)
    ws, wb = fresh()
    wb.step(args.warmup, graph=False)
    wb.step(2, graph=True)
    wb.sync()
    wb.close()
    ws.close()

    # timed region: steps 1..K of the workload from its seeded initial state,
    # eager launches with HIP events between phases
    sim, batch = fresh()
    batch.profile(True)
    batch.sync()
    cu0 = cell_updates(sim)
    t0 = time.perf_counter()
    batch.step(args.steps, graph=False)
    batch.sync()
    t1 = time.perf_counter()
    cu1 = cell_updates(sim)
    phases = batch.phase_times()
    batch.profile(False)
    wall = t1 - t0
    value = (cu1 - cu0) / wall
    n_mean = (cu1 - cu0) / max(1, args.steps)
    n_final, _, steps_done = sim.counts()

    # the same steps replayed as one captured hipGraph per step (the first
    # step captures and is not timed)
    gsim, gbatch = fresh()
    gbatch.step(1, graph=True)
    gbatch.sync()
    gcu0 = cell_updates(gsim)
    g0 = time.perf_counter()
    gsteps = max(1, args.steps - 1)
    gbatch.step(gsteps, graph=True)
    gbatch.sync()
    g1 = time.perf_counter()
    gcu1 = cell_updates(gsim)
    gbatch.close()
    gsim.close()

    # occupied bins of the final state (for the permute/bins byte count)
    import numpy as np
    cols = sim.download_cells()
    dims = np.array(list(desc.dims), dtype=np.int64)
    vox = np.minimum(np.floor(np.stack(cols[:3], 1) / desc.dx).astype(np.int64), dims - 1)
    occ_final = len(np.unique(vox[:, 0] + dims[0] * (vox[:, 1] + dims[1] * vox[:, 2])))
    occ = n_mean * occ_final / max(1, n_final)  # occupied bins per step, scaled to the mean population

    table = {}
    for name, (us, nl) in phases.items():
        if nl == 0:
            continue
        byts = algorithmic_bytes(name, n_mean, w.nvox, len(w.fields), W.CELL_BYTES, occ)
        avg_us = us / nl
        gbs = (byts / (avg_us * 1e-6) / 1e9) if byts else None
        table[name] = {"total_ms": us / 1000.0, "launches": nl, "avg_us": avg_us,
                       "alg_bytes": byts, "GBps": gbs}
        _log(f"{name:13s} {us / 1000.0:9.2f} ms {nl:6d} launches {avg_us:9.2f} us/launch"
             + (f" {gbs:8.1f} GB/s ({100 * gbs / HBM_PEAK_GBS:5.1f}% of peak)" if gbs else ""))
    dom = max((k for k in table if table[k]["alg_bytes"]), key=lambda k: table[k]["total_ms"])
    traffic = None
    tf = os.path.join(ROOT, "profiles", "traffic_large_1m_256.json")
    if os.path.exists(tf):
        try:
            traffic = json.load(open(tf)).get(dom)
        except Exception:
            traffic = None
    roof = {"bound": "hbm", "kernel": dom, "achieved": table[dom]["GBps"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": table[dom]["GBps"] / HBM_PEAK_GBS, "traffic": traffic,
            "alg_bytes_per_launch": table[dom]["alg_bytes"], "avg_launch_us": table[dom]["avg_us"]}

    cpu = None if args.no_cpu_baseline else cpu_baseline(w, args.cpu_steps)

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1000.0 * wall / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": w.name, "cells_initial": w.n_cells, "grid": list(w.dims), "fields": len(w.fields),
                   "key_format": f"morton{desc.key_format}", "mean_live_cells": n_mean, "final_live_cells": n_final,
                   "steps_completed": steps_done, "device": info.name.decode(), "cus": info.compute_units},
        "graph_ms_per_step": 1000.0 * (g1 - g0) / gsteps,
        "graph_steps": gsteps,
        "graph_value": None,
        "roofline": roof,
        "phases": table,
        "cpu_baseline": cpu,
    }
    out["graph_value"] = (gcu1 - gcu0) / (g1 - g0) if gsteps else None
    print(json.dumps(out), flush=True)
    batch.close()
    sim.close()
    ctx.close()


if __name__ == "__main__":
    main()
