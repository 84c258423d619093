"""Full-size parity and binary64 drift of one step, shared by the GPU tests
and the drift report (tests/drift_report.py).

A checkpoint steps a fresh sim on the GPU to step k, downloads its state,
steps it once more and downloads again; the oracle takes one step from the
same step-k state in binary32 (the GPU's step must be bit-identical) and in
binary64 (the per-step drift of fields and positions). Positions are matched
by cell id over the cells that existed at step k: a threshold decision can
differ between precisions, and then the ids of the step's daughters, assigned
by rank, no longer name the same cells.

The tolerance is 1e-5 relative per step (relative to max(|x64|, 1e-3)).
Fields meet it everywhere. A position can exceed it only where the force law
jumps: at the contact distance S the repulsion k_rep (S - d) ends at 0 while
the adhesion starts at k_adh (the reference law jumps the same way,
engine/assets/shaders/tumor_sim.wgsl:556-569), so a pair whose distance lies
within rounding of S takes a different branch in the two precisions.
"""
from __future__ import annotations

import numpy as np

from oracle import oracle as O
from oncolytica_amd.hip._sim import Batch, Sim

TOL = 1e-5


def identical(gcols, gfields, st) -> str | None:
    """None when the GPU columns and fields equal the oracle state's bit for
    bit, else a description of the first difference."""
    ocols = st.columns()
    if len(gcols[0]) != len(ocols[0]):
        return f"live cells {len(gcols[0])} vs oracle {len(ocols[0])}"
    for k, (g, o) in enumerate(zip(gcols, ocols)):
        bad = np.nonzero(g.view(np.uint32) != o.view(np.uint32))[0]
        if bad.size:
            return f"column {k} differs at {bad[:5]}"
    for f, (g, o) in enumerate(zip(gfields, st.fields)):
        bad = np.nonzero(g.view(np.uint32) != o.view(np.uint32))[0]
        if bad.size:
            return f"field {f} differs at {bad[:5]}"
    return None


def rel(a64, b32) -> np.ndarray:
    a = np.asarray(a64, np.float64)
    return np.abs(a - np.asarray(b32, np.float64)) / np.maximum(np.abs(a), 1e-3)


def stats(r: np.ndarray) -> dict:
    if r.size == 0:
        return {"max_rel": 0.0, "rms_rel": 0.0, "p999_rel": 0.0, "above_tol": 0, "elements": 0}
    return {"max_rel": float(r.max()), "rms_rel": float(np.sqrt(np.mean(r * r))),
            "p999_rel": float(np.percentile(r, 99.9)), "above_tol": int(np.sum(r > TOL)), "elements": int(r.size)}


def near_contact(cols, idx, S, R, eps) -> np.ndarray:
    """For each cell index in idx: does it have a partner within R whose
    distance lies within eps of S?"""
    x, y, z = (np.asarray(c, np.float64) for c in cols[:3])
    order = np.argsort(x, kind="stable")
    xs = x[order]
    out = []
    for i in idx:
        lo, hi = np.searchsorted(xs, [x[i] - R, x[i] + R])
        j = order[lo:hi]
        j = j[j != i]
        d = np.sqrt((x[j] - x[i]) ** 2 + (y[j] - y[i]) ** 2 + (z[j] - z[i]) ** 2)
        out.append(bool(np.any(np.abs(d[d < R] - S) < eps)))
    return np.array(out, dtype=bool)


def checkpoint(ctx, w, cols0, k: int, params_at=None, log=print) -> dict:
    """One step from step k of workload w (seeded initial cells cols0), on the
    GPU and in the oracle. params_at(step) gives the rule parameters in force
    at a step (default: w.params throughout); steps are taken as graph
    replays between parameter changes."""
    params_at = params_at or (lambda s: list(w.params))
    sim = Sim(ctx, w.desc())
    b = Batch(ctx)
    try:
        sim.upload_cells(cols0)
        b.add(sim)
        done = 0
        while done < k:
            e = done + 1
            while e < k and params_at(e) == params_at(done):
                e += 1
            sim.set_params(params_at(done))
            b.step(e - done, graph=True)
            done = e
        b.sync()
        params = params_at(k)
        sim.set_params(params)
        ck = sim.download_cells()
        fk = [sim.download_field(f) for f in range(len(w.fields))]
        nid = sim.counts()[1]
        log(f"drift: {w.name} step {k}, {len(ck[0]):,} cells; stepping once")
        b.step(1, graph=False)
        b.sync()
        g1 = sim.download_cells()
        gf1 = [sim.download_field(f) for f in range(len(w.fields))]
    finally:
        b.close()
        sim.close()
    d = w.desc()
    for i, p in enumerate(params):
        d.params[i] = p
    o32 = O.TumorState(d, ck, fields=fk, real="f32", next_id=nid)
    o32.step.value = k
    o32.run(1)
    mismatch = identical(g1, gf1, o32)
    del o32
    o64 = O.TumorState(d, ck, fields=fk, real="f64", next_id=nid)
    o64.step.value = k
    o64.run(1)
    c64 = o64.columns()
    fields = [stats(rel(o64.fields[f], gf1[f])) for f in range(len(w.fields))]
    del o64
    common, i32, i64 = np.intersect1d(g1[3], c64[3], assume_unique=True, return_indices=True)
    keep = common < nid
    i32, i64, common = i32[keep], i64[keep], common[keep]
    r = np.max(np.stack([rel(c64[c][i64], g1[c][i32]) for c in range(3)]), axis=0)
    over = np.nonzero(r > TOL)[0]
    by_id = np.argsort(ck[3])
    idx = by_id[np.searchsorted(ck[3][by_id], common[over])]
    S = 2 * float(np.float32(w.cell_radius))
    at_jump = near_contact(ck, idx, S, w.interaction_radius, TOL)
    out = {"workload": w.name, "step": k, "cells": int(len(ck[0])), "bit_identical_to_oracle_f32": mismatch is None,
           "f32_mismatch": mismatch, "live_after_gpu": int(len(g1[0])), "live_after_f64": int(len(c64[0])),
           "position_drift_vs_f64": stats(r), "positions_above_tol_at_contact_jump": int(at_jump.sum()),
           "field_drift_vs_f64": fields}
    log(f"drift: {w.name} step {k}: bit-identical to binary32 {mismatch is None}; binary64 drift: positions "
        f"{out['position_drift_vs_f64']}, fields {fields}")
    return out
