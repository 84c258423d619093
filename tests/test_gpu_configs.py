"""Full-size parity of the stress, 1M-cell and served configurations.

Stress (16M cells in 32 Gaussian clusters, 512^3, three fields, mitotic
burst with division age 1 for steps [100, 110)): the step from the initial
state and the step from step 105, inside the burst, are bit-identical to the
binary32 oracle taking one step from the same state. The binary64 oracle
gives the per-step drift: fields stay within 1e-5 relative everywhere, and a
position drifts beyond 1e-5 only where the force law jumps, i.e. where a
pair's distance is within rounding of the contact distance S (repulsion
k_rep (S - d) ends at 0 there while the adhesion starts at k_adh, as the
reference law does, engine/assets/shaders/tumor_sim.wgsl:556-569).

1M cells (256^3, two fields): the steps from the initial state and from step
20 meet the same bounds.

Served (log-normal cells, median 100k, capped at 4M; grid side
roundup32(max(64, 2 cbrt(cells)))): jobs arriving over time and admitted into
a running batch end with exactly the state the oracle reaches in their
horizon.
"""
import numpy as np
import pytest

from _drift import TOL, checkpoint, identical
from oracle import oracle as O
from oncolytica_amd import serving
from oncolytica_amd import workloads as W

pytestmark = pytest.mark.gpu


def _check(ck):
    """Bit-identical to binary32; fields within 1e-5 of binary64 everywhere;
    at most 1e-5 of the positions above 1e-5, each next to the contact jump."""
    assert ck["bit_identical_to_oracle_f32"], f"{ck['workload']} step {ck['step']}: {ck['f32_mismatch']}"
    for f, fd in enumerate(ck["field_drift_vs_f64"]):
        assert fd["max_rel"] <= TOL, f"field {f}: {fd}"
    pd = ck["position_drift_vs_f64"]
    assert pd["above_tol"] <= TOL * pd["elements"], pd
    assert ck["positions_above_tol_at_contact_jump"] == pd["above_tol"], ck


@pytest.mark.timeout(1200)
def test_stress_16m_512_steps_bit_exact_and_drift(ctx):
    w = W.stress_c5()
    base = list(w.params)
    hot = [1.0] + base[1:]
    cols = W.initial_cells(w)
    for k in (0, 105):
        _check(checkpoint(ctx, w, cols, k, lambda s: hot if 100 <= s < 110 else base))


def test_large_1m_256_bit_exact_and_drift(ctx):
    """The 1M-cell workload (256^3, O2 + drug, repulsion + adhesion): the step
    from the initial state and the step from step 20 (1.8M cells)."""
    w = W.large_c2()
    cols = W.initial_cells(w)
    for k in (0, 20):
        _check(checkpoint(ctx, w, cols, k))


def test_served_jobs_full_size_match_the_oracle(ctx):
    jobs = serving.poisson_jobs(4.0, 3.0, seed=3)[:8]
    for i, j in enumerate(jobs):
        # arrivals 20 ms apart and horizons of 20..48 steps: later jobs join a running batch
        j.arrival = 0.02 * i
        j.horizon = 20 + 4 * i
    assert max(j.workload.n_cells for j in jobs) > 2_000_000
    rep = serving.Server(ctx, hbm_budget=200 << 30, max_tick_steps=4, keep_results=True,
                         log=lambda m: print(m, flush=True)).run(jobs)
    assert rep.summary()["completed"] == len(jobs)
    assert rep.peak_resident > 1  # jobs shared the batch
    for j in jobs:
        st = O.TumorState(j.workload.desc(), W.initial_cells(j.workload))
        st.run(j.horizon)
        gcols, gfields = j.result
        msg = identical(gcols, gfields, st)
        print(f"served job {j.id}: {j.workload.n_cells:,} cells, side {j.workload.dims[0]}, "
              f"{j.horizon} steps: {'identical' if msg is None else msg}", flush=True)
        assert msg is None, f"job {j.id}: {msg}"
