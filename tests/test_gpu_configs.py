"""Full-size parity of the stress and served configurations.

Stress (16M cells in 32 Gaussian clusters, 512^3, three fields, mitotic
burst with division age 1 for steps [100, 110)): the step from the initial
state and the step from step 105, inside the burst, are bit-identical to the
binary32 oracle taking one step from the same state. The binary64 oracle
gives the per-step drift: fields stay within 1e-5 relative everywhere, and a
position drifts beyond 1e-5 only where the force law jumps, i.e. where a
pair's distance is within rounding of the contact distance S (repulsion
k_rep (S - d) ends at 0 there while the adhesion starts at k_adh, as the
reference law does, engine/assets/shaders/tumor_sim.wgsl:556-569).

Served (log-normal cells, median 100k, capped at 4M; grid side
roundup32(max(64, 2 cbrt(cells)))): jobs arriving over time and admitted into
a running batch end with exactly the state the oracle reaches in their
horizon.
"""
import numpy as np
import pytest

from oracle import oracle as O
from oncolytica_amd import serving
from oncolytica_amd import workloads as W
from oncolytica_amd.hip._sim import Batch, Sim

pytestmark = pytest.mark.gpu


def _identical(gcols, gfields, st):
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


def _rel(a64, b32):
    a = np.asarray(a64, np.float64)
    return np.abs(a - np.asarray(b32, np.float64)) / np.maximum(np.abs(a), 1e-3)


def _near_contact(cols, idx, S, R, eps):
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


@pytest.mark.timeout(1200)
def test_stress_16m_512_steps_bit_exact_and_drift(ctx):
    w = W.stress_c5()
    base = list(w.params)
    hot = [1.0] + base[1:]
    cols = W.initial_cells(w)
    sim = Sim(ctx, w.desc())
    b = Batch(ctx)
    done = 0
    try:
        sim.upload_cells(cols)
        b.add(sim)
        for k in (0, 105):
            if k > done:
                for edge in (100, k):
                    if edge > done:
                        sim.set_params(hot if done >= 100 else base)
                        b.step(edge - done, graph=True)
                        done = edge
                b.sync()
            params = hot if 100 <= k < 110 else base
            sim.set_params(params)
            ck = sim.download_cells()
            fk = [sim.download_field(f) for f in range(len(w.fields))]
            nid = sim.counts()[1]
            print(f"stress: step {k}, {len(ck[0]):,} cells; stepping once", flush=True)
            b.step(1, graph=False)
            b.sync()
            done += 1
            g1 = sim.download_cells()
            gf1 = [sim.download_field(f) for f in range(len(w.fields))]
            d = w.desc()
            for i, p in enumerate(params):
                d.params[i] = p
            o32 = O.TumorState(d, ck, fields=fk, next_id=nid)
            o32.step.value = k
            o32.run(1)
            msg = _identical(g1, gf1, o32)
            assert msg is None, f"stress step {k} -> {k + 1}: {msg}"
            del o32
            o64 = O.TumorState(d, ck, fields=fk, real="f64", next_id=nid)
            o64.step.value = k
            o64.run(1)
            c64 = o64.columns()
            for f in range(len(w.fields)):
                r = _rel(o64.fields[f], gf1[f])
                print(f"stress: step {k} field {f} drift vs binary64: max {r.max():.2e}", flush=True)
                assert r.max() <= 1e-5
            # cells present at step k in both runs, matched by id
            common, i32, i64 = np.intersect1d(g1[3], c64[3], assume_unique=True, return_indices=True)
            keep = common < nid
            i32, i64 = i32[keep], i64[keep]
            r = np.max(np.stack([_rel(c64[c][i64], g1[c][i32]) for c in range(3)]), axis=0)
            over = np.nonzero(r > 1e-5)[0]
            print(f"stress: step {k} position drift vs binary64: rms {np.sqrt(np.mean(r * r)):.2e}, "
                  f"p99.9 {np.percentile(r, 99.9):.2e}, max {r.max():.2e}, {over.size} cells above 1e-5",
                  flush=True)
            assert over.size <= 1e-5 * r.size
            # every cell above the tolerance has a pair at the contact jump (in the step-k state)
            by_id = np.argsort(ck[3])
            idx = by_id[np.searchsorted(ck[3][by_id], common[over])]
            S = 2 * float(np.float32(w.cell_radius))
            assert np.all(_near_contact(ck, idx, S, w.interaction_radius, 1e-5))
            del o64
    finally:
        b.close()
        sim.close()


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
        msg = _identical(gcols, gfields, st)
        print(f"served job {j.id}: {j.workload.n_cells:,} cells, side {j.workload.dims[0]}, "
              f"{j.horizon} steps: {'identical' if msg is None else msg}", flush=True)
        assert msg is None, f"job {j.id}: {msg}"
