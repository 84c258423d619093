"""Batched ensemble, autotuned tile shapes and the continuous-batching server
on the GPU: results are bit-identical whatever the tile shapes, and a served
job's result equals the same job stepped alone."""
import numpy as np
import pytest

from oncolytica_amd import workloads as W
from oncolytica_amd.hip import _abi
from oncolytica_amd.hip._autotune import Tuning, tune
from oncolytica_amd.hip._sim import Batch, Sim
from oncolytica_amd.serving import Job, Server, footprint, run_alone

pytestmark = pytest.mark.gpu


def _same(a_cols, a_fields, b_cols, b_fields):
    assert len(a_cols[0]) == len(b_cols[0])
    for a, b in zip(a_cols, b_cols):
        np.testing.assert_array_equal(a.view(np.uint32), b.view(np.uint32))
    for a, b in zip(a_fields, b_fields):
        np.testing.assert_array_equal(a.view(np.uint32), b.view(np.uint32))


def _small_ensemble():
    ws = W.ensemble_c3(n_sims=4, scale=0.02)
    odd = W.plain_c1()
    odd.dims = (50, 42, 37)   # rows not a multiple of 4: the scalar stencil path
    odd.name = "plain_odd"
    odd.n_cells = 1500
    return ws + [odd]


def _run_batch(ctx, ws, tuning, steps, fused=1, forces_waves=0, fused_rule=True):
    sims = [Sim(ctx, w.desc()) for w in ws]
    b = Batch(ctx)
    if tuning:
        tuning.apply(b)
    b.set_fused(fused)
    b.set_fused_rule(fused_rule)
    if forces_waves:
        b.set_forces_waves(forces_waves)
    for s, w in zip(sims, ws):
        s.upload_cells(W.initial_cells(w))
        b.add(s)
    b.step(steps // 2, graph=True)
    b.step(steps - steps // 2, graph=False)
    b.sync()
    out = [(s.download_cells(), [s.download_field(f) for f in range(s.desc.n_fields)]) for s in sims]
    b.close()
    for s in sims:
        s.close()
    return out


@pytest.mark.parametrize("tuning", [Tuning(4, 16, 16, 16, 8), Tuning(12, 8, 128, 2, 64), Tuning(16, 4, 256, 1, 1),
                                    Tuning(8, 12, 64, 4, 32)])
def test_tile_shapes_do_not_change_results(ctx, tuning):
    ws = _small_ensemble()
    ref = _run_batch(ctx, ws, None, 10)
    got = _run_batch(ctx, ws, tuning, 10)
    for (rc, rf), (gc, gf) in zip(ref, got):
        _same(rc, rf, gc, gf)


def test_autotuner_returns_a_valid_shape_and_caches(ctx, tmp_path):
    ws = W.ensemble_c3(n_sims=3, scale=0.02)
    members = [(w.desc(), W.initial_cells(w)) for w in ws]
    cache = str(tmp_path / "tune.json")
    t = tune(ctx, members, steps=1, cache_file=cache)
    assert t.items in (4, 8, 12, 16) and t.citems in (4, 8, 12, 16) and t.sbx * t.sby == 256 and t.sbz >= 1
    assert tune(ctx, members, steps=1, cache_file=cache) == t


@pytest.mark.parametrize("tuning", [None, Tuning(8, 8, 16, 16, 8), Tuning(16, 4, 256, 1, 64), Tuning(4, 16, 32, 8, 8)])
def test_fused_forces_stencil_equals_separate_launches(ctx, tuning):
    """Forces + sweep in one launch of one-wave blocks gives the bits of the
    two separate launches (float4 rows of 64, 128, 192 and a scalar-path
    grid of 50 x 42 x 37)."""
    ws = _small_ensemble()
    ref = _run_batch(ctx, ws, tuning, 10, fused=0)
    got = _run_batch(ctx, ws, tuning, 10, fused=1)
    for (rc, rf), (gc, gf) in zip(ref, got):
        _same(rc, rf, gc, gf)


@pytest.mark.parametrize("fused,tuning", [(0, Tuning(16, 4, 64, 4, 32)), (1, Tuning(8, 8, 64, 4, 32)),
                                          (1, Tuning(16, 12, 64, 4, 32)), (1, Tuning(4, 16, 16, 16, 8))])
def test_rule_inside_compaction_equals_rule_then_compaction(ctx, fused, tuning):
    """The built-in tumour rule evaluated inside the compaction kernel
    (survivors written in place, daughters straight to their ranks) gives the
    bits of the rule kernel followed by compaction, for every tile size."""
    ws = _small_ensemble()
    ref = _run_batch(ctx, ws, tuning, 10, fused=0, fused_rule=False)
    got = _run_batch(ctx, ws, tuning, 10, fused=fused, fused_rule=True)
    for (rc, rf), (gc, gf) in zip(ref, got):
        _same(rc, rf, gc, gf)


@pytest.mark.parametrize("fused,waves", [(0, 8), (1, 64), (1, 256)])
def test_capped_force_waves_give_the_same_bits(ctx, fused, waves):
    """Fewer force walkers than 64-cell chunks (each strides through its
    XCD's chunks) gives the bits of one walker per chunk."""
    ws = _small_ensemble()
    ref = _run_batch(ctx, ws, None, 6, fused=0)
    got = _run_batch(ctx, ws, None, 6, fused=fused, forces_waves=waves)
    for (rc, rf), (gc, gf) in zip(ref, got):
        _same(rc, rf, gc, gf)


def test_set_tuning_rejects_bad_shapes(ctx):
    b = Batch(ctx)
    try:
        for args in ((10, 4, 64, 4, 32), (16, 5, 64, 4, 32), (32, 4, 64, 4, 32), (16, 32, 64, 4, 32),
                     (16, 4, 64, 3, 32), (16, 4, 64, 4, 0)):
            with pytest.raises(_abi.OncError):
                b.set_tuning(*args)
        for f in (-1, 2):
            with pytest.raises(_abi.OncError):
                _abi.call("onc_batch_set_fused", b.h, f)
        for wv in (-8, 4, 12):
            with pytest.raises(_abi.OncError):
                b.set_forces_waves(wv)
        with pytest.raises(_abi.OncError):
            _abi.call("onc_batch_set_fused_rule", b.h, 2)
    finally:
        b.close()


def _jobs(n, seed=7):
    rng = np.random.Generator(np.random.PCG64(seed))
    jobs = []
    t = 0.0
    for i in range(n):
        t += rng.exponential(0.02)
        w, _ = W.served_job(rng, i, seed=seed, scale=0.01)
        jobs.append(Job(id=i, workload=w, horizon=int(rng.integers(5, 60)), arrival=t))
    return jobs


def test_served_results_equal_alone(ctx):
    jobs = _jobs(12)
    rep = Server(ctx, hbm_budget=64 << 30, keep_results=True, max_tick_steps=16).run(jobs)
    s = rep.summary()
    assert s["completed"] == len(jobs)
    assert rep.peak_resident > 1, "jobs should overlap on the device"
    for j in jobs:
        cols, fields = run_alone(ctx, j)
        _same(cols, fields, *j.result)
    assert rep.cell_updates == sum(j.cell_updates for j in jobs) > 0


def test_admission_respects_hbm_budget(ctx):
    jobs = _jobs(8, seed=3)
    for j in jobs:
        j.arrival = 0.0
    fps = [footprint(j.workload.desc()) for j in jobs]
    budget = max(fps) + min(fps)   # never room for every job at once
    rep = Server(ctx, hbm_budget=budget, max_tick_steps=8).run(jobs)
    assert rep.summary()["completed"] == len(jobs)
    assert rep.peak_bytes <= budget
    assert rep.peak_resident < len(jobs)
    assert all(j.admitted >= j.arrival for j in jobs)


def test_oversized_job_rejected(ctx):
    j = _jobs(1)[0]
    with pytest.raises(ValueError, match="budget"):
        Server(ctx, hbm_budget=1024).run([j])


def test_stop_step_freezes_a_sim_in_its_batch(ctx):
    """A sim whose completed steps reach its stop is left untouched by later
    batch steps (eager and graph); its state equals that many steps alone."""
    ws = W.ensemble_c3(n_sims=3, scale=0.02)
    stops = [5, 12, 1000]
    alone = []
    for w, k in zip(ws, stops):
        s = Sim(ctx, w.desc())
        s.upload_cells(W.initial_cells(w))
        s.step(min(k, 15))
        s.sync()
        alone.append((s.download_cells(), [s.download_field(0)]))
        s.close()
    sims = [Sim(ctx, w.desc()) for w in ws]
    b = Batch(ctx)
    for s, w, k in zip(sims, ws, stops):
        s.upload_cells(W.initial_cells(w))
        s.set_stop(k)
        b.add(s)
    b.step(8, graph=True)
    b.step(7, graph=False)
    b.sync()
    for s, k, (cols, fields) in zip(sims, stops, alone):
        assert s.counts()[2] == min(k, 15)
        _same(cols, fields, s.download_cells(), [s.download_field(0)])
    b.close()
    for s in sims:
        s.close()


def test_profile_every_samples_steps(ctx):
    """Phase timing on every 3rd of 10 eager steps times steps 0, 3, 6, 9;
    a stride below 1 is rejected."""
    ws = _small_ensemble()[:2]
    sims = [Sim(ctx, w.desc()) for w in ws]
    b = Batch(ctx)
    try:
        for s, w in zip(sims, ws):
            s.upload_cells(W.initial_cells(w))
            b.add(s)
        b.profile(True, phases=["plan", "forces_stencil"], every=3)
        b.step(10, graph=False)
        b.sync()
        pt = b.phase_times()
        assert pt["plan"][1] == 4 and pt["forces_stencil"][1] == 4
        assert pt["plan"][0] > 0.0 and pt["morton"][1] == 0
        with pytest.raises(_abi.OncError):
            _abi.call("onc_batch_profile_every", b.h, 1, 0)
    finally:
        b.close()
        for s in sims:
            s.close()


def test_preempted_jobs_resume_bit_exact(ctx, tmp_path):
    """With room for one job at a time and preemption after 0 s, every newly
    arrived job evicts the running one to a bundle; evicted jobs are restored
    later and every job's result equals the job stepped alone."""
    jobs = _jobs(3, seed=5)
    for j, h in zip(jobs, (40, 30, 20)):
        j.arrival = 0.0
        j.horizon = h
    budget = max(footprint(j.workload.desc()) for j in jobs)
    rep = Server(ctx, hbm_budget=budget, max_tick_steps=8, keep_results=True, evict_dir=str(tmp_path),
                 preempt_after=0.0).run(jobs)
    s = rep.summary()
    assert s["completed"] == 3 and s["evictions"] >= 2 and rep.peak_resident == 1
    assert sum(j.evictions for j in jobs) == s["evictions"]
    for j in jobs:
        cols, fields = run_alone(ctx, j)
        _same(cols, fields, *j.result)
    assert not any(p.name.startswith("job") for p in tmp_path.iterdir())  # bundles removed on restore
    assert rep.cell_updates == sum(j.cell_updates for j in jobs) > 0


def test_preemption_needs_a_directory(ctx):
    with pytest.raises(ValueError, match="evict_dir"):
        Server(ctx, hbm_budget=1 << 30, preempt_after=1.0)
