"""Host logic of the continuous-batching server and the autotuner, on the
CPU: the device objects are replaced by stand-ins that record what the
server asks of them."""
import math

import numpy as np
import pytest

from oncolytica_amd import serving
from oncolytica_amd import workloads as W
from oncolytica_amd.hip import _autotune


class FakeClock:
    def __init__(self):
        self.t = 0.0

    def __call__(self):
        return self.t

    def sleep(self, dt):
        self.t += max(dt, 1e-3)


class FakeSim:
    live = 0

    def __init__(self, ctx, desc):
        self.desc = desc
        self.steps = 0
        FakeSim.live += 1

    def upload_cells(self, cols):
        pass

    def set_stop(self, stop):
        self.stop = stop

    def close(self):
        FakeSim.live -= 1


class FakeBatch:
    def __init__(self, ctx, clock, step_cost=1e-3):
        self.sims = []
        self.log = []
        self.clock = clock
        self.cost = step_cost

    def add(self, s):
        self.sims.append(s)

    def remove(self, s):
        self.sims.remove(s)

    def step(self, k, graph=True):
        for s in self.sims:
            s.steps = min(s.steps + k, s.stop)
        self.log.append((k, len(self.sims), graph))
        self.clock.t += k * self.cost

    def sync(self):
        pass

    def close(self):
        pass


@pytest.fixture
def fake(monkeypatch):
    clock = FakeClock()
    batches = []

    def mk_batch(ctx):
        b = FakeBatch(ctx, clock)
        batches.append(b)
        return b

    monkeypatch.setattr(serving, "Sim", FakeSim)
    monkeypatch.setattr(serving, "Batch", mk_batch)
    monkeypatch.setattr(serving, "footprint", lambda d: int(d.capacity) * 100)
    monkeypatch.setattr(serving, "_cu", lambda sim: sim.steps)
    monkeypatch.setattr(serving.W, "initial_cells", lambda w, n=None: [])
    FakeSim.live = 0
    return clock, batches


def _jobs(spec):
    out = []
    for i, (arrival, cells, horizon) in enumerate(spec):
        w = W.plain_c1()
        w.capacity = cells
        out.append(serving.Job(id=i, workload=w, horizon=horizon, arrival=arrival))
    return out


def test_every_job_runs_exactly_its_horizon(fake):
    clock, batches = fake
    jobs = _jobs([(0.0, 10, 7), (0.0, 10, 30), (0.005, 10, 3), (0.5, 10, 100)])
    rep = serving.Server(None, hbm_budget=10**9, max_tick_steps=16).run(jobs, clock=clock, sleep=clock.sleep)
    assert [j.cell_updates for j in jobs] == [7, 30, 3, 100]
    assert all(j.finished >= j.admitted >= j.arrival for j in jobs)
    assert rep.summary()["completed"] == 4
    assert max(k for k, _, _ in batches[0].log) <= 16
    assert FakeSim.live == 0


def test_budget_and_first_fit_admission(fake):
    clock, batches = fake
    # job 0 takes the budget; job 1 (large) must wait; job 2 (small) fits beside job 0
    jobs = _jobs([(0.0, 6, 50), (0.0, 6, 10), (0.0, 3, 10)])
    rep = serving.Server(None, hbm_budget=900, max_tick_steps=64).run(jobs, clock=clock, sleep=clock.sleep)
    assert rep.peak_bytes <= 900
    assert jobs[2].admitted == 0.0 and jobs[1].admitted > 0.0
    assert jobs[1].admitted >= jobs[2].finished
    assert [j.cell_updates for j in jobs] == [50, 10, 10]


def test_ticks_run_to_the_longest_horizon_and_graph_only_when_long(fake):
    clock, batches = fake
    jobs = _jobs([(0.0, 10, 2), (0.0, 10, 40)])
    serving.Server(None, hbm_budget=10**9, graph_min_steps=4, max_tick_steps=16).run(
        jobs, clock=clock, sleep=clock.sleep)
    assert batches[0].log[0] == (16, 2, True)   # the short job stops on the device after 2 steps
    assert [j.cell_updates for j in jobs] == [2, 40]
    jobs = _jobs([(0.0, 10, 2), (0.0, 10, 3)])
    serving.Server(None, hbm_budget=10**9, graph_min_steps=4).run(jobs, clock=clock, sleep=clock.sleep)
    assert batches[1].log == [(3, 2, False)]


def test_rejects_oversized_and_empty_horizon(fake):
    clock, _ = fake
    with pytest.raises(ValueError, match="budget"):
        serving.Server(None, hbm_budget=10).run(_jobs([(0.0, 10, 5)]), clock=clock, sleep=clock.sleep)
    with pytest.raises(ValueError, match="horizon"):
        serving.Server(None, hbm_budget=10**9).run(_jobs([(0.0, 10, 0)]), clock=clock, sleep=clock.sleep)


def test_poisson_generator_shapes():
    jobs = serving.poisson_jobs(4.0, 30.0, seed=1)
    assert 60 <= len(jobs) <= 180
    assert all(100 <= j.horizon <= 2000 for j in jobs)
    assert all(j.workload.dims[0] % 32 == 0 and j.workload.dims[0] >= 64 for j in jobs)
    assert all(math.isfinite(j.arrival) for j in jobs)
    gaps = np.diff([0.0] + [j.arrival for j in jobs])
    assert 0.15 < gaps.mean() < 0.4


def test_autotune_signature_groups_capacities():
    a = W.plain_c1().desc()
    b = W.plain_c1().desc()
    b.capacity = a.capacity - 5
    assert _autotune.signature("gfx950", [a]) == _autotune.signature("gfx950", [b])
    c = W.large_c2().desc()
    assert _autotune.signature("gfx950", [a, c]) == _autotune.signature("gfx950", [c, a])
    assert _autotune.signature("gfx950", [a]) != _autotune.signature("gfx950", [c])


class FakeBundles:
    """Stand-in for oncolytica_amd.bundle: remembers an evicted sim's steps."""

    def __init__(self):
        self.saved = {}
        self.log = []

    def export_bundle(self, sim, path):
        self.saved[path] = sim.steps
        self.log.append(("export", path))

    def load_bundle(self, ctx, path):
        s = FakeSim(ctx, None)
        s.steps = s.base = self.saved.pop(path)
        self.log.append(("load", path))
        return s, None, {}


def test_preemption_admits_the_job_it_evicted_for(fake, monkeypatch, tmp_path):
    """Job 0 holds the budget; job 1 arrives later and cannot fit. After the
    wait limit job 0 is evicted, and job 1 (not job 0, which arrived first
    and fits again) takes the freed memory: one eviction, no ping-pong."""
    clock, batches = fake
    bundles = FakeBundles()
    monkeypatch.setattr(serving, "B", bundles)
    monkeypatch.setattr(FakeSim, "base", 0, raising=False)
    monkeypatch.setattr(serving, "_cu", lambda sim: sim.steps - sim.base)
    jobs = _jobs([(0.0, 6, 300), (0.001, 6, 100)])
    rep = serving.Server(None, hbm_budget=1000, max_tick_steps=16, evict_dir=str(tmp_path),
                         preempt_after=0.01).run(jobs, clock=clock, sleep=clock.sleep)
    assert rep.evictions == 1 and jobs[0].evictions == 1 and jobs[1].evictions == 0
    assert [k for k, _ in bundles.log] == ["export", "load"]
    # job 1 ran to its end before job 0 came back
    assert jobs[1].finished <= jobs[0].finished
    assert [j.cell_updates for j in jobs] == [300, 100]
    assert FakeSim.live == 0
