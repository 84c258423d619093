"""Continuous-batching server for simulation jobs on one GPU.

Jobs arrive over time (`Job.arrival`, seconds from the start of `run`). The
server admits waiting jobs, in arrival order with first fit, while their
device footprint (`onc_sim_footprint`: cell state, sort buffers, fields and
bin table) fits in the HBM budget and a slot is free. Admitted jobs stay
resident and are folded into one batch (`onc_batch_*`: one launch per phase
over all resident sims, one captured hipGraph per timestep replayed for the
whole tick). Each job's horizon is set on the device (`onc_sim_set_stop`):
a job that reaches it is left untouched by the rest of the tick, so a tick
can run k = min(max_tick_steps, most steps left of any resident job) steps
with no host round trip, and every job still ends exactly at its horizon.
Jobs that finished during a tick leave the batch at its end and free their
memory; jobs that arrived meanwhile join at the next tick. While the GPU runs
a tick, the host builds the initial state of the jobs that will be admitted
next. Jobs admitted during a tick are resident (allocated, uploaded) while it
runs and join the batch when it ends.

A job's result does not depend on what else is in the batch: a batched step
is bit-identical to stepping the sim alone.

Preemption (optional, `evict_dir` and `preempt_after`): when the oldest
waiting job has waited longer than `preempt_after` seconds and does not fit,
the resident jobs with the most steps left are evicted between ticks, each
to a bundle (oncolytica_amd.bundle: committed cells, fields, step counter),
until it fits; they wait again in arrival order and are restored from their
bundles when admitted, continuing bit for bit. The memory freed is held for
the job it was freed for: it is admitted before any evicted job.
"""
from __future__ import annotations

import math
import os
import shutil
import time
from dataclasses import dataclass, field
from typing import Callable, Optional

import numpy as np

from . import bundle as B
from . import workloads as W
from .hip import _abi
from .hip._sim import Batch, Sim


@dataclass
class Job:
    id: int
    workload: W.Workload
    horizon: int
    arrival: float
    cols0: Optional[list] = None
    sim: Optional[Sim] = None
    footprint: int = 0
    left: int = 0
    admitted: float = math.nan
    finished: float = math.nan
    cell_updates: int = 0
    result: Optional[tuple] = None
    bundle: Optional[str] = None   # where the job waits while evicted
    evictions: int = 0

    @property
    def latency(self) -> float:
        return self.finished - self.arrival


@dataclass
class ServeReport:
    jobs: list
    wall_s: float
    cell_updates: int
    ticks: int
    graph_ticks: int
    peak_resident: int
    peak_bytes: int
    budget_bytes: int
    host_times: dict = field(default_factory=dict)
    evictions: int = 0

    def summary(self) -> dict:
        lat = np.array([j.latency for j in self.jobs if not math.isnan(j.finished)])
        wait = np.array([j.admitted - j.arrival for j in self.jobs if not math.isnan(j.admitted)])
        return {
            "jobs": len(self.jobs),
            "completed": int(lat.size),
            "p50_latency_s": float(np.percentile(lat, 50)) if lat.size else None,
            "p99_latency_s": float(np.percentile(lat, 99)) if lat.size else None,
            "mean_wait_s": float(wait.mean()) if wait.size else None,
            "cell_updates": self.cell_updates,
            "wall_s": self.wall_s,
            "cell_updates_per_s": self.cell_updates / self.wall_s if self.wall_s > 0 else None,
            "ticks": self.ticks,
            "graph_ticks": self.graph_ticks,
            "peak_resident_jobs": self.peak_resident,
            "peak_resident_bytes": self.peak_bytes,
            "hbm_budget_bytes": self.budget_bytes,
            "evictions": self.evictions,
            "host_times_s": self.host_times,
        }


def footprint(desc: _abi.SimDesc) -> int:
    import ctypes as C
    v = C.c_uint64()
    _abi.call("onc_sim_footprint", C.byref(desc), C.byref(v))
    return int(v.value)


def poisson_jobs(rate: float, duration: float, seed: int = 42, scale: float = 1.0,
                 max_cells: int = 4_000_000) -> list[Job]:
    """Poisson arrivals at `rate` jobs/s over `duration` seconds; each job a
    served tumour sim (log-normal cells, median 100k, sigma 1, capped; grid
    side roundup32(max(64, 2 cbrt(cells))); horizon U[100, 2000])."""
    rng = np.random.Generator(np.random.PCG64(seed))
    t, jobs = 0.0, []
    while True:
        t += rng.exponential(1.0 / rate)
        if t >= duration:
            break
        w, h = W.served_job(rng, len(jobs), seed=seed, scale=scale, max_cells=max_cells)
        jobs.append(Job(id=len(jobs), workload=w, horizon=h, arrival=t))
    return jobs


class Server:
    def __init__(self, ctx: _abi.Context, hbm_budget: int, max_jobs: int = 256, max_tick_steps: int = 64,
                 graph_min_steps: int = 4, keep_results: bool = False, tuning=None,
                 log: Optional[Callable[[str], None]] = None, evict_dir: Optional[str] = None,
                 preempt_after: float = math.inf):
        self.ctx = ctx
        self.budget = int(hbm_budget)
        self.max_jobs = max_jobs
        self.max_tick_steps = max_tick_steps
        self.graph_min_steps = graph_min_steps
        self.keep_results = keep_results
        self.tuning = tuning
        self.log = log
        self.evict_dir = evict_dir
        self.preempt_after = preempt_after
        if evict_dir is None and math.isfinite(preempt_after):
            raise ValueError("preemption needs evict_dir")

    @staticmethod
    def _prepare(job: Job) -> None:
        if job.cols0 is None:
            job.cols0 = W.initial_cells(job.workload)
            job.footprint = footprint(job.workload.desc())

    def run(self, jobs: list[Job], clock: Callable[[], float] = time.perf_counter,
            sleep: Callable[[float], None] = time.sleep) -> ServeReport:
        jobs = sorted(jobs, key=lambda j: j.arrival)
        for j in jobs:
            if j.horizon < 1:
                raise ValueError(f"job {j.id}: horizon must be >= 1")
            j.footprint = footprint(j.workload.desc())
            if j.footprint > self.budget:
                raise ValueError(f"job {j.id} needs {j.footprint} B, more than the HBM budget {self.budget} B")
        batch = Batch(self.ctx)
        if self.tuning is not None:
            self.tuning.apply(batch)
        pending = list(jobs)       # not yet arrived
        waiting: list[Job] = []    # arrived, not admitted
        staged: list[Job] = []     # admitted: resident on the device, joining the batch at the next tick
        active: list[Job] = []     # in the batch
        used = 0
        reserved: Optional[Job] = None  # the job that preemption freed memory for
        ticks = graph_ticks = peak_res = peak_bytes = evictions = 0
        total_cu = 0
        times = {"admit_s": 0.0, "retire_s": 0.0, "launch_s": 0.0, "gpu_wait_s": 0.0}
        t0 = clock()

        def arrivals():
            now = clock() - t0
            while pending and pending[0].arrival <= now:
                waiting.append(pending.pop(0))

        def admit():
            # arrival order, first fit within the budget and the slots; memory
            # freed by preemption is held for the job it was freed for, which
            # is admitted before any other (else an evicted job that arrived
            # earlier would take the space back, be evicted again at the next
            # tick, and so on, while the job it was evicted for still waits)
            nonlocal used, reserved
            ta = clock()
            order = list(waiting)
            if reserved is not None:
                order = [reserved] + [w for w in order if w is not reserved]
            for j in order:
                if len(active) + len(staged) >= self.max_jobs:
                    break
                if used + j.footprint > self.budget:
                    if j is reserved:
                        break
                    continue
                if j is reserved:
                    reserved = None
                if j.bundle is not None:
                    # an evicted job: restored with its cells, fields and step counter
                    j.sim, _, _ = B.load_bundle(self.ctx, j.bundle)
                    j.sim.set_stop(j.horizon)
                    shutil.rmtree(j.bundle, ignore_errors=True)
                    j.bundle = None
                else:
                    self._prepare(j)
                    j.sim = Sim(self.ctx, j.workload.desc())
                    j.sim.upload_cells(j.cols0)
                    j.sim.set_stop(j.horizon)
                    j.cols0 = None
                    j.left = j.horizon
                    j.admitted = clock() - t0
                used += j.footprint
                waiting.remove(j)
                staged.append(j)
            times["admit_s"] += clock() - ta

        try:
            while pending or waiting or staged or active:
                arrivals()
                for j in staged:
                    batch.add(j.sim)
                    active.append(j)
                if staged and self.log:
                    self.log(f"t={clock() - t0:7.2f}s resident={len(active)} waiting={len(waiting)} "
                             f"used={used / 2**30:.1f} GiB")
                staged.clear()
                if not active:
                    admit()
                    if not staged and pending:
                        sleep(max(0.0, min(0.05, pending[0].arrival - (clock() - t0))))
                    continue
                peak_res = max(peak_res, len(active))
                peak_bytes = max(peak_bytes, used)
                k = min(max(j.left for j in active), self.max_tick_steps)
                use_graph = k >= self.graph_min_steps
                tl = clock()
                batch.step(k, graph=use_graph)
                times["launch_s"] += clock() - tl
                ticks += 1
                graph_ticks += int(use_graph)
                # host work overlapping the tick: admit arrived jobs (device
                # allocation, upload) and build the initial state of the next
                arrivals()
                admit()
                horizon_t = clock() - t0
                for j in waiting + [p for p in pending[:4] if p.arrival <= horizon_t + 0.5]:
                    self._prepare(j)
                tw = clock()
                batch.sync()
                times["gpu_wait_s"] += clock() - tw
                done_t = clock() - t0
                tr = clock()
                for j in list(active):
                    j.left = max(0, j.left - k)
                    if j.left == 0:
                        j.finished = done_t
                        cu = _cu(j.sim)
                        j.cell_updates += cu
                        total_cu += cu
                        if self.keep_results:
                            j.result = (j.sim.download_cells(),
                                        [j.sim.download_field(f) for f in range(j.sim.desc.n_fields)])
                        batch.remove(j.sim)
                        j.sim.close()
                        j.sim = None
                        used -= j.footprint
                        active.remove(j)
                times["retire_s"] += clock() - tr
                # preemption: the oldest job that has not run yet has waited
                # too long and does not fit; evict the jobs with the most steps
                # left (evicted jobs never preempt: no ping-pong)
                fresh = [w for w in waiting if math.isnan(w.admitted)]
                if fresh and self.evict_dir is not None:
                    head = min(fresh, key=lambda w: w.arrival)
                    if (clock() - t0) - head.arrival > self.preempt_after and used + head.footprint > self.budget:
                        reserved = head
                        for v in sorted(active, key=lambda a: (-a.left, a.id)):
                            if used + head.footprint <= self.budget:
                                break
                            v.evictions += 1
                            evictions += 1
                            v.bundle = os.path.join(self.evict_dir, f"job{v.id}_{v.evictions}")
                            B.export_bundle(v.sim, v.bundle)
                            cu = _cu(v.sim)
                            v.cell_updates += cu
                            total_cu += cu
                            batch.remove(v.sim)
                            v.sim.close()
                            v.sim = None
                            used -= v.footprint
                            active.remove(v)
                            waiting.append(v)
                            waiting.sort(key=lambda w: w.arrival)
                            if self.log:
                                self.log(f"t={clock() - t0:7.2f}s evicted job {v.id} ({v.left} steps left)")
        finally:
            for j in active + staged:
                if j.sim is not None:
                    j.sim.close()
            for j in jobs:
                if j.bundle is not None:
                    shutil.rmtree(j.bundle, ignore_errors=True)
            batch.close()
        wall = clock() - t0
        return ServeReport(jobs=jobs, wall_s=wall, cell_updates=total_cu, ticks=ticks, graph_ticks=graph_ticks,
                           peak_resident=peak_res, peak_bytes=peak_bytes, budget_bytes=self.budget,
                           host_times=times, evictions=evictions)


def _cu(sim: Sim) -> int:
    import ctypes as C
    v = C.c_uint64()
    _abi.call("onc_sim_cell_updates", sim.h, C.byref(v))
    return int(v.value)


def run_alone(ctx: _abi.Context, job: Job):
    """The job stepped alone (its own sim, eager launches): the reference
    result a served job must reproduce."""
    sim = Sim(ctx, job.workload.desc())
    try:
        sim.upload_cells(W.initial_cells(job.workload))
        sim.step(job.horizon)
        sim.sync()
        return sim.download_cells(), [sim.download_field(f) for f in range(sim.desc.n_fields)]
    finally:
        sim.close()
