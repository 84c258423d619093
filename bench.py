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
             events around the dominant kernel (the kernel with the largest
             total time in a separate pass that times every phase);
  roofline   the dominant kernel's algorithmic bytes per launch / its mean
             event-timed duration in the timed region, against 8 TB/s;
  cpu_baseline  the CPU oracle (binary32 C restatement, one thread) on a
             bounded sample of the same workload;
  long_run   the same workload stepped --long-steps (1000) times from its
             initial state as graph replays, timed whole: the population
             grows, then collapses to ~0.1M cells once the spheroid's core
             turns hypoxic.
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


def algorithmic_bytes(phase: str, n_mean: float, nvox: int, n_fields: int, cell_bytes: int, occ_bins: float,
                      sweep_bytes: float = None):
    """Algorithmic HBM bytes of ONE launch of a phase (units: cells/voxels of
    the timed steps). Per-unit figures: Morton 20 B/cell; radix pass 16 B/cell;
    permute 2S + key 4 B/cell + 8 B per occupied bin; forces 12 B read + 12 B
    written + 4 B key per cell; stencil 8 B per swept voxel (read and
    written) + 4 B per voxel filled with its quiet tile's value + 0 per voxel
    kept (`sweep_bytes`, measured; 8 B/voxel/field when not given); forces and
    stencil in one launch: the sum of the two; rule 2S + 4F + 4 B flags per
    cell; compaction 4 B flag + 8 B scan + 2S per cell; the rule inside
    compaction 2S + 4F per cell (state and fields read, survivors written)."""
    S = cell_bytes
    sw = 8.0 * nvox * n_fields if sweep_bytes is None else sweep_bytes
    return {
        "morton": 20.0 * n_mean,
        "sort": 16.0 * n_mean,
        "permute_bins": (2 * S + 4) * n_mean + 8.0 * occ_bins,
        "forces": 28.0 * n_mean,
        "stencil": sw,
        "forces_stencil": 28.0 * n_mean + sw,
        "rule_compact": (2 * S + 4 * n_fields) * n_mean,
        "rule": (2 * S + 4 * n_fields + 4) * n_mean,
        "compact": (4 + 8 + 2 * S) * n_mean,
    }.get(phase)


def cpu_baseline(w, steps: int, steps_1t: int):
    """The oracle (binary32 C restatement; its per-cell and per-voxel loops
    run on all host cores with OpenMP) on `steps` steps of the workload from
    its seeded initial state, and the same on one thread for `steps_1t`
    steps."""
    from oracle import oracle as O
    from oncolytica_amd import workloads as W
    cols = W.initial_cells(w)

    def rate(nsteps):
        st = O.TumorState(w.desc(), cols)
        cu = 0
        t0 = time.perf_counter()
        for _ in range(nsteps):
            cu += st.n.value
            st.run(1)
        return cu / (time.perf_counter() - t0), time.perf_counter() - t0

    threads = O.threads()
    v, dt = rate(steps)
    O.set_threads(1)
    v1, dt1 = rate(steps_1t)
    O.set_threads(threads)
    return {"value": v, "unit": "cell-updates/s", "cores": threads, "kind": "port",
            "sample": f"{steps} steps of {w.name} from its seeded initial state "
                      f"({w.n_cells:,} cells, {w.dims[0]}^3, {len(w.fields)} fields), "
                      f"oracle/onco_oracle_real.c binary32 on {threads} OpenMP threads, {dt:.1f} s",
            "value_1thread": v1, "sample_1thread": f"{steps_1t} steps on 1 thread, {dt1:.1f} s",
            "host_cpu": _cpu_model(), "host": _host_cpus()}


def _cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _host_cpus() -> dict:
    """Logical CPUs of the host (nproc: those this process may run on; the
    machine's total) and lscpu's socket / core / thread counts."""
    import subprocess
    out = {"nproc": len(os.sched_getaffinity(0)), "logical_cpus": os.cpu_count()}
    try:
        txt = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in txt.splitlines():
            k, _, v = line.partition(":")
            k = k.strip()
            if k in ("Model name", "Socket(s)", "Core(s) per socket", "Thread(s) per core", "CPU(s)",
                     "On-line CPU(s) list", "Vendor ID", "CPU max MHz"):
                out["lscpu " + k] = v.strip()
    except (OSError, subprocess.SubprocessError):
        pass
    return out


def _sim_cu(sim) -> int:
    import ctypes as C
    from oncolytica_amd.hip import _abi
    v = C.c_uint64()
    _abi.call("onc_sim_cell_updates", sim.h, C.byref(v))
    return int(v.value)


class Marker:
    """A one-element Philox launch (k_philox, which nothing else in the step
    workload runs) before and after the timed region, so that a kernel trace
    of this command can pick out the timed region's dispatches."""

    def __init__(self, ctx):
        import numpy as np
        self.ctx = ctx
        self.a = ctx.upload(np.zeros(4, np.uint32))
        self.b = ctx.alloc(16)

    def __call__(self):
        from oncolytica_amd.hip import _abi
        _abi.call("onc_philox", self.ctx.h, self.a.p, 1, 1, self.b.p)

    def free(self):
        self.a.free()
        self.b.free()


def kernel_rooflines(ctx):
    """The pure kernels at large sizes, each the median of five calls timed
    with HIP events around its launches (onc_ctx_timing), against 8 TB/s:
    plain copy of 256 MiB (the reference rate), stable radix sort of 16M
    (u32 key, u32 value) pairs at 24 and 27 key bits (3 and 4 passes of
    16 B/pair; the 256^3 and 512^3 key widths), exclusive scan of 64M u32
    (8 B/element), one 7-point sweep of a 512^3 binary32 field (8 B/voxel),
    Morton keys of 16M positions (12 B read + 4 B written per cell), raw
    neighbour forces of 1M key-sorted cells (28 B per cell; bound by VALU issue)."""
    import numpy as np
    from oncolytica_amd.hip import _abi
    from oncolytica_amd.hip._abi import BC_DIRICHLET, FieldDesc
    rng = np.random.default_rng(7)
    out = {}
    ctx.timing(True)

    def timed(name, alg_bytes, fn, **extra):
        fn()
        us = []
        for _ in range(5):
            fn()
            us.append(ctx.last_kernel_us())
        t = float(np.median(us))
        gbs = alg_bytes / (t * 1e-6) / 1e9
        out[name] = {"us": t, "alg_bytes": alg_bytes, "GBps": gbs, "frac": gbs / HBM_PEAK_GBS, **extra}
        _log(f"kernel {name:18s} {t:9.1f} us {gbs:8.0f} GB/s ({100 * gbs / HBM_PEAK_GBS:5.1f}% of 8 TB/s)")

    nw = 64 << 20
    a, b = ctx.alloc(4 * nw), ctx.alloc(4 * nw)
    timed("copy_256MiB", 8.0 * nw, lambda: _abi.call("onc_copy_u32", ctx.h, a.p, b.p, nw), words=nw)
    a.free()
    b.free()
    n = 16 << 20
    v = ctx.upload(np.arange(n, dtype=np.uint32))
    ko, vo = ctx.alloc(4 * n), ctx.alloc(4 * n)
    for bits in (24, 27):
        k = ctx.upload(rng.integers(0, 1 << bits, n, dtype=np.uint64).astype(np.uint32))
        P = (bits + 7) // 8
        timed(f"sort_pairs_16M_{bits}b", 16.0 * P * n,
              lambda: _abi.call("onc_sort_pairs", ctx.h, k.p, v.p, n, bits, ko.p, vo.p), n=n, key_bits=bits,
              passes=P)
        k.free()
    v.free()
    ko.free()
    vo.free()
    n = 64 << 20
    x = ctx.upload(rng.integers(0, 16, n, dtype=np.uint64).astype(np.uint32))
    y = ctx.alloc(4 * n)
    timed("exscan_64M", 8.0 * n, lambda: _abi.call("onc_exscan_u32", ctx.h, x.p, y.p, n, None), n=n)
    x.free()
    y.free()
    d = 512
    u = ctx.upload(rng.random(d ** 3, dtype=np.float32))
    w = ctx.alloc(4 * d ** 3)
    f = FieldDesc()
    f.rate, f.bc, f.bc_value, f.lo, f.hi = 0.1, BC_DIRICHLET, 1.0, 0.0, 1.0
    dims = np.array([d, d, d], dtype=np.int32)
    timed("stencil7_512", 8.0 * d ** 3,
          lambda: _abi.call("onc_stencil7", ctx.h, u.p, w.p, dims.ctypes.data_as(_abi._i32p), f), voxels=d ** 3)
    u.free()
    w.free()
    n = 16 << 20
    px, py, pz = (ctx.upload((rng.random(n) * 512).astype(np.float32)) for _ in range(3))
    keys = ctx.alloc(4 * n)
    timed("morton_16M", 16.0 * n,
          lambda: _abi.call("onc_morton_keys", ctx.h, px.p, py.p, pz.p, n, dims.ctypes.data_as(_abi._i32p), 1.0,
                            _abi.KEY_MORTON32, keys.p), n=n)
    for q in (px, py, pz, keys):
        q.free()
    # raw neighbour forces (onc_forces) of the 1M initial cells of the
    # 256^3 workload, sorted on the host by their voxels' Morton codes (the
    # standard 3-D interleave of 10-bit coordinates, which the 256^3 keys
    # are); 28 B per cell (packed position read, force written): the walk is
    # bound by the latency of its gathers, not by these bytes
    from oncolytica_amd import workloads as W
    w2 = W.large_c2(n_cells=1_000_000)
    d2 = w2.desc()
    c2 = W.initial_cells(w2)
    n = len(c2[0])

    def spread10(a):
        a = a.astype(np.uint32) & np.uint32(0x3FF)
        a = (a | (a << np.uint32(16))) & np.uint32(0x030000FF)
        a = (a | (a << np.uint32(8))) & np.uint32(0x0300F00F)
        a = (a | (a << np.uint32(4))) & np.uint32(0x030C30C3)
        return (a | (a << np.uint32(2))) & np.uint32(0x09249249)

    vox = [np.clip(np.floor(np.asarray(c2[k], dtype=np.float32) / np.float32(d2.dx)), 0, d2.dims[k] - 1)
           for k in range(3)]
    order = np.argsort(spread10(vox[0]) | (spread10(vox[1]) << np.uint32(1)) | (spread10(vox[2]) << np.uint32(2)),
                       kind="stable")
    ps = [ctx.upload(np.ascontiguousarray(c2[k][order], dtype=np.float32)) for k in range(3)]
    fo = [ctx.alloc(4 * n) for _ in range(3)]
    timed("forces_1M", 28.0 * n,
          lambda: _abi.call("onc_forces", ctx.h, d2, ps[0].p, ps[1].p, ps[2].p, n, fo[0].p, fo[1].p, fo[2].p),
          n=n, cells_per_s=None)
    out["forces_1M"]["cells_per_s"] = n / (out["forces_1M"]["us"] * 1e-6)
    for q in ps + fo:
        q.free()
    ctx.timing(False)
    return out


def metrics_leg(ctx, desc, cols0, Sim):
    """Metric rules on the device (onc_sim_metrics): the benchmark model's
    translated metric rule (population, summed age, summed O2 at the cells,
    cells ready to divide) over the 1M committed initial cells, ten calls
    timed on the host clock (each includes its commit check, two launches and
    the 16-byte read-back). Algorithmic bytes: position and age (16 B) plus
    one O2 gather (4 B) per cell."""
    from oncolytica_amd.hip._sim import Module
    from oncolytica_amd.hip._translator import CellSchema, TissueSchema, Translator, param_values
    from oncolytica_amd.spheroid import Spheroid, SpheroidMetrics, SpheroidParams, SpheroidTissue, TumorCell
    import numpy as np
    model = Spheroid()
    tr = Translator(model, CellSchema.of(TumorCell), TissueSchema.of(SpheroidTissue), model._params)
    src = tr.translate([model.lifecycle], [model.count], SpheroidMetrics)
    t0 = time.perf_counter()
    mod = Module(ctx, src)
    t_build = time.perf_counter() - t0
    sim = Sim(ctx, desc)
    sim.upload_cells(cols0)
    # the workload's rule parameters (built-in order: division age, division
    # O2, hypoxia, drug kill, daughter offset) in the translated module's slots
    a0, a1, a2, a3, a4 = (float(x) for x in desc.params[:5])
    sim.set_params(param_values(SpheroidParams(div_age=a0, div_o2=a1, hypoxia=a2, drug_kill=a3, daughter_offset=a4)))
    nv = len(tr.metric_fields)
    sim.metrics(mod, nv)
    ts = []
    for _ in range(10):
        t0 = time.perf_counter()
        words = sim.metrics(mod, nv)
        ts.append(time.perf_counter() - t0)
    n = len(cols0[0])
    vals = {}
    for (name, t), w in zip(tr.metric_fields, words):
        vals[name] = float(np.array([w], np.uint32).view(np.float32)[0]) if t == "f32" else int(w)
    us = float(np.median(ts)) * 1e6
    sim.close()
    mod.close()
    out = {"cells": n, "us_per_call": us, "alg_bytes": 20 * n, "GBps": 20 * n / (us * 1e-6) / 1e9,
           "hiprtc_build_s": t_build, "values": vals}
    _log(f"metrics: {n:,} cells in {us:.1f} us per call ({out['GBps']:.0f} GB/s incl. host sync), {vals}")
    return out


def bundle_leg(ctx, desc, cols0, Sim):
    """Checkpoint of the 1M-cell sim after 10 steps (oncolytica_amd.bundle):
    export to a temporary directory and restore on the same device, on the
    host clock; the restored sim's next step is compared bit for bit with
    the original's."""
    import shutil
    import tempfile
    import numpy as np
    from oncolytica_amd import bundle as B
    from oncolytica_amd import workloads as W
    sim = Sim(ctx, desc)
    sim.upload_cells(cols0)
    sim.step(10)
    sim.sync()
    d = tempfile.mkdtemp(prefix="onc_bundle_")
    try:
        t0 = time.perf_counter()
        B.export_bundle(sim, d, W.TUMOR_COL_NAMES)
        t_exp = time.perf_counter() - t0
        nbytes = sum(os.path.getsize(os.path.join(d, f)) for f in os.listdir(d))
        t0 = time.perf_counter()
        r, _, cfg = B.load_bundle(ctx, d)
        t_load = time.perf_counter() - t0
        sim.step(1)
        r.step(1)
        same = all(np.array_equal(a.view(np.uint32), b.view(np.uint32))
                   for a, b in zip(sim.download_cells(), r.download_cells()))
        r.close()
    finally:
        shutil.rmtree(d, ignore_errors=True)
        sim.close()
    out = {"cells": cfg["state"]["n"], "bytes": nbytes, "export_s": t_exp, "restore_s": t_load,
           "next_step_bit_identical": bool(same)}
    _log(f"bundle: {out['cells']:,} cells, {nbytes / 1e6:.0f} MB, export {t_exp:.2f} s, restore {t_load:.2f} s, "
         f"next step bit-identical {same}")
    return out


def ensemble(ctx, n_sims: int, steps: int):
    """Batched uneven ensemble (64 sims, cells log-uniform in [10k, 500k],
    grids from {64, 128, 192}^3, per-sim uptake / division age / hypoxia):
    sequential plain runs (each sim alone, one after the other) against one
    batch (one launch per phase, one hipGraph per step, autotuned tiles).
    Both timed from resident initial state to the end of the last step; the
    batched results are compared bit for bit with the plain ones."""
    import numpy as np
    from oncolytica_amd import workloads as W
    from oncolytica_amd.hip._autotune import tune
    from oncolytica_amd.hip._sim import Batch, Sim
    ws = W.ensemble_c3(n_sims=n_sims)
    cols = [W.initial_cells(w) for w in ws]
    _log(f"ensemble: {n_sims} sims, {sum(w.n_cells for w in ws):,} cells, {steps} steps")
    t_tune = time.perf_counter()
    tun = tune(ctx, [(w.desc(), c) for w, c in zip(ws, cols)], steps=3)
    t_tune = time.perf_counter() - t_tune
    _log(f"ensemble: autotuned {tun} in {t_tune:.1f} s")
    check = set(range(0, n_sims, max(1, n_sims // 8)))
    plain_t, plain_cu, plain_out = 0.0, 0, {}
    for i, (w, c) in enumerate(zip(ws, cols)):
        s = Sim(ctx, w.desc())
        s.upload_cells(c)
        s.sync()
        t0 = time.perf_counter()
        s.step(steps)
        s.sync()
        plain_t += time.perf_counter() - t0
        plain_cu += _sim_cu(s)
        if i in check:
            plain_out[i] = (s.download_cells(), s.download_field(0))
        s.close()
    _log(f"ensemble: sequential plain {plain_cu / plain_t / 1e6:.0f} M cell-updates/s ({plain_t:.2f} s)")
    sims = [Sim(ctx, w.desc()) for w in ws]
    b = Batch(ctx)
    tun.apply(b)
    for s, c in zip(sims, cols):
        s.upload_cells(c)
        b.add(s)
    b.sync()
    t0 = time.perf_counter()
    b.step(steps, graph=True)
    b.sync()
    bt = time.perf_counter() - t0
    bcu = sum(_sim_cu(s) for s in sims)
    same = True
    for i, (pc, pf) in plain_out.items():
        got = sims[i].download_cells()
        same &= all(np.array_equal(a.view(np.uint32), g.view(np.uint32)) for a, g in zip(pc, got))
        same &= np.array_equal(pf.view(np.uint32), sims[i].download_field(0).view(np.uint32))
    b.close()
    for s in sims:
        s.close()
    _log(f"ensemble: batched {bcu / bt / 1e6:.0f} M cell-updates/s ({bt:.2f} s), "
         f"speedup {(bcu / bt) / (plain_cu / plain_t):.2f}x, bit-identical {same}")
    return {"workload": "ensemble_64_uneven", "sims": n_sims, "steps": steps,
            "cells_initial": int(sum(w.n_cells for w in ws)),
            "batched_cell_updates_per_s": bcu / bt, "batched_s": bt,
            "plain_sequential_cell_updates_per_s": plain_cu / plain_t, "plain_sequential_s": plain_t,
            "speedup": (bcu / bt) / (plain_cu / plain_t), "cell_updates": bcu,
            "bit_identical_to_plain": bool(same), "sims_compared": len(plain_out),
            "tuning": {"sort_items": tun.items, "compact_items": tun.citems, "sbx": tun.sbx, "sby": tun.sby,
                       "sbz": tun.sbz, "fused": tun.fused},
            "autotune_s": t_tune}


def serve(ctx, rate: float, seconds: float, budget_gb: float, progress_s: float = 0.0):
    """Served workload: Poisson arrivals, HBM-budget admission, continuous
    batching; per-job latency = completion - arrival (wall clock). With
    progress_s > 0 the server's admission log is printed at most once per
    progress_s seconds (a long window then still shows progress)."""
    from oncolytica_amd.serving import Server, poisson_jobs
    jobs = poisson_jobs(rate, seconds)
    _log(f"serving: {len(jobs)} jobs over {seconds:.0f} s at {rate}/s, budget {budget_gb:.0f} GB")
    log = None
    if progress_s > 0:
        last = [-progress_s]

        def log(msg):
            now = time.perf_counter()
            if now - last[0] >= progress_s:
                last[0] = now
                _log("serving: " + msg)
    srv = Server(ctx, hbm_budget=int(budget_gb * 1e9), log=log)
    rep = srv.run(jobs)
    out = rep.summary()
    out.update({"workload": "served_poisson", "rate_jobs_per_s": rate, "arrival_window_s": seconds})
    _log(f"serving: p50 {out['p50_latency_s']:.2f} s p99 {out['p99_latency_s']:.2f} s, "
         f"{out['cell_updates_per_s'] / 1e6:.0f} M cell-updates/s, peak {out['peak_resident_jobs']} resident, "
         f"{out['ticks']} ticks, host {out['host_times_s']}")
    return out


def stress(ctx, steps: int = 300, burst=(100, 110)):
    """Stress workload: 16M cells in 32 Gaussian clusters, 512^3, three
    fields, 300 steps with a mitotic burst (division age 1 for steps
    [100, 110)). Timed as one captured graph per step (re-captured when the
    parameters change). Its full-size parity and binary64 drift are checked
    by tests/test_gpu_configs.py and reported by tests/drift_report.py."""
    from oncolytica_amd import workloads as W
    from oncolytica_amd.hip._sim import Batch, Sim
    w = W.stress_c5()
    cols0 = W.initial_cells(w)
    base = list(w.params)
    hot = [1.0] + base[1:]
    _log(f"stress: {w.n_cells:,} cells, {w.dims[0]}^3, {len(w.fields)} fields, {steps} steps, burst {burst}")

    def params_at(k):
        return hot if burst[0] <= k < burst[1] else base

    sim = Sim(ctx, w.desc())
    sim.upload_cells(cols0)
    b = Batch(ctx)
    b.add(sim)
    b.sync()
    done, ns = 0, []
    t0 = time.perf_counter()
    for edge in (burst[0], burst[1], steps):
        if edge <= done:
            continue
        sim.set_params(params_at(done))
        b.step(edge - done, graph=True)
        done = edge
        b.sync()
        ns.append((done, sim.counts()[0]))
    b.sync()
    wall = time.perf_counter() - t0
    cu = _sim_cu(sim)
    n_final = sim.counts()[0]
    b.close()
    sim.close()
    out = {"workload": w.name, "cells_initial": w.n_cells, "grid": list(w.dims), "fields": len(w.fields),
           "steps": steps, "burst_steps": list(burst), "wall_s": wall, "cell_updates": cu,
           "cell_updates_per_s": cu / wall, "live_cells_at": ns, "final_live_cells": n_final}
    _log(f"stress: {cu / wall / 1e6:.0f} M cell-updates/s over {steps} steps ({wall:.2f} s), n at {ns}")
    return out


# the driver's command under a kernel trace (`--summarize-timed DIR` reduces
# it, and `--summarize DIR` takes its kernel table from it). The serving leg
# runs after the timed region and is left out: under the kernel trace the
# process took a host-side SIGSEGV inside the HIP runtime during that leg, and
# during the 1000-step eager pass of the default command, which both run to
# completion without the profiler.
DRIVER_ARGS = "--gpus 1 --steps 20 --warmup 5"
TIMED_RUN = ("rocprofv3 --kernel-trace --stats -d {d}/kt -o run -- python3 bench.py " + DRIVER_ARGS +
             " --serve-seconds 0 > {d}/bench.json")
# rocprofv3 runs whose outputs `--summarize DIR` reduces (each under its own
# subdirectory of DIR, written with `-d DIR/<sub> -o run`): the traced driver
# command and two byte-counter runs of the same steps
_PMC_ARGS = DRIVER_ARGS + " --no-cpu-baseline --ensemble-sims 0 --serve-seconds 0 --long-steps 0 --no-kernels --calibrate"
PROFILE_RUNS = {
    "kt": TIMED_RUN,
    "FETCH_SIZE": "rocprofv3 --pmc FETCH_SIZE -d {d}/FETCH_SIZE -o run -- python3 bench.py " + _PMC_ARGS,
    "WRITE_SIZE": "rocprofv3 --pmc WRITE_SIZE -d {d}/WRITE_SIZE -o run -- python3 bench.py " + _PMC_ARGS,
}
KERNEL_PHASE = {"k_plan": "plan", "k_morton": "morton", "k_sort_pass": "sort", "k_permute_bins": "permute_bins",
                "k_forces": "forces", "k_stencil": "stencil", "k_rule_tumor": "rule", "k_compact": "compact",
                "k_forces_stencil": "forces_stencil", "k_rule_compact": "rule_compact", "k_copy_u32": "copy",
                "k_copy_v4": "copy"}
COPY_BYTES = 4 * (64 << 20)  # one --calibrate copy launch reads and writes this many bytes
# the pure kernels alone (bench.py --kernels-only): trace and byte counters
PROFILE_RUNS_KERNELS = {
    "kt": "rocprofv3 --kernel-trace --stats -d {d}/kkt -o run -- python3 bench.py --kernels-only",
    "FETCH_SIZE": "rocprofv3 --pmc FETCH_SIZE -d {d}/kFETCH_SIZE -o run -- python3 bench.py --kernels-only --calibrate",
    "WRITE_SIZE": "rocprofv3 --pmc WRITE_SIZE -d {d}/kWRITE_SIZE -o run -- python3 bench.py --kernels-only --calibrate",
}
# algorithmic bytes per dispatch of the pure kernels at the sizes kernel_rooflines runs
PURE_ALG_BYTES = {
    "k_copy_v4": (8 * (64 << 20), "256 MiB read + 256 MiB written"),
    "k_sort_up": (4 * (16 << 20) + 4 * 256 * 4096, "16M keys read, 256 x 4096 counts written"),
    "k_sort_down": (16 * (16 << 20) + 4 * 256 * 4096, "16M (key, value) pairs read and written, 1M offsets read"),
    "k_exscan": (None, "8 B per element"),
    "k_stencil": (8 * 512 ** 3, "512^3 binary32 read + written"),
    "k_morton_keys": (16 * (16 << 20), "16M positions (12 B) read, keys (4 B) written"),
    "k_forces_raw": (28 * 1_000_000, "1M packed positions (16 B) read, forces (12 B) written"),
}


def _kernel_phase(name: str):
    base = name.split("(")[0].split()[-1].split("<")[0]
    return KERNEL_PHASE.get(base)


def _rocpd(path: str):
    import glob
    import sqlite3
    dbs = sorted(glob.glob(os.path.join(path, "**", "*_results.db"), recursive=True))
    if not dbs:
        raise SystemExit(f"no rocpd database under {path}")
    return sqlite3.connect(dbs[-1])


def summarize(d: str, workload: str = "large_1m_256", outdir: str = None):
    """Reduce the PROFILE_RUNS outputs under `d` to profiles/: the kernel-trace
    statistics, and HBM traffic per launch from FETCH_SIZE / WRITE_SIZE (KiB per
    dispatch) scaled by what the same counters read for the plain copy kernel
    of known size run in the same process."""
    kt = _rocpd(os.path.join(d, "kt")).execute(
        "select name, count(*), avg(end-start), sum(end-start), min(end-start), max(end-start) "
        "from kernels group by name order by sum(end-start) desc").fetchall()
    total = sum(r[3] for r in kt)
    pmc = {}
    for ctr in ("FETCH_SIZE", "WRITE_SIZE"):
        for name, v in _rocpd(os.path.join(d, ctr)).execute(
                "select kernel_name, avg(value) from counters_collection where counter_name = ? "
                "group by kernel_name", (ctr,)).fetchall():
            ph = _kernel_phase(name)
            if ph:
                pmc.setdefault(ph, {})[ctr] = float(v)
    cal = {c: COPY_BYTES / 1024.0 / pmc["copy"][c] for c in ("FETCH_SIZE", "WRITE_SIZE")}
    traffic = {ph: 1024.0 * (v["FETCH_SIZE"] * cal["FETCH_SIZE"] + v["WRITE_SIZE"] * cal["WRITE_SIZE"])
               for ph, v in pmc.items() if ph != "copy" and len(v) == 2}
    outdir = outdir or os.path.join(ROOT, "profiles")
    os.makedirs(outdir, exist_ok=True)
    tj = os.path.join(outdir, f"traffic_{workload}.json")
    with open(tj, "w") as f:
        json.dump({**{k: round(v) for k, v in traffic.items()},
                   "_calibration": {"copy_bytes_each_way": COPY_BYTES,
                                    "fetch_factor": round(cal["FETCH_SIZE"], 4),
                                    "write_factor": round(cal["WRITE_SIZE"], 4)},
                   "_commands": [PROFILE_RUNS["FETCH_SIZE"].format(d=d),
                                 PROFILE_RUNS["WRITE_SIZE"].format(d=d)]}, f, indent=1)
    lines = [f"# Kernel trace and HBM traffic, {workload}", "",
             "Kernel trace (`--kernel-trace --stats`):", "",
             "    export TMPDIR=/tmp; " + PROFILE_RUNS["kt"].format(d=d), "",
             "| kernel | calls | avg us | min us | max us | total ms | % |", "|---|---|---|---|---|---|---|"]
    for name, n, avg, tot, mn, mx in kt:
        lines.append(f"| `{name.split('(')[0]}` | {n} | {avg / 1e3:.2f} | {mn / 1e3:.2f} | {mx / 1e3:.2f} | "
                     f"{tot / 1e6:.2f} | {100.0 * tot / total:.1f} |")
    lines += ["", "HBM traffic per launch (two `--pmc` runs, one counter each):", ""]
    lines += ["    " + PROFILE_RUNS[c].format(d=d) for c in ("FETCH_SIZE", "WRITE_SIZE")]
    lines += ["", f"Calibration: the plain copy kernel (`k_copy_v4`, {COPY_BYTES >> 20} MiB read + "
                  f"{COPY_BYTES >> 20} MiB written per launch) reads FETCH_SIZE {pmc['copy']['FETCH_SIZE']:.0f} KiB "
                  f"and WRITE_SIZE {pmc['copy']['WRITE_SIZE']:.0f} KiB, so FETCH_SIZE is scaled by "
                  f"{cal['FETCH_SIZE']:.3f} and WRITE_SIZE by {cal['WRITE_SIZE']:.3f}.", "",
              "| phase | FETCH_SIZE KiB | WRITE_SIZE KiB | HBM bytes / launch (scaled) |", "|---|---|---|---|"]
    for ph, v in sorted(pmc.items()):
        if ph in traffic:
            lines.append(f"| {ph} | {v['FETCH_SIZE']:.0f} | {v['WRITE_SIZE']:.0f} | {traffic[ph]:,.0f} |")
    md = os.path.join(outdir, f"kernels_{workload}.md")
    with open(md, "w") as f:
        f.write("\n".join(lines) + "\n")
    print(f"wrote {tj} and {md}")


def summarize_timed(d: str, outdir: str = None, workload: str = "large_1m_256"):
    """Reduce a TIMED_RUN (its trace under d/kt, its JSON line in
    d/bench.json) to profiles/timed_region_<workload>.md: every kernel
    dispatched between the two Marker launches (the timed region), and the
    dominant kernel's traced average duration beside the event-timed one the
    JSON line's roofline uses."""
    rows = _rocpd(os.path.join(d, "kt")).execute("select name, start, end from kernels order by start").fetchall()
    base = [r[0].split("(")[0].split()[-1].split("<")[0] for r in rows]
    marks = [i for i, b in enumerate(base) if b == "k_philox"]
    if len(marks) < 2:
        raise SystemExit("the trace has no pair of timed-region markers (k_philox)")
    m0, m1 = marks[0], marks[1]
    line = None
    for ln in open(os.path.join(d, "bench.json")):
        if ln.startswith("{"):
            line = json.loads(ln)
    roof = line["roofline"]
    stats = {}
    for (name, st, en), b in zip(rows[m0 + 1:m1], base[m0 + 1:m1]):
        n, tot, mn, mx = stats.get(b, (0, 0, float("inf"), 0))
        du = en - st
        stats[b] = (n + 1, tot + du, min(mn, du), max(mx, du))
    dom = "k_" + roof["kernel"]
    n, tot, _, _ = stats[dom]
    avg_us = tot / n / 1e3
    frac_trace = roof["alg_bytes_per_launch"] / (avg_us * 1e-6) / 1e9 / HBM_PEAK_GBS
    region_ms = (rows[m1][1] - rows[m0][2]) / 1e6
    lines = [f"# Timed region of the benchmark command, {workload}", "",
             "Kernel trace of the benchmark command itself:", "",
             "    export TMPDIR=/tmp; " + TIMED_RUN.format(d=d), "",
             f"The timed region is the span between the two one-element `k_philox` marker launches "
             f"({line['steps']} steps, {region_ms:.2f} ms from the end of the first marker to the start of the second; "
             f"the JSON line reports {line['ms_per_step'] * line['steps']:.2f} ms on the host clock).", "",
             "| kernel | calls | avg us | min us | max us | total ms |", "|---|---|---|---|---|---|"]
    for b, (n_, t_, mn, mx) in sorted(stats.items(), key=lambda kv: -kv[1][1]):
        lines.append(f"| `{b}` | {n_} | {t_ / n_ / 1e3:.2f} | {mn / 1e3:.2f} | {mx / 1e3:.2f} | {t_ / 1e6:.3f} |")
    lines += ["", f"Dominant kernel `{dom}`: {n} launches in the region, traced average {avg_us:.2f} us; "
                  f"the JSON line's event-timed average over {roof['launches_timed']} sampled launches is "
                  f"{roof['avg_launch_us']:.2f} us (ratio {roof['avg_launch_us'] / avg_us:.3f}).",
              f"Algorithmic bytes per launch {roof['alg_bytes_per_launch'] / 1e6:.1f} MB: "
              f"{roof['alg_bytes_per_launch'] / (avg_us * 1e-6) / 1e9:.0f} GB/s on the traced average, "
              f"frac {frac_trace:.3f} of 8 TB/s (the JSON line: {roof['frac']:.3f})."]
    for part, v in (roof.get("parts") or {}).items():
        lines.append(f"Its {part} part alone (`{v['kernel']}`, separate-launch pass): {v['avg_launch_us']:.2f} us, "
                     f"{v['achieved']:.0f} GB/s, frac {v['frac']:.3f} ({v['bound']}).")
    outdir = outdir or os.path.join(ROOT, "profiles")
    os.makedirs(outdir, exist_ok=True)
    md = os.path.join(outdir, f"timed_region_{workload}.md")
    with open(md, "w") as f:
        f.write("\n".join(lines) + "\n")
    print(f"wrote {md}")


def summarize_kernels(d: str, outdir: str = None):
    """Reduce the PROFILE_RUNS_KERNELS outputs under `d` to
    profiles/pure_kernels_16m.md: per pure kernel, its traced duration and
    its HBM bytes per dispatch from the two calibrated counter runs, against
    its algorithmic bytes."""
    kt = {name.split("(")[0]: (n, avg) for name, n, avg in _rocpd(os.path.join(d, "kkt")).execute(
        "select name, count(*), avg(end-start) from kernels group by name").fetchall()}
    pmc = {}
    for ctr in ("FETCH_SIZE", "WRITE_SIZE"):
        for name, v in _rocpd(os.path.join(d, "k" + ctr)).execute(
                "select kernel_name, avg(value) from counters_collection where counter_name = ? "
                "group by kernel_name", (ctr,)).fetchall():
            pmc.setdefault(name.split("(")[0], {})[ctr] = float(v)
    copy = pmc["k_copy_v4"]
    cal = {c: COPY_BYTES / 1024.0 / copy[c] for c in ("FETCH_SIZE", "WRITE_SIZE")}
    lines = ["# Pure kernels at large sizes: duration and HBM traffic per dispatch", "",
             "Kernel trace, then two counter runs (one counter each) of the same kernels:", ""]
    lines += ["    export TMPDIR=/tmp; " + PROFILE_RUNS_KERNELS[c].format(d=d) for c in ("kt", "FETCH_SIZE", "WRITE_SIZE")]
    lines += ["", f"Calibration: the plain copy (`k_copy_v4`, {COPY_BYTES >> 20} MiB each way) reads FETCH_SIZE "
                  f"{copy['FETCH_SIZE']:.0f} KiB and WRITE_SIZE {copy['WRITE_SIZE']:.0f} KiB, so FETCH_SIZE is scaled "
                  f"by {cal['FETCH_SIZE']:.3f} and WRITE_SIZE by {cal['WRITE_SIZE']:.3f}.", "",
              "`k_sort_up` / `k_sort_down` are one radix pass each (16M pairs); `k_exscan<64, ...>` is the scan of a "
              "pass's 1M (digit, tile) counts, `k_exscan<128, ...>` the 64M-element scan.", "",
              "| kernel | calls | avg us | HBM read MB | HBM written MB | algorithmic MB | traffic / algorithmic |",
              "|---|---|---|---|---|---|---|"]
    for name in sorted(pmc):
        v = pmc[name]
        if len(v) != 2 or name not in kt:
            continue
        rd = v["FETCH_SIZE"] * cal["FETCH_SIZE"] * 1024.0
        wr = v["WRITE_SIZE"] * cal["WRITE_SIZE"] * 1024.0
        base = name.split("<")[0].replace("void ", "")
        alg = PURE_ALG_BYTES.get(base, (None, ""))[0]
        if base == "k_exscan":
            alg = 8.0 * (64 << 20) if "<128" in name else 8.0 * 256 * 4096
        n, avg = kt[name]
        lines.append(f"| `{name}` | {n} | {avg / 1e3:.1f} | {rd / 1e6:.1f} | {wr / 1e6:.1f} | "
                     + (f"{alg / 1e6:.1f} | {(rd + wr) / alg:.2f} |" if alg else "— | — |"))
    outdir = outdir or os.path.join(ROOT, "profiles")
    os.makedirs(outdir, exist_ok=True)
    md = os.path.join(outdir, "pure_kernels_16m.md")
    with open(md, "w") as f:
        f.write("\n".join(lines) + "\n")
    print(f"wrote {md}")


def main():
    if len(sys.argv) >= 3 and sys.argv[1] == "--summarize":
        summarize(sys.argv[2], outdir=sys.argv[3] if len(sys.argv) > 3 else None)
        return
    if len(sys.argv) >= 3 and sys.argv[1] == "--summarize-timed":
        summarize_timed(sys.argv[2], outdir=sys.argv[3] if len(sys.argv) > 3 else None)
        return
    if len(sys.argv) >= 3 and sys.argv[1] == "--summarize-kernels":
        summarize_kernels(sys.argv[2], outdir=sys.argv[3] if len(sys.argv) > 3 else None)
        return
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--cells", type=int, default=1_000_000)
    ap.add_argument("--cpu-steps", type=int, default=80,
                    help="oracle steps in the multi-core CPU baseline sample (~10 s on 16 threads)")
    ap.add_argument("--cpu-steps-1t", type=int, default=10, help="oracle steps in the one-thread CPU sample (~10 s)")
    ap.add_argument("--stress-only", action="store_true",
                    help="run only the 16M-cell 512^3 stress workload and print its record as the JSON line")
    ap.add_argument("--stress", action="store_true",
                    help="also run the 16M-cell 512^3 stress workload (300 steps, mitotic burst)")
    ap.add_argument("--long-steps", type=int, default=1000,
                    help="steps of the long run of the same workload (graph replays from its initial state, "
                         "through the hypoxic collapse to ~0.1M cells; 0: skip)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--ensemble-sims", type=int, default=64, help="sims of the batched uneven ensemble (0: skip)")
    ap.add_argument("--ensemble-steps", type=int, default=500)
    ap.add_argument("--serve-seconds", type=float, default=20.0,
                    help="arrival window of the served workload (Poisson, 4 jobs/s; 0: skip)")
    ap.add_argument("--serve-rate", type=float, default=4.0)
    ap.add_argument("--serve-only", action="store_true",
                    help="run only the served workload (--serve-seconds window) and print its record as the "
                         "JSON line, with a progress line every 30 s")
    ap.add_argument("--serve-budget-gb", type=float, default=200.0)
    ap.add_argument("--no-kernels", action="store_true",
                    help="skip the pure-kernel rooflines (sort, scan, sweep, Morton at large sizes)")
    ap.add_argument("--event-every", type=int, default=0,
                    help="in the timed region, HIP events bracket the dominant kernel on every N-th step "
                         "(an event pair between launches costs GPU time of its own; 0: every 2nd step up "
                         "to 200 steps, so a short run still samples 10 or more launches, else every 10th)")
    ap.add_argument("--kernels-only", action="store_true",
                    help="run only the pure-kernel rooflines (one JSON line with their figures)")
    ap.add_argument("--calibrate", action="store_true",
                    help="run a plain 256 MiB copy kernel 5 times first (HBM byte-counter calibration)")
    args = ap.parse_args()

    if args.event_every <= 0:
        args.event_every = 2 if args.steps <= 200 else 10
    if args.gpus != 1:
        print(f"bench.py: runs on one GPU; --gpus {args.gpus} is not supported "
              "(sims and ensembles would shard by job, one process per GPU)", file=sys.stderr)
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

    if args.kernels_only:
        print(json.dumps({"kernels": kernel_rooflines(ctx)}))
        return

    if args.serve_only:
        sv = serve(ctx, args.serve_rate, args.serve_seconds, args.serve_budget_gb, progress_s=30.0)
        print(json.dumps({"metric": "p50/p99 job latency and cell-updates/s of the served workload",
                          "value": sv["cell_updates_per_s"], "unit": "cell-updates/s", "n_gpus": 1,
                          "higher_is_better": True, "dtype": "f32", "data": "synthetic",
                          "config": {"workload": sv["workload"], "rate_jobs_per_s": sv["rate_jobs_per_s"],
                                     "arrival_window_s": sv["arrival_window_s"],
                                     "hbm_budget_bytes": sv["hbm_budget_bytes"], "device": info.name.decode(),
                                     "cus": info.compute_units},
                          "serving": sv}))
        return

    if args.stress_only:
        st = stress(ctx)
        print(json.dumps({"metric": "cell-updates/s of the stress workload", "value": st["cell_updates_per_s"],
                          "unit": "cell-updates/s", "n_gpus": 1, "steps": st["steps"],
                          "higher_is_better": True, "dtype": "f32", "data": "synthetic",
                          "config": {"workload": st["workload"], "cells_initial": st["cells_initial"],
                                     "grid": st["grid"], "fields": st["fields"], "burst_steps": st["burst_steps"],
                                     "device": info.name.decode(), "cus": info.compute_units},
                          "stress": st}))
        return

    w = W.large_c2(n_cells=args.cells)
    desc = w.desc()
    cols0 = W.initial_cells(w)

    import ctypes as C

    def fresh(fused=True):
        sim = Sim(ctx, desc)
        sim.upload_cells(cols0)
        b = Batch(ctx)
        b.set_fused(fused)
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

    # profiled passes: the same K steps with HIP events around every phase
    # (per-phase tables; events cost GPU time, so these passes are not the
    # timed one), with forces and the sweep in one launch (the default) and
    # in two launches (their separate times; the sweep's own roofline)
    def profiled(fused):
        psim, pbatch = fresh(fused)
        pbatch.profile(True)
        pbatch.sync()
        pcu0 = cell_updates(psim)
        pbatch.step(args.steps, graph=False)
        pbatch.sync()
        ph = pbatch.phase_times()
        nm = (cell_updates(psim) - pcu0) / max(1, args.steps)
        pbatch.close()
        psim.close()
        return ph, nm

    phases, pn_mean = profiled(True)
    phases_sep, _ = profiled(False)

    def occupied_bins(sim, n_mean):
        import numpy as np
        cols = sim.download_cells()
        n = len(cols[0])
        dims = np.array(list(desc.dims), dtype=np.int64)
        vox = np.minimum(np.floor(np.stack(cols[:3], 1) / desc.dx).astype(np.int64), dims - 1)
        occ_final = len(np.unique(vox[:, 0] + dims[0] * (vox[:, 1] + dims[1] * vox[:, 2])))
        return n_mean * occ_final / max(1, n)  # occupied bins per step, scaled to the mean population

    table = {}
    for name, (us, nl) in phases.items():
        if nl == 0:
            continue
        table[name] = {"total_ms": us / 1000.0, "launches": nl, "avg_us": us / nl}
    table_sep = {name: {"total_ms": us / 1000.0, "launches": nl, "avg_us": us / nl}
                 for name, (us, nl) in phases_sep.items() if nl and name in ("forces", "stencil")}
    dom = max((k for k in table if algorithmic_bytes(k, 1.0, 1, 1, 1, 1.0)), key=lambda k: table[k]["total_ms"])

    # timed region: steps 1..K of the workload from its seeded initial state,
    # eager launches, HIP events around the dominant kernel on every
    # --event-every-th step (its average launch time for the roofline)
    sim, batch = fresh()
    batch.profile(True, phases=[dom], every=args.event_every)
    batch.sync()
    cu0 = cell_updates(sim)
    marker = Marker(ctx)
    marker()
    t0 = time.perf_counter()
    batch.step(args.steps, graph=False)
    batch.sync()
    t1 = time.perf_counter()
    marker()
    marker.free()
    cu1 = cell_updates(sim)
    dom_us, dom_n = batch.phase_times()[dom]
    batch.profile(False)
    wall = t1 - t0
    value = (cu1 - cu0) / wall
    n_mean = (cu1 - cu0) / max(1, args.steps)
    n_final, _, steps_done = sim.counts()
    occ = occupied_bins(sim, n_mean)

    # the same timed region with forces and the sweep as two launches
    ssim, sbatch = fresh(False)
    sbatch.sync()
    scu0 = cell_updates(ssim)
    s0 = time.perf_counter()
    sbatch.step(args.steps, graph=False)
    sbatch.sync()
    s1 = time.perf_counter()
    sep_value = (cell_updates(ssim) - scu0) / (s1 - s0)
    sbatch.close()
    ssim.close()

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

    # how the sweep treated the field voxels over the same steps (quiet tiles
    # are written with their value or left alone instead of read and written),
    # in each launch shape: the sweep's algorithmic bytes per launch
    def sweep_treatment(fused):
        qs, qb = fresh(fused)
        tot = [0, 0, 0]
        for _ in range(args.steps):
            qb.step(1, graph=False)
            for k, v in enumerate(qb.sweep_voxels()):
                tot[k] += v
        qb.close()
        qs.close()
        m = [v / max(1, args.steps) for v in tot]
        return {"swept_voxels": m[0], "filled_voxels": m[1], "kept_voxels": m[2],
                "bytes": 8.0 * m[0] + 4.0 * m[1], "dense_bytes": 8.0 * w.nvox * len(w.fields)}
    sweep_f = sweep_treatment(True)
    sweep_s = sweep_treatment(False)
    _log(f"sweep per step: {sweep_f['swept_voxels']:.0f} voxels swept, {sweep_f['filled_voxels']:.0f} filled, "
         f"{sweep_f['kept_voxels']:.0f} kept of {w.nvox * len(w.fields)} (one launch with the forces)")
    for tag, tb, swb in (("", table, sweep_f["bytes"]), ("separate launches: ", table_sep, sweep_s["bytes"])):
        for name, row in tb.items():
            byts = algorithmic_bytes(name, n_mean, w.nvox, len(w.fields), W.CELL_BYTES, occ, swb)
            gbs = (byts / (row["avg_us"] * 1e-6) / 1e9) if byts else None
            row.update({"alg_bytes": byts, "GBps": gbs})
            _log(f"{tag}{name:14s} {row['total_ms']:9.2f} ms {row['launches']:6d} launches "
                 f"{row['avg_us']:9.2f} us/launch"
                 + (f" {gbs:8.1f} GB/s ({100 * gbs / HBM_PEAK_GBS:5.1f}% of peak)" if gbs else ""))
    dom_avg = dom_us / max(1, dom_n)
    dom_bytes = algorithmic_bytes(dom, n_mean, w.nvox, len(w.fields), W.CELL_BYTES, occ, sweep_f["bytes"])
    dom_gbs = dom_bytes / (dom_avg * 1e-6) / 1e9
    _log(f"timed region: {value / 1e6:.0f} M cell-updates/s, dominant kernel {dom} {dom_avg:.2f} us/launch "
         f"{dom_gbs:.0f} GB/s; graph replay {(gcu1 - gcu0) / (g1 - g0) / 1e6:.0f} M cell-updates/s; "
         f"forces and sweep as two launches {sep_value / 1e6:.0f} M cell-updates/s")
    traffic = None
    tf = os.path.join(ROOT, "profiles", "traffic_large_1m_256.json")
    if os.path.exists(tf):
        try:
            traffic = json.load(open(tf)).get(dom)
        except Exception:
            traffic = None
    roof = {"bound": "hbm", "kernel": dom, "achieved": dom_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": dom_gbs / HBM_PEAK_GBS, "traffic": traffic,
            "alg_bytes_per_launch": dom_bytes, "avg_launch_us": dom_avg,
            "events_every_steps": args.event_every, "launches_timed": dom_n}
    # the fused launch's two parts, each timed as its own launch in the
    # separate-launch profiled pass (same steps): the HBM-bound sweep and the
    # force walk, bound by VALU issue (its bytes are reported against HBM for
    # reference)
    parts = {}
    for part, ph in (("sweep", "stencil"), ("walk", "forces")):
        row = table_sep.get(ph)
        if row and row.get("GBps"):
            parts[part] = {"kernel": "k_stencil" if ph == "stencil" else "k_forces", "avg_launch_us": row["avg_us"],
                           "alg_bytes_per_launch": row["alg_bytes"], "achieved": row["GBps"],
                           "frac": row["GBps"] / HBM_PEAK_GBS,
                           "bound": "hbm" if ph == "stencil" else "valu issue"}
    roof["parts"] = parts
    roof["sweep_treatment"] = {"one_launch": sweep_f, "separate_launch": sweep_s}

    # the long run: the workload stepped --long-steps times from its initial
    # state as graph replays, the whole run timed (the population grows, then
    # collapses once the spheroid's core turns hypoxic; most of the run is
    # spent at ~0.1M cells, where every launch sits on its latency floor)
    long_run = None
    if args.long_steps > 0:
        lsim, lbatch = fresh()
        lbatch.step(1, graph=True)
        lbatch.sync()
        lcu0 = cell_updates(lsim)
        l0 = time.perf_counter()
        lbatch.step(args.long_steps - 1, graph=True)
        lbatch.sync()
        l1 = time.perf_counter()
        lcu = cell_updates(lsim) - lcu0
        ln_final, _, _ = lsim.counts()
        long_run = {"steps": args.long_steps, "timed_steps": args.long_steps - 1,
                    "value": lcu / (l1 - l0), "unit": "cell-updates/s",
                    "ms_per_step": 1000.0 * (l1 - l0) / max(1, args.long_steps - 1),
                    "mean_live_cells": lcu / max(1, args.long_steps - 1), "final_live_cells": ln_final}
        _log(f"long run: {args.long_steps} steps, {long_run['value'] / 1e6:.0f} M cell-updates/s, "
             f"{long_run['ms_per_step'] * 1000:.1f} us/step, mean {long_run['mean_live_cells']:.0f} cells")
        lbatch.close()
        lsim.close()

    met = metrics_leg(ctx, desc, cols0, Sim)
    bun = bundle_leg(ctx, desc, cols0, Sim)
    kern = None if args.no_kernels else kernel_rooflines(ctx)
    cpu = None if args.no_cpu_baseline else cpu_baseline(w, args.cpu_steps, args.cpu_steps_1t)
    ens = ensemble(ctx, args.ensemble_sims, args.ensemble_steps) if args.ensemble_sims > 0 else None
    served = serve(ctx, args.serve_rate, args.serve_seconds, args.serve_budget_gb) if args.serve_seconds > 0 else None
    stress_out = stress(ctx) if args.stress else None

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
        "phases_separate_launches": table_sep,
        "kernels": kern,
        "metrics": met,
        "bundle": bun,
        "value_separate_launches": sep_value,
        "cpu_baseline": cpu,
        "ensemble": ens,
        "serving": served,
        "stress": stress_out,
        "long_run": long_run,
    }
    out["graph_value"] = (gcu1 - gcu0) / (g1 - g0) if gsteps else None
    print(json.dumps(out), flush=True)
    batch.close()
    sim.close()
    ctx.close()


if __name__ == "__main__":
    main()
