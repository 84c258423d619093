"""Autotuner for a batch's tile shapes.

Four knobs, three of them of `onc_batch_set_tuning`:

* cells per thread of a radix-sort tile (the onesweep passes): 4, 8, 12 or
  16, i.e. tiles of 1024 to 4096 cells. Larger tiles mean fewer look-back
  predecessors and fewer status words, smaller ones more tiles in flight
  (each tile's block works on one CU, so a small population needs small
  tiles to use the chip).
* cells per thread of a compaction tile, the same choice for the compaction
  scan and its column copies.
* forces and the 7-point sweep in one launch or two (`onc_batch_set_fused`);
* the stencil block: `sbx` float4 columns (a power of two up to 256) by
  256 / sbx rows, marching `sbz` planes along z (in 64-thread tiles of at
  most 64 columns when the sweep shares a launch with the forces).

Each candidate runs a few eager steps of throwaway copies of the batch's
sims, re-uploaded to the same initial state before every candidate, with HIP
events around every phase; each knob is chosen by the time of the phases it
shapes (the fused launches count for both of their parts). Results are cached per device
architecture and batch signature, in memory and optionally in a JSON file.
Every shape computes the same result bit for bit (tested); only time moves.
"""
from __future__ import annotations

import json
import math
import os
from dataclasses import asdict, dataclass
from typing import Optional, Sequence

import numpy as np

from . import _abi
from ._sim import Batch, Sim

ITEMS = (4, 8, 16)
STENCIL_SBX = (16, 32, 64, 128)
STENCIL_SBZ = (8, 16, 32, 64)

_memo: dict[str, "Tuning"] = {}


@dataclass(frozen=True)
class Tuning:
    items: int = 8
    citems: int = 4
    sbx: int = 64
    sby: int = 4
    sbz: int = 16
    fused: int = 1

    def apply(self, batch: Batch) -> None:
        batch.set_tuning(self.items, self.citems, self.sbx, self.sby, self.sbz)
        batch.set_fused(bool(self.fused))


def signature(arch: str, descs: Sequence[_abi.SimDesc]) -> str:
    """Batch signature: the architecture plus, per sim (sorted), its grid,
    field count and capacity rounded to a power of two."""
    parts = sorted(f"{d.dims[0]}x{d.dims[1]}x{d.dims[2]}f{d.n_fields}c{1 << max(0, int(d.capacity) - 1).bit_length()}"
                   for d in descs)
    return arch + "|" + ",".join(parts)


def _phase_us(batch: Batch, names: Sequence[str]) -> float:
    pt = batch.phase_times()
    return sum(pt[n][0] for n in names if n in pt)


def tune(ctx: _abi.Context, members: Sequence[tuple[_abi.SimDesc, Sequence[np.ndarray]]], steps: int = 3,
         cache_file: Optional[str] = None, log=None) -> Tuning:
    """Tile shapes for a batch of `members` = [(descriptor, initial cell
    columns)]. Runs 2 len(ITEMS) + 2 + len(STENCIL_SBX) len(STENCIL_SBZ)
    candidates of `steps` eager steps each on throwaway sims."""
    arch = ctx.info().arch.decode()
    key = signature(arch, [d for d, _ in members])
    if key in _memo:
        return _memo[key]
    if cache_file and os.path.exists(cache_file):
        try:
            disk = json.load(open(cache_file))
            if key in disk:
                t = Tuning(**disk[key])
                _memo[key] = t
                return t
        except (OSError, ValueError, TypeError):
            pass
    sims = [Sim(ctx, d) for d, _ in members]
    batch = Batch(ctx)
    try:
        for s in sims:
            batch.add(s)

        def run(t: Tuning, names) -> float:
            t.apply(batch)
            for s, (_, cols) in zip(sims, members):
                s.upload_cells(cols)
            batch.step(1, graph=False)  # first launches after a rebuild
            batch.sync()
            batch.profile(True, phases=names)
            batch.step(steps, graph=False)
            batch.sync()
            us = _phase_us(batch, names)
            batch.profile(False)
            return us

        # each knob is timed on the phases it shapes (the rule runs inside
        # compaction, and forces beside the sweep, when those launches are fused)
        base = Tuning()
        best_items, best_t = base.items, math.inf
        for it in ITEMS:
            us = run(Tuning(it, base.citems, base.sbx, base.sby, base.sbz, base.fused), ("sort",))
            if log:
                log(f"autotune sort items={it}: {us / steps:.1f} us/step")
            if us < best_t:
                best_items, best_t = it, us
        best_citems, best_t = base.citems, math.inf
        for it in ITEMS:
            us = run(Tuning(best_items, it, base.sbx, base.sby, base.sbz, base.fused), ("compact", "rule_compact"))
            if log:
                log(f"autotune compaction items={it}: {us / steps:.1f} us/step")
            if us < best_t:
                best_citems, best_t = it, us
        sweep = ("forces", "stencil", "forces_stencil")
        best_f, best_t = base.fused, math.inf
        for f in (1, 0):
            us = run(Tuning(best_items, best_citems, base.sbx, base.sby, base.sbz, f), sweep)
            if log:
                log(f"autotune forces + sweep fused={f}: {us / steps:.1f} us/step")
            if us < best_t:
                best_f, best_t = f, us
        best_s, best_t = (base.sbx, base.sbz), math.inf
        for sbx in STENCIL_SBX:
            for sbz in STENCIL_SBZ:
                us = run(Tuning(best_items, best_citems, sbx, 256 // sbx, sbz, best_f), sweep)
                if log:
                    log(f"autotune stencil sbx={sbx} sby={256 // sbx} sbz={sbz}: {us / steps:.1f} us/step")
                if us < best_t:
                    best_s, best_t = (sbx, sbz), us
        t = Tuning(best_items, best_citems, best_s[0], 256 // best_s[0], best_s[1], best_f)
    finally:
        batch.close()
        for s in sims:
            s.close()
    _memo[key] = t
    if cache_file:
        try:
            disk = json.load(open(cache_file)) if os.path.exists(cache_file) else {}
        except (OSError, ValueError):
            disk = {}
        disk[key] = asdict(t)
        with open(cache_file, "w") as f:
            json.dump(disk, f, indent=1)
    return t
