"""On-disk bundle of one sim: everything needed to restore it on a device and
continue bit for bit (checkpoint, and evict / restore of served jobs).

Replaces the reference's GpuBackend.get_config / export_bundle
(framework/src/oncolytica/gpu/runtime/_gpu_backend.py:196-334), which writes
config.json (struct layouts, uniforms, bind groups, execution order),
shader.wgsl and raw dumps of the uniform, state, cell, tissue and chemistry
buffers, and has no loader. Here a bundle directory holds:

  config.json  format tag and version; the sim descriptor (grid, voxel size,
               key format, capacity, seed, mechanics, per-field diffusion and
               boundary constants, column types, rule parameters); the
               committed state counters (live cells, next daughter id,
               completed steps); the column names in word order; the
               execution order of one step; the rule module's file and kernel
               (absent: the built-in tumour rule) and its metric fields.
  cells.bin    the live cells as structure of arrays: column 0 for all cells,
               then column 1, ... (32-bit words, little-endian).
  fields.bin   each field's nx * ny * nz binary32 values, x fastest.
  chem.bin     (when the model has a chemistry grid) each chemistry column's
               32-bit words over the coarse grid, x fastest.
  rule.hip     the translated HIP source of the cell (and metric) rules,
               rebuilt with hiprtc on load.

Restoring uploads the cells with their next id, the fields and the step
counter (which keys the Philox draws), so the restored sim's next steps are
the ones the original would have taken.
"""
from __future__ import annotations

import json
import os
from typing import Optional, Sequence

import numpy as np

from .hip import _abi
from .hip._abi import FieldDesc, SimDesc
from .hip._sim import Module, Sim

FORMAT = "oncolytica-amd-bundle"
VERSION = 1

_FIELD_KEYS = ("rate", "decay", "init", "bc", "bc_value", "src_radius", "uptake", "lo", "hi")
_DESC_SCALARS = ("dx", "key_format", "capacity", "seed", "cell_radius", "interaction_radius", "k_rep", "k_adh",
                 "mobility", "max_disp")
EXECUTION_ORDER = [
    {"pass": "plan", "kernels": ["k_plan"]},
    {"pass": "keys", "kernels": ["k_morton"]},
    {"pass": "sort", "kernels": ["k_sort_pass"], "iterations": "radix passes"},
    {"pass": "permute", "kernels": ["k_permute_bins"]},
    {"pass": "forces_stencil", "kernels": ["k_forces_stencil"]},
    {"pass": "rule_compact", "kernels": ["k_rule_compact | onc_rule_main + k_compact"]},
]


def desc_to_dict(d: SimDesc) -> dict:
    out = {"dims": list(d.dims)}
    for k in _DESC_SCALARS:
        out[k] = getattr(d, k)
    out["fields"] = [{k: getattr(d.fields[f], k) for k in _FIELD_KEYS} for f in range(d.n_fields)]
    out["col_types"] = [int(d.col_types[c]) for c in range(d.n_cols)]
    out["params"] = [float(d.params[k]) for k in range(d.n_params)]
    out["chem_col_types"] = [int(d.chem_col_types[c]) for c in range(d.n_chem_cols)]
    return out


def desc_from_dict(o: dict) -> SimDesc:
    d = SimDesc()
    d.dims[:] = [int(v) for v in o["dims"]]
    for k in _DESC_SCALARS:
        setattr(d, k, o[k])
    fs = o["fields"]
    if len(fs) > _abi.MAX_FIELDS:
        raise ValueError(f"a bundle holds at most {_abi.MAX_FIELDS} fields")
    d.n_fields = len(fs)
    for f, fo in enumerate(fs):
        fd = FieldDesc()
        for k in _FIELD_KEYS:
            setattr(fd, k, fo[k])
        d.fields[f] = fd
    cts = o["col_types"]
    if len(cts) > _abi.MAX_COLS:
        raise ValueError(f"a bundle holds at most {_abi.MAX_COLS} columns")
    d.n_cols = len(cts)
    for c, t in enumerate(cts):
        d.col_types[c] = int(t)
    ps = o["params"]
    d.n_params = len(ps)
    for k, v in enumerate(ps):
        d.params[k] = v
    ch = o.get("chem_col_types", [])
    if len(ch) > _abi.MAX_CHEM_COLS:
        raise ValueError(f"a bundle holds at most {_abi.MAX_CHEM_COLS} chemistry columns")
    d.n_chem_cols = len(ch)
    for c, t in enumerate(ch):
        d.chem_col_types[c] = int(t)
    return d


def get_config(sim: Sim, column_names: Optional[Sequence[str]] = None, kernel: Optional[str] = None,
               metric_fields: Sequence = ()) -> dict:
    """The bundle's config.json contents for the sim's committed state."""
    n, next_id, step = sim.counts()
    names = list(column_names) if column_names else [f"c{k}" for k in range(sim.desc.n_cols)]
    return {
        "format": FORMAT,
        "version": VERSION,
        "desc": desc_to_dict(sim.desc),
        "state": {"n": n, "next_id": next_id, "step": step},
        "columns": names,
        "execution_order": EXECUTION_ORDER,
        "module": None if kernel is None else {"path": "rule.hip", "kernel": kernel,
                                               "metric_fields": [list(m) for m in metric_fields]},
        "buffers": {"cells": "cells.bin", "fields": "fields.bin",
                    **({"chem": "chem.bin"} if sim.desc.n_chem_cols else {})},
        "layout": layout(sim.desc, len(metric_fields)),
    }


def layout(desc: SimDesc, n_metrics: int = 0) -> dict:
    """The `layout` section of the reference's launch config
    (gpu/runtime/_gpu_backend.py:196-293; framework/tests/launch_config.json
    "layout"), for this build's storage: every record is stored as 32-bit
    columns (structure of arrays), so a record's stride is 4 bytes per column
    (the reference's strides are those of WGSL-aligned structs)."""
    dims = [int(v) for v in desc.dims[:3]]
    return {"tissue_shape": dims, "chemistry_shape": [max(1, d // 2) for d in dims],
            "max_agents": int(desc.capacity), "cell_stride": 4 * int(desc.n_cols),
            "tissue_stride": 4 * int(desc.n_fields), "chem_stride": 4 * int(desc.n_chem_cols),
            "metric_stride": 4 * int(n_metrics), "storage": "structure of arrays, 32-bit columns"}


def export_bundle(sim: Sim, export_dir: str, column_names: Optional[Sequence[str]] = None,
                  module_src: Optional[str] = None, kernel: Optional[str] = None,
                  metric_fields: Sequence = ()) -> str:
    """Writes the sim's committed state as a bundle into export_dir (created
    if missing); returns the path of config.json."""
    if (module_src is None) != (kernel is None):
        raise ValueError("module_src and kernel go together")
    os.makedirs(export_dir, exist_ok=True)
    cfg = get_config(sim, column_names, kernel, metric_fields)
    cols = sim.download_cells()
    if len(cols) and len(cols[0]) != cfg["state"]["n"]:
        raise RuntimeError("cell count changed while exporting")
    with open(os.path.join(export_dir, "cells.bin"), "wb") as f:
        for c in cols:
            f.write(np.ascontiguousarray(c).view(np.uint32).astype("<u4").tobytes())
    with open(os.path.join(export_dir, "fields.bin"), "wb") as f:
        for k in range(sim.desc.n_fields):
            f.write(sim.download_field(k).astype("<f4").tobytes())
    if sim.desc.n_chem_cols:
        with open(os.path.join(export_dir, "chem.bin"), "wb") as f:
            for k in range(sim.desc.n_chem_cols):
                f.write(sim.download_chem(k).view(np.uint32).astype("<u4").tobytes())
    if module_src is not None:
        with open(os.path.join(export_dir, "rule.hip"), "w", encoding="utf-8") as f:
            f.write(module_src)
    path = os.path.join(export_dir, "config.json")
    with open(path, "w", encoding="utf-8") as f:
        json.dump(cfg, f, indent=1)
    return path


def load_bundle(ctx: _abi.Context, bundle: str) -> tuple[Sim, Optional[Module], dict]:
    """Restores a bundle (its directory or config.json) on ctx's device:
    returns the sim (its rule set), the rebuilt rule module (None for the
    built-in rule) and the config."""
    d = bundle if os.path.isdir(bundle) else os.path.dirname(bundle)
    with open(os.path.join(d, "config.json"), encoding="utf-8") as f:
        cfg = json.load(f)
    if cfg.get("format") != FORMAT or cfg.get("version") != VERSION:
        raise ValueError(f"not a version-{VERSION} {FORMAT} bundle: {d}")
    desc = desc_from_dict(cfg["desc"])
    st = cfg["state"]
    n = int(st["n"])
    if n > desc.capacity:
        raise ValueError("the bundle holds more cells than its capacity")
    raw = np.fromfile(os.path.join(d, cfg["buffers"]["cells"]), dtype="<u4")
    if raw.size != n * desc.n_cols:
        raise ValueError("cells.bin does not hold n x columns words")
    nvox = int(np.prod(desc.dims[:3]))
    fld = np.fromfile(os.path.join(d, cfg["buffers"]["fields"]), dtype="<f4")
    if fld.size != nvox * desc.n_fields:
        raise ValueError("fields.bin does not hold one grid per field")
    chem = None
    if desc.n_chem_cols:
        chem = np.fromfile(os.path.join(d, cfg["buffers"]["chem"]), dtype="<u4")
        nchem = int(np.prod([max(1, int(v) // 2) for v in desc.dims[:3]]))
        if chem.size != nchem * desc.n_chem_cols:
            raise ValueError("chem.bin does not hold one grid per chemistry column")
    mod = None
    sim = Sim(ctx, desc)
    try:
        m = cfg.get("module")
        if m:
            with open(os.path.join(d, m["path"]), encoding="utf-8") as f:
                mod = Module(ctx, f.read())
            sim.set_rule(mod, m["kernel"])
        dt = {_abi.COL_F32: np.float32, _abi.COL_I32: np.int32, _abi.COL_U32: np.uint32}
        cols = [raw[c * n:(c + 1) * n].astype(np.uint32).view(dt[int(desc.col_types[c])])
                for c in range(desc.n_cols)]
        sim.upload_cells(cols, next_id=int(st["next_id"]))
        for k in range(desc.n_fields):
            sim.upload_field(k, np.ascontiguousarray(fld[k * nvox:(k + 1) * nvox], dtype=np.float32))
        for k in range(desc.n_chem_cols):
            nchem = sim.nchem
            sim.upload_chem(k, chem[k * nchem:(k + 1) * nchem].astype(np.uint32).view(dt[int(desc.chem_col_types[k])]))
        sim.set_step(int(st["step"]))
    except Exception:
        sim.close()
        if mod is not None:
            mod.close()
        raise
    return sim, mod, cfg
