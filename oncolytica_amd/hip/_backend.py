"""HipBackend: the reference's backend plugin protocol
(framework/src/oncolytica/core/runtime/_backend.py:6-36) on liboncohip.so.

compile()      translates the cell rules to HIP C++ (hip/_translator.py),
               builds them with hiprtc (onc_module_build), creates the device
               sim (onc_sim_create) and uploads cells and tissue fields
               (replaces GpuBackend.compile, gpu/runtime/_gpu_backend.py:42-104);
run_step()     one asynchronous timestep (onc_sim_step; replaces
               GpuBackend.run_step, _gpu_backend.py:113-171);
sync_to_host() downloads SoA columns and fields (replaces the per-object
               unpack of _gpu_backend.py:373-405);
sync_to_device() uploads them (replaces _gpu_backend.py:353-371).

Metric rules are translated with the cell rules into two more kernels of
the same module; with collect_metrics set, run_step reduces them over the
committed cells on the device (onc_sim_metrics: fixed summation order, so
binary32 metrics are reproducible) and sets the Metrics fields from the
result (replaces the atomicAdd metric kernels, _shader_builder.py:193-204,
and the host loop of cpu/_cpu_backend.py:65-70).
"""
from __future__ import annotations

import atexit
import json
import os
import typing
from typing import Any, Optional

import numpy as np

from .. import _types as T
from . import _abi
from ._abi import FieldDesc, SimDesc
from ._sim import Module, Sim
from ._translator import (BOOL, COL_F32, COL_I32, COL_U32, F32, I32, U32, ChemSchema, CellSchema, TissueSchema,
                          Translator, param_values)

_CONTEXTS: dict[int, _abi.Context] = {}


def context(device: int = 0) -> _abi.Context:
    """The process-wide HIP context of a device."""
    c = _CONTEXTS.get(device)
    if c is None:
        c = _abi.Context(device)
        _CONTEXTS[device] = c
    return c


@atexit.register
def _close_contexts() -> None:
    for c in _CONTEXTS.values():
        try:
            c.close()
        except Exception:
            pass
    _CONTEXTS.clear()


_NP = {COL_F32: np.float32, COL_I32: np.int32, COL_U32: np.uint32}

# 32-bit word type of a chemistry column of each field type
_CHEM_NP = {F32: np.float32, I32: np.int32, U32: np.uint32, BOOL: np.uint32}


def _default(cls, name, t):
    v = getattr(cls, name, None)
    if v is None or callable(v):
        return 0
    return float(v) if t == F32 else int(v)


# no built-in forces unless the engine sets them (Engine.setup_mechanics)
NO_MECHANICS = {"k_rep": 0.0, "k_adh": 0.0, "mobility": 0.25, "max_disp": 0.25, "interaction_radius": None,
                "cell_radius": None}


def _get_path(obj, path: str):
    for p in path.split("."):
        obj = getattr(obj, p)
    return obj


def _grid_field(grid, name: str) -> np.ndarray:
    """One field of a tissue grid, flattened x fastest (the device layout).
    The package's Grid stores fields as arrays (Grid.flat); the reference's
    Grid holds one record object per voxel in an [x, y, z] object array
    (core/utils/_containers.py:17-51)."""
    flat = getattr(grid, "flat", None)
    if callable(flat):
        return flat(name)
    objs = grid._data
    vals = np.empty(objs.shape, dtype=np.float64)
    for idx, o in np.ndenumerate(objs):
        vals[idx] = getattr(o, name)
    return np.ascontiguousarray(vals.transpose(2, 1, 0)).reshape(-1)


def _set_grid_field(grid, name: str, values: np.ndarray) -> None:
    flat = getattr(grid, "flat", None)
    if callable(flat):
        dst = flat(name)
        dst[:] = values.astype(dst.dtype)
        return
    objs = grid._data
    X, Y, Z = objs.shape
    v = values.reshape(Z, Y, X).transpose(2, 1, 0)
    for idx, o in np.ndenumerate(objs):
        setattr(o, name, type(getattr(o, name))(v[idx]))


def _columnar(cells) -> Optional[dict]:
    """The columns of a columnar AgentList (this package's set_columns), or
    None for a list of cell objects (the reference's AgentList,
    core/utils/_containers.py:85-123)."""
    fn = getattr(cells, "columns", None)
    cols = fn() if callable(fn) else None
    return cols if cols is not None and not cells._data else None


class HipBackend:
    """Constructed as the reference constructs GpuBackend, (engine, ctx)
    (core/runtime/_engine.py:153-161): `ctx` is the validator's context,
    which the reference's GPU compiler reads; the translator here does its own
    checks and keeps it only for callers that look for it. Of the engine it
    reads the reference Engine's attributes (core/runtime/_engine.py:26-55):
    max_agents, tissue_shape, tissue_voxel_size, cell_diameter, cells,
    tissue, _sim, _metrics and the rule lists. device, seed, key_format and
    mechanics are this package's Engine's additions and default to device 0,
    seed 0, the 30-bit key format where the grid allows it, and no built-in
    forces (the reference has none)."""

    def __init__(self, engine: Any, ctx: Any = None) -> None:
        self.engine = engine
        self.validation_ctx = ctx
        self.device = int(getattr(engine, "device", 0))
        self._hctx: Optional[_abi.Context] = None
        self.sim: Optional[Sim] = None
        self.module: Optional[Module] = None
        self.schema: Optional[CellSchema] = None
        self.tissue: Optional[TissueSchema] = None
        self.chem: Optional[ChemSchema] = None
        self.source: str = ""
        self._next_id = 0
        self.metric_fields: list = []  # (name, "f32" / "i32" / "u32") per metric word

    @property
    def ctx(self) -> _abi.Context:
        """The HIP context of the engine's device (created on first use)."""
        if self._hctx is None:
            self._hctx = context(self.device)
        return self._hctx

    # ---- compile ---------------------------------------------------------------------
    def compile(self, sim_instance: Any) -> None:
        e = self.engine
        spec = type(sim_instance)._spec
        self.schema = CellSchema.of(spec.cell_class)
        self.tissue = TissueSchema.of(spec.tissue_class if e.tissue is not None else None)
        has_chem = getattr(e, "chemistry", None) is not None or bool(e._chemistry_rules)
        self.chem = ChemSchema.of(spec.chemistry_class if has_chem else None)
        tr = Translator(sim_instance, self.schema, self.tissue, sim_instance._params, chem=self.chem)
        self.source = tr.translate(e._cell_rules, e._metric_rules, spec.metrics_class,
                                   tissue_rules=e._tissue_rules, chem_rules=e._chemistry_rules)
        self.metric_fields = list(tr.metric_fields)
        self.module = Module(self.ctx, self.source)
        self.sim = Sim(self.ctx, self._desc(sim_instance))
        self.sim.set_rule(self.module, Translator.KERNEL)
        self.sync_to_device()

    def _desc(self, sim_instance) -> SimDesc:
        e = self.engine
        d = SimDesc()
        shape = tuple(int(v) for v in e.tissue_shape)
        d.dims[:] = list(shape)
        d.dx = e.tissue_voxel_size
        d.key_format = int(getattr(e, "key_format", 30 if max(shape) <= 1024 else 32))
        d.capacity = e.max_agents
        d.seed = int(getattr(e, "seed", 0))
        m = dict(NO_MECHANICS)
        m.update(getattr(e, "mechanics", None) or {})
        d.interaction_radius = m["interaction_radius"] if m["interaction_radius"] is not None else \
            min(e.tissue_voxel_size, e.cell_diameter)
        d.cell_radius = m["cell_radius"] if m["cell_radius"] is not None else 0.5 * e.cell_diameter
        d.k_rep, d.k_adh, d.mobility, d.max_disp = m["k_rep"], m["k_adh"], m["mobility"], m["max_disp"]
        d.n_fields = len(self.tissue.fields)
        for f, (name, t, dflt) in enumerate(self.tissue.fields):
            fd = FieldDesc()
            if isinstance(dflt, T.Diffusing):
                b = dflt.boundary
                fd.rate, fd.decay, fd.init = dflt.rate, dflt.decay, dflt.init
                fd.bc = _abi.BC_DIRICHLET if b.kind == "dirichlet" else _abi.BC_POINT_SOURCE
                fd.bc_value, fd.src_radius = b.value, b.radius
                fd.uptake, fd.lo, fd.hi = dflt.uptake, dflt.lo, dflt.hi
            else:
                # a static per-voxel value: swept with rate 0 (unchanged), no fixed voxels
                fd.rate, fd.decay, fd.init = 0.0, 0.0, float(dflt or 0.0)
                fd.bc, fd.bc_value, fd.src_radius, fd.uptake = _abi.BC_NONE, 0.0, 0.0, 0.0
                fd.lo, fd.hi = -3.0e38, 3.0e38
            d.fields[f] = fd
        d.n_cols = len(self.schema.columns)
        for k, col in enumerate(self.schema.columns):
            d.col_types[k] = col.coltype
        ct = self.chem.col_types if self.chem is not None else []
        d.n_chem_cols = len(ct)
        for k, t in enumerate(ct):
            d.chem_col_types[k] = t
        vals = param_values(sim_instance._params)
        d.n_params = len(vals)
        for k, v in enumerate(vals):
            d.params[k] = v
        return d

    def update_params(self) -> None:
        if self.sim is not None:
            self.sim.set_params(param_values(self.engine._sim._params))

    # ---- execution ---------------------------------------------------------------------
    def run_step(self, collect_metrics: bool = False) -> None:
        self.sim.step(1)
        if collect_metrics and self.engine._metric_rules:
            self.collect_metrics()

    def collect_metrics(self) -> None:
        """Reduce the metric rules over the committed cells on the device and
        set the engine's Metrics fields."""
        words = self.sim.metrics(self.module, len(self.metric_fields))
        m = self.engine._metrics
        m._clear()
        for (name, t), w in zip(self.metric_fields, words):
            if t == "f32":
                v = float(np.array([w], dtype=np.uint32).view(np.float32)[0])
            elif t == "i32":
                v = int(np.array([w], dtype=np.uint32).view(np.int32)[0])
            else:
                v = int(w)
            setattr(m, name, v)

    def run_steps(self, n: int) -> None:
        self.sim.step(n)

    # ---- bundle (checkpoint) ---------------------------------------------------------
    def get_config(self) -> dict:
        """The bundle config of the compiled sim's committed state
        (replaces GpuBackend.get_config, _gpu_backend.py:196-293)."""
        if self.sim is None:
            raise RuntimeError("Simulation must be compiled before generating config.")
        from .. import bundle
        return bundle.get_config(self.sim, [c.name for c in self.schema.columns], Translator.KERNEL,
                                 self.metric_fields)

    def export_bundle(self, export_dir: str, export_initial_state: bool = True) -> str:
        """Writes config.json and rule.hip, and with export_initial_state the
        cells and fields as the device holds them now (replaces
        GpuBackend.export_bundle, _gpu_backend.py:295-334, which re-uploads
        the host objects first; here the device state is the current one).
        Returns the path of config.json; oncolytica_amd.bundle.load_bundle
        restores it."""
        if self.sim is None:
            raise RuntimeError("Simulation must be compiled before exporting a bundle.")
        from .. import bundle
        names = [c.name for c in self.schema.columns]
        if export_initial_state:
            return bundle.export_bundle(self.sim, export_dir, names, self.source, Translator.KERNEL,
                                        self.metric_fields)
        os.makedirs(export_dir, exist_ok=True)
        cfg = self.get_config()
        cfg["buffers"] = {}
        with open(os.path.join(export_dir, "rule.hip"), "w", encoding="utf-8") as f:
            f.write(self.source)
        path = os.path.join(export_dir, "config.json")
        with open(path, "w", encoding="utf-8") as f:
            json.dump(cfg, f, indent=1)
        return path

    def get_metrics(self) -> Any:
        return self.engine._metrics

    # ---- host <-> device -------------------------------------------------------------------
    def _columns_from_objects(self, objs) -> list[np.ndarray]:
        n = len(objs)
        cols = []
        for col in self.schema.columns:
            dt = _NP[col.coltype]
            if col.name == "_id":
                ids = [getattr(o, "_id", None) for o in objs]
                if any(v is None for v in ids):
                    base = self._next_id
                    ids = [v if v is not None else base + k for k, v in enumerate(ids)]
                cols.append(np.asarray(ids, dtype=dt))
                continue
            vals = [_get_path(o, col.name) for o in objs]
            cols.append(np.asarray(vals, dtype=np.float64).astype(dt) if n else np.zeros(0, dt))
        return cols

    def _columns_from_arrays(self, c: dict) -> list[np.ndarray]:
        pos = np.asarray(c["pos"], dtype=np.float32)
        n = len(pos)
        out = []
        for col in self.schema.columns:
            dt = _NP[col.coltype]
            if col.name.startswith("pos."):
                out.append(np.ascontiguousarray(pos[:, "xyz".index(col.name[-1])]))
            elif col.name == "_id":
                out.append(np.asarray(c.get("_id", np.arange(n) + self._next_id), dtype=dt))
            elif "." in col.name:
                base, ax = col.name.split(".")
                out.append(np.ascontiguousarray(np.asarray(c[base])[:, "xyz".index(ax)], dtype=dt))
            elif col.name in c:
                out.append(np.asarray(c[col.name], dtype=dt))
            else:
                dflt = getattr(self.schema.cls, col.name, 0)
                out.append(np.full(n, dflt, dtype=dt))
        return out

    def sync_to_device(self) -> None:
        e = self.engine
        cells = e.cells
        columnar = _columnar(cells)
        if columnar is not None:
            cols = self._columns_from_arrays(columnar)
        else:
            cells._data = [a for a in cells._data if getattr(a, "_alive", True)]
            cols = self._columns_from_objects(cells._data)
        n = len(cols[0])
        if n > e.max_agents:
            raise RuntimeError(f"{n} cells exceed max_agents={e.max_agents}")
        next_id = max(self._next_id, int(cols[3].max()) + 1 if n else 0)
        self.sim.upload_cells(cols, next_id=next_id)
        self._next_id = next_id
        if e.tissue is not None:
            for f, (name, t, _) in enumerate(self.tissue.fields):
                self.sim.upload_field(f, _grid_field(e.tissue, name).astype(np.float32))
        if self.chem is not None and self.chem.cls is not None:
            grid = getattr(e, "chemistry", None)
            for k, (name, t) in enumerate(self.chem.fields):
                dt = _CHEM_NP[t]
                if grid is not None:
                    vals = np.asarray(_grid_field(grid, name))
                    vals = vals.astype(np.float32) if t == F32 else vals.astype(np.int64).astype(dt)
                else:
                    vals = np.full(self.sim.nchem, _default(self.chem.cls, name, t), dtype=dt)
                self.sim.upload_chem(k, vals)

    def download_columns(self) -> dict:
        """Cells as numpy columns keyed by field name ('pos' is (n, 3))."""
        raw = self.sim.download_cells()
        out: dict = {"pos": np.stack(raw[:3], axis=1), "_id": raw[3]}
        k = 4
        for name, t in self.schema.fields:
            if name == "pos":
                continue
            if t in ("vec3", "ivec3"):
                out[name] = np.stack(raw[k:k + 3], axis=1)
                k += 3
            else:
                out[name] = raw[k].astype(bool) if t == "bool" else raw[k]
                k += 1
        return out

    def sync_to_host(self) -> None:
        e = self.engine
        cols = self.download_columns()
        n_, nid, _ = self.sim.counts()
        self._next_id = nid
        if _columnar(e.cells) is not None:
            e.cells.set_columns(cols)
        else:
            cls = self.schema.cls
            objs = []
            pos = cols["pos"]
            others = [(name, t) for name, t in self.schema.fields if name != "pos"]
            for i in range(len(pos)):
                o = cls.__new__(cls)
                object.__setattr__(o, "pos", T.vec3(*pos[i]))
                for name, t in others:
                    v = cols[name][i]
                    if t == "vec3":
                        v = T.vec3(*v)
                    elif t == "ivec3":
                        v = T.ivec3(*v)
                    elif t == F32:
                        v = float(v)
                    elif t == "bool":
                        v = bool(v)
                    else:
                        v = int(v)
                    object.__setattr__(o, name, v)
                o._id = int(cols["_id"][i])
                o._alive = True
                o._agent_list = e.cells
                objs.append(o)
            e.cells._data = objs
        if e.tissue is not None:
            for f, (name, t, _) in enumerate(self.tissue.fields):
                _set_grid_field(e.tissue, name, self.sim.download_field(f))
        grid = getattr(e, "chemistry", None)
        if grid is not None and self.chem is not None and self.chem.cls is not None:
            for k, (name, t) in enumerate(self.chem.fields):
                v = self.sim.download_chem(k)
                _set_grid_field(grid, name, v.astype(bool) if t == BOOL else v)

    def close(self) -> None:
        if self.sim is not None:
            self.sim.close()
            self.sim = None
        if self.module is not None:
            self.module.close()
            self.module = None
