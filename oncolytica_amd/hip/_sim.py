"""Python handles over the C ABI: Sim (onc_sim_*), Batch (onc_batch_*) and
Module (onc_module_*). Cell state is a list of 32-bit SoA columns:
pos.x, pos.y, pos.z (f32), id (u32), then the model's own columns."""
from __future__ import annotations

import ctypes as C
from typing import Optional, Sequence

import numpy as np

from . import _abi
from ._abi import SimDesc, call, ptr

_DT = {_abi.COL_F32: np.float32, _abi.COL_I32: np.int32, _abi.COL_U32: np.uint32}


class Module:
    """A hiprtc-compiled rule module (replaces WGSL create_shader_module,
    framework/src/oncolytica/gpu/runtime/_device_manager.py:81)."""

    def __init__(self, ctx: _abi.Context, src: str, opts: Sequence[str] = ()):
        self.ctx = ctx
        self.src = src
        h = C.c_void_p()
        arr = (C.c_char_p * max(1, len(opts)))(*[o.encode() for o in opts])
        call("onc_module_build", ctx.h, src.encode(), arr, len(opts), C.byref(h))
        self.h = h

    def log(self) -> str:
        buf = C.create_string_buffer(1 << 16)
        call("onc_module_log", self.h, buf, len(buf))
        return buf.value.decode(errors="replace")

    def close(self) -> None:
        if self.h:
            _abi.lib().onc_module_destroy(self.h)
            self.h = None


class Sim:
    """One simulation resident in device memory."""

    def __init__(self, ctx: _abi.Context, desc: SimDesc):
        self.ctx = ctx
        self.desc = desc
        h = C.c_void_p()
        call("onc_sim_create", ctx.h, C.byref(desc), C.byref(h))
        self.h = h
        self.col_types = [desc.col_types[i] for i in range(desc.n_cols)]
        self.chem_types = [desc.chem_col_types[i] for i in range(desc.n_chem_cols)]
        self.dims = tuple(desc.dims)
        self.module: Optional[Module] = None

    @property
    def n_cols(self) -> int:
        return len(self.col_types)

    @property
    def nvox(self) -> int:
        return int(self.dims[0]) * int(self.dims[1]) * int(self.dims[2])

    def set_rule(self, module: Optional[Module], name: Optional[str] = None) -> None:
        call("onc_sim_set_rule", self.h, module.h if module else None, name.encode() if name else None)
        self.module = module

    def set_params(self, params: Sequence[float]) -> None:
        """Rule parameters for the next steps. The host descriptor follows the
        device copy (as onc_sim_set_params updates it), so a bundle exported
        afterwards holds the parameters the sim runs with."""
        arr = (C.c_float * len(params))(*params)
        call("onc_sim_set_params", self.h, arr, len(params))
        for k, v in enumerate(arr):
            self.desc.params[k] = v
        self.desc.n_params = max(int(self.desc.n_params), len(params))

    def upload_cells(self, cols: Sequence[np.ndarray], next_id: Optional[int] = None) -> None:
        if len(cols) != self.n_cols:
            raise ValueError(f"expected {self.n_cols} columns, got {len(cols)}")
        n = len(cols[0]) if cols else 0
        arrs = []
        for t, c in zip(self.col_types, cols):
            a = np.ascontiguousarray(c, dtype=_DT[t])
            if len(a) != n:
                raise ValueError("columns differ in length")
            arrs.append(a)
        if next_id is None:
            next_id = int(arrs[3].max()) + 1 if n else 0
        pa = (C.c_void_p * self.n_cols)(*[ptr(a) for a in arrs])
        call("onc_sim_upload_cells", self.h, pa, n, next_id)

    def counts(self) -> tuple[int, int, int]:
        n, nid, st = C.c_uint32(), C.c_uint32(), C.c_uint32()
        call("onc_sim_counts", self.h, C.byref(n), C.byref(nid), C.byref(st))
        return int(n.value), int(nid.value), int(st.value)

    def download_cells(self) -> list[np.ndarray]:
        n = self.counts()[0]
        arrs = [np.empty(n, dtype=_DT[t]) for t in self.col_types]
        pa = (C.c_void_p * self.n_cols)(*[ptr(a) for a in arrs])
        got = C.c_uint32()
        call("onc_sim_download_cells", self.h, pa, n, C.byref(got))
        assert got.value == n
        return arrs

    def upload_field(self, f: int, a: np.ndarray) -> None:
        a = np.ascontiguousarray(a, dtype=np.float32).reshape(-1)
        if a.size != self.nvox:
            raise ValueError("field size mismatch")
        call("onc_sim_upload_field", self.h, f, ptr(a))

    def download_field(self, f: int) -> np.ndarray:
        out = np.empty(self.nvox, dtype=np.float32)
        call("onc_sim_download_field", self.h, f, ptr(out))
        return out

    @property
    def chem_dims(self) -> tuple[int, int, int]:
        """Voxels per axis of the coarse chemistry grid: max(1, dims / 2)."""
        return tuple(max(1, int(d) // 2) for d in self.dims)

    @property
    def nchem(self) -> int:
        cx, cy, cz = self.chem_dims
        return cx * cy * cz

    def upload_chem(self, col: int, a: np.ndarray) -> None:
        a = np.ascontiguousarray(a, dtype=_DT[self.chem_types[col]]).reshape(-1)
        if a.size != self.nchem:
            raise ValueError("chemistry column size mismatch")
        call("onc_sim_upload_chem", self.h, col, ptr(a))

    def download_chem(self, col: int) -> np.ndarray:
        out = np.empty(self.nchem, dtype=_DT[self.chem_types[col]])
        call("onc_sim_download_chem", self.h, col, ptr(out))
        return out

    def set_stop(self, stop: int) -> None:
        """Advance only while completed steps < stop (later batch steps leave
        the sim untouched)."""
        call("onc_sim_set_stop", self.h, int(stop) & 0xFFFFFFFF)

    def step(self, n: int = 1) -> None:
        call("onc_sim_step", self.h, n)

    def sync(self) -> None:
        call("onc_sim_sync", self.h)

    def set_step(self, step: int) -> None:
        """Set the completed-step counter (restoring from a bundle)."""
        call("onc_sim_set_step", self.h, int(step))

    def metrics(self, module: "Module", nvals: int) -> np.ndarray:
        """The metric words (uint32 bits) of the module's metric kernels over
        the committed cells."""
        out = np.zeros(nvals, dtype=np.uint32)
        call("onc_sim_metrics", self.h, module.h, nvals, out.ctypes.data_as(C.POINTER(C.c_uint32)))
        return out

    def close(self) -> None:
        if self.h:
            _abi.lib().onc_sim_destroy(self.h)
            self.h = None


class Batch:
    """Many sims stepped together: one launch per phase over a packed segment
    table, optionally replaying one captured hipGraph per timestep."""

    def __init__(self, ctx: _abi.Context):
        self.ctx = ctx
        h = C.c_void_p()
        call("onc_batch_create", ctx.h, C.byref(h))
        self.h = h
        self.sims: list[Sim] = []

    def add(self, sim: Sim) -> None:
        call("onc_batch_add", self.h, sim.h)
        self.sims.append(sim)

    def remove(self, sim: Sim) -> None:
        call("onc_batch_remove", self.h, sim.h)
        self.sims.remove(sim)

    def set_tuning(self, items: int, citems: int, bx: int, by: int, bz: int) -> None:
        call("onc_batch_set_tuning", self.h, items, citems, bx, by, bz)

    def set_tile_policy(self, big_from: int, small_below: int) -> None:
        """Tile sizes by population: the tuned sizes below `big_from` live
        cells, 16 / 8 cells per thread from there until the population falls
        below `small_below` (big_from = 0: the tuned sizes always)."""
        call("onc_batch_set_tile_policy", self.h, int(big_from), int(small_below))

    def tile_items(self) -> tuple[int, int]:
        """Cells per thread of the sort and compaction tiles of the next steps."""
        a, c = C.c_int(), C.c_int()
        call("onc_batch_tile_items", self.h, C.byref(a), C.byref(c))
        return a.value, c.value

    def set_fused(self, fused: bool) -> None:
        """Forces and the 7-point sweep in one launch (True, the default) or
        in two (False); results are identical."""
        call("onc_batch_set_fused", self.h, 1 if fused else 0)

    def set_fused_rule(self, fused: bool) -> None:
        """The built-in tumour rule inside the compaction kernel (True, the
        default, where the schema allows) or as its own launch (False);
        results are identical."""
        call("onc_batch_set_fused_rule", self.h, 1 if fused else 0)

    def set_forces_waves(self, waves: int) -> None:
        """Cap on the one-wave blocks of the force walk (0: one per 64 cells
        of capacity; else a positive multiple of 8); results are identical."""
        call("onc_batch_set_forces_waves", self.h, waves)

    def step(self, n: int = 1, graph: bool = True) -> None:
        call("onc_batch_step", self.h, n, 1 if graph else 0)

    def sync(self) -> None:
        call("onc_batch_sync", self.h)

    def profile(self, enable: bool = True, phases=None, every: int = 1) -> None:
        """Time phases of eager steps with HIP events: all of them, or only
        the named ones (each event costs GPU time of its own), on every
        `every`-th step."""
        if not enable:
            mask = 0
        elif phases is None:
            mask = (1 << _abi.PHASES) - 1
        else:
            mask = 0
            for p in phases:
                mask |= 1 << _abi.PHASE_NAMES.index(p)
        call("onc_batch_profile_every", self.h, mask, int(every))

    def phase_times(self) -> dict[str, tuple[float, int]]:
        us = (C.c_double * _abi.PHASES)()
        nl = (C.c_uint64 * _abi.PHASES)()
        call("onc_batch_phase_times", self.h, us, nl, _abi.PHASES)
        return {_abi.PHASE_NAMES[i]: (us[i], int(nl[i])) for i in range(_abi.PHASES)}

    def sweep_voxels(self) -> tuple[int, int, int]:
        """(swept, filled, kept) field voxels of the last step's sweep: read
        and written, written with their tile's one value, left as they were."""
        v = [C.c_uint64() for _ in range(3)]
        call("onc_batch_sweep_voxels", self.h, *[C.byref(x) for x in v])
        return tuple(int(x.value) for x in v)

    def close(self) -> None:
        if self.h:
            _abi.lib().onc_batch_destroy(self.h)
            self.h = None
            self.sims.clear()
