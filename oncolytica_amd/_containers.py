"""Data containers of the rule API.

Grid: a 3-D grid of data records (tissue voxels), stored as one numpy array
per field in the device layout (x fastest), with record-style access
grid[x, y, z].field, grid.sample(x, y, z) and iteration as in
framework/src/oncolytica/core/utils/_containers.py:17-82, which held one
Python object per voxel.

AgentList: the cell agents; Python objects as in the reference
(_containers.py:85-123) plus a columnar fast path for large populations.
"""
from __future__ import annotations

import typing
from typing import Any, Iterator, Optional

import numpy as np

from . import _types as T


def _np_dtype(t) -> np.dtype:
    if t in T.FLOAT_TYPES:
        return np.dtype(np.float32)
    if t is T.u32:
        return np.dtype(np.uint32)
    if t in T.INT_TYPES:
        return np.dtype(np.int32)
    if t in T.BOOL_TYPES:
        return np.dtype(np.bool_)
    raise TypeError(f"grid fields must be scalars, got {t!r}")


def _grid_fields(cls) -> list[tuple[str, Any]]:
    hints = typing.get_type_hints(cls)
    out = []
    for klass in reversed(cls.__mro__):
        for name in klass.__dict__.get("__annotations__", {}):
            if not name.startswith("_") and name not in [n for n, _ in out]:
                out.append((name, hints[name]))
    return out


def fixed_mask(shape, boundary) -> np.ndarray:
    """Voxels a diffusing field holds at its boundary value, flattened x
    fastest: every boundary-layer voxel (Dirichlet), or the voxels of face
    z=0 whose centre lies within the source radius of the face centre."""
    X, Y, Z = shape
    m = np.zeros((Z, Y, X), dtype=bool)
    if boundary.kind == "dirichlet":
        m[0, :, :] = m[-1, :, :] = True
        m[:, 0, :] = m[:, -1, :] = True
        m[:, :, 0] = m[:, :, -1] = True
    else:
        fx = (np.arange(X, dtype=np.float32) + np.float32(0.5)) - np.float32(X) * np.float32(0.5)
        fy = (np.arange(Y, dtype=np.float32) + np.float32(0.5)) - np.float32(Y) * np.float32(0.5)
        r = np.float32(boundary.radius)
        m[0] = (fx[None, :] * fx[None, :] + fy[:, None] * fy[:, None]) <= r * r
    return m.reshape(-1)


class _RecordView:
    """A voxel of a Grid seen as a record of its class (attribute access goes
    to the grid's arrays)."""

    __slots__ = ("_grid", "_index", "_coord")

    def __getattr__(self, name):
        g = object.__getattribute__(self, "_grid")
        if name in g._arrays:
            v = g._arrays[name][object.__getattribute__(self, "_index")]
            return v.item()
        return getattr(g._struct_class, name)

    def __setattr__(self, name, value):
        if name in _RecordView.__slots__:
            object.__setattr__(self, name, value)
            return
        g = object.__getattribute__(self, "_grid")
        if name not in g._arrays:
            raise AttributeError(f"{g._struct_class.__name__} has no field '{name}'")
        g._arrays[name][object.__getattribute__(self, "_index")] = value

    @property
    def coord(self) -> T.ivec3:
        return T.ivec3(*object.__getattribute__(self, "_coord"))

    def __repr__(self):
        g = object.__getattribute__(self, "_grid")
        vals = ", ".join(f"{n}={getattr(self, n)!r}" for n in g._arrays)
        return f"{g._struct_class.__name__}@{tuple(object.__getattribute__(self, '_coord'))}({vals})"


class Grid:
    """3-D grid of `struct_class` records, one numpy array per field."""

    def __init__(self, struct_class, shape: tuple, engine: Any = None) -> None:
        self._struct_class = struct_class
        self._engine = engine
        self._shape = tuple(int(s) for s in shape)
        n = int(np.prod(self._shape))
        self._arrays: dict[str, np.ndarray] = {}
        self._fields = _grid_fields(struct_class)
        for name, t in self._fields:
            d = getattr(struct_class, name, None)
            a = np.zeros(n, dtype=_np_dtype(t))
            if isinstance(d, T.Diffusing):
                a[:] = d.init
                a[fixed_mask(self._shape, d.boundary)] = d.boundary.value
            elif d is not None:
                a[:] = d
            self._arrays[name] = a

    @property
    def shape(self) -> tuple:
        return self._shape

    def _index(self, x, y, z) -> int:
        X, Y, Z = self._shape
        if not (0 <= x < X and 0 <= y < Y and 0 <= z < Z):
            raise IndexError(f"voxel {(x, y, z)} outside grid {self._shape}")
        return x + X * (y + Y * z)

    def _view(self, x, y, z) -> _RecordView:
        v = object.__new__(_RecordView)
        object.__setattr__(v, "_grid", self)
        object.__setattr__(v, "_index", self._index(x, y, z))
        object.__setattr__(v, "_coord", (x, y, z))
        return v

    def __getitem__(self, idx):
        x, y, z = idx
        return self._view(int(x), int(y), int(z))

    def __setitem__(self, idx, value) -> None:
        view = self[idx]
        for name, _ in self._fields:
            setattr(view, name, getattr(value, name))

    def sample(self, x: int, y: int, z: int):
        return self._view(x, y, z)

    def __iter__(self) -> Iterator:
        X, Y, Z = self._shape
        for x in range(X):
            for y in range(Y):
                for z in range(Z):
                    yield self._view(x, y, z)

    def __len__(self) -> int:
        return int(np.prod(self._shape))

    def array(self, name: str) -> np.ndarray:
        """The field as a [x, y, z]-indexed view of its storage."""
        X, Y, Z = self._shape
        return self._arrays[name].reshape(Z, Y, X).transpose(2, 1, 0)

    def flat(self, name: str) -> np.ndarray:
        """The field's storage, x fastest (the device layout)."""
        return self._arrays[name]

    def __repr__(self):
        return f"Grid(struct={self._struct_class.__name__!r}, shape={self._shape})"


class AgentList:
    """The cell agents of an Engine.

    Cells are Python objects (`_data`) as in the reference; `set_columns`
    installs a columnar population instead (numpy arrays keyed by field name,
    'pos' as an (n, 3) array), which the HIP backend uploads without building
    objects. `columns()` returns the last downloaded columns."""

    def __init__(self, engine: Any) -> None:
        self._engine = engine
        self._data: list = []
        self._columns: Optional[dict] = None

    def __iter__(self):
        return iter(self._data)

    def __len__(self) -> int:
        if self._columns is not None and not self._data:
            return int(len(self._columns["pos"]))
        return len(self._data)

    def add(self, agent) -> None:
        if self._columns is not None:
            raise RuntimeError("this AgentList holds columns; use set_columns() or clear() first")
        agent._alive = True
        agent._agent_list = self
        self._data.append(agent)

    def set_columns(self, columns: dict) -> None:
        pos = np.asarray(columns["pos"], dtype=np.float32)
        if pos.ndim != 2 or pos.shape[1] != 3:
            raise ValueError("columns['pos'] must have shape (n, 3)")
        self._data = []
        self._columns = {k: np.asarray(v) for k, v in columns.items()}

    def columns(self) -> Optional[dict]:
        return self._columns

    def clear(self) -> None:
        self._data = []
        self._columns = None

    def _remove_dead(self) -> int:
        before = len(self._data)
        self._data = [a for a in self._data if getattr(a, "_alive", True)]
        return before - len(self._data)
