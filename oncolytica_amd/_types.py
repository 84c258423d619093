"""DSL data types of the rule API: scalar aliases, vec3 / ivec3, the data
layout bases (Tissue, Chemistry, Cell, Metrics, Params), Simulation, the RNG
entry points, and the declarative diffusing-field description.

Mirrors the surface of framework/src/oncolytica/core/utils/_types.py (same
names, constructor keywords, generic specialisation and validation errors),
so user models written for the reference import unchanged.
"""
from __future__ import annotations

import builtins as _builtins
import copy as _copy
import dataclasses
import math as _math
import random as _random
import sys as _sys
from dataclasses import dataclass
from types import SimpleNamespace
from typing import Annotated, Any, Generic, NewType, TypeVar, get_args, get_origin, get_type_hints

S = TypeVar("S")
T = TypeVar("T")
C = TypeVar("C")
A = TypeVar("A")
M = TypeVar("M")
P = TypeVar("P")
U = TypeVar("U")

i32 = NewType("i32", int)
u32 = NewType("u32", int)
f32 = NewType("f32", float)
bool = _builtins.bool  # noqa: A001 - the DSL spells it ol.bool

FLOAT_TYPES = {_builtins.float, f32}
INT_TYPES = {_builtins.int, i32, u32}
BOOL_TYPES = {_builtins.bool}


class vec3:
    """Three binary32-intended components (Python floats on the host)."""

    __slots__ = ("x", "y", "z")

    def __init__(self, x: float = 0.0, y: float = 0.0, z: float = 0.0) -> None:
        object.__setattr__(self, "x", float(x))
        object.__setattr__(self, "y", float(y))
        object.__setattr__(self, "z", float(z))

    def __setattr__(self, name, value):
        if name not in self.__slots__:
            raise AttributeError(f"'vec3' object has no attribute '{name}'")
        object.__setattr__(self, name, float(value))

    def __class_getitem__(cls, _item):
        return cls

    def copy(self) -> "vec3":
        return vec3(self.x, self.y, self.z)

    def _bin(self, o, op):
        if isinstance(o, vec3):
            return vec3(op(self.x, o.x), op(self.y, o.y), op(self.z, o.z))
        if isinstance(o, (int, float)) and not isinstance(o, _builtins.bool):
            return vec3(op(self.x, o), op(self.y, o), op(self.z, o))
        return NotImplemented

    def __add__(self, o):
        return self._bin(o, lambda a, b: a + b) if isinstance(o, vec3) else NotImplemented

    __radd__ = __add__

    def __sub__(self, o):
        return self._bin(o, lambda a, b: a - b) if isinstance(o, vec3) else NotImplemented

    def __mul__(self, o):
        return self._bin(o, lambda a, b: a * b)

    def __rmul__(self, o):
        return self._bin(o, lambda a, b: b * a) if not isinstance(o, vec3) else NotImplemented

    def __truediv__(self, o):
        return self._bin(o, lambda a, b: a / b)

    def __neg__(self):
        return vec3(-self.x, -self.y, -self.z)

    def __eq__(self, o):
        return isinstance(o, vec3) and (self.x, self.y, self.z) == (o.x, o.y, o.z)

    def __iter__(self):
        return iter((self.x, self.y, self.z))

    def __repr__(self):
        return f"vec3({self.x:.4f}f, {self.y:.4f}f, {self.z:.4f}f)"


class ivec3:
    """Three integer components."""

    __slots__ = ("x", "y", "z")

    def __init__(self, x: int = 0, y: int = 0, z: int = 0) -> None:
        object.__setattr__(self, "x", int(x))
        object.__setattr__(self, "y", int(y))
        object.__setattr__(self, "z", int(z))

    def __setattr__(self, name, value):
        if name not in self.__slots__:
            raise AttributeError(f"'ivec3' object has no attribute '{name}'")
        object.__setattr__(self, name, int(value))

    def __class_getitem__(cls, _item):
        return cls

    def copy(self) -> "ivec3":
        return ivec3(self.x, self.y, self.z)

    def __add__(self, o):
        return ivec3(self.x + o.x, self.y + o.y, self.z + o.z) if isinstance(o, ivec3) else NotImplemented

    __radd__ = __add__

    def __sub__(self, o):
        return ivec3(self.x - o.x, self.y - o.y, self.z - o.z) if isinstance(o, ivec3) else NotImplemented

    def __mul__(self, o):
        if isinstance(o, ivec3):
            return ivec3(self.x * o.x, self.y * o.y, self.z * o.z)
        if isinstance(o, (int, float)):
            s = int(o)
            return ivec3(self.x * s, self.y * s, self.z * s)
        return NotImplemented

    __rmul__ = __mul__

    def __neg__(self):
        return ivec3(-self.x, -self.y, -self.z)

    def __eq__(self, o):
        return isinstance(o, ivec3) and (self.x, self.y, self.z) == (o.x, o.y, o.z)

    def __iter__(self):
        return iter((self.x, self.y, self.z))

    def __repr__(self):
        return f"ivec3({self.x:d}, {self.y:d}, {self.z:d})"


VEC_TYPES = {vec3, ivec3}
PRIMITIVE_TYPES = FLOAT_TYPES | INT_TYPES | BOOL_TYPES | VEC_TYPES
ZERO = {float: 0.0, f32: 0.0, int: 0, i32: 0, u32: 0, _builtins.bool: False}


def _zero(t):
    if t is vec3:
        return vec3()
    if t is ivec3:
        return ivec3()
    return ZERO.get(t)


# ---- declarative diffusing fields ---------------------------------------------

@dataclass(frozen=True)
class Boundary:
    kind: str           # "dirichlet" | "point_source"
    value: float = 0.0
    radius: float = 0.0


def dirichlet(value: float) -> Boundary:
    """Every boundary-layer voxel is held at `value`."""
    return Boundary("dirichlet", float(value))


def point_source(value: float, radius: float) -> Boundary:
    """Zero-flux faces; the voxels of face z=0 within `radius` voxels of the
    face centre are held at `value`."""
    return Boundary("point_source", float(value), float(radius))


@dataclass(frozen=True)
class Diffusing:
    """Default value of a Tissue field that makes it a diffusing substrate
    (one explicit 7-point sweep per step): rate = D dt / dx^2."""

    rate: float
    init: float = 0.0
    decay: float = 0.0
    boundary: Boundary = Boundary("dirichlet", 0.0)
    uptake: float = 0.0
    lo: float = 0.0
    hi: float = 1.0


def diffusing(rate: float, init: float = 0.0, decay: float = 0.0, boundary: Boundary | None = None,
              uptake: float = 0.0, lo: float = 0.0, hi: float = 1.0) -> Diffusing:
    """Declare a Tissue field as a diffusing substrate. `uptake` is removed
    from a voxel per resident cell per step (the reaction term of
    engine/assets/shaders/tumor_sim.wgsl:387-412, here as a declaration so
    rules keep mutating only their own cell)."""
    return Diffusing(float(rate), float(init), float(decay), boundary or dirichlet(init), float(uptake),
                     float(lo), float(hi))


# ---- simulation ------------------------------------------------------------------

def _resolve_own_hints(cls) -> dict:
    own = cls.__dict__.get("__annotations__", {})
    mod = _sys.modules.get(cls.__module__)
    g = vars(mod) if mod is not None else {}
    out = {}
    for k, v in own.items():
        if isinstance(v, str):
            try:
                v = eval(v, g)  # noqa: S307 - annotations of user classes
            except Exception:
                pass
        out[k] = v
    return out


class Simulation(Generic[T, C, A, M, P]):
    """Base of user models: class MySim(ol.Simulation[Tissue, Chem, Cell, Metrics, Params])."""

    _engine: Any = None

    def __init__(self):
        self._params = self._spec.params_class()
        self._metrics = self._spec.metrics_class(self)
        self._params._validate()

    def __init_subclass__(cls, **kwargs):
        super().__init_subclass__(**kwargs)
        gb = next((b for b in getattr(cls, "__orig_bases__", []) if get_origin(b) is Simulation), None)
        if gb is None:
            raise RuntimeError(
                "Explicit type specialization is required for Simulation class.\n"
                "Usage: class MySimulation(ol.Simulation[MyTissue, MyChem, MyCell, MyMetrics, MyParams]):")
        args = get_args(gb)
        cls._spec = SimpleNamespace(tissue_class=args[0], chemistry_class=args[1], cell_class=args[2],
                                    metrics_class=args[3], params_class=args[4])
        for i, klass in enumerate(args):
            base = BASE_CLASSES[i]
            if isinstance(klass, str) or not isinstance(klass, type):
                raise TypeError(f"Parameter {i + 1} of ol.Simulation is a forward reference "
                                f"(got string '{klass}' instead of a class).")
            if not issubclass(klass, base):
                raise TypeError(f"Validation failed for parameter {i + 1} in ol.Simulation.\n"
                                f"The type '{klass.__name__}' must be a subclass of 'ol.{base.__name__}'.\n")
            if sum(issubclass(klass, b) for b in BASE_CLASSES) > 1:
                raise TypeError(f"Validation failed for parameter {i + 1} in ol.Simulation.\n"
                                f"The type '{klass.__name__}' must be a subclass of exactly one base class: "
                                f"'ol.{base.__name__}'.\n")
            if base is Cell:
                hints = get_type_hints(klass)
                if hints.get("pos") is not vec3:
                    raise TypeError(f"Class '{klass.__name__}' must have attribute 'pos: ol.vec3'.")

    @property
    def params(self):
        return self._params

    @property
    def engine(self):
        return self._engine

    def tissue_at(self, pos: vec3):
        """Tissue record of the voxel containing pos (clamped to the grid);
        index = floor(pos / tissue_voxel_size) (core/utils/_types.py:314-330)."""
        e = self._engine
        if e is None or e.tissue is None:
            raise RuntimeError("Simulation.tissue_at() called but engine/tissue is not initialized.")
        vs, sh = e.tissue_voxel_size, e.tissue_shape
        ix = max(0, min(int(pos.x / vs), sh[0] - 1))
        iy = max(0, min(int(pos.y / vs), sh[1] - 1))
        iz = max(0, min(int(pos.z / vs), sh[2] - 1))
        return e.tissue.sample(ix, iy, iz)

    def chemistry_at(self, pos: vec3):
        e = self._engine
        if e is None or e.chemistry is None:
            raise RuntimeError("Simulation.chemistry_at() called but engine/chemistry is not initialized.")
        vs, sh = e.chemistry_voxel_size, e.chemistry_shape
        ix = max(0, min(int(pos.x / vs), sh[0] - 1))
        iy = max(0, min(int(pos.y / vs), sh[1] - 1))
        iz = max(0, min(int(pos.z / vs), sh[2] - 1))
        return e.chemistry.sample(ix, iy, iz)

    def __setattr__(self, name, value):
        super().__setattr__(name, value)
        if name == "params":
            if not isinstance(value, self._spec.params_class):
                raise TypeError(f"Invalid type for 'params'.\nExpected an instance of "
                                f"'{self._spec.params_class.__name__}', but got '{type(value).__name__}'.")
            value._validate()

    def __repr__(self):
        return f"{type(self).__name__}()"


# ---- data layouts ---------------------------------------------------------------

class BaseData(Generic[U]):
    """Base of every data layout: annotated fields with defaults."""

    def __init_subclass__(cls, **kwargs):
        super().__init_subclass__(**kwargs)
        hints = get_type_hints(cls, include_extras=False)
        defaults = {}
        for n in hints:
            if hasattr(cls, n):
                v = getattr(cls, n)
                defaults[n] = v.init if isinstance(v, Diffusing) else v

        def __init__(self, **kw):
            for name, dtype in hints.items():
                if name.startswith("_"):
                    continue
                if name in kw:
                    val = kw[name]
                elif name in defaults:
                    val = defaults[name]
                    if isinstance(val, (vec3, ivec3)):
                        val = val.copy()
                else:
                    val = _zero(dtype)
                object.__setattr__(self, name, val)
            for k in kw:
                if k not in hints:
                    raise TypeError(f"{type(self).__name__} has no field '{k}'")

        cls.__init__ = __init__

    def copy(self):
        new = _copy.copy(self)
        for k, v in list(new.__dict__.items()):
            if isinstance(v, (vec3, ivec3)):
                new.__dict__[k] = v.copy()
        return new

    def copy_from(self, other) -> None:
        for name in get_type_hints(type(self)):
            if not name.startswith("_"):
                v = getattr(other, name)
                setattr(self, name, v.copy() if isinstance(v, (vec3, ivec3)) else v)


class Tissue(BaseData, Generic[S, T, C, A, M, P]):
    """Per-voxel data of the tissue grid (voxel = bin = tissue_voxel_size).
    A field whose default is ol.diffusing(...) is a diffusing substrate."""


class Chemistry(BaseData, Generic[S, T, C, A, M, P]):
    """Per-voxel data of the coarse (2x) chemistry grid."""


class Cell(BaseData, Generic[S, T, C, A, M, P]):
    """A cell agent; must declare pos: ol.vec3."""

    _alive: Annotated[_builtins.bool, "cpu_only"] = True

    def divide(self, cell) -> None:
        lst = getattr(self, "_agent_list", None)
        if lst is None:
            raise RuntimeError("divide() outside a simulation")
        lst.add(cell)

    def die(self) -> None:
        self._alive = False

    @property
    def neighbors(self):
        raise RuntimeError("cell.neighbors is evaluated inside compiled rules only")


class Metrics(Generic[S, T, C, A, M, P]):
    """Per-step reductions filled by @ol.metric_rule."""

    def __init__(self, sim=None):
        self._sim = sim
        for name, t in _resolve_own_hints(type(self)).items():
            if not name.startswith("_") and not hasattr(self, name):
                object.__setattr__(self, name, _zero(t))

    def __init_subclass__(cls, **kwargs):
        super().__init_subclass__(**kwargs)
        for name, t in _resolve_own_hints(cls).items():
            if name.startswith("_") or name in cls.__dict__:
                continue
            z = _zero(t)
            if z is not None:
                setattr(cls, name, z)

    def _fields(self):
        out = []
        for klass in reversed(type(self).__mro__):
            for name, t in _resolve_own_hints(klass).items():
                if not name.startswith("_") and name not in [n for n, _ in out]:
                    out.append((name, t))
        return out

    def _clear(self) -> None:
        for name, t in self._fields():
            z = _zero(t)
            if z is not None:
                object.__setattr__(self, name, z)

    @property
    def sim(self):
        return self._sim

    def __repr__(self):
        return f"{type(self).__name__}(" + ", ".join(f"{n}={getattr(self, n)!r}" for n, _ in self._fields()) + ")"


@dataclass(frozen=True, init=False)
class Params:
    """Frozen simulation parameters (packed sorted by name for the device)."""

    def __init__(self, **kw):
        hints = get_type_hints(type(self))
        for k in kw:
            if k not in hints:
                raise TypeError(f"Missing field '{k}' in {type(self).__name__}")
        for name, ann in hints.items():
            if name in kw:
                v = kw[name]
            elif hasattr(type(self), name):
                v = getattr(type(self), name)
            elif _zero(ann) is not None:
                v = _zero(ann)
            else:
                raise TypeError(f"Missing value for field '{name}' in {type(self).__name__}")
            object.__setattr__(self, name, v)
        self._validate()

    def __init_subclass__(cls, **kwargs):
        super().__init_subclass__(**kwargs)
        dataclass(frozen=True, init=False)(cls)
        bad = {k for k, v in cls.__dict__.items()
               if not k.startswith("_") and not callable(v) and k not in getattr(cls, "__annotations__", {})}
        if bad:
            raise TypeError(f"All fields in {cls.__name__} must be annotated\nUnannotated: {', '.join(sorted(bad))}")
        cls()._validate()

    def _validate_field(self, name, val) -> None:
        hints = get_type_hints(type(self))
        if name not in hints:
            raise TypeError(f"Missing field '{name}' in {type(self).__name__}")
        t = hints[name]
        if (t in FLOAT_TYPES and not isinstance(val, float)
                or t in INT_TYPES and (not isinstance(val, int) or isinstance(val, _builtins.bool))
                or t in BOOL_TYPES and not isinstance(val, _builtins.bool)
                or t in VEC_TYPES and not isinstance(val, t)):
            raise TypeError(f"Type mismatch for '{name}' in {type(self).__name__}: "
                            f"expected {t.__name__}, got {type(val).__name__} ({val!r})")

    def _validate(self) -> None:
        for name, t in get_type_hints(type(self)).items():
            if t not in PRIMITIVE_TYPES:
                raise TypeError(f"Unsupported type '{t}' for '{name}' in {type(self).__name__}")
            if hasattr(self, name):
                self._validate_field(name, getattr(self, name))

    def _sorted_items(self):
        return sorted(((n, getattr(self, n)) for n in get_type_hints(type(self))), key=lambda kv: kv[0])


# ---- RNG (host-side draws; compiled rules use Philox4x32-10) --------------------

def random() -> float:
    return _random.random()


def random_dir() -> vec3:
    z = random() * 2.0 - 1.0
    phi = random() * _math.tau
    r = _math.sqrt(max(0.0, 1.0 - z * z))
    return vec3(r * _math.cos(phi), r * _math.sin(phi), z)


BASE_CLASSES = [Tissue, Chemistry, Cell, Metrics, Params]
