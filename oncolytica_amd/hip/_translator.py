"""Python-AST -> HIP C++ translator for @ol.cell_rule, @ol.tissue_rule,
@ol.chemistry_rule and @ol.metric_rule methods.

Replaces the reference's AST -> WGSL translator
(framework/src/oncolytica/gpu/compiler/_compiler.py:233-376,
gpu/compiler/translator/_stmt_translator.py, _expr_translator.py) with one that
emits HIP C++ for hiprtc. The emitted kernel:

  * loads the cell's state from the permuted buffer (`var cell = Cells_In[i]`,
    gpu/compiler/_shader_builder.py:207-221), with its integrated position;
  * runs every cell rule in declaration order on that cell (each sees the
    previous rule's writes, as the CPU backend does, cpu/_cpu_backend.py:46-49);
  * writes the cell back, its alive / divide flags and its daughter.

Semantics (deterministic spec):
  * cell.die() marks the cell; it is removed by the stable compaction
    (core/utils/_types.py:578-579 with _containers.py:114-117).
  * cell.divide(d) records one daughter per cell per step (the last call
    wins); daughters are appended after the survivors in parent order.
  * ol.random() / ol.random_dir() draw from Philox4x32-10 keyed by the seed,
    counter (cell id, step, call index) instead of xorshift32
    (gpu/compiler/template.wgsl:75-98).
  * for nb in cell.neighbors visits the cells of the 27 Moore voxels in
    template.wgsl:11-21 order (x fastest), out-of-range voxels skipped,
    the cell itself skipped (_stmt_translator.py:459-491), reading the state
    as it was before the rules ran.
  * self.tissue_at(pos) reads the tissue fields of the voxel of pos after this
    step's diffusion sweep (core/utils/_types.py:314-330).
  * Arithmetic follows Python's left-to-right evaluation order in binary32
    with no contraction; int / int is true division, // and % floor.
"""
from __future__ import annotations

import ast
import inspect
import textwrap
import typing
from dataclasses import dataclass, field
from typing import Any, Optional

from .. import _types as T
from .._errors import CompilationError

# ---- types ---------------------------------------------------------------------
F32, I32, U32, BOOL, VEC3, IVEC3, VOID = "f32", "i32", "u32", "bool", "vec3", "ivec3", "void"
SCALARS = (F32, I32, U32, BOOL)
CTYPE = {F32: "float", I32: "int", U32: "unsigned", BOOL: "bool", VEC3: "onc_v3", IVEC3: "onc_iv3", VOID: "void"}


@dataclass(frozen=True)
class StructT:
    kind: str   # "cell" | "tissue" | "chem" | "metrics"
    cls: type

    @property
    def cname(self) -> str:
        return f"{self.kind.capitalize()}_{self.cls.__name__}"


@dataclass(frozen=True)
class TupleT:
    elems: tuple

    @property
    def cname(self) -> str:
        return "Tup_" + "_".join(_tname(e) for e in self.elems)


def _tname(t) -> str:
    if isinstance(t, StructT):
        return t.cname
    if isinstance(t, TupleT):
        return t.cname
    return t


def ctype(t) -> str:
    if isinstance(t, (StructT, TupleT)):
        return t.cname
    return CTYPE[t]


def py_to_dsl(ann) -> Any:
    """Map an annotation (class or DSL alias) to a DSL type tag."""
    if ann in T.FLOAT_TYPES:
        return F32
    if ann is T.i32 or ann is int:
        return I32
    if ann is T.u32:
        return U32
    if ann is bool:
        return BOOL
    if ann is T.vec3:
        return VEC3
    if ann is T.ivec3:
        return IVEC3
    if ann is None or ann is type(None):
        return VOID
    if isinstance(ann, type) and issubclass(ann, T.Cell):
        return StructT("cell", ann)
    if isinstance(ann, type) and issubclass(ann, T.Tissue):
        return StructT("tissue", ann)
    if isinstance(ann, type) and issubclass(ann, T.Chemistry):
        return StructT("chem", ann)
    origin = typing.get_origin(ann)
    if origin is tuple:
        return TupleT(tuple(py_to_dsl(a) for a in typing.get_args(ann)))
    raise TypeError(f"unsupported type in rule code: {ann!r}")


def data_fields(cls) -> list[tuple[str, Any]]:
    """Device-visible fields of a data class, in declaration order (base
    classes first); names starting with '_' are host-only."""
    out = []
    seen = set()
    for klass in reversed(cls.__mro__):
        ann = klass.__dict__.get("__annotations__", {})
        if not ann:
            continue
        hints = typing.get_type_hints(klass)
        for name in ann:
            if name.startswith("_") or name in seen:
                continue
            seen.add(name)
            out.append((name, hints[name]))
    return out


# ---- column schema ------------------------------------------------------------------
COL_F32, COL_I32, COL_U32 = 0, 1, 2


@dataclass
class Column:
    name: str        # struct member path: "pos.x", "_id", "health"
    coltype: int


@dataclass
class CellSchema:
    cls: type
    fields: list     # (name, dsl type)
    columns: list    # Column

    @classmethod
    def of(cls_, cell_cls) -> "CellSchema":
        fields = [(n, py_to_dsl(t)) for n, t in data_fields(cell_cls)]
        names = [n for n, _ in fields]
        if "pos" not in names or dict(fields)["pos"] != VEC3:
            raise CompilationError(f"Class '{cell_cls.__name__}' must have attribute 'pos: ol.vec3'.")
        cols = [Column("pos.x", COL_F32), Column("pos.y", COL_F32), Column("pos.z", COL_F32),
                Column("_id", COL_U32)]
        for n, t in fields:
            if n == "pos":
                continue
            if t == F32:
                cols.append(Column(n, COL_F32))
            elif t == I32:
                cols.append(Column(n, COL_I32))
            elif t in (U32, BOOL):
                cols.append(Column(n, COL_U32))
            elif t == VEC3:
                cols += [Column(f"{n}.{a}", COL_F32) for a in "xyz"]
            elif t == IVEC3:
                cols += [Column(f"{n}.{a}", COL_I32) for a in "xyz"]
            else:
                raise CompilationError(f"Cell field '{n}' has unsupported type {t}")
        if len(cols) > 48:
            raise CompilationError("a cell may hold at most 48 scalar columns")
        return cls_(cell_cls, fields, cols)


@dataclass
class TissueSchema:
    cls: Optional[type]
    fields: list      # (name, dsl type, Diffusing | default)

    @classmethod
    def of(cls_, tissue_cls) -> "TissueSchema":
        if tissue_cls is None:
            return cls_(None, [])
        out = []
        for n, t in data_fields(tissue_cls):
            dt = py_to_dsl(t)
            if dt not in SCALARS:
                raise CompilationError(f"Tissue field '{n}': only scalar fields are supported on the HIP backend")
            out.append((n, dt, getattr(tissue_cls, n, None)))
        if len(out) > 4:
            raise CompilationError("a tissue may hold at most 4 fields on the HIP backend")
        return cls_(tissue_cls, out)


@dataclass
class ChemSchema:
    """The Chemistry record of the coarse grid: scalar fields, one 32-bit
    column each (f32, i32, u32; bool as u32)."""
    cls: Optional[type]
    fields: list      # (name, dsl type)

    @classmethod
    def of(cls_, chem_cls) -> "ChemSchema":
        if chem_cls is None:
            return cls_(None, [])
        out = []
        for n, t in data_fields(chem_cls):
            dt = py_to_dsl(t)
            if dt not in SCALARS:
                raise CompilationError(f"Chemistry field '{n}': only scalar fields are supported on the HIP backend")
            out.append((n, dt))
        if len(out) > 16:
            raise CompilationError("a chemistry record may hold at most 16 fields on the HIP backend")
        return cls_(chem_cls, out)

    @property
    def col_types(self) -> list[int]:
        return [COL_F32 if t == F32 else (COL_I32 if t == I32 else COL_U32) for _, t in self.fields]


@dataclass
class ParamSlot:
    name: str
    dsl: Any
    slot: int


def param_slots(params) -> list[ParamSlot]:
    """Params packed sorted by name (gpu/runtime/_device_manager.py:180-204);
    every scalar takes one float slot, a vec3 three."""
    out, k = [], 0
    if params is None:
        return out
    hints = typing.get_type_hints(type(params))
    for name in sorted(hints):
        t = py_to_dsl(hints[name])
        out.append(ParamSlot(name, t, k))
        k += 3 if t in (VEC3, IVEC3) else 1
    if k > 32:
        raise CompilationError("Params may hold at most 32 scalar slots")
    return out


def param_values(params) -> list[float]:
    vals = []
    if params is None:
        return vals
    for ps in param_slots(params):
        v = getattr(params, ps.name)
        if ps.dsl in (VEC3, IVEC3):
            vals += [float(v.x), float(v.y), float(v.z)]
        else:
            vals.append(float(v))
    return vals


# ---- the translator -------------------------------------------------------------

_MATH1 = {  # ol.math name -> device expression template on a float
    "sqrt": "__builtin_sqrtf({0})", "exp": "__expf_ieee({0})", "log": "__logf_ieee({0})",
    "log2": "log2f({0})", "exp2": "exp2f({0})", "sin": "sinf({0})", "cos": "cosf({0})", "tan": "tanf({0})",
    "asin": "asinf({0})", "acos": "acosf({0})", "atan": "atanf({0})", "sinh": "sinhf({0})",
    "cosh": "coshf({0})", "tanh": "tanhf({0})", "floor": "__builtin_floorf({0})",
    "ceil": "__builtin_ceilf({0})", "trunc": "__builtin_truncf({0})", "round": "roundf({0})",
    "fabs": "__builtin_fabsf({0})", "fract": "onc_fract({0})", "sign": "onc_sign({0})",
    "inverse_sqrt": "(1.0f / __builtin_sqrtf({0}))",
}

PRELUDE = r'''
#include "onc_device.h"

struct onc_v3 { float x, y, z; };
struct onc_iv3 { int x, y, z; };
__device__ __forceinline__ onc_v3 onc_v3_make(float x, float y, float z) { onc_v3 r; r.x = x; r.y = y; r.z = z; return r; }
__device__ __forceinline__ onc_iv3 onc_iv3_make(int x, int y, int z) { onc_iv3 r; r.x = x; r.y = y; r.z = z; return r; }
__device__ __forceinline__ onc_v3 operator+(onc_v3 a, onc_v3 b) { return onc_v3_make(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ onc_v3 operator-(onc_v3 a, onc_v3 b) { return onc_v3_make(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ onc_v3 operator*(onc_v3 a, onc_v3 b) { return onc_v3_make(a.x * b.x, a.y * b.y, a.z * b.z); }
__device__ __forceinline__ onc_v3 operator*(onc_v3 a, float s) { return onc_v3_make(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ onc_v3 operator*(float s, onc_v3 a) { return onc_v3_make(s * a.x, s * a.y, s * a.z); }
__device__ __forceinline__ onc_v3 operator/(onc_v3 a, float s) { return onc_v3_make(a.x / s, a.y / s, a.z / s); }
__device__ __forceinline__ onc_v3 operator-(onc_v3 a) { return onc_v3_make(-a.x, -a.y, -a.z); }
__device__ __forceinline__ bool operator==(onc_v3 a, onc_v3 b) { return a.x == b.x && a.y == b.y && a.z == b.z; }
__device__ __forceinline__ bool operator!=(onc_v3 a, onc_v3 b) { return !(a == b); }
__device__ __forceinline__ onc_iv3 operator+(onc_iv3 a, onc_iv3 b) { return onc_iv3_make(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ onc_iv3 operator-(onc_iv3 a, onc_iv3 b) { return onc_iv3_make(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ onc_iv3 operator*(onc_iv3 a, int s) { return onc_iv3_make(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ bool operator==(onc_iv3 a, onc_iv3 b) { return a.x == b.x && a.y == b.y && a.z == b.z; }
__device__ __forceinline__ float onc_dot(onc_v3 a, onc_v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ float onc_length_sq(onc_v3 a) { return onc_dot(a, a); }
__device__ __forceinline__ float onc_length(onc_v3 a) { return __builtin_sqrtf(onc_dot(a, a)); }
__device__ __forceinline__ float onc_distance(onc_v3 a, onc_v3 b) { return onc_length(a - b); }
__device__ __forceinline__ float onc_distance_sq(onc_v3 a, onc_v3 b) { return onc_length_sq(a - b); }
__device__ __forceinline__ onc_v3 onc_normalize(onc_v3 a) { float n = onc_length(a); return n > 0.0f ? a / n : onc_v3_make(0.0f, 0.0f, 0.0f); }
__device__ __forceinline__ onc_v3 onc_cross(onc_v3 a, onc_v3 b) { return onc_v3_make(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
__device__ __forceinline__ onc_v3 onc_reflect(onc_v3 v, onc_v3 nrm) { float d = 2.0f * onc_dot(v, nrm); return v - nrm * d; }
__device__ __forceinline__ float onc_lerp(float a, float b, float t) { return a + (b - a) * t; }
__device__ __forceinline__ onc_v3 onc_lerp_v(onc_v3 a, onc_v3 b, float t) { return a + (b - a) * t; }
__device__ __forceinline__ float onc_clamp_f(float v, float lo, float hi) { return v < lo ? lo : (v > hi ? hi : v); }
__device__ __forceinline__ int onc_clamp_i(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }
__device__ __forceinline__ onc_v3 onc_clamp_v(onc_v3 v, onc_v3 lo, onc_v3 hi) { return onc_v3_make(onc_clamp_f(v.x, lo.x, hi.x), onc_clamp_f(v.y, lo.y, hi.y), onc_clamp_f(v.z, lo.z, hi.z)); }
__device__ __forceinline__ float onc_min_f(float a, float b) { return b < a ? b : a; }
__device__ __forceinline__ float onc_max_f(float a, float b) { return b > a ? b : a; }
__device__ __forceinline__ int onc_min_i(int a, int b) { return b < a ? b : a; }
__device__ __forceinline__ int onc_max_i(int a, int b) { return b > a ? b : a; }
__device__ __forceinline__ float onc_sign(float x) { return x > 0.0f ? 1.0f : (x < 0.0f ? -1.0f : 0.0f); }
__device__ __forceinline__ float onc_step(float e, float x) { return x < e ? 0.0f : 1.0f; }
__device__ __forceinline__ float onc_smoothstep(float e0, float e1, float x) { float t = onc_clamp_f((x - e0) / (e1 - e0), 0.0f, 1.0f); return t * t * (3.0f - 2.0f * t); }
__device__ __forceinline__ float onc_fract(float x) { return x - __builtin_floorf(x); }
__device__ __forceinline__ int onc_floordiv_i(int a, int b) { int q = a / b; return (q * b != a && ((a < 0) != (b < 0))) ? q - 1 : q; }
__device__ __forceinline__ int onc_mod_i(int a, int b) { int m = a % b; return (m != 0 && ((m < 0) != (b < 0))) ? m + b : m; }
__device__ __forceinline__ float onc_floordiv_f(float a, float b) { return __builtin_floorf(a / b); }
__device__ __forceinline__ float onc_mod_f(float a, float b) { float m = fmodf(a, b); return (m != 0.0f && ((m < 0.0f) != (b < 0.0f))) ? m + b : m; }
__device__ __forceinline__ int onc_abs_i(int a) { return a < 0 ? -a : a; }
__device__ __forceinline__ float __expf_ieee(float x) { return expf(x); }
__device__ __forceinline__ float __logf_ieee(float x) { return logf(x); }
__device__ __forceinline__ onc_v3 onc_v3_rand_dir(onc_rng& r) { onc_v3 v; onc_rand_dir(r, &v.x, &v.y, &v.z); return v; }
'''


def _namespace(fn) -> dict:
    """Names a rule body can see: the module globals plus the variables it
    closes over (models declared inside a function)."""
    ns = dict(fn.__globals__)
    if fn.__closure__:
        for name, cell in zip(fn.__code__.co_freevars, fn.__closure__):
            try:
                ns[name] = cell.cell_contents
            except ValueError:
                pass
    return ns


@dataclass
class Local:
    cname: str
    t: Any


class Env:
    def __init__(self, parent: Optional["Env"] = None):
        self.vars: dict[str, Local] = {}
        self.parent = parent

    def get(self, name):
        e = self
        while e is not None:
            if name in e.vars:
                return e.vars[name]
            e = e.parent
        return None


@dataclass
class FuncInfo:
    key: str
    cname: str
    node: ast.FunctionDef
    globals: dict
    self_t: Any          # StructT of the receiver, or "sim"
    ret: Any
    params: list         # (name, type)
    code: str = ""


class Translator:
    """Translates the cell rules of one Simulation subclass into a HIP
    module with one kernel, `onc_rule_main`, and its metric rules (if any)
    into `onc_metric_main` + `onc_metric_final` (onc_sim_metrics)."""

    KERNEL = "onc_rule_main"
    METRIC_KERNELS = ("onc_metric_main", "onc_metric_final")

    TISSUE_KERNEL = "onc_tissue_pass"
    CHEM_KERNEL = "onc_chem_pass"

    def __init__(self, sim, cell_schema: CellSchema, tissue: TissueSchema, params,
                 chem: Optional[ChemSchema] = None):
        self.sim = sim
        self.sim_cls = type(sim)
        self.cell = cell_schema
        self.tissue = tissue
        self.chem = chem if chem is not None else ChemSchema(None, [])
        self.rule_kind = "cell"  # kind of the rule being translated: cell, tissue, chem, metric
        self.tissue_passes = 0
        self.chem_pass_rules: list[str] = []  # rule function of each chemistry pass of a step
        self.pslots = {p.name: p for p in param_slots(params)}
        self.structs: dict[str, StructT] = {}
        self.tuples: dict[str, TupleT] = {}
        self.funcs: dict[str, FuncInfo] = {}
        self.order: list[str] = []
        self.tmp = 0
        self.lines: list[str] = []
        self.indent = 0
        self.cur: Optional[FuncInfo] = None
        self.in_metric = False
        self.metric_fields: list[tuple[str, str]] = []  # (name, f32 / i32 / u32), in word order

    # ---- entry ---------------------------------------------------------------------
    def translate(self, rules: list, metric_rules: list = (), metrics_cls=None, tissue_rules: list = (),
                  chem_rules: list = ()) -> str:
        self.struct_of(StructT("cell", self.cell.cls))
        if self.tissue.cls is not None:
            self.struct_of(StructT("tissue", self.tissue.cls))
        if self.chem.cls is not None:
            self.struct_of(StructT("chem", self.chem.cls))
        tissue_names, chem_names = [], []
        for kind, lst, out in (("tissue", tissue_rules, tissue_names), ("chem", chem_rules, chem_names)):
            for r in lst:
                fn = r.__func__ if hasattr(r, "__func__") else r
                node = self._parse(fn)
                args = node.args.args
                rec = self.tissue.cls if kind == "tissue" else self.chem.cls
                what = "tissue" if kind == "tissue" else "chemistry"
                if rec is None:
                    raise CompilationError(f"{what} rule '{fn.__name__}' needs the engine's {what} grid", node)
                if len(args) != 2:
                    raise CompilationError(f"{what} rule '{fn.__name__}' must take (self, voxel)", node)
                key = f"{kind}rule:{fn.__name__}"
                info = FuncInfo(key, f"{kind}rule_{fn.__name__}", node, _namespace(fn), "sim", VOID,
                                [(args[1].arg, StructT(kind, rec))])
                self.funcs[key] = info
                self.rule_kind = kind
                try:
                    self._emit_func(info, is_rule=True)
                finally:
                    self.rule_kind = "cell"
                out.append(info.cname)
                if kind == "chem":
                    iters = int(getattr(fn, "_iterations", 1))
                    if iters < 0:
                        raise CompilationError(f"chemistry rule '{fn.__name__}': iterations must be >= 0", node)
                    self.chem_pass_rules += [info.cname] * iters
        self.tissue_passes = len(tissue_names)
        self.grid_rules = (tissue_names, chem_names)
        rule_names = []
        for r in rules:
            fn = r.__func__ if hasattr(r, "__func__") else r
            node = self._parse(fn)
            args = node.args.args
            if len(args) != 2:
                raise CompilationError(f"cell rule '{fn.__name__}' must take (self, cell)", node)
            key = f"rule:{fn.__name__}"
            info = FuncInfo(key, f"rule_{fn.__name__}", node, _namespace(fn), "sim", VOID,
                            [(args[1].arg, StructT("cell", self.cell.cls))])
            self.funcs[key] = info
            self._emit_func(info, is_rule=True)
            rule_names.append(info.cname)
        metric_names = []
        if metric_rules:
            if metrics_cls is None:
                raise CompilationError("metric rules need a Metrics class")
            mt = self.struct_of(StructT("metrics", metrics_cls))
            self.metric_fields = [(n, t) for n, t in self.struct_fields(mt)]
            if not self.metric_fields:
                raise CompilationError(f"'{metrics_cls.__name__}' has no metric fields")
            if len(self.metric_fields) > 64:
                raise CompilationError("at most 64 metric fields")
        for r in metric_rules:
            fn = r.__func__ if hasattr(r, "__func__") else r
            node = self._parse(fn)
            args = node.args.args
            if len(args) != 3:
                raise CompilationError(f"metric rule '{fn.__name__}' must take (self, cell, metrics)", node)
            key = f"metric:{fn.__name__}"
            info = FuncInfo(key, f"metric_{fn.__name__}", node, _namespace(fn), "sim", VOID,
                            [(args[1].arg, StructT("cell", self.cell.cls)), (args[2].arg, mt)])
            self.funcs[key] = info
            self.in_metric = True
            self.rule_kind = "metric"
            try:
                self._emit_func(info, is_rule=True)
            finally:
                self.in_metric = False
                self.rule_kind = "cell"
            metric_names.append(info.cname)
        return self._assemble(rule_names, metric_names)

    def _parse(self, fn) -> ast.FunctionDef:
        try:
            src = textwrap.dedent(inspect.getsource(fn))
        except (OSError, TypeError) as e:
            raise CompilationError(f"cannot read the source of '{fn.__name__}': {e}")
        mod = ast.parse(src)
        node = mod.body[0]
        if not isinstance(node, ast.FunctionDef):
            raise CompilationError(f"'{fn.__name__}' is not a plain function")
        return node

    # ---- structs -------------------------------------------------------------------
    def struct_of(self, st: StructT) -> StructT:
        self.structs.setdefault(st.cname, st)
        return st

    def struct_fields(self, st: StructT) -> list:
        if st.kind == "metrics":
            out = []
            for n, ann in data_fields(st.cls):
                try:
                    t = py_to_dsl(ann)
                except TypeError:
                    t = None
                if t not in (F32, I32, U32):
                    raise CompilationError(f"metric field '{n}' must be ol.f32, ol.i32 or ol.u32")
                out.append((n, t))
            return out
        if st.kind == "cell":
            return list(self.cell.fields) if st.cls is self.cell.cls else [(n, py_to_dsl(t)) for n, t in data_fields(st.cls)]
        if st.kind == "chem":
            return list(self.chem.fields)
        return [(n, t) for n, t, _ in self.tissue.fields]

    def field_type(self, st: StructT, name: str, node):
        for n, t in self.struct_fields(st):
            if n == name:
                return t
        raise CompilationError(f"'{st.cls.__name__}' has no field '{name}'", node)

    # ---- emission helpers ---------------------------------------------------------------
    def emit(self, s: str) -> None:
        self.lines.append("    " * self.indent + s)

    def fresh(self, base: str) -> str:
        self.tmp += 1
        return f"_{base}{self.tmp}"

    # ---- functions -------------------------------------------------------------------
    def _emit_func(self, info: FuncInfo, is_rule: bool = False) -> None:
        saved = (self.lines, self.indent, self.cur, getattr(self, "_nb_break", None),
                 getattr(self, "_loop_depth_at_nb", None), getattr(self, "_loop_depth", 0))
        self.lines, self.indent, self.cur = [], 1, info
        self._nb_break, self._loop_depth_at_nb, self._loop_depth = None, None, 0
        env = Env()
        sig = []
        if info.self_t != "sim":
            sig.append(f"{ctype(info.self_t)}& self")
            env.vars[info.node.args.args[0].arg] = Local("self", info.self_t)
        for pname, pt in info.params:
            ref = "&" if (is_rule and isinstance(pt, StructT)) else ""
            sig.append(f"{ctype(pt)}{ref} {pname}_")
            env.vars[pname] = Local(f"{pname}_", pt)
        sig.append("RuleCtx& ctx")
        self.env = env
        self._declare_locals(info.node.body, env)
        self.stmts(info.node.body, env)
        body = self.lines
        (self.lines, self.indent, self.cur, self._nb_break, self._loop_depth_at_nb, self._loop_depth) = saved
        head = f"__device__ {ctype(info.ret)} {info.cname}({', '.join(sig)})"
        tail = []
        if info.ret != VOID:
            tail.append(f"    return {self.zero(info.ret)};")
        info.code = head + " {\n" + "\n".join(body + tail) + "\n}\n"
        info.proto = head + ";\n"
        self.order.append(info.key)

    def _declare_locals(self, body, env: Env) -> None:
        """Hoist every local assigned anywhere in the body (Python scoping) to
        the top of the C++ function, typed by its annotation or first value."""
        for node in self._walk_stmts(body):
            if isinstance(node, ast.AnnAssign) and isinstance(node.target, ast.Name):
                name = node.target.id
                if env.get(name) is None:
                    t = self.ann_type(node.annotation, node)
                    self._declare(env, name, t)
            elif isinstance(node, ast.Assign):
                for tgt in node.targets:
                    names = [tgt] if isinstance(tgt, ast.Name) else (tgt.elts if isinstance(tgt, ast.Tuple) else [])
                    if isinstance(tgt, ast.Tuple):
                        rt = self.expr_type(node.value, env)
                        if not isinstance(rt, TupleT) or len(rt.elems) != len(tgt.elts):
                            raise CompilationError("tuple assignment needs a call returning a tuple of the same size", node)
                        for el, et in zip(tgt.elts, rt.elems):
                            if isinstance(el, ast.Name) and env.get(el.id) is None:
                                self._declare(env, el.id, et)
                    elif isinstance(tgt, ast.Name) and env.get(tgt.id) is None:
                        self._declare(env, tgt.id, self.expr_type(node.value, env))

    def _declare(self, env: Env, name: str, t) -> None:
        cname = f"{name}_"
        env.vars[name] = Local(cname, t)
        self.emit(f"{ctype(t)} {cname} = {self.zero(t)};")

    def _walk_stmts(self, body):
        for st in body:
            yield st
            if isinstance(st, (ast.If, ast.While)):
                yield from self._walk_stmts(st.body)
                yield from self._walk_stmts(st.orelse)
            elif isinstance(st, ast.For):
                if self._is_macro_loop(st.iter):
                    continue  # the loop body has its own scope (the loop variable)
                yield from self._walk_stmts(st.body)

    def zero(self, t) -> str:
        if t == F32:
            return "0.0f"
        if t in (I32, U32):
            return "0"
        if t == BOOL:
            return "false"
        if t == VEC3:
            return "onc_v3_make(0.0f, 0.0f, 0.0f)"
        if t == IVEC3:
            return "onc_iv3_make(0, 0, 0)"
        if isinstance(t, StructT):
            if t.kind == "cell":
                return f"{t.cname}_defaults()"
            return f"{t.cname}{{}}"
        if isinstance(t, TupleT):
            self.tuples.setdefault(t.cname, t)
            return f"{t.cname}{{}}"
        return ""

    def ann_type(self, ann_node, node):
        g = self.cur.globals if self.cur else {}
        try:
            ann = eval(compile(ast.Expression(ann_node), "<ann>", "eval"), dict(g))  # noqa: S307 - user annotation
        except Exception as e:
            raise CompilationError(f"cannot resolve annotation: {e}", node)
        try:
            t = py_to_dsl(ann)
        except TypeError as e:
            raise CompilationError(str(e), node)
        if isinstance(t, StructT):
            self.struct_of(t)
        if isinstance(t, TupleT):
            self.tuples.setdefault(t.cname, t)
        return t

    # ---- statements -------------------------------------------------------------------
    def stmts(self, body, env: Env) -> None:
        for st in body:
            self.stmt(st, env)

    def stmt(self, node, env: Env) -> None:
        m = getattr(self, f"st_{type(node).__name__}", None)
        if m is None:
            raise CompilationError(f"unsupported statement '{type(node).__name__}' in rule code", node)
        m(node, env)

    def st_Pass(self, node, env):
        pass

    def st_Break(self, node, env):
        lbl = getattr(self, "_nb_break", None)
        if lbl and self._loop_depth_at_nb == self._loop_depth:
            self.emit(f"goto {lbl};")
        else:
            self.emit("break;")

    def st_Continue(self, node, env):
        self.emit("continue;")

    def st_Return(self, node, env):
        ret = self.cur.ret
        if node.value is None:
            if ret != VOID:
                raise CompilationError("missing return value", node)
            self.emit("return;")
            return
        if isinstance(node.value, ast.Tuple):
            if not isinstance(ret, TupleT):
                raise CompilationError("returning a tuple from a function not annotated -> tuple[...]", node)
            parts = [self.conv(*self.expr(e, env), et, e) for e, et in zip(node.value.elts, ret.elems)]
            self.emit(f"return {ret.cname}{{{', '.join(parts)}}};")
            return
        c, t = self.expr(node.value, env)
        self.emit(f"return {self.conv(c, t, ret, node)};")

    def st_Expr(self, node, env):
        v = node.value
        if isinstance(v, ast.Constant) and isinstance(v.value, str):
            return  # docstring
        if isinstance(v, ast.Call) and isinstance(v.func, ast.Attribute):
            recv_t = self.expr_type(v.func.value, env) if not self._is_self(v.func.value) else "sim"
            if isinstance(recv_t, StructT) and recv_t.kind == "cell":
                if v.func.attr == "die":
                    if not self._is_rule_cell(v.func.value, env):
                        raise CompilationError("only the rule's own cell may die()", node)
                    self.emit("ctx.alive = false;")
                    return
                if v.func.attr == "divide":
                    if not self._is_rule_cell(v.func.value, env):
                        raise CompilationError("only the rule's own cell may divide()", node)
                    if len(v.args) != 1:
                        raise CompilationError("divide() takes one cell", node)
                    c, t = self.expr(v.args[0], env)
                    if not (isinstance(t, StructT) and t.kind == "cell" and t.cls is self.cell.cls):
                        raise CompilationError("divide() needs a cell of the simulation's cell class", node)
                    self.emit(f"ctx.daughter = {c};")
                    self.emit("ctx.div = true;")
                    return
                if v.func.attr == "copy_from":
                    tgt, tt = self.expr(v.func.value, env)
                    c, t = self.expr(v.args[0], env)
                    self.emit(f"{tgt} = {c};")
                    return
        c, t = self.expr(v, env)
        self.emit(f"(void)({c});")

    def _is_self(self, node) -> bool:
        return isinstance(node, ast.Name) and self.cur is not None and self.cur.self_t == "sim" and \
            node.id == self.cur.node.args.args[0].arg

    def _is_rule_cell(self, node, env) -> bool:
        """The rule's own cell: the cell parameter of a rule, or `self` of a
        cell-class helper (helpers that kill or divide are called on it)."""
        if not isinstance(node, ast.Name):
            return False
        loc = env.get(node.id)
        if loc is None or not (isinstance(loc.t, StructT) and loc.t.kind == "cell"):
            return False
        if isinstance(self.cur.self_t, StructT):
            return loc.cname == "self"
        if self.cur.key.startswith("rule:"):
            return loc.cname == f"{self.cur.params[0][0]}_"
        return False

    def st_AnnAssign(self, node, env):
        if node.value is None:
            return
        self._assign_to(node.target, node.value, env, node)

    def st_Assign(self, node, env):
        if len(node.targets) != 1:
            raise CompilationError("chained assignment is not supported", node)
        tgt = node.targets[0]
        if isinstance(tgt, ast.Tuple):
            c, t = self.expr(node.value, env)
            tmp = self.fresh("tup")
            self.emit(f"{ctype(t)} {tmp} = {c};")
            for k, el in enumerate(tgt.elts):
                lc, lt = self.lvalue(el, env)
                self.emit(f"{lc} = {self.conv(f'{tmp}.f{k}', t.elems[k], lt, node)};")
            return
        self._assign_to(tgt, node.value, env, node)

    def _assign_to(self, tgt, value, env, node):
        lc, lt = self.lvalue(tgt, env)
        c, t = self.expr(value, env)
        self.emit(f"{lc} = {self.conv(c, t, lt, node)};")

    def _metric_target(self, node, env):
        """(C expression, type) of `m.field` when node is a metric field."""
        if not (isinstance(node, ast.Attribute) and isinstance(node.value, ast.Name)):
            return None
        loc = env.get(node.value.id)
        if loc is None or not (isinstance(loc.t, StructT) and loc.t.kind == "metrics"):
            return None
        return f"{loc.cname}.{node.attr}", self.field_type(loc.t, node.attr, node)

    def st_AugAssign(self, node, env):
        mtgt = self._metric_target(node.target, env)
        if mtgt is not None:
            if not isinstance(node.op, (ast.Add, ast.Sub)):
                raise CompilationError("metric fields take += and -= only", node)
            lc, lt = mtgt
            sign = "+" if isinstance(node.op, ast.Add) else "-"
            rc = self.conv(*self.expr(node.value, env), lt, node)
            if lt == I32:  # wraps mod 2^32, as the reference's atomicAdd / atomicSub
                self.emit(f"{lc} = (int)((unsigned){lc} {sign} (unsigned)({rc}));")
            else:
                self.emit(f"{lc} = {lc} {sign} {rc};")
            return
        lc, lt = self.lvalue(node.target, env)
        rc, rt = self.binop(node.op, lc, lt, *self.expr(node.value, env), node)
        self.emit(f"{lc} = {self.conv(rc, rt, lt, node)};")

    def lvalue(self, node, env):
        if isinstance(node, ast.Name):
            loc = env.get(node.id)
            if loc is None:
                raise CompilationError(f"unknown variable '{node.id}'", node)
            if loc.cname.startswith("_nb"):
                raise CompilationError("neighbour cells are read-only", node)
            return loc.cname, loc.t
        if isinstance(node, ast.Attribute):
            if self._metric_target(node, env) is not None:
                raise CompilationError("metric fields take += and -= only", node)
            bc, bt = self.expr(node.value, env)
            if isinstance(bt, StructT):
                if bt.kind in ("tissue", "chem"):
                    what = "tissue" if bt.kind == "tissue" else "chemistry"
                    if self.rule_kind in ("cell", "metric") and self.cur.key.split(":")[0] in ("rule", "metric"):
                        raise CompilationError(f"cell rules may not write {what} fields", node)
                    if bc.startswith("_g"):
                        raise CompilationError(f"neighbouring {what} voxels are read-only", node)
                    return f"{bc}.{node.attr}", self.field_type(bt, node.attr, node)
                if self.in_metric and bt.kind == "cell":
                    raise CompilationError("metric rules read cells; they may not write them", node)
                if bc.startswith("_nb"):
                    raise CompilationError("neighbour cells are read-only", node)
                return f"{bc}.{node.attr}", self.field_type(bt, node.attr, node)
            if bt in (VEC3, IVEC3) and node.attr in ("x", "y", "z"):
                return f"{bc}.{node.attr}", F32 if bt == VEC3 else I32
        raise CompilationError("unsupported assignment target", node)

    def st_If(self, node, env):
        c, t = self.expr(node.test, env)
        self.emit(f"if ({self.truth(c, t)}) {{")
        self.indent += 1
        self.stmts(node.body, env)
        self.indent -= 1
        if node.orelse:
            self.emit("} else {")
            self.indent += 1
            self.stmts(node.orelse, env)
            self.indent -= 1
        self.emit("}")

    def st_While(self, node, env):
        c, t = self.expr(node.test, env)
        self.emit(f"while ({self.truth(c, t)}) {{")
        self.indent += 1
        self._loop_depth = getattr(self, "_loop_depth", 0) + 1
        self.stmts(node.body, env)
        self._loop_depth -= 1
        self.indent -= 1
        self.emit("}")

    def _is_neighbors(self, it) -> bool:
        return isinstance(it, ast.Attribute) and it.attr == "neighbors"

    def _is_macro_loop(self, it) -> bool:
        return isinstance(it, ast.Attribute) and it.attr in ("neighbors", "cells", "tissues")

    def st_For(self, node, env):
        if node.orelse:
            raise CompilationError("for ... else is not supported", node)
        if self._is_macro_loop(node.iter):
            st = self.expr_type(node.iter.value, env)
            kind = st.kind if isinstance(st, StructT) else None
            if node.iter.attr == "neighbors" and kind == "cell":
                return self._for_neighbors(node, env)
            if node.iter.attr == "neighbors" and kind in ("tissue", "chem"):
                return self._for_grid(node, env, "neighbors", st)
            if node.iter.attr == "cells" and kind in ("tissue", "chem"):
                return self._for_grid(node, env, "cells", st)
            if node.iter.attr == "tissues" and kind == "chem":
                return self._for_grid(node, env, "tissues", st)
            raise CompilationError(f"'.{node.iter.attr}' is not defined on this value", node)
        it = node.iter
        if isinstance(it, ast.Call) and isinstance(it.func, ast.Name) and it.func.id == "range":
            args = [self.conv(*self.expr(a, env), I32, a) for a in it.args]
            if len(args) == 1:
                a0, a1, a2 = "0", args[0], "1"
            elif len(args) == 2:
                a0, a1, a2 = args[0], args[1], "1"
            else:
                a0, a1, a2 = args
            if not isinstance(node.target, ast.Name):
                raise CompilationError("range loop needs a simple variable", node)
            loc = env.get(node.target.id)
            if loc is None:
                self._declare(env, node.target.id, I32)
                loc = env.get(node.target.id)
            stp = self.fresh("step")
            self.emit(f"{{ const int {stp} = {a2};")
            self.emit(f"for ({loc.cname} = {a0}; {stp} > 0 ? {loc.cname} < {a1} : {loc.cname} > {a1}; "
                      f"{loc.cname} += {stp}) {{")
            self.indent += 1
            self._loop_depth = getattr(self, "_loop_depth", 0) + 1
            self.stmts(node.body, env)
            self._loop_depth -= 1
            self.indent -= 1
            self.emit("} }")
            return
        raise CompilationError("only 'for x in range(...)' and 'for nb in <cell>.neighbors' loops are supported", node)

    def _for_neighbors(self, node, env):
        if self.in_metric:
            raise CompilationError("metric rules may not walk neighbours (the bins index the step's cells, "
                                   "not the committed ones)", node)
        src = node.iter.value
        sc, st = self.expr(src, env)
        if not (isinstance(st, StructT) and st.kind == "cell"):
            raise CompilationError("'.neighbors' is only defined on cells in cell rules", node)
        if not isinstance(node.target, ast.Name):
            raise CompilationError("neighbour loop needs a simple variable", node)
        var = self.fresh("nb")
        end = self.fresh("nbend")
        sv, q, j, b = self.fresh("src"), self.fresh("q"), self.fresh("j"), self.fresh("b")
        self.emit("{")
        self.indent += 1
        self.emit(f"const {ctype(st)}& {sv} = {sc};")
        self.emit(f"const onc_dsim& _S = *ctx.S;")
        # the bins hold this step's sorted positions, so a stored cell is
        # located by its sorted position (a constructed cell by its own)
        self.emit(f"const float4 _c = {sv}._idx != 0xFFFFFFFFu ? ((const float4*)_S.pos4)[{sv}._idx] : "
                  f"make_float4({sv}.pos.x, {sv}.pos.y, {sv}.pos.z, 0.0f);")
        self.emit("const int _vx = onc_vox1(_c.x, _S.dx, _S.dims[0]), _vy = onc_vox1(_c.y, _S.dx, _S.dims[1]), "
                  "_vz = onc_vox1(_c.z, _S.dx, _S.dims[2]);")
        self.emit("const onc_nkeys _nk = onc_neighbor_keys(onc_morton((onc_u32)_vx, (onc_u32)_vy, (onc_u32)_vz, "
                  "_S.key_format), _S.key_format);")
        self.emit("for (int _oz = -1; _oz <= 1; _oz++) {")
        self.indent += 1
        self.emit("const int _nz = _vz + _oz;")
        self.emit("if (_nz < 0 || _nz >= _S.dims[2]) continue;")
        self.emit("const onc_u32 _zk = _oz < 0 ? _nk.z[0] : (_oz == 0 ? _nk.z[1] : _nk.z[2]);")
        self.emit(f"for (int {q} = 0; {q} < 9; {q}++) {{")
        self.indent += 1
        self.emit(f"const int _ny = _vy + {q} / 3 - 1, _nx = _vx + {q} % 3 - 1;")
        self.emit("if (_nx < 0 || _ny < 0 || _nx >= _S.dims[0] || _ny >= _S.dims[1]) continue;")
        self.emit(f"const uint2 {b} = ((const uint2*)_S.bins)[_nk.x[{q} % 3] | _nk.y[{q} / 3] | _zk];")
        self.emit(f"const onc_u32 {b}_n = onc_bin_len({b}, ctx.bin_tag);")
        self.emit(f"if ({b}_n == 0u) continue;")
        self.emit(f"for (onc_u32 {j} = {b}.x; {j} < {b}.x + {b}_n; {j}++) {{")
        self.indent += 1
        self.emit(f"if ({j} == {sv}._idx) continue;")
        self.emit(f"{ctype(st)} {var} = {st.cname}_load(_S, _S.st[ctx.par ^ 1u], {j});")
        inner = Env(env)
        inner.vars[node.target.id] = Local(var, st)
        saved = (getattr(self, "_nb_break", None), getattr(self, "_loop_depth_at_nb", None))
        self._loop_depth = getattr(self, "_loop_depth", 0) + 1
        self._nb_break, self._loop_depth_at_nb = end, self._loop_depth
        self._declare_locals(node.body, inner)
        self.stmts(node.body, inner)
        self._nb_break, self._loop_depth_at_nb = saved
        self._loop_depth -= 1
        self.indent -= 1
        self.emit("}")
        self.indent -= 1
        self.emit("}")
        self.indent -= 1
        self.emit("}")
        self.indent -= 1
        self.emit(f"}} {end}:;")

    def _for_grid(self, node, env, what, st):
        """Voxel macros (_stmt_translator.py:493-608, restated):
        `.neighbors` of a tissue or chemistry voxel: its 26 Moore neighbours in
        template.wgsl:11-21 order (x fastest), out-of-range ones skipped, read
        from the pass's input buffer; `.cells` of a tissue voxel: the cells of
        its bin, and of a chemistry voxel: those of its 8 tissue voxels, read
        as the cell rules' neighbours are (this step's sorted state);
        `.tissues` of a chemistry voxel: its 8 tissue voxels (dz, dy, dx
        order, x fastest, clipped at the grid, core/utils/_types.py:508-524)."""
        if self.in_metric:
            raise CompilationError("metric rules may not walk voxels", node)
        if not isinstance(node.target, ast.Name):
            raise CompilationError("voxel loop needs a simple variable", node)
        sc, _ = self.expr(node.iter.value, env)
        var, end, sv = self.fresh("gv"), self.fresh("gend"), self.fresh("gsrc")
        self.emit("{")
        self.indent += 1
        self.emit(f"const {ctype(st)}& {sv} = {sc};")
        self.emit("const onc_dsim& _S = *ctx.S;")
        chem = st.kind == "chem"
        dims = "_S.cdims" if chem else "_S.dims"
        if what == "neighbors":
            vt = st
            self.emit(f"for (int _k = 0; _k < 27; _k++) {{")
            self.indent += 1
            self.emit("if (_k == 13) continue;")
            self.emit(f"const int _nx = {sv}._coord.x + _k % 3 - 1, _ny = {sv}._coord.y + (_k / 3) % 3 - 1, "
                      f"_nz = {sv}._coord.z + _k / 9 - 1;")
            self.emit(f"if (_nx < 0 || _ny < 0 || _nz < 0 || _nx >= {dims}[0] || _ny >= {dims}[1] || "
                      f"_nz >= {dims}[2]) continue;")
            buf = "ctx.cin" if chem else "ctx.tin"
            self.emit(f"{ctype(vt)} {var} = {vt.cname}_loadv(ctx, {buf}, _nx, _ny, _nz);")
            closers = 1
        elif what == "tissues":
            if self.tissue.cls is None:
                raise CompilationError("'.tissues' needs the engine's tissue grid", node)
            vt = self.struct_of(StructT("tissue", self.tissue.cls))
            self.emit("for (int _k = 0; _k < 8; _k++) {")
            self.indent += 1
            self.emit(f"const int _nx = 2 * {sv}._coord.x + (_k & 1), _ny = 2 * {sv}._coord.y + ((_k >> 1) & 1), "
                      f"_nz = 2 * {sv}._coord.z + (_k >> 2);")
            self.emit("if (_nx >= _S.dims[0] || _ny >= _S.dims[1] || _nz >= _S.dims[2]) continue;")
            self.emit(f"{ctype(vt)} {var} = {vt.cname}_loadv(ctx, ctx.tin, _nx, _ny, _nz);")
            closers = 1
        else:  # cells
            vt = self.struct_of(StructT("cell", self.cell.cls))
            if chem:
                self.emit("for (int _k = 0; _k < 8; _k++) {")
                self.indent += 1
                self.emit(f"const int _nx = 2 * {sv}._coord.x + (_k & 1), _ny = 2 * {sv}._coord.y + ((_k >> 1) & 1), "
                          f"_nz = 2 * {sv}._coord.z + (_k >> 2);")
                self.emit("if (_nx >= _S.dims[0] || _ny >= _S.dims[1] || _nz >= _S.dims[2]) continue;")
                closers = 1
            else:
                self.emit(f"const int _nx = {sv}._coord.x, _ny = {sv}._coord.y, _nz = {sv}._coord.z;")
                self.emit("{")
                self.indent += 1
                closers = 1
            self.emit("const uint2 _b = ((const uint2*)_S.bins)[onc_morton((onc_u32)_nx, (onc_u32)_ny, "
                      "(onc_u32)_nz, _S.key_format)];")
            self.emit("const onc_u32 _bn = onc_bin_len(_b, ctx.bin_tag);")
            self.emit("if (_bn == 0u) continue;" if chem else "if (_bn != 0u)")
            j = self.fresh("j")
            self.emit(f"for (onc_u32 {j} = _b.x; {j} < _b.x + _bn; {j}++) {{")
            self.indent += 1
            self.emit(f"{ctype(vt)} {var} = {vt.cname}_load(_S, _S.st[ctx.par ^ 1u], {j});")
            closers += 1
        inner = Env(env)
        inner.vars[node.target.id] = Local(var, vt)
        saved = (getattr(self, "_nb_break", None), getattr(self, "_loop_depth_at_nb", None))
        self._loop_depth = getattr(self, "_loop_depth", 0) + 1
        self._nb_break, self._loop_depth_at_nb = end, self._loop_depth
        self._declare_locals(node.body, inner)
        self.stmts(node.body, inner)
        self._nb_break, self._loop_depth_at_nb = saved
        self._loop_depth -= 1
        for _ in range(closers):
            self.indent -= 1
            self.emit("}")
        self.indent -= 1
        self.emit(f"}} {end}:;")

    # ---- expressions ---------------------------------------------------------------
    def expr_type(self, node, env):
        saved = (self.lines, self.indent)
        self.lines = []
        try:
            return self.expr(node, env)[1]
        finally:
            self.lines, self.indent = saved

    def truth(self, c, t) -> str:
        if t == BOOL:
            return c
        if t in (F32, I32, U32):
            return f"(({c}) != 0)"
        raise CompilationError(f"a {t} value has no truth value")

    def conv(self, c, t, to, node=None) -> str:
        if t == to:
            return c
        if to == F32 and t in (I32, U32, BOOL):
            return f"(float)({c})"
        if to == I32 and t in (F32, U32, BOOL):
            return f"(int)({c})"
        if to == U32 and t in (F32, I32, BOOL):
            return f"(unsigned)({c})"
        if to == BOOL and t in (F32, I32, U32):
            return f"(({c}) != 0)"
        if to == VEC3 and t == IVEC3:
            return f"onc_v3_make((float)({c}).x, (float)({c}).y, (float)({c}).z)"
        raise CompilationError(f"cannot convert {_tname(t)} to {_tname(to)}", node)

    def expr(self, node, env):
        m = getattr(self, f"ex_{type(node).__name__}", None)
        if m is None:
            raise CompilationError(f"unsupported expression '{type(node).__name__}' in rule code", node)
        return m(node, env)

    def literal(self, v, node):
        if isinstance(v, bool):
            return ("true" if v else "false"), BOOL
        if isinstance(v, int):
            if not -(2 ** 31) <= v < 2 ** 31:
                return f"{v}u", U32
            return str(v), I32
        if isinstance(v, float):
            if v != v:
                return "__builtin_nanf(\"\")", F32
            if v in (float("inf"), float("-inf")):
                return ("__builtin_inff()" if v > 0 else "(-__builtin_inff())"), F32
            return f"{v!r}f", F32
        if isinstance(v, T.vec3):
            return f"onc_v3_make({v.x!r}f, {v.y!r}f, {v.z!r}f)", VEC3
        raise CompilationError(f"unsupported constant {v!r}", node)

    def ex_Constant(self, node, env):
        return self.literal(node.value, node)

    def ex_Name(self, node, env):
        loc = env.get(node.id)
        if loc is not None:
            return loc.cname, loc.t
        g = self.cur.globals
        if node.id in g:
            v = g[node.id]
            if isinstance(v, (bool, int, float, T.vec3)):
                return self.literal(v, node)
        if node.id in ("True", "False"):
            return ("true" if node.id == "True" else "false"), BOOL
        raise CompilationError(f"unknown name '{node.id}'", node)

    def _module_attr(self, node):
        """Resolve ol.math.xxx / ol.xxx / math.xxx chains to (module path, attr)."""
        parts = []
        n = node
        while isinstance(n, ast.Attribute):
            parts.append(n.attr)
            n = n.value
        if isinstance(n, ast.Name):
            parts.append(n.id)
            return list(reversed(parts))
        return None

    def ex_Attribute(self, node, env):
        # self.params.X / self.CONST
        if isinstance(node.value, ast.Attribute) and self._is_self(node.value.value) and node.value.attr == "params":
            ps = self.pslots.get(node.attr)
            if ps is None:
                raise CompilationError(f"Params has no field '{node.attr}'", node)
            k = ps.slot
            if ps.dsl == F32:
                return f"ctx.S->params[{k}]", F32
            if ps.dsl in (I32, U32):
                return f"(({ctype(ps.dsl)})ctx.S->params[{k}])", ps.dsl
            if ps.dsl == BOOL:
                return f"(ctx.S->params[{k}] != 0.0f)", BOOL
            if ps.dsl == VEC3:
                return f"onc_v3_make(ctx.S->params[{k}], ctx.S->params[{k + 1}], ctx.S->params[{k + 2}])", VEC3
            return (f"onc_iv3_make((int)ctx.S->params[{k}], (int)ctx.S->params[{k + 1}], (int)ctx.S->params[{k + 2}])",
                    IVEC3)
        if self._is_self(node.value):
            v = getattr(self.sim_cls, node.attr, None)
            if isinstance(v, (bool, int, float, T.vec3)):
                return self.literal(v, node)
            raise CompilationError(f"'{node.attr}' is not a constant of the simulation class", node)
        path = self._module_attr(node)
        if path is not None and env.get(path[0]) is None:
            root = self.cur.globals.get(path[0])
            obj = root
            try:
                for p in path[1:]:
                    obj = getattr(obj, p)
            except AttributeError:
                obj = None
            if isinstance(obj, (bool, int, float)):
                return self.literal(obj, node)
        bc, bt = self.expr(node.value, env)
        if isinstance(bt, StructT):
            if bt.kind == "metrics":
                raise CompilationError("metric fields can only be added to inside a metric rule", node)
            if bt.kind in ("tissue", "chem") and node.attr in ("coord", "pos") and \
                    node.attr not in [n for n, _ in self.struct_fields(bt)]:
                if node.attr == "coord":
                    return f"{bc}._coord", IVEC3
                # voxel centre: coord * size + size / 2 (core/utils/_types.py:465-467, 496-498)
                vs = "(2.0f * ctx.S->dx)" if bt.kind == "chem" else "ctx.S->dx"
                return (f"onc_v3_make((float){bc}._coord.x * {vs} + {vs} / 2.0f, (float){bc}._coord.y * {vs} + "
                        f"{vs} / 2.0f, (float){bc}._coord.z * {vs} + {vs} / 2.0f)"), VEC3
            return f"{bc}.{node.attr}", self.field_type(bt, node.attr, node)
        if bt in (VEC3, IVEC3) and node.attr in ("x", "y", "z"):
            return f"{bc}.{node.attr}", F32 if bt == VEC3 else I32
        raise CompilationError(f"unsupported attribute '{node.attr}'", node)

    def ex_UnaryOp(self, node, env):
        c, t = self.expr(node.operand, env)
        if isinstance(node.op, ast.Not):
            return f"(!{self.truth(c, t)})", BOOL
        if isinstance(node.op, ast.USub):
            if t == BOOL:
                return f"(-(int)({c}))", I32
            return f"(-{c})", t
        if isinstance(node.op, ast.UAdd):
            return c, t
        raise CompilationError("unsupported unary operator", node)

    def ex_BinOp(self, node, env):
        return self.binop(node.op, *self.expr(node.left, env), *self.expr(node.right, env), node)

    def binop(self, op, lc, lt, rc, rt, node):
        if lt == BOOL:
            lc, lt = f"(int)({lc})", I32
        if rt == BOOL:
            rc, rt = f"(int)({rc})", I32
        sym = {ast.Add: "+", ast.Sub: "-", ast.Mult: "*", ast.Div: "/"}.get(type(op))
        if lt in (VEC3, IVEC3) or rt in (VEC3, IVEC3):
            if lt == VEC3 and rt == VEC3 and sym in ("+", "-", "*"):
                return f"({lc} {sym} {rc})", VEC3
            if lt == IVEC3 and rt == IVEC3 and sym in ("+", "-"):
                return f"({lc} {sym} {rc})", IVEC3
            if lt == VEC3 and rt in (F32, I32, U32) and sym in ("*", "/"):
                return f"({lc} {sym} {self.conv(rc, rt, F32)})", VEC3
            if rt == VEC3 and lt in (F32, I32, U32) and sym == "*":
                return f"({self.conv(lc, lt, F32)} * {rc})", VEC3
            raise CompilationError(f"unsupported vector operation: operand types {_tname(lt)} and {_tname(rt)} do not match", node)
        isf = F32 in (lt, rt)
        if isinstance(op, ast.Div):
            return f"({self.conv(lc, lt, F32)} / {self.conv(rc, rt, F32)})", F32
        if isinstance(op, ast.FloorDiv):
            if isf:
                return f"onc_floordiv_f({self.conv(lc, lt, F32)}, {self.conv(rc, rt, F32)})", F32
            return f"onc_floordiv_i({self.conv(lc, lt, I32)}, {self.conv(rc, rt, I32)})", I32
        if isinstance(op, ast.Mod):
            if isf:
                return f"onc_mod_f({self.conv(lc, lt, F32)}, {self.conv(rc, rt, F32)})", F32
            return f"onc_mod_i({self.conv(lc, lt, I32)}, {self.conv(rc, rt, I32)})", I32
        if isinstance(op, ast.Pow):
            return f"powf({self.conv(lc, lt, F32)}, {self.conv(rc, rt, F32)})", F32
        if sym is None:
            if isinstance(op, (ast.BitAnd, ast.BitOr, ast.BitXor, ast.LShift, ast.RShift)) and not isf:
                s2 = {ast.BitAnd: "&", ast.BitOr: "|", ast.BitXor: "^", ast.LShift: "<<", ast.RShift: ">>"}[type(op)]
                rt2 = U32 if U32 in (lt, rt) else I32
                return f"({lc} {s2} {rc})", rt2
            raise CompilationError("unsupported binary operator", node)
        if isf:
            return f"({self.conv(lc, lt, F32)} {sym} {self.conv(rc, rt, F32)})", F32
        rt2 = U32 if U32 in (lt, rt) else I32
        return f"({self.conv(lc, lt, rt2)} {sym} {self.conv(rc, rt, rt2)})", rt2

    def ex_BoolOp(self, node, env):
        sym = "&&" if isinstance(node.op, ast.And) else "||"
        parts = [self.truth(*self.expr(v, env)) for v in node.values]
        return "(" + f" {sym} ".join(parts) + ")", BOOL

    def ex_Compare(self, node, env):
        parts = []
        left = self.expr(node.left, env)
        for op, comp in zip(node.ops, node.comparators):
            right = self.expr(comp, env)
            sym = {ast.Lt: "<", ast.LtE: "<=", ast.Gt: ">", ast.GtE: ">=", ast.Eq: "==", ast.NotEq: "!="}.get(type(op))
            if sym is None:
                raise CompilationError("unsupported comparison", node)
            (lc, lt), (rc, rt) = left, right
            if lt in (VEC3, IVEC3) or rt in (VEC3, IVEC3):
                if sym not in ("==", "!=") or lt != rt:
                    raise CompilationError("vectors compare only with == / !=", node)
                parts.append(f"({lc} {sym} {rc})")
            else:
                if F32 in (lt, rt):
                    lc, rc = self.conv(lc, lt, F32), self.conv(rc, rt, F32)
                elif lt != rt:
                    lc, rc = f"(long long)({lc})", f"(long long)({rc})"
                parts.append(f"({lc} {sym} {rc})")
            left = right
        return "(" + " && ".join(parts) + ")", BOOL

    def ex_IfExp(self, node, env):
        tc, tt = self.expr(node.test, env)
        ac, at = self.expr(node.body, env)
        bc, bt = self.expr(node.orelse, env)
        rt = at if at == bt else (F32 if F32 in (at, bt) else at)
        return f"({self.truth(tc, tt)} ? {self.conv(ac, at, rt, node)} : {self.conv(bc, bt, rt, node)})", rt

    def ex_Call(self, node, env):
        f = node.func
        path = self._module_attr(f) if isinstance(f, (ast.Attribute, ast.Name)) else None
        if isinstance(f, ast.Attribute) and self._is_self(f.value):
            if f.attr == "tissue_at":
                if self.tissue.cls is None:
                    raise CompilationError("tissue_at() needs engine.setup_tissue(...)", node)
                pc, pt = self.expr(node.args[0], env)
                st = StructT("tissue", self.tissue.cls)
                return f"{st.cname}_at(ctx, {self.conv(pc, pt, VEC3, node)})", st
            if f.attr == "chemistry_at":
                if self.chem.cls is None:
                    raise CompilationError("chemistry_at() needs engine.setup_chemistry(...)", node)
                if self.in_metric:
                    raise CompilationError("metric rules may not read the chemistry grid", node)
                pc, pt = self.expr(node.args[0], env)
                st = self.struct_of(StructT("chem", self.chem.cls))
                return f"{st.cname}_at(ctx, {self.conv(pc, pt, VEC3, node)})", st
            meth = getattr(self.sim_cls, f.attr, None)
            if meth is None or not callable(meth):
                raise CompilationError(f"simulation has no method '{f.attr}'", node)
            return self._call_helper(f"sim:{f.attr}", meth, "sim", None, node, env)
        # methods on cells / tissues
        if isinstance(f, ast.Attribute) and not (path and env.get(path[0]) is None and path[0] in self.cur.globals):
            rc, rt = self.expr(f.value, env)
            if isinstance(rt, StructT):
                meth = getattr(rt.cls, f.attr, None)
                if meth is None or not callable(meth):
                    raise CompilationError(f"'{rt.cls.__name__}' has no method '{f.attr}'", node)
                if f.attr in ("die", "divide"):
                    raise CompilationError(f"{f.attr}() is a statement", node)
                return self._call_helper(f"{rt.kind}:{rt.cls.__name__}.{f.attr}", meth, rt, rc, node, env)
            raise CompilationError("unsupported method call", node)
        name = path[-1] if path else None
        root = self.cur.globals.get(path[0]) if path else None
        obj = root
        if path:
            try:
                for p in path[1:]:
                    obj = getattr(obj, p)
            except AttributeError:
                obj = None
        args = [self.expr(a, env) for a in node.args]
        # data-class constructors
        if isinstance(obj, type) and issubclass(obj, T.Cell):
            return self._construct_cell(obj, node, env)
        if obj is T.vec3 or name == "vec3":
            vals = [self.conv(c, t, F32, node) for c, t in args] + ["0.0f"] * (3 - len(args))
            if len(args) == 1 and args[0][1] == F32:
                vals = [args[0][0]] * 3
            return f"onc_v3_make({', '.join(vals[:3])})", VEC3
        if obj is T.ivec3 or name == "ivec3":
            vals = [self.conv(c, t, I32, node) for c, t in args] + ["0"] * (3 - len(args))
            return f"onc_iv3_make({', '.join(vals[:3])})", IVEC3
        if name in ("f32", "float"):
            return self.conv(*args[0], F32, node), F32
        if name in ("i32", "int"):
            return self.conv(*args[0], I32, node), I32
        if name == "u32":
            return self.conv(*args[0], U32, node), U32
        if name == "bool":
            return self.conv(*args[0], BOOL, node), BOOL
        if name in ("random", "random_dir") and self.in_metric:
            raise CompilationError("metric rules may not draw random numbers", node)
        if name == "random" and (obj is T.random or root is not None):
            return "onc_rand(ctx.rng)", F32
        if name == "random_dir":
            return "onc_v3_rand_dir(ctx.rng)", VEC3
        return self._math(name, args, node)

    def _math(self, name, args, node):
        ts = [t for _, t in args]
        cs = [c for c, _ in args]
        if name in _MATH1:
            return _MATH1[name].format(self.conv(cs[0], ts[0], F32, node)), F32
        if name in ("abs",):
            if ts[0] in (I32,):
                return f"onc_abs_i({cs[0]})", I32
            return f"__builtin_fabsf({self.conv(cs[0], ts[0], F32, node)})", F32
        if name in ("min", "max"):
            if all(t in (I32, BOOL) for t in ts):
                return f"onc_{name}_i({self.conv(cs[0], ts[0], I32)}, {self.conv(cs[1], ts[1], I32)})", I32
            return f"onc_{name}_f({self.conv(cs[0], ts[0], F32)}, {self.conv(cs[1], ts[1], F32)})", F32
        if name == "clamp":
            if ts[0] == VEC3:
                return f"onc_clamp_v({cs[0]}, {self.conv(cs[1], ts[1], VEC3)}, {self.conv(cs[2], ts[2], VEC3)})", VEC3
            if all(t == I32 for t in ts):
                return f"onc_clamp_i({cs[0]}, {cs[1]}, {cs[2]})", I32
            return f"onc_clamp_f({', '.join(self.conv(c, t, F32) for c, t in args)})", F32
        if name in ("pow",):
            return f"powf({self.conv(cs[0], ts[0], F32)}, {self.conv(cs[1], ts[1], F32)})", F32
        if name == "atan2":
            return f"atan2f({self.conv(cs[0], ts[0], F32)}, {self.conv(cs[1], ts[1], F32)})", F32
        if name in ("lerp", "step", "smoothstep"):
            return f"onc_{name}({', '.join(self.conv(c, t, F32) for c, t in args)})", F32
        vec = {"length": F32, "length_sq": F32, "distance": F32, "distance_sq": F32, "dot": F32,
               "normalize": VEC3, "cross": VEC3, "reflect": VEC3}
        if name in vec:
            return f"onc_{name}({', '.join(cs)})", vec[name]
        if name == "lerp_vec":
            return f"onc_lerp_v({cs[0]}, {cs[1]}, {self.conv(cs[2], ts[2], F32)})", VEC3
        raise CompilationError(f"unsupported call '{name}' in rule code", node)

    def _construct_cell(self, cls, node, env):
        if cls is not self.cell.cls:
            raise CompilationError(f"only '{self.cell.cls.__name__}' cells can be created here", node)
        st = self.struct_of(StructT("cell", cls))
        tmp = self.fresh("cell")
        self.emit(f"{st.cname} {tmp} = {st.cname}_defaults();")
        if node.args:
            raise CompilationError("cell constructors take keyword arguments", node)
        for kw in node.keywords:
            ft = self.field_type(st, kw.arg, node)
            c, t = self.expr(kw.value, env)
            self.emit(f"{tmp}.{kw.arg} = {self.conv(c, t, ft, node)};")
        return tmp, st

    def _call_helper(self, key, meth, recv_t, recv_c, node, env):
        fn = getattr(meth, "__func__", meth)
        info = self.funcs.get(key)
        if info is None:
            fnode = self._parse(fn)
            hints = typing.get_type_hints(fn)
            params = []
            for a in fnode.args.args[1:]:
                if a.arg not in hints:
                    raise CompilationError(f"parameter '{a.arg}' of '{fn.__name__}' needs an annotation", fnode)
                pt = py_to_dsl(hints[a.arg])
                if isinstance(pt, StructT):
                    self.struct_of(pt)
                params.append((a.arg, pt))
            ret = py_to_dsl(hints.get("return", None))
            if isinstance(ret, TupleT):
                self.tuples.setdefault(ret.cname, ret)
            cname = "h_" + key.replace(":", "_").replace(".", "_")
            info = FuncInfo(key, cname, fnode, _namespace(fn), recv_t, ret, params)
            self.funcs[key] = info
            saved_env = getattr(self, "env", None)
            self._emit_func(info)
            self.env = saved_env
        if len(node.args) != len(info.params):
            raise CompilationError(f"'{key}' takes {len(info.params)} arguments", node)
        args = [self.conv(*self.expr(a, env), pt, a) for a, (_, pt) in zip(node.args, info.params)]
        if recv_t != "sim":
            args = [recv_c] + args
        return f"{info.cname}({', '.join(args + ['ctx'])})", info.ret

    # ---- assembly ---------------------------------------------------------------------
    def _member_lines(self, st: StructT) -> list[str]:
        out = []
        for n, t in self.struct_fields(st):
            out.append(f"    {ctype(t)} {n};")
        if st.kind == "cell":
            out.append("    onc_u32 _id;")
            out.append("    onc_u32 _idx;")
        if st.kind in ("tissue", "chem"):
            out.append("    onc_iv3 _coord;")
            out.append("    onc_u32 _v;")
        return out

    def _defaults_expr(self, st: StructT) -> list[str]:
        lines = []
        for n, t in self.struct_fields(st):
            d = getattr(st.cls, n, None)
            if isinstance(d, T.Diffusing):
                d = d.init
            if d is None or isinstance(d, (property,)) or callable(d):
                lines.append(f"    r.{n} = {self.zero(t)};")
                continue
            if t == VEC3:
                lines.append(f"    r.{n} = onc_v3_make({float(d.x)!r}f, {float(d.y)!r}f, {float(d.z)!r}f);")
            elif t == IVEC3:
                lines.append(f"    r.{n} = onc_iv3_make({int(d.x)}, {int(d.y)}, {int(d.z)});")
            else:
                c, lt = self.literal(d, None)
                lines.append(f"    r.{n} = {self.conv(c, lt, t)};")
        return lines

    def _assemble(self, rule_names: list[str], metric_names: list[str] = ()) -> str:
        cs = StructT("cell", self.cell.cls)
        out = [PRELUDE]
        out.append("struct RuleCtx;\n")
        # structs
        for st in self.structs.values():
            out.append(f"struct {st.cname} {{\n" + "\n".join(self._member_lines(st)) + "\n};\n")
        for tt in self.tuples.values():
            mem = "\n".join(f"    {ctype(e)} f{k};" for k, e in enumerate(tt.elems))
            out.append(f"struct {tt.cname} {{\n{mem}\n}};\n")
        out.append(f"""struct RuleCtx {{
    const onc_dsim* S;
    onc_u32 par, i, bin_tag;
    onc_u32 tin;   // tissue field buffer read by tissue_at / voxel loads
    onc_u32 cin;   // chemistry buffer read by chemistry_at / voxel loads
    bool alive, div;
    onc_rng rng;
    {cs.cname} daughter;
}};
""")
        # cell load / store / defaults
        ld, stl = [], []
        for k, col in enumerate(self.cell.columns):
            if col.coltype == COL_F32:
                ld.append(f"    r.{col.name} = __uint_as_float(c[{k}][i]);")
                stl.append(f"    c[{k}][i] = __float_as_uint(v.{col.name});")
            elif col.coltype == COL_I32:
                ld.append(f"    r.{col.name} = (int)c[{k}][i];")
                stl.append(f"    c[{k}][i] = (onc_u32)v.{col.name};")
            else:
                fld = col.name
                ft = dict(self.cell.fields).get(fld)
                if ft == BOOL:
                    ld.append(f"    r.{fld} = c[{k}][i] != 0u;")
                    stl.append(f"    c[{k}][i] = v.{fld} ? 1u : 0u;")
                else:
                    ld.append(f"    r.{col.name} = c[{k}][i];")
                    stl.append(f"    c[{k}][i] = v.{col.name};")
        out.append(f"__device__ __forceinline__ {cs.cname} {cs.cname}_defaults() {{\n    {cs.cname} r;\n"
                   + "\n".join(self._defaults_expr(cs)) + "\n    r._id = 0u;\n    r._idx = 0xFFFFFFFFu;\n    return r;\n}\n")
        out.append(f"__device__ __forceinline__ {cs.cname} {cs.cname}_load(const onc_dsim& S, onc_gcol const* c, onc_u32 i) {{\n"
                   f"    {cs.cname} r;\n" + "\n".join(ld) + "\n    r._idx = i;\n    return r;\n}\n")
        out.append(f"__device__ __forceinline__ void {cs.cname}_store(const onc_dsim& S, onc_gcol const* c, onc_u32 i, {cs.cname} v) {{\n"
                   "    v.pos.x = onc_clampf(v.pos.x, 0.0f, (float)S.dims[0] * S.dx);\n"
                   "    v.pos.y = onc_clampf(v.pos.y, 0.0f, (float)S.dims[1] * S.dx);\n"
                   "    v.pos.z = onc_clampf(v.pos.z, 0.0f, (float)S.dims[2] * S.dx);\n"
                   + "\n".join(stl) + "\n}\n")
        if self.tissue.cls is not None:
            ts = StructT("tissue", self.tissue.cls)
            tl, tw = [], []
            for f, (n, t, _) in enumerate(self.tissue.fields):
                tl.append(f"    r.{n} = {self.conv(f'S.fld[buf][{f}][r._v]', F32, t)};")
                tw.append(f"    S.fld[buf][{f}][r._v] = {self.conv(f'r.{n}', t, F32)};")
            out.append(f"__device__ __forceinline__ {ts.cname} {ts.cname}_loadv(RuleCtx& ctx, onc_u32 buf, int x, int y, int z) {{\n"
                       f"    const onc_dsim& S = *ctx.S;\n    {ts.cname} r;\n    r._coord = onc_iv3_make(x, y, z);\n"
                       f"    r._v = (onc_u32)x + (onc_u32)S.dims[0] * ((onc_u32)y + (onc_u32)S.dims[1] * (onc_u32)z);\n"
                       + "\n".join(tl) + "\n    return r;\n}\n")
            out.append(f"__device__ __forceinline__ void {ts.cname}_storev(RuleCtx& ctx, onc_u32 buf, const {ts.cname}& r) {{\n"
                       f"    const onc_dsim& S = *ctx.S;\n" + "\n".join(tw) + "\n}\n")
            out.append(f"__device__ __forceinline__ {ts.cname} {ts.cname}_at(RuleCtx& ctx, onc_v3 p) {{\n"
                       f"    const onc_dsim& S = *ctx.S;\n"
                       f"    return {ts.cname}_loadv(ctx, ctx.tin, onc_vox1(p.x, S.dx, S.dims[0]), "
                       f"onc_vox1(p.y, S.dx, S.dims[1]), onc_vox1(p.z, S.dx, S.dims[2]));\n}}\n")
        if self.chem.cls is not None:
            ch = StructT("chem", self.chem.cls)
            cl, cw = [], []
            for k, (n, t) in enumerate(self.chem.fields):
                w = f"S.chem[buf][{k}][r._v]"
                if t == F32:
                    cl.append(f"    r.{n} = __uint_as_float({w});")
                    cw.append(f"    {w} = __float_as_uint(r.{n});")
                elif t == I32:
                    cl.append(f"    r.{n} = (int){w};")
                    cw.append(f"    {w} = (onc_u32)r.{n};")
                elif t == BOOL:
                    cl.append(f"    r.{n} = {w} != 0u;")
                    cw.append(f"    {w} = r.{n} ? 1u : 0u;")
                else:
                    cl.append(f"    r.{n} = {w};")
                    cw.append(f"    {w} = r.{n};")
            out.append(f"__device__ __forceinline__ {ch.cname} {ch.cname}_loadv(RuleCtx& ctx, onc_u32 buf, int x, int y, int z) {{\n"
                       f"    const onc_dsim& S = *ctx.S;\n    {ch.cname} r;\n    r._coord = onc_iv3_make(x, y, z);\n"
                       f"    r._v = (onc_u32)x + (onc_u32)S.cdims[0] * ((onc_u32)y + (onc_u32)S.cdims[1] * (onc_u32)z);\n"
                       + "\n".join(cl) + "\n    return r;\n}\n")
            out.append(f"__device__ __forceinline__ void {ch.cname}_storev(RuleCtx& ctx, onc_u32 buf, const {ch.cname}& r) {{\n"
                       f"    const onc_dsim& S = *ctx.S;\n" + "\n".join(cw) + "\n}\n")
            # voxel of pos: floor(pos / (2 dx)) clamped (core/utils/_types.py:332-346)
            out.append(f"__device__ __forceinline__ {ch.cname} {ch.cname}_at(RuleCtx& ctx, onc_v3 p) {{\n"
                       f"    const onc_dsim& S = *ctx.S;\n    const float cdx = 2.0f * S.dx;\n"
                       f"    return {ch.cname}_loadv(ctx, ctx.cin, onc_vox1(p.x, cdx, S.cdims[0]), "
                       f"onc_vox1(p.y, cdx, S.cdims[1]), onc_vox1(p.z, cdx, S.cdims[2]));\n}}\n")
        for key in self.order:
            out.append(self.funcs[key].proto)
        for key in self.order:
            out.append(self.funcs[key].code)
        calls = "\n".join(f"        {r}(cell, ctx);" for r in rule_names)
        out.append(f"""extern "C" __global__ __launch_bounds__(256) void {self.KERNEL}(onc_rule_args r) {{
    onc_for_cells(r.cell_prefix, r.nsims, [&](int s, onc_u32 i) {{
        const onc_dsim& S = r.sims[s];
        RuleCtx ctx;
        ctx.S = &S;
        ctx.par = S.par;
        ctx.i = i;
        ctx.bin_tag = onc_bin_tag(S.step);
        ctx.tin = ctx.par ^ 1u;
        ctx.cin = (S.step * S.chem_passes + S.chem_passes) & 1u;
        ctx.alive = true;
        ctx.div = false;
        {cs.cname} cell = {cs.cname}_load(S, S.st[ctx.par ^ 1u], i);
        cell.pos.x = __uint_as_float(S.st[ctx.par][0][i]);
        cell.pos.y = __uint_as_float(S.st[ctx.par][1][i]);
        cell.pos.z = __uint_as_float(S.st[ctx.par][2][i]);
        onc_rng_init(ctx.rng, S, cell._id, S.step);
{calls}
        {cs.cname}_store(S, S.st[ctx.par], i, cell);
        S.flags[i] = (ctx.alive ? 1u : 0u) | (ctx.div ? 2u : 0u);
        if (ctx.div) {cs.cname}_store(S, S.dau, i, ctx.daughter);
    }});
}}
""")
        out.append(self._grid_kernels())
        if metric_names:
            out.append(self._metric_kernels(metric_names))
        return "\n".join(out)

    def _grid_kernels(self) -> str:
        """onc_tissue_pass / onc_chem_pass and the pass counts the runtime
        reads (onc_sim_set_rule). Tissue pass j runs tissue rule j; chemistry
        passes run each chemistry rule `iterations` times in declaration order
        (gpu/runtime/_gpu_backend.py:127-142). Every pass is Jacobi: voxels read
        one buffer and write the other. Random draws of voxel v: Philox
        ctr = {v, step, call >> 2, tag}, tag 0x10000 + j (tissue pass j) or
        0x20000 + p (chemistry pass p); cells draw with tag 0."""
        tissue_names, chem_names = getattr(self, "grid_rules", ([], []))
        passes = self.chem_pass_rules
        out = [f'extern "C" __device__ onc_u32 onc_grid_tissue_passes = {len(tissue_names)}u;',
               f'extern "C" __device__ onc_u32 onc_grid_chem_passes = {len(passes)}u;', ""]
        if tissue_names:
            ts = StructT("tissue", self.tissue.cls)
            cases = "\n".join(f"            case {j}u: {n}(t, ctx); break;" for j, n in enumerate(tissue_names))
            out.append(f"""extern "C" __global__ __launch_bounds__(256) void {self.TISSUE_KERNEL}(onc_grid_args a) {{
    const onc_dsim& S = a.sims[blockIdx.y];
    if (S.frozen) return;
    const onc_u32 X = (onc_u32)S.dims[0], Y = (onc_u32)S.dims[1], Z = (onc_u32)S.dims[2];
    const onc_u32 nvox = X * Y * Z;
    RuleCtx ctx;
    ctx.S = &S;
    ctx.par = S.par;
    ctx.i = 0xFFFFFFFFu;
    ctx.bin_tag = onc_bin_tag(S.step);
    ctx.tin = S.par ^ 1u ^ (a.pass & 1u);
    ctx.cin = (S.step * S.chem_passes) & 1u;
    ctx.alive = true;
    ctx.div = false;
    const onc_u32 tout = S.par ^ (a.pass & 1u);
    for (onc_u32 v = blockIdx.x * 256u + threadIdx.x; v < nvox; v += gridDim.x * 256u) {{
        const onc_u32 x = v % X, y = (v / X) % Y, z = v / (X * Y);
        {ts.cname} t = {ts.cname}_loadv(ctx, ctx.tin, (int)x, (int)y, (int)z);
        onc_rng_init(ctx.rng, S, v, S.step);
        ctx.rng.tag = 0x10000u + a.pass;
        switch (a.pass) {{
{cases}
            default: break;
        }}
        {ts.cname}_storev(ctx, tout, t);
    }}
}}
""")
        if passes:
            chs = StructT("chem", self.chem.cls)
            cases = "\n".join(f"            case {p}u: {n}(c, ctx); break;" for p, n in enumerate(passes))
            out.append(f"""extern "C" __global__ __launch_bounds__(256) void {self.CHEM_KERNEL}(onc_grid_args a) {{
    const onc_dsim& S = a.sims[blockIdx.y];
    if (S.frozen) return;
    const onc_u32 X = (onc_u32)S.cdims[0], Y = (onc_u32)S.cdims[1], Z = (onc_u32)S.cdims[2];
    const onc_u32 nvox = X * Y * Z;
    RuleCtx ctx;
    ctx.S = &S;
    ctx.par = S.par;
    ctx.i = 0xFFFFFFFFu;
    ctx.bin_tag = onc_bin_tag(S.step);
    ctx.tin = S.par ^ 1u;
    ctx.cin = (S.step * S.chem_passes + a.pass) & 1u;
    ctx.alive = true;
    ctx.div = false;
    const onc_u32 cout = ctx.cin ^ 1u;
    for (onc_u32 v = blockIdx.x * 256u + threadIdx.x; v < nvox; v += gridDim.x * 256u) {{
        const onc_u32 x = v % X, y = (v / X) % Y, z = v / (X * Y);
        {chs.cname} c = {chs.cname}_loadv(ctx, ctx.cin, (int)x, (int)y, (int)z);
        onc_rng_init(ctx.rng, S, v, S.step);
        ctx.rng.tag = 0x20000u + a.pass;
        switch (a.pass) {{
{cases}
            default: break;
        }}
        {chs.cname}_storev(ctx, cout, c);
    }}
}}
""")
        return "\n".join(out)

    def _metric_kernels(self, metric_names: list[str]) -> str:
        """onc_metric_main / onc_metric_final (onc_sim_metrics): per-thread
        sums in cell order, then fixed trees (the order in onc_device.h)."""
        cs = StructT("cell", self.cell.cls)
        mt = next(st for st in self.structs.values() if st.kind == "metrics")
        nv = len(self.metric_fields)

        def bits(t, v):
            return f"__float_as_uint({v})" if t == F32 else f"(onc_u32)({v})"

        def add(t, a, b):
            if t == F32:
                return f"__float_as_uint(__uint_as_float({a}) + __uint_as_float({b}))"
            return f"{a} + {b}"

        zero = "\n".join(f"    m.{n} = {self.zero(t)};" for n, t in self.metric_fields)
        calls = "\n".join(f"        {r}(cell, m, ctx);" for r in metric_names)
        tree = []
        for k, (n, t) in enumerate(self.metric_fields):
            tree.append(f"""    s_v[t] = {bits(t, f"m.{n}")};
    __syncthreads();
    for (onc_u32 h = 128u; h > 0u; h >>= 1) {{
        if (t < h) s_v[t] = {add(t, "s_v[t]", "s_v[t + h]")};
        __syncthreads();
    }}
    if (t == 0u) a.partials[(size_t)blockIdx.x * {nv}u + {k}u] = s_v[0];
    __syncthreads();""")
        fin = []
        for k, (n, t) in enumerate(self.metric_fields):
            fin.append(f"""    {{
        onc_u32 v = 0u;
        for (onc_u32 q = t; q < a.nblocks; q += 256u) v = {add(t, "v", f"a.partials[(size_t)q * {nv}u + {k}u]")};
        s_v[t] = v;
        __syncthreads();
        for (onc_u32 h = 128u; h > 0u; h >>= 1) {{
            if (t < h) s_v[t] = {add(t, "s_v[t]", "s_v[t + h]")};
            __syncthreads();
        }}
        if (t == 0u) a.out[{k}u] = s_v[0];
        __syncthreads();
    }}""")
        return (f"""extern "C" __device__ onc_u32 onc_metric_nvals = {nv}u;  // words per block and in out (onc_sim_metrics checks)

extern "C" __global__ __launch_bounds__(256) void onc_metric_main(onc_metric_args a) {{
    __shared__ onc_u32 s_v[256];
    const onc_dsim& S = *a.S;
    const onc_u32 t = threadIdx.x;
    {mt.cname} m;
{zero}
    for (onc_u32 j = 0; j < ONC_METRIC_CPT; j++) {{
        const onc_u32 i = (blockIdx.x * ONC_METRIC_CPT + j) * 256u + t;
        if (i >= a.n) break;
        RuleCtx ctx;
        ctx.S = &S;
        ctx.par = a.parity ^ 1u;
        ctx.i = i;
        ctx.bin_tag = 0u;
        ctx.tin = a.parity;
        ctx.cin = 0u;
        ctx.alive = true;
        ctx.div = false;
        {cs.cname} cell = i < a.split ? {cs.cname}_load(S, S.st[a.parity], i) : {cs.cname}_load(S, S.dau2, i - a.split);
{calls}
    }}
{chr(10).join(tree)}
}}

extern "C" __global__ __launch_bounds__(256) void onc_metric_final(onc_metric_fin_args a) {{
    __shared__ onc_u32 s_v[256];
    const onc_u32 t = threadIdx.x;
{chr(10).join(fin)}
}}
""")
