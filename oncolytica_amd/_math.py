"""ol.math: math intrinsics usable in rules (compiled to HIP device math) and
on the host (Python implementations with the same meaning), mirroring
framework/src/oncolytica/core/utils/_math.py."""
from __future__ import annotations

import builtins as _b
import math as _m

from ._types import vec3


def length(v: vec3) -> float:
    return _m.sqrt(v.x * v.x + v.y * v.y + v.z * v.z)


def length_sq(v: vec3) -> float:
    return v.x * v.x + v.y * v.y + v.z * v.z


def distance(a: vec3, b: vec3) -> float:
    return length(a - b)


def distance_sq(a: vec3, b: vec3) -> float:
    return length_sq(a - b)


def normalize(v: vec3) -> vec3:
    n = length(v)
    return vec3(v.x / n, v.y / n, v.z / n) if n > 0.0 else vec3()


def dot(a: vec3, b: vec3) -> float:
    return a.x * b.x + a.y * b.y + a.z * b.z


def cross(a: vec3, b: vec3) -> vec3:
    return vec3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x)


def reflect(v: vec3, normal: vec3) -> vec3:
    d = 2.0 * dot(v, normal)
    return vec3(v.x - d * normal.x, v.y - d * normal.y, v.z - d * normal.z)


def lerp_vec(a: vec3, b: vec3, t: float) -> vec3:
    return vec3(a.x + (b.x - a.x) * t, a.y + (b.y - a.y) * t, a.z + (b.z - a.z) * t)


def clamp(value, min_val, max_val):
    if isinstance(value, vec3):
        return vec3(*(_b.max(lo, _b.min(hi, x)) for x, lo, hi in zip(value, min_val, max_val)))
    return _b.max(min_val, _b.min(max_val, value))


def lerp(a: float, b: float, t: float) -> float:
    return a + (b - a) * t


def sign(x: float) -> float:
    return 1.0 if x > 0 else (-1.0 if x < 0 else 0.0)


def step(edge: float, x: float) -> float:
    return 0.0 if x < edge else 1.0


def smoothstep(edge0: float, edge1: float, x: float) -> float:
    t = clamp((x - edge0) / (edge1 - edge0), 0.0, 1.0)
    return t * t * (3.0 - 2.0 * t)


def fract(x: float) -> float:
    return x - _m.floor(x)


def inverse_sqrt(x: float) -> float:
    return 1.0 / _m.sqrt(x)


sqrt = _m.sqrt
exp = _m.exp
exp2 = lambda x: 2.0 ** x  # noqa: E731
log = _m.log
log2 = _m.log2
pow = _m.pow  # noqa: A001
sin = _m.sin
cos = _m.cos
tan = _m.tan
asin = _m.asin
acos = _m.acos
atan = _m.atan
atan2 = _m.atan2
sinh = _m.sinh
cosh = _m.cosh
tanh = _m.tanh
floor = _m.floor
ceil = _m.ceil
trunc = _m.trunc
round = _b.round  # noqa: A001
fabs = _m.fabs
abs = _b.abs  # noqa: A001
min = _b.min  # noqa: A001
max = _b.max  # noqa: A001
pi = _m.pi
tau = _m.tau
inf = _m.inf
