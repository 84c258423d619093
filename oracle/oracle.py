"""ctypes wrapper of the CPU oracle (liboncooracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg as the checker; the product package never
imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboncooracle.so")
# A planted-mistake build (tests/test_oracle_mutants.py) is loaded in place of
# the oracle through this variable; it is never set otherwise.
_PLANTED = os.environ.get("ONCO_ORACLE_PLANTED_LIB")

_lib = None


def build() -> str:
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if _PLANTED:
            L = C.CDLL(_PLANTED)
        else:
            if not os.path.exists(LIB_PATH):
                build()
            L = C.CDLL(LIB_PATH)
        P = C.c_void_p
        u32 = C.c_uint32
        L.oracle_morton30.restype = u32
        L.oracle_morton30.argtypes = [u32, u32, u32]
        L.oracle_morton32.restype = u32
        L.oracle_morton32.argtypes = [u32, u32, u32]
        L.oracle_morton_batch.argtypes = [P, P, P, u32, C.c_int, P]
        L.oracle_key_bits.restype = C.c_int
        L.oracle_key_bits.argtypes = [P]
        L.oracle_radix_sort_pairs.argtypes = [P, P, u32, C.c_int]
        L.oracle_exscan_u32.restype = u32
        L.oracle_exscan_u32.argtypes = [P, P, u32]
        L.oracle_bins.argtypes = [P, u32, P, P]
        L.oracle_philox_batch.argtypes = [P, u32, C.c_uint64, P]
        L.oracle_threads.restype = C.c_int
        L.oracle_threads.argtypes = []
        L.oracle_set_threads.argtypes = [C.c_int]
        for sfx, real in (("f32", C.c_float), ("f64", C.c_double)):
            getattr(L, f"oracle_keys_{sfx}").argtypes = [P, P, P, u32, P, real, C.c_int, P]
            getattr(L, f"oracle_sincos_q_{sfx}").argtypes = [real, P, P]
            getattr(L, f"oracle_random_dir_{sfx}").argtypes = [u32, u32, C.c_uint64, u32, P]
            getattr(L, f"oracle_field_init_{sfx}").argtypes = [P, P, P]
            getattr(L, f"oracle_stencil7_{sfx}").argtypes = [P, P, P, P]
            getattr(L, f"oracle_forces_raw_{sfx}").argtypes = [P, u32, P, P, P, P, P, P]
            getattr(L, f"oracle_tumor_run_{sfx}").argtypes = [P, P, P, P, P, P, P, P, P, P, P, C.c_int]
            getattr(L, f"oracle_tumor_run_{sfx}").restype = C.c_int
            getattr(L, f"oracle_pre_rule_{sfx}").argtypes = [P, u32, P, P, P, P, P, P, P, P, P, P, P, P]
            getattr(L, f"oracle_pre_rule_{sfx}").restype = C.c_int
            getattr(L, f"oracle_dir_from_uniforms_{sfx}").argtypes = [real, real, P]
        _lib = L
    return _lib


def _p(a: np.ndarray) -> int:
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data


def _dims(d) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(d, dtype=np.int32))


def _real(sfx: str):
    return np.float32 if sfx == "f32" else np.float64


def morton(x, y, z, fmt: int = 30) -> np.ndarray:
    x = np.ascontiguousarray(x, dtype=np.uint32)
    y = np.ascontiguousarray(y, dtype=np.uint32)
    z = np.ascontiguousarray(z, dtype=np.uint32)
    out = np.empty(len(x), dtype=np.uint32)
    lib().oracle_morton_batch(_p(x), _p(y), _p(z), len(x), fmt, _p(out))
    return out


def key_bits(dims) -> int:
    return int(lib().oracle_key_bits(_p(_dims(dims))))


def keys(px, py, pz, dims, dx: float, fmt: int = 30, real: str = "f32") -> np.ndarray:
    t = _real(real)
    px, py, pz = (np.ascontiguousarray(a, dtype=t) for a in (px, py, pz))
    out = np.empty(len(px), dtype=np.uint32)
    getattr(lib(), f"oracle_keys_{real}")(_p(px), _p(py), _p(pz), len(px), _p(_dims(dims)), dx, fmt, _p(out))
    return out


def radix_sort_pairs(k, v, key_bits_: int) -> tuple[np.ndarray, np.ndarray]:
    k = np.array(k, dtype=np.uint32, copy=True)
    v = np.array(v, dtype=np.uint32, copy=True)
    lib().oracle_radix_sort_pairs(_p(k), _p(v), len(k), key_bits_)
    return k, v


def exscan(a) -> tuple[np.ndarray, int]:
    a = np.ascontiguousarray(a, dtype=np.uint32)
    out = np.empty_like(a)
    tot = lib().oracle_exscan_u32(_p(a), _p(out), len(a))
    return out, int(tot)


def metric_reduce(contrib, kind: str, cpt: int = 8) -> int:
    """Sum of per-cell metric contributions in the order of the metric
    kernels (onc_metric_main / onc_metric_final); returns the 32-bit word.

    The reference adds metrics with atomicAdd / atomicSub
    (gpu/compiler/translator/_stmt_translator.py:240-268; its f32 helper is
    never defined) after clearing them (gpu/runtime/_gpu_backend.py:118-120,
    157-166), and the CPU backend loops over the cells
    (cpu/_cpu_backend.py:65-70). Integer sums wrap mod 2^32 in any order;
    binary32 sums here follow the fixed order: block b holds cells
    [b, b+1) x 256 cpt, thread t adds cells (b cpt + j) 256 + t for
    j = 0..cpt-1 from +0.0, the block adds thread values in a tree (t += t + h
    for h = 128, 64, ..., 1), and the last stage adds block partials
    t, t + 256, ... per thread from +0.0, then the same tree."""
    c = np.asarray(contrib)
    if kind == "f32":
        dt = np.float32
        c = c.astype(np.float32)
    elif kind in ("i32", "u32"):
        dt = np.uint32
        c = (c.astype(np.int64) & 0xFFFFFFFF).astype(np.uint32)
    else:
        raise ValueError(kind)
    n = len(c)
    per = 256 * cpt
    nb = max(1, -(-n // per))
    x = np.zeros(nb * per, dtype=dt)
    x[:n] = c
    x = x.reshape(nb, cpt, 256)
    acc = np.zeros((nb, 256), dtype=dt)
    with np.errstate(over="ignore"):
        for j in range(cpt):
            acc = acc + x[:, j, :]

        def tree(v):
            h = v.shape[-1] // 2
            while h >= 1:
                v[..., :h] = v[..., :h] + v[..., h:2 * h]
                h //= 2
            return v[..., 0]

        part = tree(acc)
        v = np.zeros(256, dtype=dt)
        for q0 in range(0, nb, 256):
            ch = part[q0:q0 + 256]
            v[:len(ch)] = v[:len(ch)] + ch
        r = tree(v)
    return int(np.asarray(r, dtype=dt).view(np.uint32))


def bins(sorted_keys, table: int, fill: int = 0xFFFFFFFF) -> tuple[np.ndarray, np.ndarray]:
    k = np.ascontiguousarray(sorted_keys, dtype=np.uint32)
    s = np.full(table, fill, dtype=np.uint32)
    e = np.full(table, fill, dtype=np.uint32)
    lib().oracle_bins(_p(k), len(k), _p(s), _p(e))
    return s, e


def philox(ctr: np.ndarray, key: int) -> np.ndarray:
    ctr = np.ascontiguousarray(ctr, dtype=np.uint32).reshape(-1, 4)
    out = np.empty_like(ctr)
    lib().oracle_philox_batch(_p(ctr), len(ctr), key, _p(out))
    return out


def sincos_q(a: float, real: str = "f32") -> tuple[float, float]:
    t = C.c_float if real == "f32" else C.c_double
    s, c = t(), t()
    getattr(lib(), f"oracle_sincos_q_{real}")(a, C.byref(s), C.byref(c))
    return s.value, c.value


def random_dir(cell_id: int, step: int, seed: int, c0: int = 0, real: str = "f32") -> np.ndarray:
    out = np.empty(3, dtype=_real(real))
    getattr(lib(), f"oracle_random_dir_{real}")(cell_id, step, seed, c0, _p(out))
    return out


def field_init(fd, dims, real: str = "f32") -> np.ndarray:
    out = np.empty(int(np.prod(dims)), dtype=_real(real))
    getattr(lib(), f"oracle_field_init_{real}")(C.byref(fd), _p(_dims(dims)), _p(out))
    return out


def stencil7(u: np.ndarray, dims, fd, real: str = "f32") -> np.ndarray:
    u = np.ascontiguousarray(u, dtype=_real(real)).reshape(-1)
    out = np.empty_like(u)
    getattr(lib(), f"oracle_stencil7_{real}")(_p(u), _p(out), _p(_dims(dims)), C.byref(fd))
    return out


def forces_raw(desc, px, py, pz, real: str = "f32"):
    t = _real(real)
    px, py, pz = (np.ascontiguousarray(a, dtype=t) for a in (px, py, pz))
    n = len(px)
    fx, fy, fz = (np.empty(n, dtype=t) for _ in range(3))
    getattr(lib(), f"oracle_forces_raw_{real}")(C.byref(desc), n, _p(px), _p(py), _p(pz), _p(fx), _p(fy), _p(fz))
    return fx, fy, fz


def pre_rule(desc, px, py, pz, fields, real: str = "f32"):
    """Steps 1-6 of the timestep (sort, uptake, forces + integration, one
    sweep per field) for a rule of any form. fields: list of arrays, swept in
    place. Returns perm, sorted keys, sorted positions before the forces and
    integrated positions (sorted order)."""
    t = _real(real)
    px, py, pz = (np.ascontiguousarray(a, dtype=t) for a in (px, py, pz))
    n = len(px)
    perm = np.empty(n, np.uint32)
    keys_ = np.empty(n, np.uint32)
    sp = [np.empty(n, t) for _ in range(3)]
    ip = [np.empty(n, t) for _ in range(3)]
    for f in fields:
        assert f.dtype == t and f.flags["C_CONTIGUOUS"]
    fp = (C.c_void_p * max(1, len(fields)))(*[_p(f) for f in fields])
    rc = getattr(lib(), f"oracle_pre_rule_{real}")(C.byref(desc), n, _p(px), _p(py), _p(pz), _p(perm), _p(keys_),
                                                     *[_p(a) for a in sp], *[_p(a) for a in ip], fp)
    assert rc == 0
    return perm, keys_, sp, ip


def dir_from_uniforms(u1, u2, real: str = "f32") -> np.ndarray:
    out = np.empty(3, dtype=_real(real))
    getattr(lib(), f"oracle_dir_from_uniforms_{real}")(u1, u2, _p(out))
    return out


class TumorState:
    """Oracle-side state of the built-in tumour model (capacity-sized arrays)."""

    def __init__(self, desc, cols, fields=None, real: str = "f32", next_id=None):
        t = _real(real)
        self.desc = desc
        self.real = real
        cap = int(desc.capacity)
        n = len(cols[0])
        self.n = C.c_uint32(n)
        self.next_id = C.c_uint32(int(cols[3].max()) + 1 if next_id is None and n else (next_id or 0))
        self.step = C.c_uint32(0)
        self.px, self.py, self.pz, self.age = (np.zeros(cap, dtype=t) for _ in range(4))
        self.id = np.zeros(cap, dtype=np.uint32)
        self.gen = np.zeros(cap, dtype=np.uint32)
        self.px[:n], self.py[:n], self.pz[:n] = cols[0], cols[1], cols[2]
        self.id[:n] = cols[3]
        self.age[:n] = cols[4]
        self.gen[:n] = cols[5]
        dims = tuple(desc.dims)
        self.fields = []
        for f in range(desc.n_fields):
            if fields is not None:
                self.fields.append(np.ascontiguousarray(fields[f], dtype=t).reshape(-1).copy())
            else:
                self.fields.append(field_init(desc.fields[f], dims, real))
        self._fp = (C.c_void_p * max(1, len(self.fields)))(*[_p(f) for f in self.fields])

    def run(self, nsteps: int) -> None:
        fn = getattr(lib(), f"oracle_tumor_run_{self.real}")
        rc = fn(C.byref(self.desc), C.byref(self.n), C.byref(self.next_id), C.byref(self.step),
                _p(self.px), _p(self.py), _p(self.pz), _p(self.id), _p(self.age), _p(self.gen),
                self._fp, nsteps)
        assert rc == 0

    def columns(self) -> list[np.ndarray]:
        n = self.n.value
        return [self.px[:n].copy(), self.py[:n].copy(), self.pz[:n].copy(), self.id[:n].copy(),
                self.age[:n].copy(), self.gen[:n].copy()]


def threads() -> int:
    """Threads the oracle's per-cell and per-voxel loops use (OpenMP)."""
    return int(lib().oracle_threads())


def set_threads(n: int) -> None:
    lib().oracle_set_threads(int(n))
