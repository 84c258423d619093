"""ctypes binding of liboncohip.so (the C ABI declared in include/onco.h).

The library is mandatory: if it is missing or fails to load, every call
raises. There is no CPU fallback in the product path.
"""
from __future__ import annotations

import ctypes as C
import os
import threading
import weakref

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "liboncohip.so")

MAX_FIELDS = 4
MAX_COLS = 48
MAX_PARAMS = 32
MAX_CHEM_COLS = 16
PHASES = 10

OK = 0
ERR_INVALID = -1
ERR_HIP = -2
ERR_NOMEM = -3
ERR_COMPILE = -4
ERR_CAPACITY = -5
ERR_STATE = -6
ERR_DEVICE = -7

KEY_MORTON30 = 30
KEY_MORTON32 = 32
BC_DIRICHLET = 0
BC_POINT_SOURCE = 1
BC_NONE = 2
COL_F32 = 0
COL_I32 = 1
COL_U32 = 2

PHASE_NAMES = ["plan", "morton", "sort", "permute_bins", "forces", "stencil", "rule", "compact",
               "forces_stencil", "rule_compact"]


class OncError(RuntimeError):
    """A liboncohip call returned a negative status."""

    def __init__(self, code: int, msg: str):
        super().__init__(f"liboncohip error {code}: {msg}")
        self.code = code


class FieldDesc(C.Structure):
    _fields_ = [
        ("rate", C.c_float), ("decay", C.c_float), ("init", C.c_float), ("bc", C.c_int32),
        ("bc_value", C.c_float), ("src_radius", C.c_float), ("uptake", C.c_float),
        ("lo", C.c_float), ("hi", C.c_float),
    ]


class SimDesc(C.Structure):
    _fields_ = [
        ("dims", C.c_int32 * 3), ("dx", C.c_float), ("key_format", C.c_int32),
        ("capacity", C.c_uint32), ("seed", C.c_uint64),
        ("cell_radius", C.c_float), ("interaction_radius", C.c_float),
        ("k_rep", C.c_float), ("k_adh", C.c_float), ("mobility", C.c_float), ("max_disp", C.c_float),
        ("n_fields", C.c_int32), ("fields", FieldDesc * MAX_FIELDS),
        ("n_cols", C.c_int32), ("col_types", C.c_int32 * MAX_COLS),
        ("n_params", C.c_int32), ("params", C.c_float * MAX_PARAMS),
        ("n_chem_cols", C.c_int32), ("chem_col_types", C.c_int32 * MAX_CHEM_COLS),
    ]


class DeviceInfo(C.Structure):
    _fields_ = [
        ("name", C.c_char * 64), ("arch", C.c_char * 32), ("compute_units", C.c_int32),
        ("xcds", C.c_int32), ("hbm_bytes", C.c_uint64), ("free_bytes", C.c_uint64),
    ]


_P = C.c_void_p
_u32 = C.c_uint32
_i32p = C.POINTER(C.c_int32)

_PROTOS = {
    "onc_last_error": (C.c_char_p, []),
    "onc_abi_version": (C.c_int, []),
    "onc_ctx_create": (C.c_int, [C.c_int, C.POINTER(_P)]),
    "onc_ctx_destroy": (C.c_int, [_P]),
    "onc_ctx_info": (C.c_int, [_P, C.POINTER(DeviceInfo)]),
    "onc_dev_alloc": (C.c_int, [_P, C.c_size_t, C.POINTER(_P)]),
    "onc_dev_free": (C.c_int, [_P, _P]),
    "onc_host_alloc": (C.c_int, [_P, C.c_size_t, C.POINTER(_P)]),
    "onc_host_free": (C.c_int, [_P, _P]),
    "onc_h2d": (C.c_int, [_P, _P, _P, C.c_size_t]),
    "onc_d2h": (C.c_int, [_P, _P, _P, C.c_size_t]),
    "onc_ctx_sync": (C.c_int, [_P]),
    "onc_module_build": (C.c_int, [_P, C.c_char_p, C.POINTER(C.c_char_p), C.c_int, C.POINTER(_P)]),
    "onc_module_log": (C.c_int, [_P, C.c_char_p, C.c_size_t]),
    "onc_module_destroy": (C.c_int, [_P]),
    "onc_sim_create": (C.c_int, [_P, C.POINTER(SimDesc), C.POINTER(_P)]),
    "onc_sim_destroy": (C.c_int, [_P]),
    "onc_sim_set_rule": (C.c_int, [_P, _P, C.c_char_p]),
    "onc_sim_set_params": (C.c_int, [_P, C.POINTER(C.c_float), C.c_int]),
    "onc_sim_upload_cells": (C.c_int, [_P, C.POINTER(_P), _u32, _u32]),
    "onc_sim_download_cells": (C.c_int, [_P, C.POINTER(_P), _u32, C.POINTER(_u32)]),
    "onc_sim_upload_field": (C.c_int, [_P, C.c_int, _P]),
    "onc_sim_download_field": (C.c_int, [_P, C.c_int, _P]),
    "onc_sim_upload_chem": (C.c_int, [_P, C.c_int, _P]),
    "onc_sim_download_chem": (C.c_int, [_P, C.c_int, _P]),
    "onc_sim_step": (C.c_int, [_P, C.c_int]),
    "onc_sim_sync": (C.c_int, [_P]),
    "onc_sim_counts": (C.c_int, [_P, C.POINTER(_u32), C.POINTER(_u32), C.POINTER(_u32)]),
    "onc_sim_set_stop": (C.c_int, [_P, C.c_uint32]),
    "onc_sim_set_step": (C.c_int, [_P, C.c_uint32]),
    "onc_sim_metrics": (C.c_int, [_P, _P, C.c_int, C.POINTER(_u32)]),
    "onc_sim_footprint": (C.c_int, [C.POINTER(SimDesc), C.POINTER(C.c_uint64)]),
    "onc_sim_cell_updates": (C.c_int, [_P, C.POINTER(C.c_uint64)]),
    "onc_batch_create": (C.c_int, [_P, C.POINTER(_P)]),
    "onc_batch_destroy": (C.c_int, [_P]),
    "onc_batch_add": (C.c_int, [_P, _P]),
    "onc_batch_remove": (C.c_int, [_P, _P]),
    "onc_batch_size": (C.c_int, [_P, C.POINTER(C.c_int)]),
    "onc_batch_step": (C.c_int, [_P, C.c_int, C.c_int]),
    "onc_batch_sync": (C.c_int, [_P]),
    "onc_batch_set_tuning": (C.c_int, [_P, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]),
    "onc_batch_set_tile_policy": (C.c_int, [_P, C.c_uint32, C.c_uint32]),
    "onc_batch_tile_items": (C.c_int, [_P, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "onc_batch_set_fused": (C.c_int, [_P, C.c_int]),
    "onc_batch_set_fused_rule": (C.c_int, [_P, C.c_int]),
    "onc_batch_set_forces_waves": (C.c_int, [_P, C.c_int]),
    "onc_batch_profile": (C.c_int, [_P, C.c_int]),
    "onc_batch_profile_every": (C.c_int, [_P, C.c_int, C.c_int]),
    "onc_batch_phase_times": (C.c_int, [_P, C.POINTER(C.c_double), C.POINTER(C.c_uint64), C.c_int]),
    "onc_batch_sweep_voxels": (C.c_int, [_P, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
    "onc_morton_keys": (C.c_int, [_P, _P, _P, _P, _u32, _i32p, C.c_float, C.c_int, _P]),
    "onc_sort_pairs": (C.c_int, [_P, _P, _P, _u32, C.c_int, _P, _P]),
    "onc_forces": (C.c_int, [_P, C.POINTER(SimDesc), _P, _P, _P, _u32, _P, _P, _P]),
    "onc_exscan_u32": (C.c_int, [_P, _P, _P, _u32, _P]),
    "onc_exscan_u32_noalloc": (C.c_int, [_P, _P, _P, _u32, _P]),
    "onc_bins": (C.c_int, [_P, _P, _u32, _P, _P]),
    "onc_stencil7": (C.c_int, [_P, _P, _P, _i32p, C.POINTER(FieldDesc)]),
    "onc_philox": (C.c_int, [_P, _P, _u32, C.c_uint64, _P]),
    "onc_copy_u32": (C.c_int, [_P, _P, _P, C.c_uint64]),
    "onc_ctx_timing": (C.c_int, [_P, C.c_int]),
    "onc_ctx_last_kernel_us": (C.c_int, [_P, C.POINTER(C.c_double)]),
}

_lib = None
_lock = threading.Lock()


def lib() -> C.CDLL:
    """Load liboncohip.so once; raises if it is absent (build it first)."""
    global _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                raise OSError(f"{LIB_PATH} is missing: build it with oncolytica_amd._build.build()")
            L = C.CDLL(LIB_PATH)
            for name, (res, args) in _PROTOS.items():
                fn = getattr(L, name)
                fn.restype = res
                fn.argtypes = args
            _lib = L
    return _lib


def declared_symbols() -> list[str]:
    return sorted(_PROTOS)


def check(rc: int) -> None:
    if rc != OK:
        msg = lib().onc_last_error()
        raise OncError(rc, msg.decode(errors="replace") if msg else "")


def call(name: str, *args) -> None:
    check(getattr(lib(), name)(*args))


def ptr(a: np.ndarray) -> int:
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data


class Context:
    """One HIP device (onc_ctx)."""

    def __init__(self, device: int = 0):
        h = _P()
        call("onc_ctx_create", device, C.byref(h))
        self.h = h
        self.device = device

    def info(self) -> DeviceInfo:
        out = DeviceInfo()
        call("onc_ctx_info", self.h, C.byref(out))
        return out

    def close(self) -> None:
        if self.h:
            lib().onc_ctx_destroy(self.h)
            self.h = None

    # device buffers for the pure-kernel entry points
    def alloc(self, nbytes: int) -> "DeviceBuffer":
        return DeviceBuffer(self, nbytes)

    def upload(self, a: np.ndarray) -> "DeviceBuffer":
        a = np.ascontiguousarray(a)
        b = DeviceBuffer(self, a.nbytes)
        call("onc_h2d", self.h, b.p, ptr(a), a.nbytes)
        return b

    def sync(self) -> None:
        call("onc_ctx_sync", self.h)

    def timing(self, on: bool = True) -> None:
        """GPU timing (HIP events) of the pure-kernel calls that follow."""
        call("onc_ctx_timing", self.h, 1 if on else 0)

    def last_kernel_us(self) -> float:
        """GPU time of the last pure-kernel call's launches (timing on)."""
        v = C.c_double()
        call("onc_ctx_last_kernel_us", self.h, C.byref(v))
        return v.value


def _free_pinned(ctx: "Context", p: int) -> None:
    # After the context is destroyed the pages stay mapped until the process ends.
    if ctx.h:
        lib().onc_host_free(ctx.h, _P(p))


class PinnedBuffer:
    """Page-locked host memory (onc_host_alloc) viewed as a numpy array.

    The memory lives as long as any numpy view of it: `array` and every view
    taken from it keep the underlying ctypes buffer alive, and the memory is
    released by a finalizer of that buffer. free() drops this handle's own
    references only, so a view kept elsewhere never points at freed memory."""

    def __init__(self, ctx: Context, count: int, dtype=np.uint32):
        self.ctx = ctx
        self.dtype = np.dtype(dtype)
        self.count = int(count)
        nbytes = max(4, self.count * self.dtype.itemsize)
        p = _P()
        call("onc_host_alloc", ctx.h, nbytes, C.byref(p))
        buf = (C.c_char * nbytes).from_address(p.value)
        weakref.finalize(buf, _free_pinned, ctx, p.value)
        self.p = p
        self.array = np.frombuffer(buf, dtype=self.dtype, count=self.count)

    def free(self) -> None:
        self.array = None
        self.p = None


class DeviceBuffer:
    def __init__(self, ctx: Context, nbytes: int):
        self.ctx = ctx
        self.nbytes = int(nbytes)
        p = _P()
        call("onc_dev_alloc", ctx.h, max(4, self.nbytes), C.byref(p))
        self.p = p

    def download(self, dtype, count: int) -> np.ndarray:
        out = np.empty(count, dtype=dtype)
        if out.nbytes:
            call("onc_d2h", self.ctx.h, ptr(out), self.p, out.nbytes)
        return out

    def write(self, a: np.ndarray) -> None:
        a = np.ascontiguousarray(a)
        call("onc_h2d", self.ctx.h, self.p, ptr(a), a.nbytes)

    def free(self) -> None:
        if self.p:
            lib().onc_dev_free(self.ctx.h, self.p)
            self.p = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def footprint(desc: SimDesc) -> int:
    out = C.c_uint64()
    call("onc_sim_footprint", C.byref(desc), C.byref(out))
    return int(out.value)
