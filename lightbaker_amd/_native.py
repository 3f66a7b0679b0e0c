"""ctypes binding of include/lightbaker.h (liblightbaker_amd.so, built in-tree).

The library is required: importing this module without it raises, so no Python path can stand in
for the HIP kernels.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "liblightbaker_amd.so"

# Records, byte-identical to the reference's Pod types (bvh.rs:12-33, raster.rs:13-22,
# gpu/mod.rs:11-23).
NODE = np.dtype([("min", "<f4", 3), ("first", "<u4"), ("max", "<f4", 3), ("count", "<u4")])
TRI = np.dtype([("v0", "<f4", 4), ("e1", "<f4", 4), ("e2", "<f4", 4)])
TEXEL = np.dtype([("pos", "<f4", 4), ("normal", "<f4", 4), ("dx", "<f4", 4), ("dy", "<f4", 4)])
PARAMS = np.dtype([
    ("sun_dir", "<f4", 4), ("sun_irr", "<f4", 4), ("zenith", "<f4", 4), ("horizon", "<f4", 4),
    ("ground", "<f4", 4), ("fallback", "<f4", 4), ("counts", "<u4", 4), ("reach", "<f4", 4),
    ("atlas", "<u4", 4),
])
assert NODE.itemsize == 32 and TRI.itemsize == 48 and TEXEL.itemsize == 64
assert PARAMS.itemsize == 144

PASS_DIRECT, PASS_GATHER, PASS_VOLUME = 0, 1, 2
FORM_PLAIN, FORM_TUNED = 0, 1
NODES_WIDE, NODES_QUANT = 0, 1


class Stats(C.Structure):
    _fields_ = [("rays", C.c_uint64), ("inner_visits", C.c_uint64), ("tri_tests", C.c_uint64),
                ("material_loads", C.c_uint64), ("light_gathers", C.c_uint64)]


# Every symbol include/lightbaker.h declares, with its ctypes signature.
_VP, _SZ, _U32, _I = C.c_void_p, C.c_size_t, C.c_uint32, C.c_int
SIGNATURES = {
    "lb_open": (_I, [C.POINTER(_VP)]),
    "lb_adapter_name": (C.c_char_p, [_VP]),
    "lb_close": (None, [_VP]),
    "lb_last_error": (C.c_char_p, []),
    "lb_scene_create": (_I, [_VP, _VP, _SZ, _VP, _SZ, _VP, _SZ, _VP, _SZ, C.POINTER(_VP)]),
    "lb_scene_destroy": (None, [_VP]),
    "lb_run": (_I, [_VP, _I, _VP, _VP, _VP, _SZ, _VP, _SZ, _VP, _SZ]),
    "lb_buffer_alloc": (_I, [_VP, _SZ, C.POINTER(_VP)]),
    "lb_buffer_free": (_I, [_VP, _VP]),
    "lb_upload": (_I, [_VP, _VP, _VP, _SZ]),
    "lb_download": (_I, [_VP, _VP, _VP, _SZ]),
    "lb_sync": (_I, [_VP]),
    "lb_run_device": (_I, [_VP, _I, _I, _VP, _VP, _VP, _SZ, _U32, _VP, _SZ, _VP]),
    "lb_set_tuning": (_I, [_VP, _I, _I, _I, _I]),
    "lb_get_tuning": (_I, [_VP, C.POINTER(_I), C.POINTER(_I), C.POINTER(_I), C.POINTER(_I)]),
    "lb_run_stats": (_I, [_VP, _I, _I, _VP, _VP, _VP, _SZ, _U32, _VP, _SZ, _VP, C.POINTER(Stats)]),
    "lb_timer_begin": (_I, [_VP]),
    "lb_timer_end": (_I, [_VP, C.POINTER(C.c_float)]),
    "lb_set_form": (_I, [_VP, _I]),
    "lb_bvh_build": (_I, [_VP, _SZ, _VP, _VP, _VP, _SZ, C.POINTER(_VP)]),
    "lb_bvh_nodes": (_SZ, [_VP, C.POINTER(_VP)]),
    "lb_bvh_tris": (_SZ, [_VP, C.POINTER(_VP)]),
    "lb_bvh_destroy": (None, [_VP]),
}


class NativeError(RuntimeError):
    pass


def load() -> C.CDLL:
    if not LIB_PATH.exists():
        raise NativeError(
            f"{LIB_PATH} is missing: build it with `python -m lightbaker_amd.build` "
            "(the HIP kernels have no fallback)")
    lib = C.CDLL(str(LIB_PATH))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


_lib: C.CDLL | None = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        _lib = load()
    return _lib


def check(rc: int) -> None:
    if rc != 0:
        raise NativeError(lib().lb_last_error().decode())


def ptr(a: np.ndarray | None) -> int | None:
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data
