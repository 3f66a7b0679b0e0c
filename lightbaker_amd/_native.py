"""ctypes binding of include/lightbaker.h (liblightbaker_amd.so, built in-tree).

The library is required: importing this module without it raises, so no Python path can stand in
for the HIP kernels.
"""
from __future__ import annotations

import atexit
import ctypes as C
import weakref
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
FORM_PLAIN, FORM_TUNED, FORM_WATERTIGHT = 0, 1, 2
NODES_WIDE, NODES_QUANT = 0, 1


class Stats(C.Structure):
    _fields_ = [("rays", C.c_uint64), ("inner_visits", C.c_uint64), ("tri_tests", C.c_uint64),
                ("material_loads", C.c_uint64), ("light_gathers", C.c_uint64),
                ("leaf_visits", C.c_uint64)]


class Tuning(C.Structure):
    _fields_ = [("workgroup", C.c_int), ("rays_per_lane", C.c_int), ("lds_stack", C.c_int),
                ("nodes", C.c_int), ("refill", C.c_int), ("schedule", C.c_int),
                ("chunk", C.c_int), ("leaf_batch", C.c_int), ("waves", C.c_int),
                ("xcd_split", C.c_int), ("gather_order", C.c_int)]


class MaterialDesc(C.Structure):
    _fields_ = [("albedo", C.c_float * 3), ("opacity", C.c_float), ("receives", C.c_uint32),
                ("casts", C.c_uint32)]


class MeshDesc(C.Structure):
    _fields_ = [("material", C.c_uint32), ("triangles", C.c_uint32),
                ("positions", C.POINTER(C.c_float))]


class TileDensity(C.Structure):
    """lb_tile_density: a tile's own texels per metre."""
    _fields_ = [("x", C.c_int32), ("z", C.c_int32), ("texels_per_meter", C.c_float)]


class SceneDesc(C.Structure):
    _fields_ = [
        ("sun_direction", C.c_float * 3), ("sun_radius_deg", C.c_float),
        ("sun_irradiance", C.c_float * 3), ("zenith", C.c_float * 3), ("horizon", C.c_float * 3),
        ("ground", C.c_float * 3), ("materials", C.POINTER(MaterialDesc)),
        ("n_materials", C.c_size_t), ("meshes", C.POINTER(MeshDesc)), ("n_meshes", C.c_size_t),
        ("texels_per_meter", C.c_float), ("tile_size", C.c_float), ("chart_angle_deg", C.c_float),
        ("padding", C.c_uint32), ("atlas_width", C.c_uint32), ("indirect_downscale", C.c_uint32),
        ("sun_samples", C.c_uint32), ("ao_rays", C.c_uint32), ("ao_reach", C.c_float),
        ("ao_strength", C.c_float), ("ao_floor", C.c_float), ("ao_contrast", C.c_float),
        ("sdf_radius", C.c_float), ("gi_rays", C.c_uint32), ("gi_bounces", C.c_uint32),
        ("gi_reach", C.c_float), ("inside_share", C.c_float), ("gi_filter", C.c_uint32),
        ("has_volume", C.c_uint32), ("volume_min", C.c_float * 3), ("volume_max", C.c_float * 3),
        ("volume_step", C.c_float), ("volume_rays", C.c_uint32),
        ("tile_density", C.POINTER(TileDensity)), ("n_tile_density", C.c_size_t),
        ("sky_sh", C.c_float * 12),
    ]


class BakeCache(C.Structure):
    """lb_bake_cache: which tiles and probes a rebake traces, and the clean ones' values."""
    _fields_ = [
        ("tile_dirty", C.c_void_p), ("known", C.c_void_p), ("state", C.c_void_p),
        ("field", C.c_void_p), ("ao", C.c_void_p), ("direct", C.c_void_p),
        ("indirect", C.c_void_p), ("n_texels", C.c_size_t), ("probe_dirty", C.c_void_p),
        ("probes", C.c_void_p), ("n_probes", C.c_size_t),
    ]


class BakeOptions(C.Structure):
    _fields_ = [("form", C.c_int), ("graph", C.c_int), ("threads", C.c_int),
                ("cache", C.POINTER(BakeCache)), ("replay", C.c_int)]


class BakeInfo(C.Structure):
    _fields_ = [
        ("atlas_width", C.c_uint32), ("atlas_height", C.c_uint32),
        ("indirect_width", C.c_uint32), ("indirect_height", C.c_uint32),
        ("tiles", C.c_uint32), ("tiles_baked", C.c_uint32), ("charts", C.c_uint32),
        ("texels", C.c_uint64), ("texels_baked", C.c_uint64), ("gather_items", C.c_uint64),
        ("probes", C.c_uint64), ("probes_baked", C.c_uint64), ("volume_size", C.c_uint32 * 3), ("volume_origin", C.c_float * 3),
        ("volume_step", C.c_float), ("bvh_nodes", C.c_uint32), ("bvh_tris", C.c_uint32),
        ("ms_direct", C.c_float), ("ms_direct_texel", C.c_float), ("ms_bounces", C.c_float),
        ("ms_volume", C.c_float), ("s_layout", C.c_double), ("s_bvh", C.c_double),
        ("s_raster", C.c_double), ("s_gather_set", C.c_double), ("s_execute", C.c_double),
        ("s_finish", C.c_double), ("s_total", C.c_double), ("volume_scale", C.c_float),
        ("ms_finish", C.c_float), ("graph_captures", C.c_uint32),
    ]


class LightParams(C.Structure):
    """lb_light_params: the parameters the adjoint differentiates by."""
    _fields_ = [("albedo", C.c_void_p), ("n_materials", C.c_size_t),
                ("zenith", C.c_float * 3), ("horizon", C.c_float * 3), ("ground", C.c_float * 3),
                ("sun", C.c_float * 3), ("sky_sh", C.c_float * 12)]


BAKE_DIRECT, BAKE_INDIRECT, BAKE_VALIDITY, BAKE_VOLUME, BAKE_CHARTS = 0, 1, 2, 3, 4
BAKE_TRI_CHARTS, BAKE_LIGHT, BAKE_CHART_TILES, BAKE_TILES = 5, 6, 7, 8
BAKE_NODES, BAKE_TRIS, BAKE_MATERIALS, BAKE_PARAMS, BAKE_TEXELS, BAKE_GATHER_INDEX = 9, 10, 11, 12, 13, 14
BAKE_TILE_KNOWN, BAKE_TILE_STATE, BAKE_TILE_FIELD, BAKE_TILE_AO = 15, 16, 17, 18
BAKE_TILE_DIRECT, BAKE_TILE_INDIRECT, BAKE_PROBES, BAKE_BAKED_TILES, BAKE_LIGHT0 = 19, 20, 21, 22, 23
BAKE_DIRECT_BC5, BAKE_INDIRECT_RGB9E5, BAKE_VOLUME_BYTES = 24, 25, 26
BAKE_GATHER_COST, BAKE_GATHER_ORDER = 27, 28

# Every symbol include/lightbaker.h declares, with its ctypes signature.
_VP, _SZ, _U32, _I = C.c_void_p, C.c_size_t, C.c_uint32, C.c_int
SIGNATURES = {
    "lb_open": (_I, [C.POINTER(_VP)]),
    "lb_adapter_name": (C.c_char_p, [_VP]),
    "lb_close": (None, [_VP]),
    "lb_last_error": (C.c_char_p, []),
    "lb_scene_create": (_I, [_VP, _VP, _SZ, _VP, _SZ, _VP, _SZ, _VP, _SZ, C.POINTER(_VP)]),
    "lb_scene_destroy": (None, [_VP]),
    "lb_scene_set_sky": (_I, [_VP, _VP]),
    "lb_run": (_I, [_VP, _I, _VP, _VP, _VP, _SZ, _VP, _SZ, _VP, _SZ]),
    "lb_buffer_alloc": (_I, [_VP, _SZ, C.POINTER(_VP)]),
    "lb_buffer_free": (_I, [_VP, _VP]),
    "lb_upload": (_I, [_VP, _VP, _VP, _SZ]),
    "lb_download": (_I, [_VP, _VP, _VP, _SZ]),
    "lb_sync": (_I, [_VP]),
    "lb_run_device": (_I, [_VP, _I, _I, _VP, _VP, _VP, _SZ, _U32, _VP, _SZ, _VP]),
    "lb_run_device_noalloc": (_I, [_VP, _I, _I, _VP, _VP, _VP, _SZ, _U32, _VP, _SZ, _VP]),
    "lb_tuned_nodes": (_I, [_VP, _SZ, _VP, _VP]),
    "lb_set_tuning": (_I, [_VP, C.POINTER(Tuning)]),
    "lb_set_pass_tuning": (_I, [_VP, C.c_int, C.POINTER(Tuning)]),
    "lb_get_pass_tuning": (_I, [_VP, C.c_int, C.POINTER(Tuning)]),
    "lb_get_tuning": (_I, [_VP, C.POINTER(Tuning)]),
    "lb_run_stats": (_I, [_VP, _I, _I, _VP, _VP, _VP, _SZ, _U32, _VP, _SZ, _VP, C.POINTER(Stats)]),
    "lb_timer_begin": (_I, [_VP]),
    "lb_timer_end": (_I, [_VP, C.POINTER(C.c_float)]),
    "lb_set_form": (_I, [_VP, _I]),
    "lb_bvh_build": (_I, [_VP, _SZ, _VP, _VP, _VP, _SZ, C.POINTER(_VP)]),
    "lb_bvh_build_threads": (_I, [_VP, _SZ, _VP, _VP, _VP, _SZ, _I, C.POINTER(_VP)]),
    "lb_bvh_nodes": (_SZ, [_VP, C.POINTER(_VP)]),
    "lb_bvh_tris": (_SZ, [_VP, C.POINTER(_VP)]),
    "lb_bvh_destroy": (None, [_VP]),
    "lb_bake_run": (_I, [_VP, C.POINTER(SceneDesc), C.POINTER(BakeOptions), C.POINTER(_VP)]),
    "lb_bake_destroy": (None, [_VP]),
    "lb_bake_info_get": (_I, [_VP, C.POINTER(BakeInfo)]),
    "lb_bake_array": (_SZ, [_VP, _I, C.POINTER(_VP)]),
    "lb_bake_prepare": (_I, [_VP, C.POINTER(SceneDesc), C.POINTER(BakeOptions), C.POINTER(_VP)]),
    "lb_bake_execute": (_I, [_VP, _VP]),
    "lb_adjoint_create": (_I, [_VP, C.POINTER(_VP)]),
    "lb_adjoint_destroy": (None, [_VP]),
    "lb_adjoint_info": (_I, [_VP, C.POINTER(_SZ), C.POINTER(_SZ), C.POINTER(_SZ),
                             C.POINTER(_SZ)]),
    "lb_adjoint_records": (_I, [_VP, _VP, _SZ]),
    "lb_adjoint_forward": (_I, [_VP, C.POINTER(LightParams), _VP]),
    "lb_adjoint_backward": (_I, [_VP, _VP, _VP]),
    "lb_bake_finish": (_I, [_VP]),
    "lb_bake_time_pass": (_I, [_VP, _I, _I, C.POINTER(C.c_float)]),
    "lb_layout_build": (_I, [C.POINTER(SceneDesc), C.POINTER(_VP)]),
    "lb_layout_destroy": (None, [_VP]),
    "lb_layout_size": (_I, [_VP, C.POINTER(_U32), C.POINTER(_U32), C.POINTER(_SZ), C.POINTER(_SZ)]),
    "lb_layout_get": (_I, [_VP, _VP, _VP, _VP, _VP]),
    "lb_layout_rasterize": (_I, [C.POINTER(SceneDesc), _VP, _SZ, _VP, _SZ]),
    "lb_encode_bc5": (_I, [_VP, _VP, _U32, _U32, _VP]),
    "lb_encode_rgb9e5": (_I, [_VP, _VP, _SZ, _VP]),
    "lb_encode_volume": (_I, [_VP, _VP, _SZ, _VP, C.POINTER(C.c_float)]),
    "lb_blake3": (_I, [_VP, _SZ, _VP]),
    "lb_deps_swept": (_I, [_VP, _VP, _SZ, _VP, _VP, _SZ, _VP, C.c_float, _VP]),
    "lb_deps_tile_hashes": (_I, [_VP, C.POINTER(SceneDesc), _VP, _VP, _VP]),
    "lb_deps_probe_hashes": (_I, [_VP, C.POINTER(SceneDesc), _VP, _VP, _SZ, _VP, _VP, _SZ,
                                  C.POINTER(_SZ)]),
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


# Objects holding native handles, by the order they are closed in: an adjoint before its bake,
# bakes and scenes and device buffers before their Gpu. Gpu.close closes its own first; at exit
# every one left open is closed here, while the HIP runtime is still up (Python's exit handlers
# run before the C library's), so no handle is released from a finalizer during teardown.
ADJOINT, BAKE, SCENE, ARRAY, GPU = range(5)
_OWNED = [weakref.WeakSet() for _ in range(5)]


def own(obj, rank: int) -> None:
    """Registers an object with a close() method and a `gpu` attribute (a Gpu: itself)."""
    _OWNED[rank].add(obj)


def close_owned(gpu=None) -> None:
    """Closes the registered objects of `gpu` (every object when None), in rank order."""
    for rank, objs in enumerate(_OWNED):
        for obj in list(objs):
            if gpu is not None and (rank == GPU or getattr(obj, "gpu", None) is not gpu):
                continue
            try:
                obj.close()
            except Exception:
                pass


atexit.register(close_owned)
