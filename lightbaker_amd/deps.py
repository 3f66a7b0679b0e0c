"""What each tile's light depends on (crates/lightbaker/src/deps.rs, bake.rs:124-136,
volume.rs:36-65): the hashes the reference keys its cache with, and the dirty flags a rebake
takes. The triangle filter runs on the device (lb_deps_swept); the hashes are BLAKE3 over the
reference's byte streams, assembled by the library's host code.

    h0 = tile_hashes(gpu, scene)             # after a bake, kept with its TileCache
    h1 = tile_hashes(gpu, edited)            # the scene edited
    bake(edited, gpu, cache=b0.cache, tile_dirty=dirty(h0, h1), probe_dirty=...)
"""
from __future__ import annotations

import ctypes as C
import json

import numpy as np

from . import _native as N
from .bake import scene_desc
from .gpu import Gpu

BAKER_VERSION = 11  # deps.rs:10: bumped when the baker's results change


def blake3(data: bytes) -> bytes:
    out = (C.c_uint8 * 32)()
    N.check(N.lib().lb_blake3(data, len(data), out))
    return bytes(out)


def _f32(v) -> str:
    """An f32 as serde_json prints it: the shortest digits that read back as the same f32, laid
    out as the ryu crate does (decimal with a ".0" for integers, or exponent notation)."""
    x = np.float32(v)
    if x == 0:
        return "-0.0" if np.signbit(x) else "0.0"
    sign = "-" if x < 0 else ""
    m, e = np.format_float_scientific(abs(x), unique=True, trim="-").split("e")
    digits = m.replace(".", "")
    k = int(e) - (len(digits) - 1)  # value = digits x 10^k
    kk = len(digits) + k
    if 0 <= k and kk <= 13:
        s = digits + "0" * k + ".0"
    elif 0 < kk <= 13:
        s = digits[:kk] + "." + digits[kk:]
    elif -6 < kk <= 0:
        s = "0." + "0" * (-kk) + digits
    elif len(digits) == 1:
        s = digits + "e" + str(kk - 1)
    else:
        s = digits[0] + "." + digits[1:] + "e" + str(kk - 1)
    return sign + s


def _json(v) -> str:
    """serde_json's compact form of the scene file's settings: floats as f32, fields in order."""
    if isinstance(v, bool):
        return "true" if v else "false"
    if isinstance(v, int):
        return str(v)
    if isinstance(v, float):
        return _f32(v)
    if isinstance(v, str):
        return json.dumps(v, ensure_ascii=False)
    if isinstance(v, (list, tuple)):
        return "[" + ",".join(_json(x) for x in v) + "]"
    if isinstance(v, dict):
        return "{" + ",".join(json.dumps(k) + ":" + _json(x) for k, x in v.items()) + "}"
    raise TypeError(type(v))


def _floats(t):
    return [float(x) for x in t]


def settings_json(scene) -> list[str]:
    """scene.rs's Sun, Sky, materials, LightmapSettings and BakeSettings as the scene file holds
    them (serde field order, camelCase)."""
    s = scene.settings
    sun = {"direction": _floats(scene.sun_direction), "radiusDeg": float(scene.sun_radius_deg),
           "irradiance": _floats(scene.sun_irradiance)}
    sky = {"zenith": _floats(scene.zenith), "horizon": _floats(scene.horizon),
           "ground": _floats(scene.ground)}
    mats = [{"name": m.name, "albedo": _floats(m.albedo), "receives": bool(m.receives),
             "casts": bool(m.casts), "opacity": float(m.opacity)} for m in scene.materials]
    lightmap = {"texelsPerMeter": float(s.texels_per_meter), "tileSize": float(s.tile_size),
                "chartAngleDeg": float(s.chart_angle_deg), "padding": int(s.padding),
                "atlasWidth": int(s.atlas_width), "indirectDownscale": int(s.indirect_downscale)}
    bake = {"sunSamples": int(s.sun_samples), "aoRays": int(s.ao_rays),
            "aoReach": float(s.ao_reach), "aoStrength": float(s.ao_strength),
            "aoFloor": float(s.ao_floor), "aoContrast": float(s.ao_contrast),
            "sdfRadius": float(s.sdf_radius), "giRays": int(s.gi_rays),
            "giBounces": int(s.gi_bounces), "giReach": float(s.gi_reach),
            "insideShare": float(s.inside_share), "giFilter": int(s.gi_filter)}
    return [_json(x) for x in (sun, sky, mats, lightmap, bake)]


def global_hash(scene) -> bytes:
    """deps.rs:60-70: the baker's version and the settings every tile depends on."""
    data = BAKER_VERSION.to_bytes(4, "little") + b"".join(j.encode() for j in settings_json(scene))
    return blake3(data)


def volume_json(scene) -> str:
    v = scene.volume
    return _json({"min": _floats(v.min), "max": _floats(v.max), "step": float(v.step),
                  "rays": int(v.rays)})


def tile_hashes(gpu: Gpu, scene, global_: bytes | None = None) -> np.ndarray:
    """bake.rs:124-128: (tiles, 32) bytes, tiles in layout order."""
    g = global_ if global_ is not None else global_hash(scene)
    d, keep = scene_desc(scene)
    lib = N.lib()
    h = C.c_void_p()
    N.check(lib.lb_layout_build(C.byref(d), C.byref(h)))
    try:
        w, hh, nc, nt = C.c_uint32(), C.c_uint32(), C.c_size_t(), C.c_size_t()
        N.check(lib.lb_layout_size(h, C.byref(w), C.byref(hh), C.byref(nc), C.byref(nt)))
        out = np.zeros((nt.value, 32), np.uint8)
        N.check(lib.lb_deps_tile_hashes(gpu.handle, C.byref(d), h, g, N.ptr(out)))
    finally:
        lib.lb_layout_destroy(h)
        del keep
    return out


def probe_hashes(gpu: Gpu, scene, global_: bytes | None = None):
    """volume.rs:36-56: ((chunks, 2) i32 keys in order, (chunks, 32) hashes)."""
    if scene.volume is None:
        return np.zeros((0, 2), np.int32), np.zeros((0, 32), np.uint8)
    g = global_ if global_ is not None else global_hash(scene)
    vj = volume_json(scene).encode()
    d, keep = scene_desc(scene)
    n = C.c_size_t()
    lib = N.lib()
    # the count first (cap 0 fails with it), then the hashes
    lib.lb_deps_probe_hashes(gpu.handle, C.byref(d), g, vj, len(vj), None, None, 0, C.byref(n))
    keys = np.zeros((n.value, 2), np.int32)
    hashes = np.zeros((n.value, 32), np.uint8)
    N.check(lib.lb_deps_probe_hashes(gpu.handle, C.byref(d), g, vj, len(vj), N.ptr(keys),
                                     N.ptr(hashes), n.value, C.byref(n)))
    del keep
    return keys, hashes


def dirty(before: np.ndarray, now: np.ndarray) -> np.ndarray:
    """A flag per tile (or chunk): its hash changed (bake.rs:129-135: no cache entry under it)."""
    return (np.asarray(before) != np.asarray(now)).any(axis=1)


def probe_dirty(scene, keys: np.ndarray, chunk_dirty: np.ndarray) -> np.ndarray:
    """The per-probe flags of the dirty chunks (volume.rs:40-42: a probe's chunk is the ground-plan
    tile under it)."""
    v = scene.volume
    f = np.float32
    # f32::round: half away from zero (the extent is never negative)
    n = [max(int(np.floor((f(v.max[a]) - f(v.min[a])) / f(v.step) + f(0.5))) + 1, 2)
         for a in range(3)]
    k = np.arange(n[0] * n[1] * n[2])
    px = f(v.min[0]) + (k % n[0]).astype(f) * f(v.step)
    pz = f(v.min[2]) + ((k // n[0]) % n[2]).astype(f) * f(v.step)
    size = f(scene.settings.tile_size)
    kx, kz = np.floor(px / size).astype(np.int64), np.floor(pz / size).astype(np.int64)
    bad = {(int(a), int(b)) for (a, b), d in zip(keys, chunk_dirty) if d}
    return np.array([(int(a), int(b)) in bad for a, b in zip(kx, kz)], bool)
