"""ORACLE — test infrastructure, not product code.

numpy access to deps_ref.c (deps.rs:40-52,104-121: a triangle's box and `swept`), and a Python
restatement of the byte streams deps.rs:72-102 (hash_reach) and bake.rs:447-477 (tile_box,
tile_seed) feed to BLAKE3, for checking the library's hashes stream for stream.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import lib


def _bind():
    L = lib()
    if not hasattr(L, "_deps_bound"):
        VP, SZ, F = C.c_void_p, C.c_size_t, C.c_float
        L.orc_swept.argtypes = [VP, VP, F, VP]
        L.orc_swept.restype = C.c_int
        L.orc_deps_bits.argtypes = [VP, SZ, VP, VP, SZ, VP, F, VP]
        L.orc_deps_bits.restype = None
        L._deps_bound = True
    return L


def swept(zone, sun, top, box) -> bool:
    z, s, b = (np.ascontiguousarray(a, np.float32) for a in (zone, sun, box))
    return bool(_bind().orc_swept(z.ctypes.data, s.ctypes.data, float(np.float32(top)), b.ctypes.data))


def bits(pos, casts, zones, sun, top) -> np.ndarray:
    """(zones, ceil(tris / 32)) u32, as lb_deps_swept lays them out."""
    p = np.ascontiguousarray(pos, np.float32).reshape(-1, 9)
    z = np.ascontiguousarray(zones, np.float32).reshape(-1, 6)
    c = None if casts is None else np.ascontiguousarray(casts, np.uint8)
    s = np.ascontiguousarray(sun, np.float32)
    words = (len(p) + 31) // 32
    out = np.zeros((len(z), words), np.uint32)
    _bind().orc_deps_bits(p.ctypes.data, len(p), None if c is None else c.ctypes.data,
                          z.ctypes.data, len(z), s.ctypes.data, float(np.float32(top)),
                          out.ctypes.data)
    return out


def u32(v) -> bytes:
    return int(v).to_bytes(4, "little")


def i32(v) -> bytes:
    return int(v).to_bytes(4, "little", signed=True)


def hash_reach_stream(global_: bytes, seed: bytes, meshes, row_bits, first) -> bytes:
    """deps.rs:79-101's bytes: meshes = [(positions (n, 9) f32, material, casts)], row_bits the
    zone's hit bits over every triangle (meshes in order), first each mesh's first triangle."""
    out = bytearray(global_ + seed)
    for mi, (pos, material, casts) in enumerate(meshes):
        if not casts:
            continue
        hits = [t for t in range(len(pos)) if row_bits[(first[mi] + t) // 32] >> ((first[mi] + t) % 32) & 1]
        if not hits:
            continue
        out += u32(mi) + u32(material)
        for t in hits:
            out += u32(t) + np.ascontiguousarray(pos[t], "<f4").tobytes()
    return bytes(out)


def tile_seed_stream(key, charts, meshes, density=None) -> bytes:
    """bake.rs:461-477's bytes: charts = [(mesh, [its triangles in order])]; a tile given a texel
    density of its own (no counterpart in the reference) ends with the density's f32 bytes."""
    out = bytearray(i32(key[0]) + i32(key[1]))
    for m, tris in charts:
        pos, material, _ = meshes[m]
        out += u32(m) + u32(material)
        for t in tris:
            out += u32(t) + np.ascontiguousarray(pos[t], "<f4").tobytes()
    if density is not None:
        out += np.float32(density).tobytes()
    return bytes(out)
