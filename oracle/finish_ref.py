"""ORACLE — test infrastructure, not product code.

numpy access to finish_ref.c: the bake's tail after the rays (volume.rs:96-147, the probes' fill
and irradiance; bake.rs:337-421, texels lit by the volume and the atlas from every tile) and the
output codecs (encode.rs:7-74, output.rs:101,108,165-181).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import lib


def _bind():
    L = lib()
    if not hasattr(L, "_finish_bound"):
        VP, SZ, U32, F, I = C.c_void_p, C.c_size_t, C.c_uint32, C.c_float, C.c_int
        L.orc_probe_fill.argtypes = [VP, SZ, SZ, SZ]
        L.orc_probe_fill.restype = None
        L.orc_volume_irr.argtypes = [VP, VP, VP, F, VP, VP, VP, C.POINTER(F)]
        L.orc_volume_irr.restype = None
        L.orc_volume_lit.argtypes = [VP, VP, VP, F, VP, VP, SZ, VP, VP, VP, VP, VP, VP, VP]
        L.orc_volume_lit.restype = SZ
        L.orc_atlas.argtypes = [VP, SZ, VP, VP, VP, VP, VP, U32, U32, U32, VP, VP, VP]
        L.orc_atlas.restype = None
        L.orc_unorm8.argtypes = [F]
        L.orc_unorm8.restype = C.c_uint8
        L.orc_bc4_block.argtypes = [VP, VP]
        L.orc_bc4_block.restype = None
        L.orc_bc5.argtypes = [VP, SZ, SZ, VP]
        L.orc_bc5.restype = I
        L.orc_volume_bytes.argtypes = [VP, SZ, VP]
        L.orc_volume_bytes.restype = F
        L.orc_unorm8_n.argtypes = [VP, SZ, VP]
        L.orc_unorm8_n.restype = None
        L.orc_rgb9e5_n.argtypes = [VP, SZ, VP]
        L.orc_rgb9e5_n.restype = None
        L.orc_rgb9e5_decode_n.argtypes = [VP, SZ, VP]
        L.orc_rgb9e5_decode_n.restype = None
        L._finish_bound = True
    return L


def _c(a, dtype):
    return np.ascontiguousarray(a, dtype=dtype)


def probe_fill(probes, size) -> np.ndarray:
    """volume.rs:96-121 on a copy of `probes` (n, 16), size (nx, ny, nz)."""
    p = _c(probes, np.float32).reshape(-1, 16).copy()
    nx, ny, nz = (int(v) for v in size)
    assert len(p) == nx * ny * nz
    _bind().orc_probe_fill(p.ctypes.data, nx, ny, nz)
    return p


def volume_irr(probes, size, origin, step, p, n):
    """volume.rs:125-147 at one point: (e rgb, sun share)."""
    pr = _c(probes, np.float32)
    sz = _c(size, np.uint32)
    org = _c(origin, np.float32)
    pp, nn = _c(p, np.float32), _c(n, np.float32)
    e = np.zeros(3, np.float32)
    s = C.c_float()
    _bind().orc_volume_irr(pr.ctypes.data, sz.ctypes.data, org.ctypes.data, float(step),
                           pp.ctypes.data, nn.ctypes.data, e.ctypes.data, C.byref(s))
    return e, np.float32(s.value)


def volume_lit(probes, size, origin, step, sun_dir, sun_irr, texels, tl: dict) -> int:
    """bake.rs:337-356 over the texels (structured TEXEL records), updating the TileLight arrays in
    `tl` (known, state, field, ao, direct (n, 3), indirect (n, 3)) in place; returns the count."""
    pr = _c(probes, np.float32)
    sz, org = _c(size, np.uint32), _c(origin, np.float32)
    sd, si = _c(sun_dir, np.float32), _c(sun_irr, np.float32)
    tx = np.ascontiguousarray(texels)
    for k, dt in (("known", np.uint8), ("state", np.uint8), ("field", np.float32),
                  ("ao", np.float32), ("direct", np.float32), ("indirect", np.float32)):
        assert tl[k].dtype == dt and tl[k].flags.c_contiguous, k
    return int(_bind().orc_volume_lit(
        pr.ctypes.data, sz.ctypes.data, org.ctypes.data, float(step), sd.ctypes.data,
        si.ctypes.data, len(tx), tx.ctypes.data, tl["known"].ctypes.data, tl["state"].ctypes.data,
        tl["field"].ctypes.data, tl["ao"].ctypes.data, tl["direct"].ctypes.data,
        tl["indirect"].ctypes.data))


def atlas(tiles, known, state, field, ao, indirect, aw, ah, ds):
    """bake.rs:364-421: (direct (ah, aw, 2), validity (ah, aw), indirect (ih, iw, 3))."""
    t = _c(tiles, np.int32).reshape(-1, 6)
    kn, st = _c(known, np.uint8), _c(state, np.uint8)
    fd, a = _c(field, np.float32), _c(ao, np.float32)
    ind = _c(indirect, np.float32).reshape(-1, 3)
    assert len(kn) == int((t[:, 2].astype(np.int64) * t[:, 3]).sum())
    iw, ih = -(-aw // ds), -(-ah // ds)
    d = np.zeros((ah, aw, 2), np.float32)
    v = np.zeros((ah, aw), np.uint8)
    e = np.zeros((ih, iw, 3), np.float32)
    _bind().orc_atlas(t.ctypes.data, len(t), kn.ctypes.data, st.ctypes.data, fd.ctypes.data,
                      a.ctypes.data, ind.ctypes.data, aw, ah, ds, d.ctypes.data, v.ctypes.data,
                      e.ctypes.data)
    return d, v, e


def unorm8(v) -> np.ndarray:
    """output.rs:101 / :167, elementwise."""
    a = _c(v, np.float32)
    out = np.zeros(a.shape, np.uint8)
    _bind().orc_unorm8_n(a.ctypes.data, a.size, out.ctypes.data)
    return out


def bc4_block(v) -> bytes:
    a = _c(v, np.uint8)
    assert a.size == 16
    out = np.zeros(8, np.uint8)
    _bind().orc_bc4_block(a.ctypes.data, out.ctypes.data)
    return out.tobytes()


def bc5(rg, w, h) -> np.ndarray:
    """encode.rs:37-50 over (h, w, 2) bytes: w * h bytes of blocks."""
    a = _c(rg, np.uint8)
    assert a.size == w * h * 2
    out = np.zeros(w * h, np.uint8)
    if _bind().orc_bc5(a.ctypes.data, w, h, out.ctypes.data) != 0:
        raise ValueError("bc5: w and h must be multiples of 4")
    return out


def rgb9e5(rgb) -> np.ndarray:
    a = _c(rgb, np.float32).reshape(-1, 3)
    out = np.zeros(len(a), np.uint32)
    _bind().orc_rgb9e5_n(a.ctypes.data, len(a), out.ctypes.data)
    return out


def rgb9e5_decode(p) -> np.ndarray:
    a = _c(p, np.uint32).reshape(-1)
    out = np.zeros((len(a), 3), np.float32)
    _bind().orc_rgb9e5_decode_n(a.ctypes.data, len(a), out.ctypes.data)
    return out


def volume_bytes(probes):
    """output.rs:108,165-181: (bytes (n, 16), scale)."""
    p = _c(probes, np.float32).reshape(-1, 16)
    out = np.zeros((len(p), 16), np.uint8)
    s = _bind().orc_volume_bytes(p.ctypes.data, len(p), out.ctypes.data)
    return out, np.float32(s)
