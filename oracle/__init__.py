"""ORACLE — test infrastructure, not product code.

ctypes access to liboracle.so, the scalar FP32 C restatement of the reference's ray passes
(crates/lightbaker/src/gpu/trace.wgsl) and CPU texel passes (filter.rs), plus the FP64 tracer.
Only tests/, the smoke check and bench.py's cpu_baseline leg use it, as the checker.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle.so"


class Counters(C.Structure):
    _fields_ = [("rays", C.c_uint64), ("inner_visits", C.c_uint64), ("tri_tests", C.c_uint64),
                ("material_loads", C.c_uint64), ("light_gathers", C.c_uint64),
                ("leaf_visits", C.c_uint64)]


_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            from . import build as _b
            _b.build()
        L = C.CDLL(str(LIB_PATH))
        VP, SZ, U32, F, I = C.c_void_p, C.c_size_t, C.c_uint32, C.c_float, C.c_int
        L.lb_sincos.argtypes = [F, C.POINTER(F), C.POINTER(F)]
        L.lb_sincos.restype = None
        L.orc_pcg.argtypes = [U32]
        L.orc_pcg.restype = U32
        L.orc_rand.argtypes = [U32, U32, F]
        L.orc_rand.restype = F
        L.orc_radical.argtypes = [U32]
        L.orc_radical.restype = F
        L.orc_cosine.argtypes = [F, F, VP, VP]
        L.orc_cosine.restype = None
        L.orc_run.argtypes = [I, VP, SZ, VP, VP, VP, VP, VP, SZ, U32, VP, VP, C.POINTER(Counters)]
        L.orc_run.restype = I
        L.orc_run_quant.argtypes = [I, VP, SZ, VP, VP, VP, VP, VP, VP, SZ, U32, VP, VP,
                                    C.POINTER(Counters)]
        L.orc_run_quant.restype = I
        L.orc_quant_compare.argtypes = [I, VP, SZ, VP, VP, VP, VP, VP, VP, SZ, U32, VP, VP,
                                        C.POINTER(C.c_ulonglong * 3)]
        L.orc_quant_compare.restype = I
        L.orc_quant_examples.argtypes = [VP]
        L.orc_quant_examples.restype = None
        L.orc_trace.argtypes = [VP, SZ, VP, VP, VP, VP, VP, F, I, U32, C.POINTER(F),
                                C.POINTER(U32), C.POINTER(I), C.POINTER(Counters)]
        L.orc_trace.restype = None
        L.orc_trace_brute.argtypes = [VP, SZ, VP, VP, VP, VP, F, U32, C.POINTER(F),
                                      C.POINTER(U32), C.POINTER(I)]
        L.orc_trace_brute.restype = None
        L.orc64_run.argtypes = [I, VP, SZ, VP, VP, VP, VP, VP, SZ, U32, VP, VP]
        L.orc64_run.restype = I
        L.orc_gather_steps.argtypes = [VP, SZ, VP, VP, VP, VP, SZ, U32, VP]
        L.orc_gather_steps.restype = I
        L.orc_gather_terms.argtypes = [VP, SZ, VP, VP, VP, VP, VP, SZ, U32, VP, VP]
        L.orc_gather_terms.restype = I
        L.orc_set_sky_sh.argtypes = [VP]
        L.orc_set_sky_sh.restype = None
        L.orc_gather_dirs.argtypes = [VP, VP, SZ, U32, VP]
        L.orc_gather_dirs.restype = I
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data


def run(pass_: int, nodes, tris, materials, charts, params, items, key_base: int = 0,
        light=None, counters: bool = False):
    """orc_run: each item's results as (n * width, 4) float32; with counters, also the counts."""
    L = lib()
    items = np.ascontiguousarray(items)
    nodes = np.ascontiguousarray(nodes)
    tris = np.ascontiguousarray(tris)
    params = np.ascontiguousarray(params)
    n = len(items)
    width = 4 if pass_ == 2 else 1
    out = np.zeros((n * width, 4), dtype=np.float32)
    mats = np.ascontiguousarray(materials, dtype=np.float32)
    ch = None if charts is None or len(charts) == 0 else np.ascontiguousarray(charts,
                                                                               dtype=np.float32)
    lt = None if light is None or len(light) == 0 else np.ascontiguousarray(light,
                                                                             dtype=np.float32)
    cnt = Counters() if counters else None
    rc = L.orc_run(pass_, _p(nodes), len(nodes), _p(tris), _p(mats), _p(ch), _p(params),
                   _p(items), n, key_base, _p(lt), _p(out), C.byref(cnt) if cnt else None)
    assert rc == 0
    if counters:
        return out, {k: getattr(cnt, k) for k, _ in Counters._fields_}
    return out


def quant_records(nodes) -> np.ndarray:
    """The TUNED form's 8-bit node records of a BVH (lb_tuned_nodes, built on the host)."""
    from lightbaker_amd import _native as N
    nodes = np.ascontiguousarray(nodes)
    q = np.zeros((len(nodes), 8), np.uint32)
    N.check(N.lib().lb_tuned_nodes(nodes.ctypes.data, len(nodes), q.ctypes.data, None))
    return q


def run_quant(pass_: int, nodes, quant, tris, materials, charts, params, items,
              key_base: int = 0, light=None, counters: bool = False):
    """orc_run_quant: as run() with the 8-bit child boxes of `quant` (quant_records) traversed as
    the device traverses them."""
    L = lib()
    items = np.ascontiguousarray(items)
    nodes, tris = np.ascontiguousarray(nodes), np.ascontiguousarray(tris)
    quant = np.ascontiguousarray(quant, np.uint32)
    params = np.ascontiguousarray(params)
    n = len(items)
    out = np.zeros((n * (4 if pass_ == 2 else 1), 4), dtype=np.float32)
    mats = np.ascontiguousarray(materials, dtype=np.float32)
    ch = None if charts is None or len(charts) == 0 else np.ascontiguousarray(charts,
                                                                               dtype=np.float32)
    lt = None if light is None or len(light) == 0 else np.ascontiguousarray(light,
                                                                             dtype=np.float32)
    cnt = Counters() if counters else None
    assert L.orc_run_quant(pass_, _p(nodes), len(nodes), _p(quant), _p(tris), _p(mats), _p(ch),
                           _p(params), _p(items), n, key_base, _p(lt), _p(out),
                           C.byref(cnt) if cnt else None) == 0
    if counters:
        return out, {k: getattr(cnt, k) for k, _ in Counters._fields_}
    return out


def quant_compare(pass_: int, nodes, quant, tris, materials, charts, params, items,
                  key_base: int = 0, light=None):
    """orc_quant_compare: run()'s results, and (rays, rays the two node forms end differently,
    those of them that are no near-tie)."""
    L = lib()
    items = np.ascontiguousarray(items)
    nodes, tris = np.ascontiguousarray(nodes), np.ascontiguousarray(tris)
    quant = np.ascontiguousarray(quant, np.uint32)
    params = np.ascontiguousarray(params)
    n = len(items)
    out = np.zeros((n * (4 if pass_ == 2 else 1), 4), dtype=np.float32)
    mats = np.ascontiguousarray(materials, dtype=np.float32)
    ch = None if charts is None or len(charts) == 0 else np.ascontiguousarray(charts,
                                                                               dtype=np.float32)
    lt = None if light is None or len(light) == 0 else np.ascontiguousarray(light,
                                                                             dtype=np.float32)
    cnt = (C.c_ulonglong * 3)()
    assert L.orc_quant_compare(pass_, _p(nodes), len(nodes), _p(quant), _p(tris), _p(mats),
                               _p(ch), _p(params), _p(items), n, key_base, _p(lt), _p(out),
                               C.byref(cnt)) == 0
    return out, tuple(int(v) for v in cnt)


def quant_examples() -> np.ndarray:
    """The first (up to 16) rays the last quant_compare found differing beyond a near-tie: rows of
    o xyz, d xyz, tmax, any, t (reference), t (quantized), then the two triangles and the ray key
    as u32 bits."""
    buf = np.zeros((16, 13), np.float32)
    lib().orc_quant_examples(buf.ctypes.data)
    return buf


def run64(pass_: int, nodes, tris, materials, charts, params, items, key_base: int = 0,
          light=None):
    """orc64_run: the FP64 tracer's results over the same rays, (n * width, 4) float64."""
    L = lib()
    items = np.ascontiguousarray(items)
    nodes, tris = np.ascontiguousarray(nodes), np.ascontiguousarray(tris)
    params = np.ascontiguousarray(params)
    n = len(items)
    out = np.zeros((n * (4 if pass_ == 2 else 1), 4), dtype=np.float64)
    mats = np.ascontiguousarray(materials, dtype=np.float32)
    ch = None if charts is None or len(charts) == 0 else np.ascontiguousarray(charts,
                                                                               dtype=np.float32)
    lt = None if light is None or len(light) == 0 else np.ascontiguousarray(light,
                                                                             dtype=np.float32)
    assert L.orc64_run(pass_, _p(nodes), len(nodes), _p(tris), _p(mats), _p(ch), _p(params),
                       _p(items), n, key_base, _p(lt), _p(out)) == 0
    return out


def gather_steps(nodes, tris, materials, params, items, key_base: int = 0) -> np.ndarray:
    """orc_gather_steps: (items, rays) u8, each gather ray's inner visits + triangle tests."""
    L = lib()
    items = np.ascontiguousarray(items)
    nodes, tris = np.ascontiguousarray(nodes), np.ascontiguousarray(tris)
    params = np.ascontiguousarray(params)
    mats = np.ascontiguousarray(materials, dtype=np.float32)
    out = np.zeros((len(items), int(params["counts"][0][2])), np.uint8)
    assert L.orc_gather_steps(_p(nodes), len(nodes), _p(tris), _p(mats), _p(params), _p(items),
                              len(items), key_base, _p(out)) == 0
    return out


def trace(nodes, tris, materials, params, o, d, tmax, any_hit, ray):
    L = lib()
    nodes, tris = np.ascontiguousarray(nodes), np.ascontiguousarray(tris)
    params = np.ascontiguousarray(params)
    t, tri, back = C.c_float(), C.c_uint32(), C.c_int()
    o = np.ascontiguousarray(o, dtype=np.float32)
    d = np.ascontiguousarray(d, dtype=np.float32)
    mats = np.ascontiguousarray(materials, dtype=np.float32)
    L.orc_trace(_p(nodes), len(nodes), _p(tris), _p(mats), _p(params), _p(o), _p(d), tmax,
                int(any_hit), ray, C.byref(t), C.byref(tri), C.byref(back), None)
    return t.value, tri.value, bool(back.value)


def trace_brute(tris, materials, params, o, d, tmax, ray):
    L = lib()
    tris, params = np.ascontiguousarray(tris), np.ascontiguousarray(params)
    t, tri, back = C.c_float(), C.c_uint32(), C.c_int()
    o = np.ascontiguousarray(o, dtype=np.float32)
    d = np.ascontiguousarray(d, dtype=np.float32)
    mats = np.ascontiguousarray(materials, dtype=np.float32)
    L.orc_trace_brute(_p(tris), len(tris), _p(mats), _p(params), _p(o), _p(d), tmax, ray,
                      C.byref(t), C.byref(tri), C.byref(back))
    return t.value, tri.value, bool(back.value)


def gather_terms(nodes, tris, materials, charts, params, items, key_base: int = 0,
                 light=None) -> np.ndarray:
    """orc_gather_terms: (items, rays, 3) float32, each gather ray's term as the sum adds it."""
    L = lib()
    items = np.ascontiguousarray(items)
    nodes, tris = np.ascontiguousarray(nodes), np.ascontiguousarray(tris)
    params = np.ascontiguousarray(params)
    mats = np.ascontiguousarray(materials, dtype=np.float32)
    ch = None if charts is None or len(charts) == 0 else np.ascontiguousarray(charts,
                                                                               dtype=np.float32)
    lt = None if light is None or len(light) == 0 else np.ascontiguousarray(light,
                                                                             dtype=np.float32)
    out = np.zeros((len(items), int(params["counts"][0][2]), 3), np.float32)
    assert L.orc_gather_terms(_p(nodes), len(nodes), _p(tris), _p(mats), _p(ch), _p(params),
                              _p(items), len(items), key_base, _p(lt), _p(out)) == 0
    return out


def set_sky_sh(sh) -> None:
    """The sky's L1 term every later pass adds (12 values: c0 rgb, c1 rgb, c2 rgb, c3 rgb, or a
    (4, 3) array; None: the reference's sky)."""
    L = lib()
    if sh is None:
        L.orc_set_sky_sh(None)
        return
    a = np.ascontiguousarray(np.asarray(sh, np.float32).reshape(12))
    L.orc_set_sky_sh(a.ctypes.data)


def gather_dirs(params, items, key_base: int = 0) -> np.ndarray:
    """orc_gather_dirs: (items, rays, 3) float32, each gather ray's direction."""
    items = np.ascontiguousarray(items)
    params = np.ascontiguousarray(params)
    out = np.zeros((len(items), int(params["counts"][0][2]), 3), np.float32)
    assert lib().orc_gather_dirs(_p(params), _p(items), len(items), key_base, _p(out)) == 0
    return out
