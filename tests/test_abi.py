"""The C-ABI library loads and exports every entry point include/lightbaker.h declares; the
host-side entry points behave on the CPU (no GPU calls)."""
import ctypes as C
import re
from pathlib import Path

import numpy as np

from lightbaker_amd import _native as N

HEADER = Path(__file__).resolve().parent.parent / "include" / "lightbaker.h"


def declared():
    text = re.sub(r"/\*.*?\*/", "", HEADER.read_text(), flags=re.S)
    return sorted(set(re.findall(r"\b(lb_\w+)\s*\(", text)))


def test_every_declared_symbol_is_exported_and_bound():
    lib = C.CDLL(str(N.LIB_PATH))
    names = declared()
    assert len(names) >= 20
    for name in names:
        assert hasattr(lib, name), name
        assert name in N.SIGNATURES, f"{name} has no ctypes signature"


def test_record_sizes_match_the_reference():
    # bvh.rs:12-33, raster.rs:13-22, gpu/mod.rs:11-23
    assert N.NODE.itemsize == 32 and N.TRI.itemsize == 48
    assert N.TEXEL.itemsize == 64 and N.PARAMS.itemsize == 144


def test_null_handles_are_errors_not_crashes():
    lib = N.lib()
    assert lib.lb_run(None, 0, None, None, None, 0, None, 0, None, 0) != 0
    assert b"null" in lib.lb_last_error()
    t = N.Tuning(256, 2, 16, 0, 16, 0)
    assert lib.lb_set_tuning(None, C.byref(t)) != 0
    st = N.Stats()
    assert lib.lb_run_stats(None, 0, 0, None, None, None, 0, 0, None, 0, None, C.byref(st)) != 0


def test_bvh_build_rejects_a_bad_material():
    lib = N.lib()
    pos = np.zeros((1, 9), np.float32)
    mat, chart, casts = np.array([3], np.uint32), np.array([0], np.uint32), np.array([1], np.uint8)
    h = C.c_void_p()
    rc = lib.lb_bvh_build(pos.ctypes.data, 1, mat.ctypes.data, chart.ctypes.data,
                          casts.ctypes.data, 1, C.byref(h))
    assert rc != 0 and b"material 3" in lib.lb_last_error()
