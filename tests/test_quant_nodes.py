"""The TUNED form's 8-bit node records (lb_tuned_nodes, as lb_scene_create builds them) read back
and decoded from their documented layout, on the CPU: every decoded child box contains the
child's FP32 box (the corners rounded outward, so culling stays conservative), in exact
arithmetic and after the FP32 rounding of origin + byte * step alike, and each record carries
its children's references. The WIDE records hold the children's boxes exactly.
"""
from __future__ import annotations

import numpy as np
import pytest

from lightbaker_amd import _native as N
from lightbaker_amd.scenes import city_scene, map_scene, small_scene

from . import kit


def records(bvh):
    lib = N.lib()
    n = len(bvh.nodes)
    quant = np.zeros((n, 8), np.uint32)
    wide = np.zeros((n, 16), np.float32)
    nodes = np.ascontiguousarray(bvh.nodes)
    N.check(lib.lb_tuned_nodes(nodes.ctypes.data, n, quant.ctypes.data, wide.ctypes.data))
    return quant, wide


def ref_of(nodes, k):
    """A child's reference: an inner child's index, or a leaf's (first, count) with bit 31."""
    nd = nodes[k]
    if nd["count"] == 0:
        return k
    if nd["count"] <= 30 and nd["first"] < (1 << 26):
        return 0x80000000 | (int(nd["count"]) << 26) | int(nd["first"])
    return 0x80000000 | (31 << 26) | k


@pytest.mark.parametrize("gen", [lambda: small_scene(1), lambda: map_scene(2),
                                 lambda: city_scene(4, buildings=300, size=120.0)],
                         ids=["small", "map", "city-block"])
def test_quantized_boxes_contain_their_fp32_boxes(gen):
    scene = gen()
    bvh, _, _ = kit.scene_bvh(scene)
    nodes = bvh.nodes
    quant, wide = records(bvh)
    inner = np.nonzero(nodes["count"] == 0)[0]
    assert len(inner) > 0
    first = nodes["first"][inner].astype(np.int64)
    a, b = nodes[first], nodes[first + 1]
    q = quant[inner]
    origin = (q[:, :3] & 0xFFFFFF00).view(np.float32).astype(np.float64)
    e = (q[:, :3] & 0xFF).astype(np.int64)
    assert (e > 0).all() and (e < 255).all()
    step = np.ldexp(1.0, e - 127)
    corner = q[:, 4:7].view(np.uint8).reshape(-1, 12).astype(np.float64)
    exact = origin[:, [0, 1, 2, 0, 1, 2, 0, 1, 2, 0, 1, 2]] + corner * step[:, [0, 1, 2] * 4]
    # the sums are exact in double: 24-bit origins plus 8-bit multiples of a power of two a few
    # dozen binades apart at most
    assert np.array_equal(exact - origin[:, [0, 1, 2] * 4], corner * step[:, [0, 1, 2] * 4])
    rounded = (origin[:, [0, 1, 2] * 4].astype(np.float32)
               + (corner * step[:, [0, 1, 2] * 4]).astype(np.float32)).astype(np.float64)
    fp32 = np.concatenate([a["min"], a["max"], b["min"], b["max"]], axis=1).astype(np.float64)
    for dec in (exact, rounded):
        assert (dec[:, 0:3] <= fp32[:, 0:3]).all() and (dec[:, 3:6] >= fp32[:, 3:6]).all()
        assert (dec[:, 6:9] <= fp32[:, 6:9]).all() and (dec[:, 9:12] >= fp32[:, 9:12]).all()
    # the outward rounding loosens a box by a few steps at most: a byte box spans <= 256 steps
    slack = np.maximum(fp32[:, 0:3] - exact[:, 0:3], exact[:, 3:6] - fp32[:, 3:6])
    assert (slack <= 2 * step).all()
    # references, and the WIDE records: both boxes exactly, then the same references
    ra = np.array([ref_of(nodes, int(k)) for k in first], np.uint32)
    rb = np.array([ref_of(nodes, int(k) + 1) for k in first], np.uint32)
    assert np.array_equal(q[:, 3], ra) and np.array_equal(q[:, 7], rb)
    w = wide[inner]
    assert np.array_equal(w[:, 0:12], fp32.astype(np.float32))
    assert np.array_equal(w[:, 12].view(np.uint32), ra) and np.array_equal(w[:, 13].view(np.uint32), rb)
    # leaves hold no record
    assert not quant[nodes["count"] > 0].any()
    print(f"{len(inner)} inner records, slack {float((slack / step).max()):.2f} steps at most",
          flush=True)
