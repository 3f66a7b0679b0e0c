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
    # an axis's word: the origin's bits (low 9 zero) with e + 151 in bits 0-7 (the step 2^e), so
    # the word shifted left by 23 reads as 2^24 times the step
    origin = (q[:, :3] & 0xFFFFFE00).view(np.float32).astype(np.float64)
    e = (q[:, :3] & 0xFF).astype(np.int64) - 151
    assert (e > -151).all() and (e < 104).all()
    assert not (q[:, :3] & 0x100).any()
    step = np.ldexp(1.0, e)
    shifted = (q[:, :3].astype(np.uint64) << np.uint64(23)) & np.uint64(0xFFFFFFFF)
    assert np.array_equal(shifted.astype(np.uint32).view(np.float32).astype(np.float64),
                          step * 2.0 ** 24)
    # the origin lies at or below both children's boxes
    lo = np.minimum(a["min"], b["min"]).astype(np.float64)
    assert (origin <= lo).all()
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


# ---- how the two node layouts' hits differ, on the CPU ------------------------------------------

def exact_hit(tri, o, d):
    """The ray o + t d against the stored triangle in float64: (hit, t)."""
    v0, e1, e2 = (tri[k][:3].astype(np.float64) for k in ("v0", "e1", "e2"))
    a = np.stack([-d, e1, e2], 1)
    if abs(np.linalg.det(a)) < 1e-300:
        return False, np.inf
    t, u, v = np.linalg.solve(a, o - v0)
    return bool(u >= 0 and v >= 0 and u + v <= 1 and t > 0), float(t)


def _scene_items(scene, tiles=None):
    from lightbaker_amd.bake import layout
    from lightbaker_amd.gpu import Bvh
    from lightbaker_amd.scenes import make_params
    L = layout(scene)
    pos, _, mat = scene.flat()
    bvh = Bvh.build(pos, mat, L.tri_charts, scene.casts())
    tiles = range(len(L.tiles)) if tiles is None else tiles
    tex = np.concatenate([L.rasterize(t).reshape(-1) for t in tiles])
    items = np.ascontiguousarray(tex[tex["pos"][:, 3] > 0.5])
    light = np.random.default_rng(1).random((L.width * L.height, 4)).astype(np.float32)
    light[:, 3] = 1.0
    return bvh, items, make_params(scene, L.width, L.height), L.affines.reshape(-1, 8), light


@pytest.mark.parametrize("which", ["small", "rebake-2-tiles"])
def test_quantized_hits_differ_only_by_near_ties_and_grazing_false_positives(which):
    """Every ray of the direct and gather passes traced over both node layouts
    (oracle.quant_compare). The 8-bit boxes contain the FP32 boxes, but the traversal visits
    children in the order of their entry distances and prunes a child whose entry lies beyond
    the nearest hit so far, and both read different boxes; so where a ray's two candidate hits
    lie a few ulps apart (a box's bottom face on the ground), the layouts may keep different
    ones (a near-tie, distances within 2^-20 t). Beyond near-ties, a ray may differ only where
    the larger boxes visit a leaf the FP32 boxes cull and the FP32 Moller-Trumbore test
    (trace.wgsl:94-104) accepts a ray that misses the triangle exactly (a grazing ray, its
    determinant near zero): the hard difference must be such a false positive, never a hit the
    FP32 traversal finds and the 8-bit one loses."""
    import oracle
    from lightbaker_amd.scenes import rebake_scene
    if which == "small":
        scene, tiles = small_scene(1), None
    else:
        scene, tiles = rebake_scene(3), [51, 52]
    bvh, items, P, charts, light = _scene_items(scene, tiles)
    q = oracle.quant_records(bvh.nodes)
    mats = scene.materials_array()
    total = [0, 0, 0]
    for pass_, it in ((0, items), (1, items[::4])):
        _, cnt = oracle.quant_compare(pass_, bvh.nodes, q, bvh.tris, mats, charts, P, it, 0, light)
        total = [a + b for a, b in zip(total, cnt)]
        ex = oracle.quant_examples()[: min(cnt[2], 16)]
        assert cnt[2] <= 16, f"pass {pass_}: {cnt[2]} hard differences"
        for row in ex:
            o, d = row[0:3].astype(np.float64), row[3:6].astype(np.float64)
            tf, tq = float(row[8]), float(row[9])
            trf, trq = (int(v) for v in row[10:12].view(np.uint32))
            # the hit one layout kept nearer than anything the other kept
            extra = trq if tq < tf else trf
            assert extra != 0xFFFFFFFF, row
            hit, _ = exact_hit(bvh.tris[extra], o, d)
            assert not hit, f"pass {pass_}: triangle {extra} is a true hit one layout lost: {row}"
    rays, differ, hard = total
    print(f"{which}: {rays} rays, {differ} differ ({differ / rays:.2e}), {hard} beyond a near-tie",
          flush=True)
    assert differ <= 1e-4 * rays
