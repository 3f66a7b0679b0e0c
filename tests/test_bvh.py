"""Bvh::build (bvh.rs:50-162) restated in C++: the invariants its construction guarantees."""
import numpy as np
import pytest

from lightbaker_amd.gpu import Bvh
from lightbaker_amd.scenes import small_scene, map_scene

from . import kit


def check_tree(bvh, n_casting):
    nodes, tris = bvh.nodes, bvh.tris
    assert len(nodes) <= 2 * max(1, len(tris))
    seen = np.zeros(len(tris), np.int32)
    stack = [0]
    reached = 0
    while stack:
        i = stack.pop()
        reached += 1
        n = nodes[i]
        if n["count"] > 0:
            seen[n["first"]:n["first"] + n["count"]] += 1
            v0 = tris["v0"][n["first"]:n["first"] + n["count"], :3]
            v1 = v0 + tris["e1"][n["first"]:n["first"] + n["count"], :3]
            v2 = v0 + tris["e2"][n["first"]:n["first"] + n["count"], :3]
            lo = np.minimum(np.minimum(v0, v1), v2).min(axis=0)
            hi = np.maximum(np.maximum(v0, v1), v2).max(axis=0)
            # a leaf's box holds its triangles (to the rounding of v0 + e)
            assert (lo >= n["min"] - 1e-4).all() and (hi <= n["max"] + 1e-4).all()
        else:
            a, b = nodes[n["first"]], nodes[n["first"] + 1]
            assert n["first"] > i
            for c in (a, b):
                assert (c["min"] >= n["min"]).all() and (c["max"] <= n["max"]).all()
            # the parent's box is the union of its children's
            assert (np.minimum(a["min"], b["min"]) == n["min"]).all()
            assert (np.maximum(a["max"], b["max"]) == n["max"]).all()
            stack += [n["first"], n["first"] + 1]
    assert reached == len(nodes)
    assert (seen == 1).all() and len(tris) == n_casting


@pytest.mark.parametrize("gen", [small_scene, map_scene])
def test_bvh_invariants(gen):
    scene = gen()
    bvh, pos, mesh = kit.scene_bvh(scene)
    check_tree(bvh, scene.triangle_count())
    leaf = bvh.nodes["count"][bvh.nodes["count"] > 0]
    assert np.percentile(leaf, 99) <= 4


def test_non_casting_and_degenerate_triangles_are_left_out():
    pos = np.array([[0, 0, 0, 1, 0, 0, 0, 0, 1],      # casts
                    [0, 0, 0, 1, 0, 0, 2, 0, 0],      # degenerate (collinear)
                    [5, 0, 5, 6, 0, 5, 5, 0, 6]], np.float32)
    mat = np.array([0, 0, 1], np.uint32)
    bvh = Bvh.build(pos, mat, np.zeros(3, np.uint32), np.array([1, 0], np.uint8))
    assert len(bvh.tris) == 1 and len(bvh.nodes) == 1 and bvh.nodes[0]["count"] == 1


def test_empty_scene_is_one_unreachable_leaf():
    # bvh.rs:76-80
    pos = np.array([[0, 0, 0, 1, 0, 0, 0, 0, 1]], np.float32)
    bvh = Bvh.build(pos, np.zeros(1, np.uint32), np.zeros(1, np.uint32), np.array([0], np.uint8))
    assert len(bvh.nodes) == 1 and bvh.nodes[0]["count"] == 0 and len(bvh.tris) == 1
    assert bvh.nodes[0]["min"][0] == np.finfo(np.float32).max


def _build_threads(scene, threads):
    import ctypes as C
    from lightbaker_amd import _native as N
    pos = np.concatenate([m.positions.reshape(-1, 9) for m in scene.meshes]).astype(np.float32)
    mat = np.concatenate([np.full(len(m.positions) // 3, m.material, np.uint32)
                          for m in scene.meshes])
    chart = np.arange(len(pos), dtype=np.uint32)
    casts = np.array([m.casts for m in scene.materials], np.uint8)
    lib = N.lib()
    h = C.c_void_p()
    N.check(lib.lb_bvh_build_threads(N.ptr(pos), len(pos), N.ptr(mat), N.ptr(chart), N.ptr(casts),
                                     len(casts), threads, C.byref(h)))
    try:
        p = C.c_void_p()
        n = lib.lb_bvh_nodes(h, C.byref(p))
        nodes = np.frombuffer((C.c_ubyte * (n * 32)).from_address(p.value), np.uint8).copy()
        n = lib.lb_bvh_tris(h, C.byref(p))
        tris = np.frombuffer((C.c_ubyte * (n * 48)).from_address(p.value), np.uint8).copy()
    finally:
        lib.lb_bvh_destroy(h)
    return nodes, tris


@pytest.mark.parametrize("gen", [lambda: small_scene(1), lambda: map_scene(2)],
                         ids=["small-scene", "map-scene"])
def test_threads_build_the_same_tree(gen):
    """Subtrees built on their own threads and numbered afterwards, and the large slices' bins
    reduced in parts, give the single loop's nodes and triangle order byte for byte (the map:
    204k triangles, its top slices past the parallel binning's 65 536)."""
    scene = gen()
    one = _build_threads(scene, 1)
    for t in (2, 3, 8):
        many = _build_threads(scene, t)
        assert np.array_equal(one[0], many[0]) and np.array_equal(one[1], many[1]), t
