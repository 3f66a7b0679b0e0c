"""The bake's device rasterizer (raster_gpu.hip) against the host restatement of raster.rs:39-114
(lb_layout_rasterize): every texel of every baked tile bit for bit, on scenes whose tiles hold
covered, conservative (thin and small faces) and empty texels; and on a rebake's dirty subset."""
from __future__ import annotations

import numpy as np
import pytest

from lightbaker_amd import _native as N
from lightbaker_amd.bake import BakeSession, layout
from lightbaker_amd.scenes import map_scene, reference_test_scene, small_scene

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("gen", [lambda: reference_test_scene(8.0), lambda: small_scene(1),
                                 lambda: map_scene(2)], ids=["reference-test-scene", "small",
                                                            "map"])
def test_device_raster_equals_host(gpu, gen):
    scene = gen()
    L = layout(scene)
    print(f"raster: {len(L.tiles)} tiles", flush=True)
    sess = BakeSession(scene, gpu, form=N.FORM_TUNED)
    try:
        got = sess.array(N.BAKE_TEXELS, N.TEXEL)
    finally:
        sess.close()
    want = np.concatenate([L.rasterize(t).reshape(-1) for t in range(len(L.tiles))])
    assert len(got) == len(want)
    bad = np.nonzero(got.view(np.uint8).reshape(len(got), -1) !=
                     want.view(np.uint8).reshape(len(want), -1))[0]
    assert len(bad) == 0, f"{len(np.unique(bad))} texels differ, first {bad[:3]}"
    assert (got["pos"][:, 3] > 0.5).sum() > 0


def test_device_raster_of_dirty_tiles(gpu):
    scene = reference_test_scene(8.0)
    L = layout(scene)
    dirty = np.zeros(len(L.tiles), bool)
    dirty[[1, 5, 6, 11]] = True
    sess = BakeSession(scene, gpu, form=N.FORM_PLAIN, tile_dirty=dirty,
                       cache=None if dirty.all() else _cache_for(L, scene))
    try:
        got = sess.array(N.BAKE_TEXELS, N.TEXEL)
    finally:
        sess.close()
    want = np.concatenate([L.rasterize(t).reshape(-1) for t in np.nonzero(dirty)[0]])
    assert np.array_equal(got.view(np.uint8), want.view(np.uint8))


def _cache_for(L, scene):
    """A cache of zeros over every tile (the rasterizer does not read it)."""
    from lightbaker_amd.bake import TileCache
    n = int((L.tiles[:, 2] * L.tiles[:, 3]).sum())
    z = np.zeros(n, np.float32)
    return TileCache(known=np.zeros(n, np.uint8), state=np.zeros(n, np.uint8), field=z, ao=z,
                     direct=np.zeros((n, 3), np.float32), indirect=np.zeros((n, 3), np.float32),
                     probes=np.zeros((_probes(scene), 16), np.float32))


def _probes(scene):
    v = scene.volume
    return int(np.prod([max(int(round((v.max[a] - v.min[a]) / v.step)) + 1, 2) for a in range(3)]))
