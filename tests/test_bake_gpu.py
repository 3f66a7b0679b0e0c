"""The reference's integration test (crates/lightbaker/tests/bake.rs:95-152) on this bake: the same
scene and the same thresholds, through every kernel form and with the bounce loop replayed from
a hipGraph. Its determinism half (a second bake gives the same `direct`, bake.rs:150) is here; its
cache half (a rebake with nothing dirty traces nothing and gives the same `direct`) is in
test_rebake_gpu.py, with the dirty tiles named by the caller: the hashes that pick them (deps.rs)
stay with the reference's host."""
from __future__ import annotations

import math

import numpy as np
import pytest

from lightbaker_amd import _native as N
from lightbaker_amd.bake import bake
from lightbaker_amd.scenes import reference_test_scene

pytestmark = pytest.mark.gpu


def ground_texel(s, b, x, z):
    """tests/bake.rs:75-89: the atlas texel a ground point falls in."""
    g = s.meshes[0].positions.reshape(-1, 3, 3)
    tri = next(t for t in range(len(g))
               if g[t, :, 0].min() <= x <= g[t, :, 0].max() and g[t, :, 2].min() <= z <= g[t, :, 2].max())
    c = b.tri_charts[tri]
    a = b.charts[c]
    p = np.array([x, 0.0, z, 1.0], np.float32)
    return int(math.floor(float(np.dot(a[1], p)))), int(math.floor(float(np.dot(a[0], p))))


@pytest.fixture(scope="module", params=[(N.FORM_PLAIN, False), (N.FORM_TUNED, False),
                                        (N.FORM_TUNED, True)],
                ids=["plain", "tuned", "tuned-graph"])
def baked(request, gpu):
    form, graph = request.param
    print(f"bake form={form} graph={graph}", flush=True)
    s = reference_test_scene(8.0)
    b = bake(s, gpu, form=form, graph=graph)
    print({k: b.info[k] for k in ("tiles", "charts", "texels_baked", "gather_items", "probes",
                                   "ms_direct", "ms_bounces", "ms_volume")}, flush=True)
    b.form, b.graph = form, graph
    return s, b


def test_shadow_east_of_the_box(baked):
    s, b = baked
    assert b.info["tiles_baked"] == b.info["tiles"]
    shade = b.direct[ground_texel(s, b, 11.0, 11.0)]
    open_ = b.direct[ground_texel(s, b, 4.0, 11.0)]
    assert shade[0] < 0.4, f"in the box's shadow: field {shade[0]}"
    assert open_[0] > 0.6, f"in the sun: field {open_[0]}"
    foot = b.direct[ground_texel(s, b, 7.85, 11.0)]
    assert foot[1] < open_[1] - 0.1, f"occlusion at the box's foot {foot[1]} vs open {open_[1]}"


def test_red_wall_bleeds(baked):
    s, b = baked
    def small(yx):
        return b.indirect[yx[0] // 2, yx[1] // 2]
    near = small(ground_texel(s, b, 38.5, 40.0))
    far = small(ground_texel(s, b, 20.0, 40.0))
    assert near[0] / near[1] > far[0] / far[1] * 1.15, f"near {near} far {far}"
    assert far[2] > 0.0


def test_volume_sees_the_sun(baked):
    s, b = baked
    v = s.volume

    def probe(x, y, z):
        i, j, k = (int((x - v.min[0]) / v.step), int((y - v.min[1]) / v.step),
                   int((z - v.min[2]) / v.step))
        return b.volume[j, k, i]
    assert probe(20.0, 0.5, 40.0)[0][3] > 0.9
    assert probe(42.0, 0.5, 40.0)[0][3] < 0.1


def test_strip_has_light_all_along(baked):
    s, b = baked
    first = sum(len(m.positions) // 3 for m in s.meshes[:3])
    c = b.tri_charts[first]
    a = b.charts[c]
    ds = b.info["atlas_width"] // b.info["indirect_width"]
    for k in range(12):
        y = 0.3 + k * 0.19
        p = np.array([30.0, y, 20.025, 1.0], np.float32)
        x = int(math.floor(float(np.dot(a[0], p)))) // ds
        yy = int(math.floor(float(np.dot(a[1], p)))) // ds
        e = b.indirect[yy, x]
        assert e.sum() > 0.1, f"strip at {y} m: {e}"


def test_rebake_is_identical(gpu, baked):
    """bake.rs:150: the same scene bakes to the same `direct`; the forms agree on it."""
    s, b = baked
    again = bake(s, gpu, form=b.form, graph=b.graph)
    assert np.array_equal(again.direct, b.direct)
    assert np.array_equal(again.indirect, b.indirect)
    plain = bake(s, gpu, form=N.FORM_PLAIN)
    # the field comes from exact shadow bits; occlusion from one compensated sum per texel
    assert np.array_equal(plain.direct[..., 0], b.direct[..., 0])
    assert np.abs(plain.direct - b.direct).max() < 1e-5
