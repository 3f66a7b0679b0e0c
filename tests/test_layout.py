"""Layout::build (layout.rs:89-261) and rasterize (raster.rs:39-114), restated in C++: what their
construction guarantees, on the CPU."""
import math

import numpy as np
import pytest

from lightbaker_amd import _native as N
from lightbaker_amd.bake import layout
from lightbaker_amd.scenes import map_scene, reference_test_scene, small_scene


def dense_small():
    """The small scene with per-tile densities of 2, 8 and 16 texels/m beside its own 4."""
    s = small_scene(1)
    s.tile_density = {(-1, -1): 2.0, (0, -1): 8.0, (-1, 0): 16.0}
    return s


def density_of(s, kx, kz):
    return (s.tile_density or {}).get((int(kx), int(kz)), s.settings.texels_per_meter)


@pytest.fixture(scope="module", params=["ref", "small", "tile-density"])
def laid(request):
    s = {"ref": lambda: reference_test_scene(8.0), "small": lambda: small_scene(1),
         "tile-density": dense_small}[request.param]()
    return s, layout(s)


def test_tiles_are_disjoint_aligned_and_in_the_atlas(laid):
    s, L = laid
    assert L.width == s.settings.atlas_width and L.height % 4 == 0
    occ = np.zeros((L.height, L.width), np.int32)
    for kx, kz, w, h, x, y in L.tiles:
        assert w % 4 == 0 and h % 4 == 0 and x % 4 == 0 and y % 4 == 0
        assert x + w <= L.width and y + h <= L.height
        occ[y:y + h, x:x + w] += 1
    assert occ.max() == 1
    # tiles in key order (layout.rs:108-112)
    keys = [tuple(t[:2]) for t in L.tiles]
    assert keys == sorted(keys)


def test_every_receiving_triangle_maps_inside_its_tile(laid):
    s, L = laid
    k = 0
    for m in s.meshes:
        p = m.positions.reshape(-1, 3, 3)
        for t in range(len(p)):
            c = L.tri_charts[k]
            k += 1
            if not s.materials[m.material].receives:
                assert c == 0xFFFFFFFF
                continue
            kx, kz, w, h, x0, y0 = L.tiles[L.chart_tiles[c]]
            a = L.affines[c]
            for v in p[t]:
                q = np.append(v, 1.0).astype(np.float32)
                u, vv = float(a[0] @ q), float(a[1] @ q)
                pad = s.settings.padding
                assert x0 + pad - 1e-3 <= u <= x0 + w - pad + 1e-3
                assert y0 + pad - 1e-3 <= vv <= y0 + h - pad + 1e-3


def test_texels_lie_on_their_chart_and_map_back_to_themselves(laid):
    s, L = laid
    pos = np.concatenate([m.positions.reshape(-1, 3, 3) for m in s.meshes])
    for t in range(len(L.tiles)):
        kx, kz, w, h, x0, y0 = L.tiles[t]
        tex = L.rasterize(t)
        ys, xs = np.nonzero(tex["pos"][..., 3] > 0.5)
        assert len(ys) > 0
        for y, x in list(zip(ys, xs))[::7]:
            tx = tex[y, x]
            c = tx["normal"][3:].view(np.uint32)[0]
            assert L.chart_tiles[c] == t
            a = L.affines[c]
            q = np.append(tx["pos"][:3], 1.0).astype(np.float32)
            u, v = float(a[0] @ q), float(a[1] @ q)
            # centre-covered texels sample their centre; conservative ones a point within half a
            # texel of it (raster.rs:78-82)
            assert abs(u - (x0 + x + 0.5)) <= 0.5 + 1e-3 and abs(v - (y0 + y + 0.5)) <= 0.5 + 1e-3
            n = tx["normal"][:3]
            assert abs(np.linalg.norm(n) - 1.0) < 1e-5
            # dx, dy step one texel: 1 / density metres (the tile's), in the face's plane
            assert abs(np.linalg.norm(tx["dx"][:3]) - 1.0 / density_of(s, kx, kz)) < 1e-3
            assert abs(float(np.dot(tx["dx"][:3], n))) < 1e-4


def test_layout_is_deterministic():
    s = small_scene(1)
    a, b = layout(s), layout(s)
    assert np.array_equal(a.affines, b.affines) and np.array_equal(a.tiles, b.tiles)


def test_reference_scene_layout_counts():
    # the reference test scene: 29 charts in 16 tiles of 16 m (GPU bake logs the same)
    L = layout(reference_test_scene(8.0))
    assert len(L.affines) == 29 and len(L.tiles) == 16


def test_a_tile_wider_than_the_atlas_is_an_error():
    s = reference_test_scene(8.0)
    s.settings.texels_per_meter = 80.0
    with pytest.raises(N.NativeError, match="texels wide, the atlas 1024"):
        layout(s)


def test_map_scene_layout():
    s = map_scene(2)
    L = layout(s)
    assert len(L.tiles) == 16  # 100 m in tiles of 25 m
    assert L.width == 4096 and 0 < L.height <= 4096


def test_tile_density_default_is_the_reference_layout():
    """No per-tile density, an empty one, or every tile listed at the scene's own density: the
    same layout byte for byte (the reference's single density, layout.rs:89-115)."""
    base = layout(small_scene(1))
    s = small_scene(1)
    s.tile_density = {}
    same = layout(s)
    assert np.array_equal(same.affines, base.affines) and np.array_equal(same.tiles, base.tiles)
    s.tile_density = {(int(t[0]), int(t[1])): 4.0 for t in base.tiles}
    same = layout(s)
    assert np.array_equal(same.affines, base.affines) and np.array_equal(same.tiles, base.tiles)
    assert np.array_equal(same.tri_charts, base.tri_charts)


def test_tile_density_scales_its_tile_only():
    """A tile at 16 texels/m against the scene's 4: each chart's affine rows (a texel a step of
    1 / density metres) have the tile's density as their length, the charts being the same; its
    texel count grows about 16-fold; a tile left at the scene's density is laid out as before
    (the packing may turn a chart another way, swapping its rows, when the sizes change)."""
    base = layout(small_scene(1))
    s = dense_small()
    L = layout(s)
    assert np.array_equal(L.tri_charts, base.tri_charts)  # the same charts, grown the same way
    for c in range(len(L.affines)):
        kx, kz = L.tiles[L.chart_tiles[c]][:2]
        norms = np.linalg.norm(L.affines[c][:, :3].astype(np.float64), axis=1)
        assert np.allclose(norms, density_of(s, kx, kz), rtol=1e-5)
        if (int(kx), int(kz)) == (0, 0):
            assert np.array_equal(L.affines[c][:, :3], base.affines[c][:, :3])
    tex = {tuple(t[:2]): int(t[2]) * int(t[3]) for t in L.tiles}
    tex0 = {tuple(t[:2]): int(t[2]) * int(t[3]) for t in base.tiles}
    assert 8 < tex[(-1, 0)] / tex0[(-1, 0)] < 20 and 0.1 < tex[(-1, -1)] / tex0[(-1, -1)] < 0.5
    assert tex[(0, 0)] == tex0[(0, 0)]


@pytest.mark.parametrize("bad,msg", [({(0, 0): 0.0}, "not positive"),
                                     ({(0, 0): float("nan")}, "not positive")])
def test_tile_density_rejects_bad_values(bad, msg):
    s = small_scene(1)
    s.tile_density = bad
    with pytest.raises(N.NativeError, match=msg):
        layout(s)
