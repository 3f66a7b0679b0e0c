"""The tile dependencies (deps.rs:72-121, bake.rs:124-136,447-477, volume.rs:36-56) with the
triangle filter on the device, through the C-ABI:

  * lb_deps_swept equals the oracle's filter (oracle/deps_ref.c) bit for bit: random triangles
    and zones, casting and not, and every tile zone of the reference test scene;
  * each tile's and probe chunk's hash equals BLAKE3 over the byte stream the reference feeds it,
    assembled here independently from the oracle's filter bits (oracle/deps_ref.py);
  * the reference's own rebake test (tests/bake.rs:140-150), restated: the same scene hashes the
    same (nothing to bake, no chunk to trace); with the box moved, some tiles but not all change;
  * a rebake from those dirty flags bakes exactly the changed tiles.
"""
from __future__ import annotations

import numpy as np
import pytest

from oracle import deps_ref as D
from lightbaker_amd import _native as N
from lightbaker_amd import deps
from lightbaker_amd.bake import bake, layout
from lightbaker_amd.scenes import reference_test_scene

pytestmark = pytest.mark.gpu


def swept_device(gpu, pos, casts, zones, sun, top):
    p = np.ascontiguousarray(pos, np.float32).reshape(-1, 9)
    z = np.ascontiguousarray(zones, np.float32).reshape(-1, 6)
    c = None if casts is None else np.ascontiguousarray(casts, np.uint8)
    s = np.ascontiguousarray(sun, np.float32)
    out = np.zeros((len(z), (len(p) + 31) // 32), np.uint32)
    N.check(N.lib().lb_deps_swept(gpu.handle, N.ptr(p), len(p), N.ptr(c), N.ptr(z), len(z),
                                  N.ptr(s), float(np.float32(top)), N.ptr(out)))
    return out


def scene_tris(scene):
    pos = np.concatenate([m.positions.reshape(-1, 9) for m in scene.meshes]).astype(np.float32)
    casts = np.concatenate([np.full(len(m.positions) // 3, scene.materials[m.material].casts,
                                    np.uint8) for m in scene.meshes])
    first = np.concatenate([[0], np.cumsum([len(m.positions) // 3 for m in scene.meshes])])
    return pos, casts, first


def sun_of(scene):
    d = np.asarray(scene.sun_direction, np.float32)
    return (d * (np.float32(1) / np.sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]))).astype(np.float32)


def test_swept_random(gpu):
    rng = np.random.default_rng(21)
    n = 3001
    c = rng.uniform(-30, 30, (n, 1, 3))
    pos = (c + rng.normal(scale=1.5, size=(n, 3, 3))).reshape(n, 9).astype(np.float32)
    casts = (rng.random(n) < 0.8).astype(np.uint8)
    zc = rng.uniform(-30, 30, (40, 3))
    zh = rng.uniform(1, 12, (40, 3))
    zones = np.concatenate([zc - zh, zc + zh], 1).astype(np.float32)
    for sun in ([0.3, 0.9, 0.2], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0], [-0.5, 0.5, -0.7]):
        s = np.asarray(sun, np.float32)
        s /= np.linalg.norm(s)
        top = np.float32(pos[:, 1::3].max())
        got = swept_device(gpu, pos, casts, zones, s, top)
        want = D.bits(pos, casts, zones, s, top)
        assert np.array_equal(got, want), sun
        assert 0 < np.unpackbits(want.view(np.uint8)).sum() < n * 40
    assert np.array_equal(swept_device(gpu, pos, None, zones, s, top), D.bits(pos, None, zones, s, top))


def tile_streams(scene):
    """Each tile's own box and tile_seed stream (bake.rs:447-477) from the layout's charts."""
    L = layout(scene)
    pos, _, first = scene_tris(scene)
    meshes = [(m.positions.reshape(-1, 9).astype(np.float32), m.material,
               scene.materials[m.material].casts) for m in scene.meshes]
    tc = L.tri_charts
    out = []
    for t in range(len(L.tiles)):
        charts = []
        lo, hi = np.full(3, np.inf, np.float32), np.full(3, -np.inf, np.float32)
        for c in np.nonzero(L.chart_tiles == t)[0]:
            g = np.nonzero(tc == c)[0]
            m = int(np.searchsorted(first, g[0], side="right") - 1)
            charts.append((m, list(g - first[m])))
            v = pos[g].reshape(-1, 3)
            lo, hi = np.minimum(lo, v.min(0)), np.maximum(hi, v.max(0))
        key = (int(L.tiles[t, 0]), int(L.tiles[t, 1]))
        dens = (scene.tile_density or {}).get(key)
        out.append((np.concatenate([lo, hi]), D.tile_seed_stream(key, charts, meshes, dens)))
    return out, meshes


@pytest.mark.parametrize("own_density", [False, True], ids=["one-density", "tile-density"])
def test_tile_hashes_match_the_streams(gpu, own_density):
    scene = reference_test_scene(8.0)
    if own_density:
        # one tile laid out at a density of its own: its seed ends with the density
        L = layout(scene)
        scene.tile_density = {(int(L.tiles[1, 0]), int(L.tiles[1, 1])): 6.0}
    s = scene.settings
    g = deps.global_hash(scene)
    got = deps.tile_hashes(gpu, scene, g)
    streams, meshes = tile_streams(scene)
    pos, casts, first = scene_tris(scene)
    reach = np.float32(max(s.gi_reach, s.ao_reach) if s.gi_bounces > 0 else s.ao_reach)
    zones = np.stack([np.concatenate([own[:3] - reach, own[3:] + reach]) for own, _ in streams])
    top = np.float32(pos[:, 1::3].max())
    bits = D.bits(pos, casts, zones, sun_of(scene), top)
    assert np.array_equal(bits, swept_device(gpu, pos, casts, zones, sun_of(scene), top))
    for t, (own, seed_stream) in enumerate(streams):
        seed = deps.blake3(seed_stream)
        want = deps.blake3(D.hash_reach_stream(g, seed, meshes, bits[t], first))
        assert bytes(got[t]) == want, t


def test_probe_hashes_match_the_streams(gpu):
    scene = reference_test_scene(8.0)
    g = deps.global_hash(scene)
    keys, got = deps.probe_hashes(gpu, scene, g)
    v = scene.volume
    f = np.float32
    n = [max(int(np.floor((f(v.max[a]) - f(v.min[a])) / f(v.step) + f(0.5))) + 1, 2) for a in range(3)]
    k = np.arange(n[0] * n[1] * n[2])
    p = np.stack([f(v.min[0]) + (k % n[0]).astype(f) * f(v.step),
                  f(v.min[1]) + (k // (n[0] * n[2])).astype(f) * f(v.step),
                  f(v.min[2]) + ((k // n[0]) % n[2]).astype(f) * f(v.step)], 1)
    size = f(scene.settings.tile_size)
    kx, kz = np.floor(p[:, 0] / size).astype(int), np.floor(p[:, 2] / size).astype(int)
    want_keys = sorted({(a, b) for a, b in zip(kx, kz)})
    assert [tuple(x) for x in keys] == want_keys
    pos, casts, first = scene_tris(scene)
    meshes = [(m.positions.reshape(-1, 9).astype(np.float32), m.material,
               scene.materials[m.material].casts) for m in scene.meshes]
    r = f(scene.settings.gi_reach)
    zones = []
    for a, b in want_keys:
        q = p[(kx == a) & (kz == b)]
        zones.append(np.concatenate([q.min(0) - r, q.max(0) + r]))
    bits = D.bits(pos, casts, np.array(zones, f), sun_of(scene), f(pos[:, 1::3].max()))
    vj = deps.volume_json(scene).encode()
    for i, (a, b) in enumerate(want_keys):
        seed = D.i32(a) + D.i32(b) + vj
        assert bytes(got[i]) == deps.blake3(D.hash_reach_stream(g, seed, meshes, bits[i], first)), i


def test_reference_rebake_behaviour(gpu):
    """tests/bake.rs:140-150: baked again unchanged, nothing to bake; with the box moved, only the
    tiles it reaches."""
    a, b = reference_test_scene(8.0), reference_test_scene(8.0)
    h0, h1 = deps.tile_hashes(gpu, a), deps.tile_hashes(gpu, b)
    assert not deps.dirty(h0, h1).any()
    k0, p0 = deps.probe_hashes(gpu, a)
    k1, p1 = deps.probe_hashes(gpu, b)
    assert np.array_equal(k0, k1) and not deps.dirty(p0, p1).any()
    moved = reference_test_scene(6.0)
    d = deps.dirty(h0, deps.tile_hashes(gpu, moved))
    print(f"deps: box moved, {d.sum()} of {len(d)} tiles dirty", flush=True)
    assert 0 < d.sum() < len(d)
    km, pm = deps.probe_hashes(gpu, moved)
    cd = deps.dirty(p0, pm)
    assert 0 < cd.sum() < len(cd)
    # a changed material's albedo changes the settings, so every tile
    c = reference_test_scene(8.0)
    c.materials[2].albedo = (0.8, 0.1, 0.1)
    assert deps.dirty(h0, deps.tile_hashes(gpu, c)).all()


def test_rebake_from_the_dirty_flags(gpu):
    a = reference_test_scene(8.0)
    print("deps: first bake", flush=True)
    b0 = bake(a, gpu, form=N.FORM_TUNED, graph=True)
    h0 = deps.tile_hashes(gpu, a)
    k0, p0 = deps.probe_hashes(gpu, a)
    moved = reference_test_scene(6.0)
    d = deps.dirty(h0, deps.tile_hashes(gpu, moved))
    km, pm = deps.probe_hashes(gpu, moved)
    pd = deps.probe_dirty(moved, km, deps.dirty(p0, pm))
    print(f"deps: rebake of {d.sum()} tiles, {pd.sum()} probes", flush=True)
    b1 = bake(moved, gpu, form=N.FORM_TUNED, graph=True, cache=b0.cache, tile_dirty=d,
              probe_dirty=pd)
    assert b1.info["tiles_baked"] == d.sum() and b1.info["probes_baked"] == pd.sum()
    assert np.array_equal(b1.cache.probes[~pd], b0.cache.probes[~pd])


def test_tile_density_change_dirties_its_tile_only(gpu):
    """A tile given a density of its own is rebaked (its hash changes); the others keep theirs:
    the density moves no casting triangle."""
    a = reference_test_scene(8.0)
    h0 = deps.tile_hashes(gpu, a)
    L = layout(a)
    key = (int(L.tiles[2, 0]), int(L.tiles[2, 1]))
    b = reference_test_scene(8.0)
    b.tile_density = {key: 5.0}
    d = deps.dirty(h0, deps.tile_hashes(gpu, b))
    assert list(np.nonzero(d)[0]) == [2]
    # a listed density equal to the scene's is still a change of the tile's inputs
    c = reference_test_scene(8.0)
    c.tile_density = {key: float(c.settings.texels_per_meter)}
    assert list(np.nonzero(deps.dirty(h0, deps.tile_hashes(gpu, c)))[0]) == [2]
