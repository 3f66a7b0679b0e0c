"""A rebake (crates/lightbaker/src/bake.rs:123-149, volume.rs:44-84): the clean tiles' light and
the clean probes come from the cache, the dirty tiles' items go into each pass as one list (dirty
tiles in layout order, each tile's texels row-major; an item's index in it is its ray key).

Bar: against the oracle's restatement of the same rebake (oracle/bake_ref.py over the dirty
tiles, its first atlas written from the cached TileLight by bake.rs:435-444), the PLAIN form and
the pooled schedule (with FP32 node records, the reference's boxes) give the atlas light, every
dirty tile's TileLight and every traced probe bit for bit; the lock-step schedule (quantized
nodes, the default) within 1e-4 relative. A rebake with nothing dirty returns the
cache and the first bake's outputs exactly (tests/bake.rs:147-150: "baked again unchanged:
nothing to bake"), and one with everything dirty equals a plain bake.
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle.bake_ref as ref
from lightbaker_amd import _native as N
from lightbaker_amd.bake import BakeSession, bake
from lightbaker_amd.scenes import rebake_scene, reference_test_scene

from .test_bake_parity_gpu import probe_items

pytestmark = pytest.mark.gpu


def dirty_masks(scene, tiles, seed, share):
    """A seeded subset of the layout's tiles, and the probes whose ground-plan chunk
    (layout.rs:347-349 `tile_of`) is one of theirs (volume.rs:39-45)."""
    rng = np.random.default_rng(seed)
    nt = len(tiles)
    k = max(1, int(round(nt * share)))
    dirty = np.zeros(nt, bool)
    dirty[rng.choice(nt, k, replace=False)] = True
    keys = {(int(tiles[t, 0]), int(tiles[t, 1])) for t in np.nonzero(dirty)[0]}
    it = probe_items(scene)
    size = np.float32(scene.settings.tile_size)
    kx = np.floor(it["pos"][:, 0] / size).astype(np.int64)
    kz = np.floor(it["pos"][:, 2] / size).astype(np.int64)
    pdirty = np.array([(int(a), int(b)) in keys for a, b in zip(kx, kz)], bool)
    return dirty, pdirty


def same(a, b):
    return np.array_equal(np.asarray(a).view(np.uint8), np.asarray(b).view(np.uint8))


@pytest.fixture(scope="module")
def first(gpu):
    print("rebake: first bake (plain)", flush=True)
    s = reference_test_scene(8.0)
    return s, bake(s, gpu, form=N.FORM_PLAIN)


@pytest.mark.parametrize("form,schedule", [(N.FORM_PLAIN, 0), (N.FORM_TUNED, 0),
                                           (N.FORM_TUNED, 1)], ids=["plain", "tuned", "pooled"])
def test_rebake_matches_oracle(gpu, first, form, schedule):
    scene, b0 = first
    dirty, pdirty = dirty_masks(scene, b0.tiles, 33, 0.3)
    print(f"rebake parity form={form} schedule={schedule}: {dirty.sum()} of {len(dirty)} tiles, "
          f"{pdirty.sum()} of {len(pdirty)} probes", flush=True)
    # the pooled schedule is bit-exact with FP32 node records (the reference's boxes)
    gpu.set_tuning(schedule=schedule, nodes=N.NODES_WIDE if schedule == 1 else N.NODES_QUANT)
    try:
        sess = BakeSession(scene, gpu, form=form, graph=form == N.FORM_TUNED, cache=b0.cache,
                           tile_dirty=dirty, probe_dirty=pdirty)
        sess.execute()
        sess.finish()
    finally:
        gpu.set_tuning(schedule=0)
    info = sess.info()
    assert info["tiles_baked"] == dirty.sum() and info["probes_baked"] == pdirty.sum()
    baked = sess.array(N.BAKE_BAKED_TILES, np.uint32)
    assert np.array_equal(baked, np.nonzero(dirty)[0])

    # the oracle over the dirty tiles, from its own first atlas
    P = sess.array(N.BAKE_PARAMS, N.PARAMS)
    PV = P.copy()
    PV["counts"][0][2] = scene.volume.rays
    tiles_all = sess.array(N.BAKE_TILES, np.int32).reshape(-1, 6)
    aw, ah = int(P["atlas"][0][0]), int(P["atlas"][0][1])
    light0 = ref.clean_light(tiles_all, dirty, b0.cache, aw, ah)
    assert same(sess.array(N.BAKE_LIGHT0, np.float32).reshape(ah, aw, 4), light0)
    want = ref.bake({
        "settings": scene.settings, "texels": sess.array(N.BAKE_TEXELS, N.TEXEL),
        "tiles": tiles_all[dirty], "light0": light0, "aw": aw, "ah": ah,
        "nodes": sess.array(N.BAKE_NODES, N.NODE), "tris": sess.array(N.BAKE_TRIS, N.TRI),
        "mats": sess.array(N.BAKE_MATERIALS, np.float32).reshape(-1, 4),
        "charts": sess.array(N.BAKE_CHARTS, np.float32).reshape(-1, 8), "params": P,
        "sun": P["sun_dir"][0][:3].copy(), "irr": np.asarray(scene.sun_irradiance, np.float32),
        "probes": probe_items(scene)[pdirty], "volume_params": PV,
    })
    assert np.array_equal(sess.array(N.BAKE_GATHER_INDEX, np.uint32), want["gathered"])

    light = sess.array(N.BAKE_LIGHT, np.float32).reshape(-1, 4)
    wl = want["light"].reshape(-1, 4)
    tc = sess.tile_cache()
    # the dirty tiles' TileLight, texels in the cache order
    offs = np.concatenate([[0], np.cumsum(tiles_all[:, 2] * tiles_all[:, 3])])
    idx = np.concatenate([np.arange(offs[t], offs[t + 1]) for t in np.nonzero(dirty)[0]])
    own = want["known"] != 0  # texels the volume does not light (bake.rs:337-359)
    probes = sess.array(N.BAKE_PROBES, np.float32).reshape(-1, 16)
    assert same(probes[~pdirty], b0.cache.probes[~pdirty]), "clean probes changed"
    clean = np.concatenate([np.arange(offs[t], offs[t + 1]) for t in np.nonzero(~dirty)[0]])
    assert same(tc.indirect[clean], b0.cache.indirect[clean]), "clean tiles changed"
    assert same(tc.known[idx][own], want["known"][own])
    if form == N.FORM_PLAIN or schedule == 1:
        bad = np.nonzero((light.view(np.uint32) != wl.view(np.uint32)).any(axis=1))[0]
        assert len(bad) == 0, f"{len(bad)} atlas texels differ: {light[bad[:3]]} vs {wl[bad[:3]]}"
        for k in ("field", "ao", "direct", "indirect"):
            assert same(getattr(tc, k)[idx][own], want[k][own]), k
        assert same(probes[pdirty], want["probes"]), "traced probes differ"
    else:
        err = np.abs(light - wl) / (np.abs(wl) + 1e-3)
        assert err.max() < 1e-4, err.max()
        for k in ("direct", "indirect"):
            a, w = getattr(tc, k)[idx][own], want[k][own]
            assert (np.abs(a - w) / (np.abs(w) + 1e-3)).max() < 1e-4, k
        pe = np.abs(probes[pdirty] - want["probes"]) / (np.abs(want["probes"]) + 1e-3)
        assert pe.max() < 1e-4, pe.max()
    sess.close()


def test_rebake_nothing_dirty(gpu, first):
    """tests/bake.rs:147-150: baked again unchanged, nothing is traced and `direct` is equal."""
    scene, b0 = first
    nt, npb = len(b0.tiles), len(b0.cache.probes)
    again = bake(scene, gpu, form=N.FORM_TUNED, cache=b0.cache, tile_dirty=np.zeros(nt, bool),
                 probe_dirty=np.zeros(npb, bool))
    assert again.info["tiles_baked"] == 0 and again.info["probes_baked"] == 0
    assert again.info["texels_baked"] == 0
    for k in ("direct", "indirect", "validity", "volume"):
        assert same(getattr(again, k), getattr(b0, k)), k
    for k in ("known", "state", "field", "ao", "direct", "indirect", "probes"):
        assert same(getattr(again.cache, k), getattr(b0.cache, k)), k


def test_rebake_everything_dirty(gpu, first):
    scene, b0 = first
    nt, npb = len(b0.tiles), len(b0.cache.probes)
    again = bake(scene, gpu, form=N.FORM_PLAIN, cache=b0.cache, tile_dirty=np.ones(nt, bool),
                 probe_dirty=np.ones(npb, bool))
    assert again.info["tiles_baked"] == nt
    for k in ("direct", "indirect", "validity", "volume", "light"):
        assert same(getattr(again, k), getattr(b0, k)), k


def test_rebake_cache_checked(gpu, first):
    scene, b0 = first
    nt = len(b0.tiles)
    dirty = np.zeros(nt, bool)
    dirty[0] = True
    short = type(b0.cache)(**{k: (v[:-1] if v is not None and k != "probes" else v)
                              for k, v in vars(b0.cache).items()})
    with pytest.raises(RuntimeError, match="bake cache"):
        bake(scene, gpu, cache=short, tile_dirty=dirty)
    with pytest.raises(RuntimeError, match="bake cache"):
        bake(scene, gpu, cache=None, tile_dirty=dirty)


def test_rebake_uneven_tiles(gpu):
    """Tiles of 4x different texel counts, a seeded dirty subset, in one list: the pooled
    schedule's rebake equals the PLAIN form's bit for bit."""
    scene = rebake_scene(3, size=100.0)
    print("rebake uneven tiles: first bake", flush=True)
    b0 = bake(scene, gpu, form=N.FORM_TUNED, graph=True)
    counts = b0.tiles[:, 2] * b0.tiles[:, 3]
    dirty, pdirty = dirty_masks(scene, b0.tiles, 33, 0.4)
    out = {}
    for name, form, schedule in (("plain", N.FORM_PLAIN, 0), ("pooled", N.FORM_TUNED, 1)):
        print(f"rebake uneven tiles: {name}", flush=True)
        gpu.set_tuning(schedule=schedule, nodes=N.NODES_WIDE)
        try:
            out[name] = bake(scene, gpu, form=form, graph=form == N.FORM_TUNED, cache=b0.cache,
                             tile_dirty=dirty, probe_dirty=pdirty)
        finally:
            gpu.set_tuning(schedule=0)
    assert out["plain"].info["tiles_baked"] == dirty.sum()
    for k in ("light", "direct", "indirect", "volume"):
        assert same(getattr(out["plain"], k), getattr(out["pooled"], k)), k
    sizes = sorted(int(c) for c in counts[dirty])
    print(f"dirty tiles' texel counts {sizes[0]}..{sizes[-1]}", flush=True)
