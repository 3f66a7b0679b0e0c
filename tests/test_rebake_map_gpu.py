"""The 64-tile rebake the batching configuration describes: lightbaker_amd.scenes.rebake_scene(3),
a 200 x 200 m map from the small scene's generator, 64 tiles of 25 m each laid out at its own
texel density (4-16 texels/m, so the tiles' texel counts differ more than tenfold), a seeded
dirty subset baked as one item list from the clean tiles' cached light (bake.rs:123-149,
volume.rs:44-84).

Bar: against the oracle's restatement of the same rebake (oracle/bake_ref.py over the dirty
tiles, its first atlas written from the cached TileLight), the PLAIN form gives the atlas light,
every dirty tile's TileLight and every traced probe bit for bit; the TUNED form (8-bit
quantized nodes, the persistent queue, the bounces in a graph) within 1e-4 relative of the same
rebake with the oracle traversing those 8-bit node records (oracle.run_quant), since on this map
the two node layouts' hits differ on a few rays: boxes stand on the ground, so a box's bottom and
the ground under it lie within an ulp of each other, and the visiting order the boxes set decides
which of the two a ray keeps (tests/test_quant_nodes.py measures it).
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle
import oracle.bake_ref as ref
from lightbaker_amd import _native as N
from lightbaker_amd.bake import BakeSession, bake
from lightbaker_amd.scenes import rebake_scene

from .test_bake_parity_gpu import probe_items
from .test_rebake_gpu import dirty_masks, same

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def first(gpu):
    scene = rebake_scene(3)
    print(f"64-tile rebake: {scene.triangle_count()} triangles; first bake (tuned)", flush=True)
    b0 = bake(scene, gpu, form=N.FORM_TUNED, graph=True)
    counts = b0.tiles[:, 2] * b0.tiles[:, 3]
    print(f"64-tile rebake: {len(b0.tiles)} tiles, {counts.min()}..{counts.max()} texels a tile",
          flush=True)
    assert len(b0.tiles) == 64 and counts.max() > 10 * counts.min()
    dirty, pdirty = dirty_masks(scene, b0.tiles, 33, 6 / 64)
    print(f"64-tile rebake: dirty tiles {list(np.nonzero(dirty)[0])}, {pdirty.sum()} probes; "
          "PLAIN rebake", flush=True)
    sess = BakeSession(scene, gpu, form=N.FORM_PLAIN, graph=False, cache=b0.cache,
                       tile_dirty=dirty, probe_dirty=pdirty)
    sess.execute()
    sess.finish()
    P = sess.array(N.BAKE_PARAMS, N.PARAMS)
    PV = P.copy()
    PV["counts"][0][2] = scene.volume.rays
    tiles_all = sess.array(N.BAKE_TILES, np.int32).reshape(-1, 6)
    aw, ah = int(P["atlas"][0][0]), int(P["atlas"][0][1])
    light0 = ref.clean_light(tiles_all, dirty, b0.cache, aw, ah)
    print("64-tile rebake: the oracle over the dirty tiles", flush=True)
    inp = {
        "settings": scene.settings, "texels": sess.array(N.BAKE_TEXELS, N.TEXEL),
        "tiles": tiles_all[dirty], "light0": light0, "aw": aw, "ah": ah,
        "nodes": sess.array(N.BAKE_NODES, N.NODE), "tris": sess.array(N.BAKE_TRIS, N.TRI),
        "mats": sess.array(N.BAKE_MATERIALS, np.float32).reshape(-1, 4),
        "charts": sess.array(N.BAKE_CHARTS, np.float32).reshape(-1, 8), "params": P,
        "sun": P["sun_dir"][0][:3].copy(), "irr": np.asarray(scene.sun_irradiance, np.float32),
        "probes": probe_items(scene)[pdirty], "volume_params": PV,
    }
    want = ref.bake(inp)
    yield dict(scene=scene, b0=b0, dirty=dirty, pdirty=pdirty, sess=sess, want=want, inp=inp,
               tiles=tiles_all, light0=light0, aw=aw, ah=ah)
    sess.close()


def _dirty_index(tiles, dirty):
    offs = np.concatenate([[0], np.cumsum(tiles[:, 2] * tiles[:, 3])])
    return np.concatenate([np.arange(offs[t], offs[t + 1]) for t in np.nonzero(dirty)[0]])


def test_rebake_64_tiles_plain_bit_exact(first):
    f = first
    sess, want = f["sess"], f["want"]
    assert sess.info()["tiles_baked"] == f["dirty"].sum()
    assert same(sess.array(N.BAKE_LIGHT0, np.float32).reshape(f["ah"], f["aw"], 4), f["light0"])
    assert np.array_equal(sess.array(N.BAKE_GATHER_INDEX, np.uint32), want["gathered"])
    light = sess.array(N.BAKE_LIGHT, np.float32).reshape(-1, 4)
    wl = want["light"].reshape(-1, 4)
    bad = np.nonzero((light.view(np.uint32) != wl.view(np.uint32)).any(axis=1))[0]
    assert len(bad) == 0, f"{len(bad)} atlas texels differ"
    tc = sess.tile_cache()
    idx = _dirty_index(f["tiles"], f["dirty"])
    own = want["known"] != 0
    assert same(tc.known[idx][own], want["known"][own])
    for k in ("field", "ao", "direct", "indirect"):
        assert same(getattr(tc, k)[idx][own], want[k][own]), k
    probes = sess.array(N.BAKE_PROBES, np.float32).reshape(-1, 16)
    assert same(probes[f["pdirty"]], want["probes"]), "traced probes differ"
    assert same(probes[~f["pdirty"]], f["b0"].cache.probes[~f["pdirty"]]), "clean probes changed"
    print(f"64-tile rebake PLAIN: {len(wl)} atlas texels, {own.sum()} dirty texels lit, "
          f"{f['pdirty'].sum()} probes, all bit-exact", flush=True)


def test_rebake_64_tiles_tuned(gpu, first):
    f = first
    print("64-tile rebake: TUNED", flush=True)
    sess = BakeSession(f["scene"], gpu, form=N.FORM_TUNED, graph=True, cache=f["b0"].cache,
                       tile_dirty=f["dirty"], probe_dirty=f["pdirty"])
    sess.execute()
    sess.finish()
    print("64-tile rebake: the oracle over the 8-bit node records", flush=True)
    want = ref.bake(dict(f["inp"], quant=oracle.quant_records(f["inp"]["nodes"])))
    light = sess.array(N.BAKE_LIGHT, np.float32).reshape(-1, 4)
    wl = want["light"].reshape(-1, 4)
    err = np.abs(light - wl) / (np.abs(wl) + 1e-3)
    assert err.max() < 1e-4, err.max()
    tc = sess.tile_cache()
    idx = _dirty_index(f["tiles"], f["dirty"])
    own = want["known"] != 0
    for k in ("direct", "indirect"):
        a, w = getattr(tc, k)[idx][own], want[k][own]
        assert (np.abs(a - w) / (np.abs(w) + 1e-3)).max() < 1e-4, k
    probes = sess.array(N.BAKE_PROBES, np.float32).reshape(-1, 16)
    pe = np.abs(probes[f["pdirty"]] - want["probes"]) / (np.abs(want["probes"]) + 1e-3)
    assert pe.max() < 1e-4
    sess.close()
