"""The bake's tail on the device (csrc/atlas.hip) and the output codecs, against the oracle's
restatement (oracle/finish_ref.c), through the C-ABI:

  * every tile's TileLight (bake.rs:337-362): the oracle's device half of the bake
    (oracle/bake_ref.py), then its probes' fill (volume.rs:96-121) and the volume's light for
    texels with none (bake.rs:337-356), equal the bake's cache arrays and volume bit for bit
    (PLAIN form: the same FP32 operations in the same order);
  * the atlas (bake.rs:364-421) from the bake's own TileLight equals the bake's direct,
    validity and indirect atlases bit for bit, in every form and in a rebake;
  * the shipped encodings (output.rs:99-110: BC5 of the quantized direct atlas, RGB9E5 of the
    indirect atlas, the volume's bytes and scale) equal the oracle's encodings of the bake's own
    float outputs bit for bit;
  * lb_encode_bc5 / lb_encode_rgb9e5 / lb_encode_volume equal the oracle on random inputs and on
    the edge cases the oracle's own tests pin (ties, uniform blocks, powers of two and their
    neighbours, clamps, NaN), and reject what the reference asserts against.
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle.bake_ref as ref
from oracle import finish_ref as F
from lightbaker_amd import _native as N
from lightbaker_amd import encode
from lightbaker_amd.bake import BakeSession, bake
from lightbaker_amd.scenes import Mesh, cube, rebake_scene, reference_test_scene

from .test_bake_parity_gpu import oracle_inputs
from .test_oracle_finish import rgb_cases
from .test_rebake_gpu import dirty_masks

pytestmark = pytest.mark.gpu


def same(a, b):
    a, b = np.ascontiguousarray(a), np.ascontiguousarray(b)
    return a.shape == b.shape and np.array_equal(a.view(np.uint8), b.view(np.uint8))


def volume_geometry(info):
    return (np.array(info["volume_size"], np.uint32), np.array(info["volume_origin"], np.float32),
            np.float32(info["volume_step"]))


def check_atlas_and_codecs(sess_or_bake, info, tiles, tc, direct, validity, indirect, volume,
                           bc5, rgb, vbytes, ds):
    aw, ah = info["atlas_width"], info["atlas_height"]
    d, v, e = F.atlas(tiles, tc.known, tc.state, tc.field, tc.ao, tc.indirect, aw, ah, ds)
    assert same(direct.reshape(ah, aw, 2), d), "direct atlas"
    assert same(validity.reshape(ah, aw), v), "validity"
    assert same(indirect.reshape(e.shape), e), "indirect atlas"
    assert same(bc5, F.bc5(F.unorm8(direct.reshape(ah, aw, 2)), aw, ah)), "BC5"
    assert same(rgb.reshape(-1), F.rgb9e5(indirect.reshape(-1, 3))), "RGB9E5"
    if volume is not None:
        want, scale = F.volume_bytes(volume.reshape(-1, 16))
        assert same(vbytes.reshape(-1, 16), want), "volume bytes"
        assert np.float32(info["volume_scale"]) == scale


def wedged_scene():
    """The reference test scene with a cube wedged inside its box, 0.1 m from every wall: the
    inner cube's texels lie inside a solid (the walls' backs within the 0.3 m inside reach), so
    where a chart keeps no light of its own the volume lights it (bake.rs:337-356)."""
    scene = reference_test_scene(8.0)
    scene.meshes.append(Mesh("wedged", 1, cube((8.1, 0.1, 10.1), (9.9, 1.9, 11.9))))
    return scene


@pytest.fixture(scope="module")
def plain(gpu):
    scene = wedged_scene()
    print("finish: PLAIN bake of the reference test scene with a wedged cube", flush=True)
    sess = BakeSession(scene, gpu, form=N.FORM_PLAIN, graph=False)
    sess.execute()
    sess.finish()
    yield scene, sess
    sess.close()


def test_tile_light_matches_oracle(plain):
    scene, sess = plain
    info = sess.info()
    inp = oracle_inputs(scene, sess)
    want = ref.bake(inp)
    size, origin, step = volume_geometry(info)
    probes = F.probe_fill(want["probes"], size)
    assert same(sess.array(N.BAKE_VOLUME, np.float32).reshape(-1, 16), probes), "filled probes"
    tl = {k: np.ascontiguousarray(want[k]) for k in ("known", "state", "field", "ao")}
    tl["direct"] = np.ascontiguousarray(want["direct"], np.float32).reshape(-1, 3)
    tl["indirect"] = np.ascontiguousarray(want["indirect"], np.float32).reshape(-1, 3)
    P = inp["params"]
    lit = F.volume_lit(probes, size, origin, step, P["sun_dir"][0][:3], inp["irr"],
                       inp["texels"], tl)
    # some of the wedged cube's texels (33 on this scene): the fallback is exercised
    assert lit > 0, "no texel lit from the volume: the fallback is not exercised"
    tc = sess.tile_cache()
    for k in ("known", "state", "field", "ao", "direct", "indirect"):
        assert same(getattr(tc, k).reshape(tl[k].shape), tl[k]), k
    assert (tc.state == 5).sum() == lit


def test_atlas_and_encodings_plain(plain):
    scene, sess = plain
    info = sess.info()
    tiles = sess.array(N.BAKE_TILES, np.int32).reshape(-1, 6)
    check_atlas_and_codecs(
        sess, info, tiles, sess.tile_cache(), sess.array(N.BAKE_DIRECT, np.float32),
        sess.array(N.BAKE_VALIDITY, np.uint8), sess.array(N.BAKE_INDIRECT, np.float32),
        sess.array(N.BAKE_VOLUME, np.float32), sess.array(N.BAKE_DIRECT_BC5, np.uint8),
        sess.array(N.BAKE_INDIRECT_RGB9E5, np.uint32), sess.array(N.BAKE_VOLUME_BYTES, np.uint8),
        scene.settings.indirect_downscale)


@pytest.mark.parametrize("ds", [1, 4])
def test_atlas_and_encodings_tuned(gpu, ds):
    scene = reference_test_scene(8.0)
    scene.settings.indirect_downscale = ds
    print(f"finish: TUNED bake, indirect downscale {ds}", flush=True)
    b = bake(scene, gpu, form=N.FORM_TUNED, graph=True)
    check_atlas_and_codecs(b, b.info, b.tiles, b.cache, b.direct, b.validity, b.indirect,
                           b.volume, b.direct_bc5, b.indirect_rgb9e5, b.volume_bytes, ds)


def test_atlas_and_encodings_rebake(gpu):
    scene = rebake_scene(3, size=100.0)
    print("finish: first bake of the rebake scene", flush=True)
    b0 = bake(scene, gpu, form=N.FORM_TUNED, graph=True)
    dirty, pdirty = dirty_masks(scene, b0.tiles, 33, 0.3)
    print(f"finish: rebake of {dirty.sum()} tiles", flush=True)
    b = bake(scene, gpu, form=N.FORM_TUNED, graph=True, cache=b0.cache, tile_dirty=dirty,
             probe_dirty=pdirty)
    offs = np.concatenate([[0], np.cumsum(b.tiles[:, 2] * b.tiles[:, 3])])
    clean = np.concatenate([np.arange(offs[t], offs[t + 1]) for t in np.nonzero(~dirty)[0]])
    for k in ("known", "state", "field", "ao", "direct", "indirect"):
        assert same(getattr(b.cache, k)[clean], getattr(b0.cache, k)[clean]), k
    check_atlas_and_codecs(b, b.info, b.tiles, b.cache, b.direct, b.validity, b.indirect,
                           b.volume, b.direct_bc5, b.indirect_rgb9e5, b.volume_bytes,
                           scene.settings.indirect_downscale)


# ---- the codecs' entry points ------------------------------------------------------------------

def test_bc5_entry_point(gpu):
    rng = np.random.default_rng(11)
    w, h = 64, 36
    rg = rng.integers(0, 256, (h, w, 2)).astype(np.uint8)
    rg[0:4, 0:4] = 77                                   # a uniform block
    rg[0:4, 4:8, 0] = np.array([14, 0] + [1] * 14).reshape(4, 4)  # the tie case
    rg[4:8, 0:4, 1] = np.array([255, 0] * 8).reshape(4, 4)
    assert same(encode.bc5(gpu, rg, w, h), F.bc5(rg, w, h))
    with pytest.raises(N.NativeError, match="multiple of 4"):
        encode.bc5(gpu, np.zeros((4, 6, 2), np.uint8), 6, 4)


def test_rgb9e5_entry_point(gpu):
    c = rgb_cases(np.random.default_rng(12))
    rng = np.random.default_rng(13)
    more = (rng.random((200000, 3)) * 10.0 ** rng.uniform(-7, 5, (200000, 1))).astype(np.float32)
    c = np.concatenate([c, more])
    got = encode.rgb9e5(gpu, c)
    want = F.rgb9e5(c)
    bad = np.nonzero(got != want)[0]
    assert len(bad) == 0, [(c[i], hex(got[i]), hex(want[i])) for i in bad[:5]]
    assert np.array_equal(encode.rgb9e5_decode(got), F.rgb9e5_decode(want))


def test_volume_entry_point(gpu):
    rng = np.random.default_rng(14)
    p = rng.normal(0, 1, (5000, 16)).astype(np.float32)
    p[:, 0:3] = np.abs(p[:, 0:3]) * 3
    p[:, 3] = rng.random(5000)
    p[:10, 0:3] = 0.0
    p[10:20, 0:3] = -1.0
    p[20, 0:3] = 1e-7
    got, scale = encode.volume_bytes(gpu, p)
    want, wscale = F.volume_bytes(p)
    assert np.float32(scale) == wscale
    assert same(got, want)
    z, zs = encode.volume_bytes(gpu, np.zeros((3, 16), np.float32))
    assert np.float32(zs) == np.float32(1e-6) and (z[:, :3] == 0).all() and (z[:, 4:13] == 128).all()


def test_lbake_file_of_a_bake(gpu):
    """output.rs:95-152 over a real bake: every chunk decodes to the device's encodings."""
    from lightbaker_amd import output
    scene = reference_test_scene(8.0)
    b = bake(scene, gpu, form=N.FORM_TUNED, graph=True)
    r = output.read(output.write(scene, b, level=3))
    h = r["header"]
    assert r["direct"]["data"] == b.direct_bc5.tobytes()
    assert (r["direct"]["width"], r["direct"]["height"]) == (b.info["atlas_width"],
                                                             b.info["atlas_height"])
    assert r["indirect"]["data"] == b.indirect_rgb9e5.astype("<u4").tobytes()
    assert r["validity"] == b.validity.tobytes() and r["volume"] == b.volume_bytes.tobytes()
    assert r["charts"] == b.charts.tobytes() and r["triangleCharts"] == b.tri_charts.tobytes()
    # the header prints the f32's shortest decimal, which reads back as the same f32
    assert np.float32(h["volume"]["scale"]) == np.float32(b.info["volume_scale"])
    assert h["stats"]["texelsBaked"] == b.info["texels_baked"]
