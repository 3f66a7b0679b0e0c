"""The whole device half of the bake against the oracle's restatement of it (oracle/bake_ref.py:
bake.rs:149-330 and volume.rs:69-117 over the oracle's C passes), on the same prepared inputs.

Bar: the PLAIN form, and the TUNED form's pooled schedule with FP32 node records (it shades in
the reference's order), give the final light atlas (direct + last bounce per texel) bit for bit.
The TUNED form's lock-step schedule (compensated sums in another order) within 1e-4 relative.
The small scene is baked again under its sky plus an L1 spherical-harmonic term, to the same bar
(the term in every escaping ray's sky and in the fallback irradiance).
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle
import oracle.bake_ref as ref
from lightbaker_amd import _native as N
from lightbaker_amd.bake import BakeSession
from lightbaker_amd.scenes import reference_test_scene, small_scene

pytestmark = pytest.mark.gpu


def probe_items(scene):
    v = scene.volume
    f = np.float32
    n = [max(int(round((f(v.max[a]) - f(v.min[a])) / f(v.step))) + 1, 2) for a in range(3)]
    nx, ny, nz = n
    k = np.arange(nx * ny * nz)
    it = np.zeros(len(k), N.TEXEL)
    it["pos"][:, 0] = f(v.min[0]) + (k % nx).astype(f) * f(v.step)
    it["pos"][:, 1] = f(v.min[1]) + (k // (nx * nz)).astype(f) * f(v.step)
    it["pos"][:, 2] = f(v.min[2]) + ((k // nx) % nz).astype(f) * f(v.step)
    it["pos"][:, 3] = 1.0
    it["normal"][:, 1] = 1.0
    it["normal"][:, 3] = k.astype(np.uint32).view(np.float32)
    return it


def oracle_inputs(scene, sess):
    P = sess.array(N.BAKE_PARAMS, N.PARAMS)
    PV = P.copy()
    PV["counts"][0][2] = scene.volume.rays
    tiles = sess.array(N.BAKE_TILES, np.int32).reshape(-1, 6)
    info = sess.info()
    return {
        "settings": scene.settings, "texels": sess.array(N.BAKE_TEXELS, N.TEXEL), "tiles": tiles,
        "aw": info["atlas_width"] or int(P["atlas"][0][0]), "ah": int(P["atlas"][0][1]),
        "nodes": sess.array(N.BAKE_NODES, N.NODE), "tris": sess.array(N.BAKE_TRIS, N.TRI),
        "mats": sess.array(N.BAKE_MATERIALS, np.float32).reshape(-1, 4),
        "charts": sess.array(N.BAKE_CHARTS, np.float32).reshape(-1, 8), "params": P,
        "sun": P["sun_dir"][0][:3].copy(), "irr": np.asarray(scene.sun_irradiance, np.float32),
        "probes": probe_items(scene), "volume_params": PV, "sky_sh": scene.sky_sh,
    }


def small_scene_l1_sky():
    """The small scene under its gradient sky plus an L1 term (lb_scene_desc.sky_sh)."""
    from .test_sky_sh import SH
    scene = small_scene(1)
    scene.sky_sh = SH
    return scene


@pytest.mark.parametrize("gen", [lambda: reference_test_scene(8.0), lambda: small_scene(1),
                                 small_scene_l1_sky],
                         ids=["reference-test-scene", "small-scene", "small-scene-l1-sky"])
@pytest.mark.parametrize("form,schedule", [(N.FORM_PLAIN, 0), (N.FORM_TUNED, 0),
                                           (N.FORM_TUNED, 1)], ids=["plain", "tuned", "pooled"])
def test_bake_matches_oracle(gpu, gen, form, schedule):
    scene = gen()
    print(f"bake parity form={form} schedule={schedule}", flush=True)
    # the pooled schedule is bit-exact with FP32 node records (the reference's boxes)
    gpu.set_tuning(schedule=schedule, nodes=N.NODES_WIDE if schedule == 1 else N.NODES_QUANT)
    try:
        sess = BakeSession(scene, gpu, form=form, graph=form == N.FORM_TUNED)
        sess.execute()
        sess.finish()
    finally:
        gpu.set_tuning(schedule=0)
    light = sess.array(N.BAKE_LIGHT, np.float32).reshape(-1, 4)
    want = ref.bake(oracle_inputs(scene, sess))
    wl = want["light"].reshape(-1, 4)
    gat = sess.array(N.BAKE_GATHER_INDEX, np.uint32)
    assert np.array_equal(gat, want["gathered"]), "gather sets differ"
    if form == N.FORM_PLAIN or schedule == 1:
        bad = np.nonzero((light.view(np.uint32) != wl.view(np.uint32)).any(axis=1))[0]
        assert len(bad) == 0, f"{len(bad)} atlas texels differ: {light[bad[:3]]} vs {wl[bad[:3]]}"
    else:
        err = np.abs(light - wl) / (np.abs(wl) + 1e-3)
        assert err.max() < 1e-4, err.max()
    sess.close()


def az_slots(total, rot):
    """The TUNED gather's azimuth slot of each ray r, per item of rotation rot (a power-of-two
    count; else r itself): with m0 = -rot * total rounded up, slot s holds the ray whose azimuth
    parameter is the s-th from 0."""
    r = np.arange(total)
    if total < 2 or total & (total - 1):
        return np.broadcast_to(r, (len(rot), total))
    b = total.bit_length() - 1
    rev = np.array([int(format(x, f"0{b}b")[::-1], 2) for x in r])
    m0 = (total - (np.asarray(rot, np.float32) * np.float32(total)).astype(np.int64)) & (total - 1)
    return (rev[None, :] - m0[:, None]) & (total - 1)


@pytest.mark.parametrize("nodes", [N.NODES_WIDE, N.NODES_QUANT], ids=["wide", "quant"])
@pytest.mark.parametrize("gen,rays", [(lambda: reference_test_scene(8.0), None),
                                      (lambda: small_scene(1), 100),
                                      (lambda: small_scene(1), 48)],
                         ids=["reference-test-scene", "small-scene-100-rays",
                              "small-scene-48-rays"])
def test_ranked_gather_order(gpu, gen, rays, nodes):
    """The TUNED lock-step gather's first bounce files each ray's traversal steps (inner visits +
    triangle tests, 255 at most) by azimuth slot, and the later bounces take each item's slots in
    (steps, slot) order. Over FP32 node records the steps equal the oracle's ray for ray
    (oracle.gather_steps). The 8-bit records' boxes are rounded outward: a ray may enter a box it
    passes by, and their entry distances may order two children the other way, so its steps differ
    (the hits do not); there only the order is pinned. The order is the steps' stable sort."""
    scene = gen()
    if rays is not None:
        scene.settings.gi_rays, scene.settings.gi_bounces = rays, 2
    print(f"ranked gather order: {scene.settings.gi_rays} rays, nodes {nodes}", flush=True)
    gpu.set_tuning(nodes=nodes)
    sess = BakeSession(scene, gpu, form=N.FORM_TUNED, graph=True)
    sess.execute()
    sess.finish()
    inp = oracle_inputs(scene, sess)
    want = ref.bake(inp)
    items = want["texels"][want["gathered"]]
    n, total = len(items), int(inp["params"]["counts"][0][2])
    assert n > 0 and total == scene.settings.gi_rays
    cost = sess.array(N.BAKE_GATHER_COST, np.uint8).reshape(n, total)
    order = sess.array(N.BAKE_GATHER_ORDER, np.uint8).reshape(n, total)
    steps = oracle.gather_steps(inp["nodes"], inp["tris"], inp["mats"], inp["params"], items)
    seed = float(inp["params"]["reach"][0][3])
    rot = [oracle.lib().orc_rand((i * 7919) & 0xFFFFFFFF, 13, seed) for i in range(n)]
    filed = np.take_along_axis(cost, az_slots(total, rot), axis=1)
    print(f"steps: {steps.mean():.2f} a ray, {filed.mean():.2f} as filed, "
          f"{(filed == steps).mean():.4f} of rays equal", flush=True)
    if nodes == N.NODES_WIDE:
        bad = np.nonzero((filed != steps).any(axis=1))[0]
        assert len(bad) == 0, f"{len(bad)} items: {filed[bad[0]][:16]} vs {steps[bad[0]][:16]}"
    else:
        assert 0.9 < filed.mean() / steps.mean() < 1.5
    assert steps.min() > 0 and len(np.unique(steps)) > 4
    assert np.array_equal(order, np.argsort(cost, axis=1, kind="stable"))
    sess.close()


def test_gather_order_empty_unless_ranked(gpu):
    """A PLAIN bake, and a TUNED one with a single bounce, rank nothing: both arrays are empty."""
    for form, bounces in ((N.FORM_PLAIN, 2), (N.FORM_TUNED, 1)):
        scene = small_scene(1)
        scene.settings.gi_bounces = bounces
        sess = BakeSession(scene, gpu, form=form, graph=form == N.FORM_TUNED)
        sess.execute()
        assert len(sess.array(N.BAKE_GATHER_COST, np.uint8)) == 0, form
        assert len(sess.array(N.BAKE_GATHER_ORDER, np.uint8)) == 0, form
        sess.close()


def test_unranked_gather_beyond_256_rays(gpu):
    """More than 256 gather rays an item (a slot no longer fits a byte): the TUNED lock-step
    bake traces every bounce unranked, within 1e-4 of the PLAIN bake, and ranks nothing."""
    def scene():
        s = small_scene(1)
        s.settings.gi_rays, s.settings.gi_bounces = 320, 2
        return s
    print("unranked gather: 320 rays", flush=True)
    want = BakeSession(scene(), gpu, form=N.FORM_PLAIN)
    want.execute()
    wl = want.array(N.BAKE_LIGHT, np.float32).reshape(-1, 4)
    want.close()
    sess = BakeSession(scene(), gpu, form=N.FORM_TUNED, graph=True)
    sess.execute()
    light = sess.array(N.BAKE_LIGHT, np.float32).reshape(-1, 4)
    assert len(sess.array(N.BAKE_GATHER_ORDER, np.uint8)) == 0
    sess.close()
    err = np.abs(light - wl) / (np.abs(wl) + 1e-3)
    assert err.max() < 1e-4, err.max()


def test_occlusion_contrast_against_oracle(gpu):
    """aoContrast 2.2 (bake.rs:238's powf): the device's powf and the host's (glibc's, as Rust's
    f32::powf) may round one value a unit in the last place apart, so the occlusion is held
    within 1e-6 relative; the light, which the occlusion does not enter, bit for bit."""
    scene = reference_test_scene(8.0)
    scene.settings.ao_contrast = 2.2
    print("bake parity: occlusion contrast 2.2", flush=True)
    sess = BakeSession(scene, gpu, form=N.FORM_PLAIN, graph=False)
    sess.execute()
    sess.finish()
    want = ref.bake(oracle_inputs(scene, sess))
    light = sess.array(N.BAKE_LIGHT, np.float32).reshape(-1, 4)
    assert np.array_equal(light.view(np.uint32), want["light"].reshape(-1, 4).view(np.uint32))
    ao = sess.tile_cache().ao
    own = want["known"] != 0
    got, w = ao[: len(own)][own], want["ao"][own]
    assert np.abs(got - w).max() <= 1e-6 * np.abs(w).max(), np.abs(got - w).max()
    assert (w < 0.999).sum() > 100  # the contrast is exercised
    sess.close()
