"""The whole device half of the bake against the oracle's restatement of it (oracle/bake_ref.py:
bake.rs:149-330 and volume.rs:69-117 over the oracle's C passes), on the same prepared inputs.

Bar: the PLAIN form, and the TUNED form's pooled schedule with FP32 node records (it shades in
the reference's order), give the final light atlas (direct + last bounce per texel) bit for bit.
The TUNED form's lock-step schedule (compensated sums in another order) within 1e-4 relative.
"""
from __future__ import annotations

import numpy as np
import pytest

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
        "probes": probe_items(scene), "volume_params": PV,
    }


@pytest.mark.parametrize("gen", [lambda: reference_test_scene(8.0), lambda: small_scene(1)],
                         ids=["reference-test-scene", "small-scene"])
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
