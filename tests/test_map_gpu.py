"""The 100 x 100 m, 3-bounce map the benchmark bakes (lightbaker_amd.scenes.map_scene(2): fBm
terrain, 400 boxes and 32-sided cylinders, 203 940 triangles, 16 materials, 256 gather rays)
against the oracle, on the bake's own scene records and its final light atlas.

Bar: on sampled direct and gather items and volume probes, the PLAIN form equals the oracle
(oracle/trace_ref.c) bit for bit; the TUNED form with 8-bit quantized nodes and the settings the
benchmark's autotuner picks on this map gives every hit-derived quantity exactly and every sum
within 2e-6 relative, against the oracle traversing the same 8-bit node records (run_quant).
The TUNED form's Kahan-compensated gather sums (over FP32 node records, so its hits are the
reference's) are at least as close to the exact sum of the same per-ray terms
(oracle.gather_terms, summed in float64) as the PLAIN form's naive FP32 sums: summation error
alone, the hits being equal.
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from lightbaker_amd import _native as N
from lightbaker_amd.bake import BakeSession
from lightbaker_amd.gpu import Bvh, Pass
from lightbaker_amd.scenes import map_scene

from . import kit
from .test_bake_parity_gpu import probe_items
from .test_parity_gpu import _compare_tuned

pytestmark = pytest.mark.gpu

# the per-pass settings the benchmark's autotuner picks on this map (lock-step, 8-bit nodes)
BENCH_TUNING = {
    Pass.DIRECT: dict(workgroup=64, rays_per_lane=8, lds_stack=16, nodes=N.NODES_QUANT,
                      schedule=0, leaf_batch=16, waves=8, xcd_split=0, gather_order=0),
    Pass.GATHER: dict(workgroup=256, rays_per_lane=8, lds_stack=16, nodes=N.NODES_QUANT,
                      schedule=0, leaf_batch=24, waves=8, xcd_split=1, gather_order=0),
    Pass.VOLUME: dict(workgroup=256, rays_per_lane=4, lds_stack=16, nodes=N.NODES_QUANT,
                      schedule=0, leaf_batch=16, waves=6, xcd_split=0, gather_order=0),
}


@pytest.fixture(scope="module")
def baked(gpu):
    scene = map_scene(2)
    print(f"map: {scene.triangle_count()} triangles, PLAIN bake for its atlas", flush=True)
    sess = BakeSession(scene, gpu, form=N.FORM_PLAIN, graph=False)
    sess.execute()
    P = sess.array(N.BAKE_PARAMS, N.PARAMS)
    PV = P.copy()
    PV["counts"][0][2] = scene.volume.rays
    aw, ah = int(P["atlas"][0][0]), int(P["atlas"][0][1])
    s = dict(
        nodes=sess.array(N.BAKE_NODES, N.NODE), tris=sess.array(N.BAKE_TRIS, N.TRI),
        mats=sess.array(N.BAKE_MATERIALS, np.float32).reshape(-1, 4),
        charts=sess.array(N.BAKE_CHARTS, np.float32).reshape(-1, 8), params=P, vparams=PV,
        light=sess.array(N.BAKE_LIGHT, np.float32).reshape(aw * ah, 4).copy(),
        texels=sess.array(N.BAKE_TEXELS, N.TEXEL).copy(),
        gathered=sess.array(N.BAKE_GATHER_INDEX, np.uint32).copy())
    sess.close()
    rng = np.random.default_rng(2024)
    cov = np.nonzero(s["texels"]["pos"][:, 3] > 0.5)[0]
    s["direct"] = s["texels"][np.sort(rng.choice(cov, 3000, replace=False))]
    s["gather"] = s["texels"][np.sort(rng.choice(s["gathered"], 2000, replace=False))]
    probes = probe_items(scene)
    s["probes"] = probes[np.sort(rng.choice(len(probes), 800, replace=False))]
    s["sb"] = gpu.scene(Bvh(s["nodes"], s["tris"]), s["mats"], s["charts"])
    s["quant"] = oracle.quant_records(s["nodes"])
    print(f"map: {len(cov)} covered texels, {len(s['gathered'])} gather items, "
          f"{len(probes)} probes", flush=True)
    yield s
    s["sb"].close()


def _items(s, pass_):
    return {Pass.DIRECT: s["direct"], Pass.GATHER: s["gather"], Pass.VOLUME: s["probes"]}[pass_]


def _params(s, pass_):
    return s["vparams"] if pass_ == Pass.VOLUME else s["params"]


def _oracle(s, pass_, quant=False):
    if quant:
        return oracle.run_quant(int(pass_), s["nodes"], s["quant"], s["tris"], s["mats"],
                                s["charts"], _params(s, pass_), _items(s, pass_), 0, s["light"])
    return oracle.run(int(pass_), s["nodes"], s["tris"], s["mats"], s["charts"], _params(s, pass_),
                      _items(s, pass_), 0, s["light"])


@pytest.mark.parametrize("pass_", [Pass.DIRECT, Pass.GATHER, Pass.VOLUME])
def test_map_plain_bit_exact(gpu, baked, pass_):
    items = _items(baked, pass_)
    print(f"map PLAIN {pass_.name}: {len(items)} items", flush=True)
    gpu.set_form(N.FORM_PLAIN)
    got = gpu.run(pass_, baked["sb"], _params(baked, pass_), items, baked["light"])
    want = _oracle(baked, pass_)
    bad = np.nonzero((kit.bits(got) != kit.bits(want)).any(axis=1))[0]
    assert len(bad) == 0, f"{len(bad)} rows differ, first {bad[:5]}"


@pytest.mark.parametrize("pass_", [Pass.DIRECT, Pass.GATHER, Pass.VOLUME])
def test_map_tuned_quantized_bench_settings(gpu, baked, pass_):
    items = _items(baked, pass_)
    print(f"map TUNED {pass_.name}: {len(items)} items, {BENCH_TUNING[pass_]}", flush=True)
    gpu.set_tuning(pass_=pass_, **BENCH_TUNING[pass_])
    gpu.set_form(N.FORM_TUNED)
    got = gpu.run(pass_, baked["sb"], _params(baked, pass_), items, baked["light"])
    _compare_tuned(pass_, got, _oracle(baked, pass_, quant=True), True)


def test_map_kahan_sums_no_worse_than_naive(gpu, baked):
    items = baked["gather"]
    print(f"map gather sums: {len(items)} items x 256 rays", flush=True)
    gpu.set_form(N.FORM_PLAIN)
    plain = gpu.run(Pass.GATHER, baked["sb"], baked["params"], items, baked["light"])[:, :3]
    gpu.set_tuning(pass_=Pass.GATHER, **dict(BENCH_TUNING[Pass.GATHER], nodes=N.NODES_WIDE))
    gpu.set_form(N.FORM_TUNED)
    tuned = gpu.run(Pass.GATHER, baked["sb"], baked["params"], items, baked["light"])[:, :3]
    terms = oracle.gather_terms(baked["nodes"], baked["tris"], baked["mats"], baked["charts"],
                                baked["params"], items, 0, baked["light"])
    rays = terms.shape[1]
    exact = terms.astype(np.float64).sum(axis=1) * np.pi / rays
    ep, et = plain - exact, tuned - exact
    rp, rt = float(np.sqrt((ep ** 2).mean())), float(np.sqrt((et ** 2).mean()))
    print(f"summation error: PLAIN rmse {rp:.3e} max {np.abs(ep).max():.3e}; "
          f"TUNED rmse {rt:.3e} max {np.abs(et).max():.3e}", flush=True)
    assert rt <= rp and np.abs(et).max() <= np.abs(ep).max()
