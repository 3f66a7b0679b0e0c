"""Quantized against FP32 node records on a city block (lightbaker_amd.scenes.city_scene: prisms
of 14 to 400 sides, tall and thin walls, many boxes sharing planes): the TUNED form with 8-bit
outward-rounded child bounds and with FP32 child bounds, each against the oracle traversing the
same node records (oracle.run for FP32 boxes, oracle.run_quant for the 8-bit ones): hit-derived
quantities exactly, sums within 2e-6. How the two layouts' hits differ (near-ties, grazing rays)
is measured on the CPU in tests/test_quant_nodes.py.
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from lightbaker_amd import _native as N
from lightbaker_amd.gpu import Pass
from lightbaker_amd.scenes import city_scene

from . import kit

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def city(gpu):
    scene = city_scene(4, buildings=300, size=120.0)
    bvh, pos, mesh = kit.scene_bvh(scene)
    rng = np.random.default_rng(11)
    items = kit.sample_items(pos, mesh, 3000, rng)
    charts = kit.charts_for(scene, lo=0.0, scale=0.75)
    light = kit.random_light(rng)
    params = kit.params_for(scene)
    sb = gpu.scene(bvh, scene.materials_array(), charts)
    print(f"city: {scene.triangle_count()} triangles, {len(bvh.nodes)} nodes", flush=True)
    yield dict(bvh=bvh, items=items, charts=charts, light=light, params=params, sb=sb,
               mats=scene.materials_array())
    sb.close()


@pytest.mark.parametrize("pass_", [Pass.DIRECT, Pass.GATHER])
def test_quantized_nodes_keep_the_hits(gpu, city, pass_):
    print(f"city {pass_.name}", flush=True)
    out = {}
    gpu.set_form(N.FORM_TUNED)
    keep = gpu.tuning(pass_)
    try:
        for name, layout in (("quant", N.NODES_QUANT), ("wide", N.NODES_WIDE)):
            gpu.set_tuning(pass_=pass_, nodes=layout)
            out[name] = gpu.run(pass_, city["sb"], city["params"], city["items"], city["light"])
    finally:
        gpu.set_tuning(pass_=pass_, **keep)
        gpu.set_form(N.FORM_PLAIN)
    want = oracle.run(int(pass_), city["bvh"].nodes, city["bvh"].tris, city["mats"],
                      city["charts"], city["params"], city["items"], 0, city["light"])
    wq = oracle.run_quant(int(pass_), city["bvh"].nodes, oracle.quant_records(city["bvh"].nodes),
                          city["bvh"].tris, city["mats"], city["charts"], city["params"],
                          city["items"], 0, city["light"])
    q, w = out["quant"], out["wide"]
    if pass_ == Pass.DIRECT:
        # hit-derived: sun bits and pick, soft and back shares exactly; occlusion sums within 2e-6
        for got, ref in ((q, wq), (w, want)):
            assert np.array_equal(kit.bits(got[:, 1]), kit.bits(ref[:, 1]))
            assert np.array_equal(got[:, [0, 3]], ref[:, [0, 3]])
            assert np.allclose(got[:, 2], ref[:, 2], rtol=2e-6, atol=1e-6)
    else:
        # each layout against the oracle over its own node records
        for got, ref in ((q, wq), (w, want)):
            assert np.array_equal(got[:, 3], ref[:, 3])
            assert np.allclose(got[:, :3], ref[:, :3], rtol=2e-6, atol=1e-6)
