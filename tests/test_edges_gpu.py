"""Edge cases of the ray passes against the oracle (oracle/trace_ref.c), through the C-ABI:

  * the traversal stack's overflow rule (trace.wgsl:120: a far child that finds the 48-entry
    stack full is dropped): a hand-built chain BVH whose only hit triangles sit in leaves the
    overflow drops, so a sun ray that would be blocked with a deeper stack sees the sun;
  * translucent materials (trace.wgsl:69-72: a hit stops the ray with probability = opacity,
    decided by the triangle and the ray's key);
  * a scene with no casting triangle (bvh.rs:76-80: one node of count 0 that no ray reaches);
  * a pass over no items (gpu/mod.rs:143-145: an empty result).

Bar, in every form (PLAIN, TUNED lock-step with quantized and FP32 nodes, TUNED pooled): the
hit-derived outputs (shadow bits, pick, sun and back shares) equal the oracle's; PLAIN and pooled
give every output bit for bit.
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from lightbaker_amd import _native as N
from lightbaker_amd.gpu import Bvh, Pass
from lightbaker_amd.scenes import Material, small_scene

from . import kit

pytestmark = pytest.mark.gpu

FORMS = [("plain", N.FORM_PLAIN, {}),
         ("tuned-quant", N.FORM_TUNED, dict(schedule=0, nodes=N.NODES_QUANT)),
         ("tuned-wide", N.FORM_TUNED, dict(schedule=0, nodes=N.NODES_WIDE, lds_stack=16)),
         ("pooled", N.FORM_TUNED, dict(schedule=1, nodes=N.NODES_WIDE))]


def run(gpu, sb, form, tuning, pass_, params, items, light):
    gpu.set_form(form)
    if tuning:
        gpu.set_tuning(**tuning)
    return gpu.run(pass_, sb, params, items, light)


def check(name, pass_, got, want):
    bits = kit.bits
    if name in ("plain", "pooled"):
        bad = np.nonzero((bits(got) != bits(want)).any(axis=1))[0]
        assert len(bad) == 0, f"{name}: rows {bad[:5]} differ: {got[bad[:2]]} vs {want[bad[:2]]}"
    elif pass_ == Pass.DIRECT:
        assert np.array_equal(bits(got[:, 1]), bits(want[:, 1])), name
        assert np.array_equal(got[:, [0, 3]], want[:, [0, 3]]), name
        assert np.allclose(got[:, 2], want[:, 2], rtol=2e-6, atol=1e-6), name
    else:
        assert np.array_equal(got[:, 3], want[:, 3]), name
        assert np.allclose(got[:, :3], want[:, :3], rtol=2e-6, atol=1e-6), name


def tri(a, b, c, material=0):
    t = np.zeros(1, N.TRI)
    t["v0"][0] = [*a, np.uint32(material).view(np.float32)]
    t["e1"][0] = [*(np.subtract(b, a)), np.uint32(0xFFFFFFFF).view(np.float32)]
    t["e2"][0] = [*(np.subtract(c, a)), 0.0]
    return t


def chain_scene():
    """Triangles k = 0..59 in the planes z = 100 - k, a chain BVH: inner node I_k holds triangles
    k..59 and has children (I_k+1, leaf k), I_k+1 entered first (its triangles lie lower). A ray
    up the z axis from the origin pushes leaf k at every I_k: leaves 0..47 fill the stack and
    leaves 48..58 are dropped. Triangles 48..58 cover the axis; every other one is a sliver whose
    box holds the axis but whose area does not, so with a deeper stack the ray would be blocked
    and with the reference's it meets nothing."""
    tris = []
    for k in range(60):
        z = 100.0 - k
        if 48 <= k <= 58:
            tris.append(tri((-1, -1, z), (1, -1, z), (-1, 1.5, z)))   # covers (0, 0)
        else:
            tris.append(tri((-1, -1, z), (1, -1, z), (-1, 0.9, z)))   # misses (0, 0)
    tris = np.concatenate(tris)
    lo = np.stack([tris["v0"][:, :3], tris["v0"][:, :3] + tris["e1"][:, :3],
                   tris["v0"][:, :3] + tris["e2"][:, :3]]).min(axis=0)
    hi = np.stack([tris["v0"][:, :3], tris["v0"][:, :3] + tris["e1"][:, :3],
                   tris["v0"][:, :3] + tris["e2"][:, :3]]).max(axis=0)
    nodes = np.zeros(119, N.NODE)

    def leaf(i, k):
        nodes[i]["min"], nodes[i]["max"] = lo[k], hi[k]
        nodes[i]["first"], nodes[i]["count"] = k, 1

    # I_0 at 0; I_k (k >= 1) at 2k - 1, its children at 2k + 1 (I_k+1) and 2k + 2 (leaf k);
    # I_58's children are leaf 59 (117) and leaf 58 (118)
    for k in range(59):
        i = 0 if k == 0 else 2 * k - 1
        nodes[i]["min"], nodes[i]["max"] = lo[k:].min(axis=0), hi[k:].max(axis=0)
        nodes[i]["first"], nodes[i]["count"] = 2 * k + 1, 0
        if k < 58:
            leaf(2 * k + 2, k)
    leaf(117, 59)
    leaf(118, 58)
    return Bvh(nodes, tris)


def texels(n, rng, spread=0.2):
    it = np.zeros(n, N.TEXEL)
    it["pos"][:, 0:2] = rng.uniform(-spread, spread, (n, 2))
    it["pos"][:, 3] = 1.0
    it["normal"][:, 2] = 1.0
    it["dx"][:, 0] = 0.01
    it["dy"][:, 1] = 0.01
    return it


def params_for(sun=(0.0, 0.0, 1.0), atlas=96):
    p = np.zeros(1, N.PARAMS)
    p["sun_dir"] = [*sun, np.float32(np.tan(np.radians(0.5)))]
    p["sun_irr"] = [3, 3, 3, 0]
    p["zenith"] = [0.5, 0.5, 0.5, 0]
    p["horizon"] = [0.5, 0.5, 0.5, 0]
    p["ground"] = [0.5, 0.5, 0.5, 0.3]
    p["fallback"] = [np.float32(0.5 * np.pi), np.float32(0.5 * np.pi), np.float32(0.5 * np.pi), 0]
    p["counts"] = [4, 64, 64, 0]
    p["reach"] = [0.8, 12.0, 0.01, 0.0]
    p["atlas"] = [atlas, atlas, 0, np.float32(0.25).view(np.uint32)]
    return p


@pytest.mark.parametrize("name,form,tuning", FORMS, ids=[f[0] for f in FORMS])
def test_stack_overflow_drops_far_children(gpu, name, form, tuning):
    bvh = chain_scene()
    mats = np.array([[0.5, 0.5, 0.5, 1.0]], np.float32)
    params = params_for()
    rng = np.random.default_rng(3)
    items = texels(64, rng)
    # half of them where the slivers leave the axis free and the covering triangles block it
    items["pos"][:32, 0:2] = rng.uniform(0.02, 0.08, (32, 2))
    light = np.zeros((96 * 96, 4), np.float32)
    sb = gpu.scene(bvh, mats, np.zeros((1, 8), np.float32))
    try:
        got = run(gpu, sb, form, tuning, Pass.DIRECT, params, items, light)
    finally:
        sb.close()
    want = oracle.run(0, bvh.nodes, bvh.tris, mats, None, params, items, 0, light)
    # their hard rays run up past the slivers: only triangles in dropped leaves would block them
    o = np.array([items["pos"][0, 0], items["pos"][0, 1], 0.01], np.float32)
    up = np.array([0.0, 0.0, 1.0], np.float32)
    assert oracle.trace_brute(bvh.tris, mats, params, o, up, 1e30, 0)[1] != 0xFFFFFFFF
    assert (kit.bits(want[:32, 1]) & 0xFFFF == 0xFFFF).all(), "the oracle keeps the drop"
    check(name, Pass.DIRECT, got, want)


@pytest.mark.parametrize("name,form,tuning", FORMS, ids=[f[0] for f in FORMS])
@pytest.mark.parametrize("pass_", [Pass.DIRECT, Pass.GATHER, Pass.VOLUME])
def test_translucent_materials(gpu, name, form, tuning, pass_):
    scene = small_scene(1)
    scene.materials = [Material("grey", (0.5, 0.5, 0.5), opacity=0.45)]
    bvh, pos, mesh = kit.scene_bvh(scene)
    rng = np.random.default_rng(5)
    items = kit.probe_items(scene, 300) if pass_ == Pass.VOLUME else kit.sample_items(pos, mesh, 600, rng)
    charts = kit.charts_for(scene)
    light = kit.random_light(rng)
    mats = scene.materials_array()
    params = kit.params_for(scene)
    sb = gpu.scene(bvh, mats, charts)
    try:
        got = run(gpu, sb, form, tuning, pass_, params, items, light)
    finally:
        sb.close()
    want = oracle.run(int(pass_), bvh.nodes, bvh.tris, mats, charts, params, items, 0, light)
    if pass_ == Pass.VOLUME and name not in ("plain", "pooled"):
        # four rows an item: L0 rgb and the sun share, L1 and the back share
        assert np.array_equal(got[0::4, 3], want[0::4, 3]) and np.array_equal(got[1::4, 3], want[1::4, 3])
        assert np.allclose(got, want, rtol=2e-6, atol=1e-6)
    else:
        check(name, pass_, got, want)


@pytest.mark.parametrize("name,form,tuning", FORMS, ids=[f[0] for f in FORMS])
def test_empty_scene_and_no_items(gpu, name, form, tuning):
    # bvh.rs:76-80: no casting triangle is one node of count 0 over the zero triangle
    nodes = np.zeros(1, N.NODE)
    tris = np.zeros(1, N.TRI)
    bvh = Bvh(nodes, tris)
    mats = np.array([[0.5, 0.5, 0.5, 1.0]], np.float32)
    params = params_for(sun=(0.6, 0.8, 0.0))
    items = texels(128, np.random.default_rng(9), spread=5.0)
    light = np.zeros((96 * 96, 4), np.float32)
    sb = gpu.scene(bvh, mats, np.zeros((1, 8), np.float32))
    try:
        for pass_ in (Pass.DIRECT, Pass.GATHER):
            got = run(gpu, sb, form, tuning, pass_, params, items, light)
            want = oracle.run(int(pass_), nodes, tris, mats, None, params, items, 0, light)
            check(name, pass_, got, want)
        assert run(gpu, sb, form, tuning, Pass.GATHER, params, items[:0], light).shape == (0, 4)
    finally:
        sb.close()
