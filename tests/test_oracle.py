"""Pins of the oracle (oracle/trace_ref.c) by what the mathematics fixes, on the CPU.

The reference holds no golden vectors for its passes (its only tests are the qualitative
thresholds of crates/lightbaker/tests/bake.rs), so each oracle function is pinned here by a
closed form, an invariant or a brute-force computation chosen so that a plausible slip (a
swapped mix, a dropped cosine factor, a wrong shift or constant) fails:
  pcg / rand / radical   integer restatement + golden table, Van der Corput values, bijectivity
  lb_sincos              libm double sin/cos within 2 ulp over the arguments the passes use
  cosine                 E[cos] = 2/3, E[cos^2] = 1/2, unit length, upper hemisphere
  trace                  equals a brute-force loop over every triangle
  gather                 an open face under a gradient sky: E = pi (h/3 + 2 z/3) (output.rs:128)
  direct AO              an infinite ceiling h above the ray origins, reach R:
                         occ = 1 - a^2 + 2 a^2 ln a, a = h/R
  volume                 constant radiance L: L0 = 0.282095 * 4 pi * L, L1 = 0, sun seen
  direct sun bits        the analytic shadow of a box on the ground
The FP64 tracer (oracle/fp64_ref.c), the yardstick of the FP32 passes' error, is held to the same
closed forms (the tests parametrized over "f32" and "f64").
"""
from __future__ import annotations

import json
import math
from pathlib import Path

import numpy as np
import pytest

import oracle
from lightbaker_amd import _native as N
from lightbaker_amd.gpu import Bvh
from lightbaker_amd.scenes import Scene, Material, Mesh, Settings, Volume, cube, make_params, quad

from . import kit

GOLDEN = Path(__file__).resolve().parent / "golden"
M32 = 0xFFFFFFFF


def pcg_int(v: int) -> int:
    s = (v * 747796405 + 2891336453) & M32
    w = (((s >> ((s >> 28) + 4)) ^ s) * 277803737) & M32
    return ((w >> 22) ^ w) & M32


def test_pcg_golden_and_integer_form():
    table = json.loads((GOLDEN / "pcg.json").read_text())
    L = oracle.lib()
    for v, want in table["pcg"]:
        assert L.orc_pcg(v) == want == pcg_int(v)
    for (a, b), want in table["rand"]:
        assert np.float32(L.orc_rand(a, b, 0.0)) == np.float32(want)


def test_pcg_is_a_bijection_on_a_block():
    # every stage (LCG step, xorshifts, odd multiply) is invertible: no two inputs collide
    L = oracle.lib()
    vals = {L.orc_pcg(v) for v in range(1 << 16)}
    assert len(vals) == 1 << 16


def test_rand_range_and_uniformity():
    L = oracle.lib()
    x = np.array([L.orc_rand(k * 7919, 19, 0.0) for k in range(20000)])
    assert x.min() >= 0.0 and x.max() <= 1.0
    hist, _ = np.histogram(x, bins=20, range=(0, 1))
    chi2 = ((hist - 1000.0) ** 2 / 1000.0).sum()
    assert chi2 < 50.0  # 19 dof, p ~ 1e-4


def test_radical_van_der_corput():
    L = oracle.lib()
    want = {0: 0.0, 1: 0.5, 2: 0.25, 3: 0.75, 4: 0.125, 5: 0.625, 6: 0.375, 7: 0.875}
    for i, v in want.items():
        assert L.orc_radical(i) == v
    assert L.orc_radical(0x80000000) == 2.0 ** -32


def test_sincos_against_libm():
    import ctypes as C
    L = oracle.lib()
    s, c = C.c_float(), C.c_float()
    xs = np.concatenate([np.linspace(0.0, 2 * math.pi, 5001), np.arange(0, 64) * 2.399963,
                         -np.linspace(0.0, 7.0, 301)]).astype(np.float32)
    for x in xs:
        L.lb_sincos(float(x), C.byref(s), C.byref(c))
        ws, wc = math.sin(float(x)), math.cos(float(x))
        assert abs(s.value - ws) <= 2.5e-7 * max(1.0, abs(ws)) + 1.2e-7, (x, s.value, ws)
        assert abs(c.value - wc) <= 2.5e-7 * max(1.0, abs(wc)) + 1.2e-7, (x, c.value, wc)


def test_cosine_distribution_moments():
    L = oracle.lib()
    n = 4096
    out = np.zeros(3, dtype=np.float32)
    for normal in ([0.0, 1.0, 0.0], [0.6, 0.0, 0.8], [0.0, -1.0, 0.0]):
        nrm = np.array(normal, dtype=np.float32)
        mu = []
        for r in range(n):
            L.orc_cosine((r + 0.5) / n, L.orc_radical(r), nrm.ctypes.data, out.ctypes.data)
            assert abs(float(np.dot(out, out)) - 1.0) < 1e-5
            mu.append(float(np.dot(out, nrm)))
        mu = np.array(mu)
        assert mu.min() >= -1e-6
        assert abs(mu.mean() - 2.0 / 3.0) < 2e-3
        assert abs((mu ** 2).mean() - 0.5) < 2e-3


def test_trace_equals_brute_force():
    from lightbaker_amd.scenes import small_scene
    scene = small_scene(1)
    bvh, pos, mesh = kit.scene_bvh(scene)
    P = kit.params_for(scene)
    mats = scene.materials_array()
    rng = np.random.default_rng(3)
    ties = 0
    hits = 0
    for k in range(3000):
        o = rng.uniform([-11, 0.01, -11], [11, 5, 11]).astype(np.float32)
        d = rng.normal(size=3)
        d = (d / np.linalg.norm(d)).astype(np.float32)
        tmax = float(rng.choice([0.8, 12.0, 1e30]))
        a = oracle.trace(bvh.nodes, bvh.tris, mats, P, o, d, tmax, False, k)
        b = oracle.trace_brute(bvh.tris, mats, P, o, d, tmax, k)
        hits += b[1] != 0xFFFFFFFF
        if a != b:
            # FP32 slab culling (trace.wgsl:79, `lo < tmax` against the best t so far) may drop a
            # coplanar face whose t is a few ulps below the one found: boxes stand on the ground
            assert a[1] != 0xFFFFFFFF and b[1] != 0xFFFFFFFF, (k, a, b)
            assert abs(a[0] - b[0]) <= 1e-6 * b[0], (k, a, b)
            ties += 1
    assert hits > 1000 and ties < 30


def _plane_scene(ceiling_h=None, extra=()):
    big = 200.0
    meshes = [Mesh("ground", 0, np.array(quad((-big, 0, -big), (-big, 0, big), (big, 0, big),
                                              (big, 0, -big)), np.float32))]
    if ceiling_h is not None:
        # facing down (towards the ground): clockwise seen from above
        meshes.append(Mesh("ceiling", 0, np.array(quad((-big, ceiling_h, -big),
                                                       (big, ceiling_h, -big),
                                                       (big, ceiling_h, big),
                                                       (-big, ceiling_h, big)), np.float32)))
    for e in extra:
        meshes.append(e)
    return Scene(sun_direction=(0.3, 1.0, 0.2), sun_radius_deg=0.5, sun_irradiance=(3, 3, 3),
                 zenith=(0.8, 0.8, 0.8), horizon=(0.2, 0.2, 0.2), ground=(0.1, 0.1, 0.1),
                 materials=[Material("g", (0.5, 0.5, 0.5))], meshes=meshes,
                 settings=Settings(ao_rays=1024, gi_rays=1024, ao_reach=0.8),
                 volume=Volume((0, 0, 0), (1, 1, 1), 1.0, 1024))


def _ground_texel(x=0.0, z=0.0):
    it = np.zeros(1, dtype=N.TEXEL)
    it["pos"] = [x, 0.0, z, 1.0]
    it["normal"] = [0.0, 1.0, 0.0, np.uint32(0).view(np.float32)]
    it["dx"] = [0.25, 0, 0, 0]
    it["dy"] = [0, 0, 0.25, 0]
    return it


def _run(scene, pass_, items, light=None, prec="f32"):
    pos, mesh, mat = scene.flat()
    bvh = Bvh.build(pos, mat, np.full(len(mat), 0xffffffff, np.uint32), scene.casts())
    P = make_params(scene, 4, 4)
    fn = oracle.run if prec == "f32" else oracle.run64
    return fn(pass_, bvh.nodes, bvh.tris, scene.materials_array(), None, P, items, 0,
              light if light is not None else np.zeros((16, 4), np.float32))


def _word(out, k, prec):
    """The direct pass's hard bits and pick word of item k (FP32: the float's bits; the FP64
    tracer stores the integer's value)."""
    if prec == "f32":
        return int(np.float32(out[k, 1]).view(np.uint32))
    return int(out[k, 1])


PREC = pytest.mark.parametrize("prec", ["f32", "f64"])


@PREC
def test_gather_open_plane_gradient_sky(prec):
    scene = _plane_scene()
    out = _run(scene, 1, _ground_texel(), prec=prec)
    z, h = 0.8, 0.2
    want = math.pi * (h / 3 + 2 * z / 3)
    assert abs(out[0, 0] - want) < 2e-3 * want, (out[0], want)
    assert out[0, 3] == 0.0


@PREC
@pytest.mark.parametrize("h", [0.4, 0.2])
def test_ao_under_infinite_ceiling(h, prec):
    scene = _plane_scene(ceiling_h=h)
    out = _run(scene, 0, _ground_texel(), prec=prec)
    a = (h - 0.01) / 0.8  # rays leave 0.01 m above the surface (bake.rs:24 LIFT)
    want = 1 - a * a + 2 * a * a * math.log(a)
    assert abs(out[0, 2] - want) < 3e-3, (out[0], want)


@PREC
def test_ao_open_plane_is_zero(prec):
    out = _run(_plane_scene(), 0, _ground_texel(), prec=prec)
    assert out[0, 2] == 0.0 and out[0, 0] == 1.0
    assert _word(out, 0, prec) == 0xFFFF  # 16 hard bits lit, pick 0


@PREC
def test_volume_constant_radiance(prec):
    far = Mesh("far", 0, np.array(quad((500, 0, 500), (500, 0, 501), (501, 0, 501), (501, 0, 500)),
                                  np.float32))
    scene = _plane_scene()
    scene.meshes = [far]
    scene.zenith = scene.horizon = scene.ground = (0.5, 0.5, 0.5)
    scene.settings.gi_rays = 1024
    probe = np.zeros(1, dtype=N.TEXEL)
    probe["pos"] = [0, 5, 0, 1]
    probe["normal"] = [0, 1, 0, np.uint32(3).view(np.float32)]
    out = _run(scene, 2, probe, prec=prec).reshape(4, 4)
    want = 0.282095 * 4 * math.pi * 0.5
    assert abs(out[0, 0] - want) < 2e-5 * want
    assert abs(out[1:, :3]).max() < 2e-3
    assert out[0, 3] == 1.0 and out[1, 3] == 0.0


@PREC
def test_sun_bits_match_the_analytic_shadow(prec):
    box = Mesh("box", 0, cube((0.0, 0.0, 0.0), (2.0, 2.0, 2.0)))
    scene = _plane_scene(extra=[box])
    scene.sun_direction = (-1.0, 0.8, 0.3)
    sun = np.array(scene.sun_direction, float)
    sun /= np.linalg.norm(sun)
    xs = np.arange(2.2, 6.0, 0.25) + 0.125
    zs = np.arange(-1.0, 3.0, 0.25) + 0.125
    items = np.concatenate([_ground_texel(x, z) for x in xs for z in zs])
    out = _run(scene, 0, items, prec=prec)

    def shadowed(p):
        # the ray p + t sun against the box [0,2]^3 (slab test in double)
        lo, hi = 0.0, 1e9
        for a in range(3):
            if abs(sun[a]) < 1e-12:
                if not (0.0 <= p[a] <= 2.0):
                    return False
                continue
            t1, t2 = (0.0 - p[a]) / sun[a], (2.0 - p[a]) / sun[a]
            lo, hi = max(lo, min(t1, t2)), min(hi, max(t1, t2))
        return lo <= hi

    checked = 0
    for k, it in enumerate(items):
        p = it["pos"][:3].astype(float) + np.array([0, 0.01, 0])
        corners = [p + np.array([dx, 0, dz]) for dx in (-0.25, 0.25) for dz in (-0.25, 0.25)]
        states = {shadowed(c) for c in corners} | {shadowed(p)}
        if len(states) != 1:
            continue  # near the shadow's edge
        bits = _word(out, k, prec) & 0xFFFF
        assert bits == (0 if states.pop() else 0xFFFF), (it["pos"], hex(bits))
        checked += 1
    assert checked > 100


@PREC
def test_texel_inside_a_slab_is_retried(prec):
    # a slab 0.2 m thick on the ground: from under it, the rays meet its top face from behind
    # within the inside reach (min(0.3, aoReach), bake.rs:115), so the texel lies inside a solid
    slab = Mesh("slab", 0, cube((-1.0, 0.0, -1.0), (1.0, 0.2, 1.0)))
    scene = _plane_scene(extra=[slab])
    out = _run(scene, 0, _ground_texel(0.0, 0.0), prec=prec)
    pick = _word(out, 0, prec) >> 24
    # every point across it (and pushed 0.03 m up) is inside: no pick; the AO rays' back share is
    # the share of cosine rays meeting the top (0.19 m up) within 0.3 m: 1 - (0.19 / 0.3)^2
    assert pick == 0
    assert abs(out[0, 3] - (1 - (0.19 / 0.3) ** 2)) < 0.02, out[0]
    # centre 0.05 m inside the slab's edge: the first quarter point (+0.3, -0.3 texel,
    # trace.wgsl:153) lies 0.025 m outside, so it is picked and the sun cells shrink round it
    edge = _run(scene, 0, _ground_texel(0.95, 0.0), prec=prec)
    assert _word(edge, 0, prec) >> 24 == 1, edge[0]
    assert edge[0, 3] == 0.0


def test_gather_terms_sum_to_the_gather():
    """oracle.gather_terms gives each gather ray's term as gather_item adds it: summed in ray
    order in FP32, times pi, over the ray count, they are the gather's output bit for bit."""
    from lightbaker_amd.scenes import small_scene
    scene = small_scene(1)
    bvh, pos, mesh = kit.scene_bvh(scene)
    rng = np.random.default_rng(17)
    items = kit.sample_items(pos, mesh, 200, rng)
    charts = kit.charts_for(scene)
    light = kit.random_light(rng)
    P = kit.params_for(scene)
    mats = scene.materials_array()
    want = oracle.run(1, bvh.nodes, bvh.tris, mats, charts, P, items, 0, light)
    terms = oracle.gather_terms(bvh.nodes, bvh.tris, mats, charts, P, items, 0, light)
    s = np.zeros((len(items), 3), np.float32)
    for r in range(terms.shape[1]):
        s = s + terms[:, r]
    got = s * np.float32(math.pi) / np.float32(terms.shape[1])
    assert np.array_equal(got.view(np.uint32), want[:, :3].view(np.uint32))
