"""Pins of the oracle's texel passes by what the mathematics fixes, on the CPU.

oracle/filter_ref.c restates filter.rs (smooth :20-53, fill :58-94, sun_field :106-141) and
oracle/bake_ref.py the direct decode, the occlusion tone and the gather set (bake.rs:207-241,
:286-299). The whole-bake GPU parity rests on them, so each is held here to a closed form or an
invariant chosen so that a plausible misreading fails:
  smooth     a lone texel's spread is the Gaussian exp(-d^2 / sigma^2), sigma^2 = 2 (0.75 r)^2,
             over the (2r+1)^2 window clipped to the tile, normalized over what the window holds;
             constant fields stay, the pass is linear, it keeps to its chart and to n.n' >= 0.9
  fill       one ring a pass (each ring reads the ring before), own chart before any chart, an
             empty texel from any chart, a chart's texel never from another chart
  sun_field  a straight shadow edge: the distance to the nearest sample of the other state less
             half a sample, clamped to the radius; the majority rule at exactly g^2/2 lit
             samples; chart borders and folds stop it
  decode     the retry point it moves a texel to is the one the direct pass found free;
             the occlusion tone with a contrast other than 1
  gather set a chart narrower than the grid gets its own texel in every block it crosses
"""
from __future__ import annotations

import math

import numpy as np
import pytest

import oracle
from lightbaker_amd import _native as N
from lightbaker_amd.scenes import Mesh, Settings, cube, make_params
from oracle import bake_ref as ref

from .test_oracle import _plane_scene

NONE = 0xFFFFFFFF
F = np.float32


def surface(w, h, chart=0, normal=(0.0, 1.0, 0.0)):
    ch = np.full(w * h, chart, np.uint32)
    nr = np.tile(np.asarray(normal, np.float32), (w * h, 1))
    return ch, np.ascontiguousarray(nr)


# ---- smooth (filter.rs:20-53) ------------------------------------------------------------------

def gaussian_spread(w, h, r, src, valid):
    """What smoothing a unit value at texel `src` gives every texel, in float64 from the
    definition: the weight of src in texel i's window over the weights of every valid texel the
    window (clipped to the tile) holds."""
    sigma2 = 2.0 * (0.75 * r) ** 2
    sx, sy = src % w, src // w
    out = np.zeros(w * h)
    for i in range(w * h):
        x, y = i % w, i // w
        if not valid[i] or abs(sx - x) > r or abs(sy - y) > r:
            continue
        tot = 0.0
        for yy in range(max(0, y - r), min(h, y + r + 1)):
            for xx in range(max(0, x - r), min(w, x + r + 1)):
                if valid[yy * w + xx]:
                    tot += math.exp(-((xx - x) ** 2 + (yy - y) ** 2) / sigma2)
        out[i] = math.exp(-((sx - x) ** 2 + (sy - y) ** 2) / sigma2) / tot
    return out


@pytest.mark.parametrize("r", [1, 2, 3])
@pytest.mark.parametrize("src", [(4, 4), (0, 0), (8, 3)], ids=["centre", "corner", "edge"])
def test_smooth_of_a_lone_texel_is_the_normalized_gaussian(r, src):
    w, h = 9, 7
    ch, nr = surface(w, h)
    valid = np.ones(w * h, np.uint8)
    valid[w * 5 + 2] = 0  # a hole in the windows: it neither weighs nor is smoothed
    s = src[1] * w + src[0]
    v = np.zeros(w * h, np.float32)
    v[s] = 1.0
    v[w * 5 + 2] = 100.0
    got = ref.smooth(w, h, ch, nr, v, valid, r)
    want = gaussian_spread(w, h, r, s, valid)
    want[w * 5 + 2] = 100.0  # an invalid texel keeps its value
    assert np.allclose(got, want, rtol=2e-6, atol=1e-9), np.abs(got - want).max()


def test_smooth_constant_linear_and_identity():
    w, h = 12, 9
    ch, nr = surface(w, h)
    rng = np.random.default_rng(5)
    valid = (rng.random(w * h) > 0.2).astype(np.uint8)
    c = np.full(w * h, 0.37, np.float32)
    out = ref.smooth(w, h, ch, nr, c, valid, 2)
    assert np.allclose(out, 0.37, rtol=1e-6)
    u = rng.random((w * h, 3)).astype(np.float32)
    v = rng.random((w * h, 3)).astype(np.float32)
    lhs = ref.smooth(w, h, ch, nr, F(2) * u + F(3) * v, valid, 2)
    rhs = F(2) * ref.smooth(w, h, ch, nr, u, valid, 2) + F(3) * ref.smooth(w, h, ch, nr, v, valid, 2)
    assert np.allclose(lhs, rhs, rtol=1e-5)
    assert np.array_equal(ref.smooth(w, h, ch, nr, u, valid, 0), u)


@pytest.mark.parametrize("ndot,joins", [(0.85, False), (0.8999, False), (0.9001, True),
                                        (0.95, True)])
def test_smooth_keeps_to_its_chart_and_fold(ndot, joins):
    """Left half chart 0, right half chart 1: nothing crosses. Within chart 0 the top rows' faces
    turn so that n . n' = ndot against the rest: they join the window only from 0.9 up."""
    w, h = 8, 6
    ch, nr = surface(w, h)
    ch[np.arange(w * h) % w >= 4] = 1
    tilt = np.array([math.sqrt(1 - ndot * ndot), ndot, 0.0], np.float32)
    nr[: 2 * w] = tilt  # rows 0 and 1
    valid = np.ones(w * h, np.uint8)
    v = np.zeros(w * h, np.float32)
    src = 2 * w + 3  # row 2, the last column of chart 0
    v[src] = 1.0
    out = ref.smooth(w, h, ch, nr, v, valid, 1)
    assert out[2 * w + 4] == 0.0, "chart 1 never sees chart 0"
    # texel (3, 1) sits on the turned face next to the source
    assert (out[1 * w + 3] > 0.0) == joins
    if joins:
        return
    same = np.array([ch[j] == ch[src] and float(np.dot(nr[j], nr[src])) >= 0.9 for j in range(w * h)])
    want = gaussian_spread(w, h, 1, src, same.astype(np.uint8))
    sel = same & (ch == 0)
    assert np.allclose(out[sel], want[sel], rtol=2e-6, atol=1e-9)


# ---- fill (filter.rs:58-94) ---------------------------------------------------------------------

@pytest.mark.parametrize("chart", [0, NONE], ids=["chart", "empty"])
def test_fill_spreads_exactly_one_ring_a_pass(chart):
    w, h = 11, 9
    ch, _ = surface(w, h, chart)
    for rings in range(0, 6):
        v = np.zeros(w * h, np.float32)
        valid = np.zeros(w * h, np.uint8)
        c = 4 * w + 5
        v[c], valid[c] = 7.0, 1
        ref.fill(w, h, ch, v, valid, rings)
        x, y = np.arange(w * h) % w, np.arange(w * h) // w
        near = np.maximum(abs(x - 5), abs(y - 4)) <= rings
        assert np.array_equal(valid.astype(bool), near), rings
        assert (v[near] == 7.0).all() and (v[~near] == 0.0).all()


def test_fill_reads_the_ring_before():
    """A row 0 . . . 8: the first ring fills the ends' neighbours (0 and 8), the second the middle
    from them, (0 + 8) / 2 = 4. Filling from texels filled in the same ring would give 0."""
    w, h = 5, 1
    ch, _ = surface(w, h)
    v = np.array([0, 0, 0, 0, 8], np.float32)
    valid = np.array([1, 0, 0, 0, 1], np.uint8)
    ref.fill(w, h, ch, v, valid, 1)
    assert list(valid) == [1, 1, 0, 1, 1] and list(v) == [0, 0, 0, 8, 8]
    ref.fill(w, h, ch, v, valid, 1)
    assert list(valid) == [1] * 5 and v[2] == 4.0


def _ring3(centre_chart, charts, values):
    """A 3 x 3 tile: the invalid centre, valid neighbours at the given corners."""
    ch = np.full(9, 5, np.uint32)
    ch[4] = centre_chart
    v = np.zeros((9, 3), np.float32)
    valid = np.zeros(9, np.uint8)
    for (i, c), val in zip(charts.items(), values):
        ch[i], v[i], valid[i] = c, val, 1
    return ch, v, valid


def test_fill_own_chart_first_and_never_another():
    # chart 1's centre: its own neighbour (value 1) wins over chart 2's (5 and 9)
    ch, v, valid = _ring3(1, {0: 1, 8: 2, 2: 2}, [(1, 1, 1), (5, 5, 5), (9, 9, 9)])
    ref.fill(3, 3, ch, v, valid, 1)
    assert valid[4] == 1 and (v[4] == 1.0).all()
    # only chart 2's neighbours: a chart's texel is not filled from another chart
    ch, v, valid = _ring3(1, {8: 2, 2: 2}, [(5, 5, 5), (9, 9, 9)])
    ref.fill(3, 3, ch, v, valid, 3)
    assert valid[4] == 0 and (v[4] == 0.0).all()
    # two of its own among others: their mean, in FP32 (1 + 2 + 4) / 3 per channel order
    ch, v, valid = _ring3(1, {0: 1, 1: 1, 3: 1, 8: 2},
                          [(1, 10, 0.5), (2, 20, 0.25), (4, 40, 0.125), (100, 100, 100)])
    ref.fill(3, 3, ch, v, valid, 1)
    want = [F(F(F(1) + F(2)) + F(4)) / F(3), F(F(F(10) + F(20)) + F(40)) / F(3),
            F(F(F(0.5) + F(0.25)) + F(0.125)) / F(3)]
    assert list(v[4]) == want


def test_fill_empty_texel_takes_any_chart():
    ch, v, valid = _ring3(NONE, {0: 1, 8: 2, 2: 3}, [(1, 1, 1), (5, 5, 5), (9, 9, 9)])
    ref.fill(3, 3, ch, v, valid, 1)
    assert valid[4] == 1 and (v[4] == 5.0).all()


# ---- sun_field (filter.rs:106-141) ---------------------------------------------------------------

G = 4


def edge_bits(w, h, x_edge):
    """Hard bits of a vertical shadow edge at x = x_edge (texels): sample c of texel (x, y) sits at
    (x + (c % g + 0.5) / g, y + (c / g + 0.5) / g) and is lit left of the edge."""
    bits = np.zeros(w * h, np.uint32)
    for i in range(w * h):
        x = i % w
        for c in range(G * G):
            if x + (c % G + 0.5) / G < x_edge:
                bits[i] |= 1 << c
    return bits


def edge_field(w, h, x_edge, radius):
    """The field of a vertical edge in closed form. A texel is lit when most of its sample columns
    lie left of the edge. The nearest sample of the other state lies in the first sample column
    past the edge (lit texels) or the last before it (shadowed ones), in one of the texel's own
    rows, half a sample (1 / 2g) above or below its centre; the edge lies half a sample short of
    it: d = sqrt(dx^2 + (1/2g)^2) - 1/2g, clamped to [0, radius], then 0.5 +- 0.5 d / radius."""
    cols = [x + (k + 0.5) / G for x in range(w) for k in range(G)]
    first_dark = min(c for c in cols if c >= x_edge)
    last_lit = max(c for c in cols if c < x_edge)
    out = np.zeros(w * h)
    half = 0.5 / G
    for i in range(w * h):
        x = i % w
        own_lit = sum(1 for k in range(G) if x + (k + 0.5) / G < x_edge)
        lit = own_lit * 2 >= G
        cx = x + 0.5
        dx = first_dark - cx if lit else cx - last_lit
        d = min(max(math.sqrt(dx * dx + half * half) - half, 0.0), radius)
        out[i] = 0.5 + 0.5 * (d if lit else -d) / radius
    return out


@pytest.mark.parametrize("x_edge", [5.3, 5.5, 5.05, 5.7, 6.0])
@pytest.mark.parametrize("radius", [2.0, 1.5, 3.0])
def test_sun_field_of_a_straight_edge(x_edge, radius):
    w, h = 12, 5
    ch, nr = surface(w, h)
    covered = np.ones(w * h, np.uint8)
    got = ref.sun_field(w, h, ch, nr, edge_bits(w, h, x_edge), covered, G, F(radius))
    want = edge_field(w, h, x_edge, radius)
    assert np.allclose(got, want, atol=3e-7), (got.reshape(h, w)[0], want.reshape(h, w)[0])
    # far from the edge the distance is the radius: exactly 1 lit, 0 shadowed
    assert got[0] == 1.0 and got[w - 1] == 0.0


@pytest.mark.parametrize("lit_bits,lit", [(8, True), (7, False), (9, True)])
def test_sun_field_majority_at_half_the_samples(lit_bits, lit):
    """A texel whose first `lit_bits` samples (rows first) are lit, among fully lit texels: 8 of
    16 counts as lit (n * 2 >= g^2). Its nearest sample of the other state is its own, one
    sample column and half a sample row away: d = sqrt(2) / 2g - 1 / 2g."""
    w, h = 5, 5
    ch, nr = surface(w, h)
    bits = np.full(w * h, 0xFFFF, np.uint32)
    c = 2 * w + 2
    bits[c] = (1 << lit_bits) - 1
    got = ref.sun_field(w, h, ch, nr, bits, np.ones(w * h, np.uint8), G, F(2.0))
    d = (math.sqrt(2.0) - 1.0) * 0.5 / G
    want = 0.5 + 0.5 * (d if lit else -d) / 2.0
    assert abs(got[c] - want) < 3e-7, (got[c], want)


@pytest.mark.parametrize("split", ["chart", "fold"])
def test_sun_field_stops_at_chart_borders_and_folds(split):
    """Lit left of column 6, shadowed right of it, the two halves on two charts (or two faces
    with n . n' = 0.85): no texel sees the other state, every field is the radius (1 or 0)."""
    w, h = 12, 4
    ch, nr = surface(w, h)
    right = np.arange(w * h) % w >= 6
    if split == "chart":
        ch[right] = 1
    else:
        nr[right] = [math.sqrt(1 - 0.85 ** 2), 0.85, 0.0]
    bits = np.where(right, 0, 0xFFFF).astype(np.uint32)
    got = ref.sun_field(w, h, ch, nr, bits, np.ones(w * h, np.uint8), G, F(2.0))
    assert (got[~right] == 1.0).all() and (got[right] == 0.0).all()
    # one surface again: the border texels see each other (closed form of the edge at 6)
    got = ref.sun_field(w, h, *surface(w, h), bits, np.ones(w * h, np.uint8), G, F(2.0))
    assert np.allclose(got, edge_field(w, h, 6.0, 2.0), atol=3e-7)


def test_sun_field_uncovered_texels():
    w, h = 6, 3
    ch, nr = surface(w, h)
    covered = np.ones(w * h, np.uint8)
    covered[4] = 0
    bits = edge_bits(w, h, 3.0)
    bits[4] = 0  # an uncovered texel's bits are not read
    got = ref.sun_field(w, h, ch, nr, bits, covered, G, F(2.0))
    assert got[4] == -1.0
    assert np.allclose(np.delete(got, 4), np.delete(edge_field(w, h, 3.0, 2.0), 4), atol=3e-7)


# ---- decode and the occlusion tone (bake.rs:207-241) ---------------------------------------------

@pytest.mark.parametrize("strength,contrast,floor", [(1.0, 1.0, 0.25), (1.3, 2.2, 0.25),
                                                     (0.7, 0.5, 0.1), (2.0, 3.0, 0.0)])
def test_ao_tone(strength, contrast, floor):
    for occ in np.linspace(0.0, 1.0, 401, dtype=np.float32):
        got = ref.ao_tone(occ, strength, contrast, floor)
        base = max(0.0, float(F(1.0) - F(strength) * occ))
        want = max(float(F(floor)), base ** contrast)
        assert abs(got - want) <= 3e-7 * max(want, 1e-30) + 1e-30, (occ, got, want)
        if contrast == 1.0:  # no power at all: FP32 max(floor, max(0, 1 - s occ))
            assert got == max(F(floor), max(F(1.0) - F(strength) * occ, F(0.0)))


def test_decode_bits_state_and_inside_share():
    st = Settings()
    tex = np.zeros(1, N.TEXEL)[0]
    for pick, state in ((0, ref.VALIDITY_CENTRE), (3, ref.VALIDITY_OFFSET),
                        (8, ref.VALIDITY_OFFSET), (9, ref.VALIDITY_PUSHED)):
        word = np.array([0xABCDEF | (pick << 24)], np.uint32).view(np.float32)[0]
        r = np.array([0.5, word, 0.25, st.inside_share], np.float32)  # at the share: not inside
        soft, hard, s, ao = ref.decode(r, tex.copy(), st)
        assert (soft, hard, s) == (0.5, 0xABCDEF, state)
        assert ao == max(F(st.ao_floor), F(1.0) - F(st.ao_strength) * F(0.25))
    r = np.array([0.5, 0.0, 0.25, np.nextafter(F(st.inside_share), F(1))], np.float32)
    assert ref.decode(r, tex.copy(), st) is None  # strictly above the share: inside a solid


def _texel():
    it = np.zeros(1, dtype=N.TEXEL)
    it["pos"] = [0.0, 0.0, 0.0, 1.0]
    it["normal"] = [0.0, 1.0, 0.0, np.uint32(0).view(np.float32)]
    it["dx"] = [1.0, 0, 0, 0]
    it["dy"] = [0, 0, 1.0, 0]
    return it


POINTS = {0: (0.0, 0.0), 1: (0.3, -0.3), 2: (-0.3, -0.3), 3: (0.3, 0.3), 4: (-0.3, 0.3),
          5: (-0.45, 0.0), 6: (0.45, 0.0), 7: (0.0, -0.45), 8: (0.0, 0.45)}


def _direct(meshes, items):
    from lightbaker_amd.gpu import Bvh
    scene = _plane_scene(extra=meshes)
    pos, mesh, mat = scene.flat()
    bvh = Bvh.build(pos, mat, np.full(len(mat), NONE, np.uint32), scene.casts())
    P = make_params(scene, 4, 4)
    out = oracle.run(0, bvh.nodes, bvh.tris, scene.materials_array(), None, P, items, 0, None)
    return out, scene.settings


@pytest.mark.parametrize("free", range(1, 10))
def test_decode_moves_to_the_point_the_direct_pass_found_free(free):
    """Small solid cubes round every sample point of a 1 m texel but one (or, for 9, a 2 cm slab
    under all of them): the direct pass (trace.wgsl:185-210) finds that point free and reports
    it; decode moves the texel there (the reference's own mirror of the pass's offsets,
    bake.rs:495-503), and the direct pass from the moved texel finds its centre outside."""
    meshes = []
    if free == 9:
        meshes.append(Mesh("slab", 0, cube((-1.0, 0.0, -1.0), (1.0, 0.02, 1.0))))
    else:
        for q, (cx, cz) in POINTS.items():
            if q != free:
                meshes.append(Mesh(f"c{q}", 0, cube((cx - 0.08, 0.0, cz - 0.08),
                                                    (cx + 0.08, 0.2, cz + 0.08))))
    it = _texel()
    out, st = _direct(meshes, it)
    assert int(out[0, 1].view(np.uint32)) >> 24 == free, out[0]
    moved = it.copy()
    d = ref.decode(out[0], moved[0], st)
    assert d is not None and d[2] == (ref.VALIDITY_PUSHED if free == 9 else ref.VALIDITY_OFFSET)
    want = (0.0, 0.03, 0.0) if free == 9 else (POINTS[free][0], 0.0, POINTS[free][1])
    assert np.allclose(moved["pos"][0, :3], want, atol=1e-7)
    again, _ = _direct(meshes, moved)
    assert int(again[0, 1].view(np.uint32)) >> 24 == 0 and again[0, 3] <= st.inside_share


# ---- the gather set (bake.rs:286-299) ------------------------------------------------------------

def test_gather_set_grid_and_narrow_charts():
    """An 8 x 6 tile, ds = 2. Chart 0 everywhere but column 3 (chart 1, one texel wide, never on
    the grid) and an invalid grid texel (4, 2). The grid texels of chart 0 are gathered; chart 1
    gets the first texel of each block it crosses, (3, 0), (3, 2), (3, 4); block (2, 1) of
    chart 0 loses its grid texel and gets its next one in row order, (5, 2)."""
    tw, th, ds = 8, 6, 2
    ch, _ = surface(tw, th)
    x, y = np.arange(tw * th) % tw, np.arange(tw * th) // tw
    ch[x == 3] = 1
    valid = np.ones(tw * th, bool)
    valid[2 * tw + 4] = False
    got = ref.gather_set(tw, tw * th, valid, ch, ds)
    want = (x % 2 == 0) & (y % 2 == 0) & valid
    for yy in (0, 2, 4):
        want[yy * tw + 3] = True
    want[2 * tw + 5] = True
    assert np.array_equal(got, want), np.nonzero(got ^ want)


def test_gather_set_partial_block_and_ds1():
    """A 5 x 8 tile with ds = 4: two block columns, the second one texel wide (the blocks a row
    are tw.div_ceil(ds) = 2). Chart 2 holds column 4, chart 3 the off-grid texels (4, 1) and
    (1, 5), in blocks (1, 0) and (0, 1): each is the first of its (block, chart) pair, so both
    are gathered; a row of blocks counted as tw / ds = 1 would merge those two blocks."""
    tw, th, ds = 5, 8, 4
    ch, _ = surface(tw, th)
    x, y = np.arange(tw * th) % tw, np.arange(tw * th) // tw
    ch[x == 4] = 2
    ch[1 * tw + 4] = 3
    ch[5 * tw + 1] = 3
    valid = np.ones(tw * th, bool)
    got = ref.gather_set(tw, tw * th, valid, ch, ds)
    want = (x % 4 == 0) & (y % 4 == 0)  # (0, 0), (4, 0), (0, 4), (4, 4)
    want[1 * tw + 4] = want[5 * tw + 1] = True
    assert np.array_equal(got, want), np.nonzero(got ^ want)
    # ds = 1: every valid texel
    valid[7] = False
    assert np.array_equal(ref.gather_set(tw, tw * th, valid, ch, 1), valid)


# ---- the whole-bake oracle's array forms equal the pinned scalar ones ---------------------------

@pytest.mark.parametrize("contrast", [1.0, 2.2])
def test_decode_all_equals_decode(contrast):
    rng = np.random.default_rng(31)
    n = 4000
    st = Settings(ao_contrast=contrast, ao_strength=1.3)
    tex = np.zeros(n, N.TEXEL)
    for f in ("pos", "normal", "dx", "dy"):
        tex[f][:, :3] = rng.normal(size=(n, 3)).astype(np.float32)
    pick = rng.integers(0, 10, n).astype(np.uint32)
    hard = rng.integers(0, 1 << 16, n).astype(np.uint32)
    res = np.stack([rng.random(n), (hard | pick << 24).view(np.float32), rng.random(n),
                    rng.choice([0.0, st.inside_share, 0.3, 0.9], n)], 1).astype(np.float32)
    cov = np.sort(rng.choice(n + 50, n, replace=False))  # a sparse subset of a longer list
    big = np.zeros(n + 50, N.TEXEL)
    big[cov] = tex
    a, b = big.copy(), big.copy()
    inside, soft, bits, state, ao = ref.decode_all(res, a, cov, st)
    for k, gi in enumerate(cov):
        d = ref.decode(res[k], b[gi], st)
        assert (d is None) == inside[k]
        if d is not None:
            assert (d[0], d[1], d[2]) == (soft[k], bits[k], state[k]) and d[3] == ao[k], k
    assert np.array_equal(a["pos"].view(np.uint32), b["pos"].view(np.uint32))


@pytest.mark.parametrize("ds", [1, 2, 4])
def test_gather_set_all_equals_gather_set(ds):
    rng = np.random.default_rng(ds)
    for tw, th in ((13, 9), (16, 16), (5, 11)):
        size = tw * th
        valid = rng.random(size) > 0.3
        chart = rng.integers(0, 4, size).astype(np.uint32)
        chart[rng.random(size) < 0.1] = NONE
        assert np.array_equal(ref.gather_set_all(tw, size, valid, chart, ds),
                              ref.gather_set(tw, size, valid, chart, ds))
