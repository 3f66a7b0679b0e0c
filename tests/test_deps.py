"""The tile dependencies' host parts, on the CPU:

  * BLAKE3 (lb_blake3, the library's host code) against the published BLAKE3 test vectors
    (tests/golden/blake3.json: inputs of 0 to 3072 bytes, one to three 1 KiB chunks, so the
    chunk chaining and the parent nodes are all exercised) and the "abc" vector;
  * the oracle's `swept` (deps.rs:104-121) against the same question in exact rational
    arithmetic (is there a t in [0, tmax] where the zone moved by t * sun overlaps the box) on
    random boxes, zones and suns, and on closed cases (the box in the zone, upwind, downwind,
    above the top, a horizontal sun, a sun along an axis);
  * the settings' JSON as the scene file writes it (serde_json's f32 layout).
"""
from __future__ import annotations

import json
from fractions import Fraction
from pathlib import Path

import numpy as np

from oracle import deps_ref as D

GOLDEN = Path(__file__).parent / "golden" / "blake3.json"


def test_blake3_vectors():
    from lightbaker_amd.deps import blake3
    v = json.loads(GOLDEN.read_text())
    data = bytes(i % 251 for i in range(4096))
    for n, want in v["hash"].items():
        assert blake3(data[:int(n)]).hex() == want, n
    assert blake3(b"abc").hex() == v["abc"]


def exact_swept(z, s, top, b):
    """The exact answer, and how far it is from flipping (the overlap's length, in t)."""
    z = [Fraction(float(x)) for x in np.float32(z)]
    b = [Fraction(float(x)) for x in np.float32(b)]
    s32 = np.float32(s)
    s = [Fraction(float(x)) for x in s32]
    top = Fraction(float(np.float32(top)))
    tmax = max((top - z[1]) / s[1], Fraction(0)) if s32[1] > np.float32(1e-4) else Fraction(0)
    lo, hi = Fraction(0), tmax
    margin = None
    for a in range(3):
        if abs(s32[a]) < np.float32(1e-6):
            gap = min(b[3 + a] - z[a], z[3 + a] - b[a])  # >= 0 when the intervals overlap
            margin = gap if margin is None else min(margin, gap)
            if gap < 0:
                return False, abs(gap)
            continue
        t1, t2 = (b[3 + a] - z[a]) / s[a], (b[a] - z[3 + a]) / s[a]
        a0, a1 = (t2, t1) if s[a] > 0 else (t1, t2)
        lo, hi = max(lo, a0), min(hi, a1)
    m = hi - lo
    margin = m if margin is None else min(margin, m)
    return hi >= lo and margin >= 0, abs(margin)


def test_swept_matches_exact_arithmetic():
    rng = np.random.default_rng(7)
    agree = {True: 0, False: 0}
    for i in range(4000):
        sun = rng.normal(size=3)
        if i % 10 == 0:
            sun[1] = abs(sun[1]) * 1e-5  # horizontal
        if i % 7 == 0:
            sun[rng.integers(0, 3)] = 0.0  # along a plane
        sun = (sun / np.linalg.norm(sun)).astype(np.float32)
        zc = rng.uniform(-20, 20, 3)
        zh = rng.uniform(1, 10, 3)
        z = np.concatenate([zc - zh, zc + zh]).astype(np.float32)
        # half the boxes near the zone's path towards (or away from) the sun
        bc = zc + (sun * rng.uniform(-10, 30) if i % 2 else 0) + rng.normal(scale=6 if i % 2 else 15, size=3)
        bh = rng.uniform(0.1, 3, 3)
        b = np.concatenate([bc - bh, bc + bh]).astype(np.float32)
        top = np.float32(rng.uniform(0, 40))
        want, margin = exact_swept(z, sun, top, b)
        got = D.swept(z, sun, top, b)
        if margin > Fraction(1, 10 ** 4):
            assert got == want, (z, sun, top, b)
            agree[want] += 1
    assert agree[True] > 500 and agree[False] > 500, agree


def test_swept_closed_cases():
    z = np.array([0, 0, 0, 2, 2, 2], np.float32)
    up = np.array([0.6, 0.8, 0.0], np.float32)
    box = lambda c, h=0.25: np.array([*(np.subtract(c, h)), *(np.add(c, h))], np.float32)
    assert D.swept(z, up, 10, box((1, 1, 1)))                 # inside the zone
    assert D.swept(z, up, 10, box((1 + 0.6 * 5, 1 + 0.8 * 5, 1)))   # towards the sun, below top
    assert not D.swept(z, up, 2.5, box((1 + 0.6 * 5, 1 + 0.8 * 5, 1)))  # past the top
    assert not D.swept(z, up, 10, box((1 - 0.6 * 5, 1 - 0.8 * 5, 1)))  # away from the sun
    assert not D.swept(z, up, 10, box((1, 1, 5)))              # beside the path
    flat = np.array([1.0, 0.0, 0.0], np.float32)                # a sun on the horizon: tmax 0
    assert D.swept(z, flat, 10, box((1.9, 1, 1))) and not D.swept(z, flat, 10, box((4, 1, 1)))
    # a sun with no z: the z extents must overlap where they stand
    assert not D.swept(z, up, 10, box((1 + 0.6 * 5, 1 + 0.8 * 5, 2.3)))
    assert D.swept(z, up, 10, box((1 + 0.6 * 5, 1 + 0.8 * 5, 2.2)))


def test_settings_json_layout():
    from lightbaker_amd.deps import _f32, _json
    assert [_f32(x) for x in (0.5, 2.0, 12.0, 0.8, 0.25, 1.0, 100.0, -1.0, 0.0)] == \
        ["0.5", "2.0", "12.0", "0.8", "0.25", "1.0", "100.0", "-1.0", "0.0"]
    assert _f32(0.1) == "0.1" and _f32(1 / 3) == "0.33333334"
    assert _json({"a": [1.5, 2], "b": True, "c": "x"}) == '{"a":[1.5,2],"b":true,"c":"x"}'


def test_settings_f32_shortest_round_trip():
    """What the f32 layout must hold whatever the value (ryu's contract): the digits read back as
    the same f32, and no fewer significant digits do; integers keep a ".0"."""
    from lightbaker_amd.deps import _f32
    rng = np.random.default_rng(11)
    bits = rng.integers(0, 0x7f800000, 20000, dtype=np.uint32)
    vals = np.concatenate([bits.view(np.float32), np.float32([1, 0.1, 0.25, 3e-7, 1e16, 123456.0])])
    for x in vals:
        s = _f32(x)
        assert np.float32(float(s)) == x, (x, s)
        mant = s.split("e")[0].replace(".", "").replace("-", "").lstrip("0").rstrip("0") or "0"
        n = len(mant)
        if n > 1:
            shorter = np.float32(float(f"{float(x):.{n - 2}e}"))
            assert shorter != x, (x, s)
        if "e" not in s:
            assert "." in s, s
