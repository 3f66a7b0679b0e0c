"""ORACLE — test infrastructure, not product code.

The device half of the reference's bake (crates/lightbaker/src/bake.rs:149-330, volume.rs:69-117)
and its host finish (bake.rs:337-421) restated over the same inputs the package's bake prepares
(its layout, texels, BVH and Params, read back from a prepared bake), with the ray passes and
the texel passes of the oracle's C restatement (trace_ref.c, filter_ref.c). A plain-form bake of
the package must equal it bit for bit.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import lib, run_quant, set_sky_sh
from . import run as run_fp32

VALIDITY_NONE, VALIDITY_CENTRE, VALIDITY_OFFSET = 0, 1, 2
VALIDITY_PUSHED, VALIDITY_FILLED, VALIDITY_VOLUME = 3, 4, 5
PUSH = np.float32(0.03)
F = np.float32


def _filters():
    L = lib()
    if not hasattr(L, "_filters_bound"):
        VP, I = C.c_void_p, C.c_int
        L.orc_smooth.argtypes = [I, I, VP, VP, VP, I, VP, I, VP]
        L.orc_smooth.restype = None
        L.orc_fill.argtypes = [I, I, VP, VP, I, VP, I]
        L.orc_fill.restype = None
        L.orc_sun_field.argtypes = [I, I, VP, VP, VP, VP, I, C.c_float, VP]
        L.orc_sun_field.restype = None
        L.orc_powf_n.argtypes = [VP, C.c_size_t, C.c_float, VP]
        L.orc_powf_n.restype = None
        L._filters_bound = True
    return L


def smooth(w, h, chart, normal, v, valid, r):
    L = _filters()
    v = np.ascontiguousarray(v, dtype=np.float32)
    C_ = v.shape[1] if v.ndim == 2 else 1
    out = np.empty_like(v)
    L.orc_smooth(w, h, chart.ctypes.data, normal.ctypes.data, v.ctypes.data, C_,
                 np.ascontiguousarray(valid, np.uint8).ctypes.data, r, out.ctypes.data)
    return out


def fill(w, h, chart, v, valid, rings):
    """In place on v and valid."""
    L = _filters()
    C_ = v.shape[1] if v.ndim == 2 else 1
    L.orc_fill(w, h, chart.ctypes.data, v.ctypes.data, C_, valid.ctypes.data, rings)


def sun_field(w, h, chart, normal, hard, covered, g, radius):
    L = _filters()
    out = np.empty(w * h, np.float32)
    L.orc_sun_field(w, h, chart.ctypes.data, normal.ctypes.data, hard.ctypes.data,
                    np.ascontiguousarray(covered, np.uint8).ctypes.data, g, radius,
                    out.ctypes.data)
    return out


def offset(q):
    """bake.rs:495-503"""
    if 1 <= q <= 4:
        return (F(0.3) if q & 1 else F(-0.3), F(0.3) if q > 2 else F(-0.3))
    return {5: (F(-0.45), F(0)), 6: (F(0.45), F(0)), 7: (F(0), F(-0.45))}.get(q, (F(0), F(0.45)))


def powf(x, y):
    """The host's powf elementwise (glibc's, as Rust's f32::powf), float32 in and out."""
    x = np.ascontiguousarray(np.atleast_1d(x), np.float32)
    out = np.empty_like(x)
    _filters().orc_powf_n(x.ctypes.data, len(x), F(y), out.ctypes.data)
    return out


def ao_tone(occ, strength, contrast, floor):
    """bake.rs:238: max(floor, max(0, 1 - strength * occ) ^ contrast), FP32."""
    v = max(F(1.0) - F(strength) * F(occ), F(0.0))
    if contrast != 1.0:
        v = powf(v, contrast)[0]
    return max(v, F(floor))


def decode(r, texel, settings):
    """bake.rs:207-241 for one covered texel: its direct-pass result `r` (vec4) and its record
    (moved in place to the point the pass sampled). Returns None when the texel lies inside a
    solid (invalid), else (soft, hard bits, validity state, occlusion tone)."""
    if r[3] > F(settings.inside_share):
        return None
    bits = int(np.asarray(r[1:2], np.float32).view(np.uint32)[0])
    pick = bits >> 24
    state = VALIDITY_CENTRE if pick == 0 else (VALIDITY_PUSHED if pick == 9 else VALIDITY_OFFSET)
    if pick == 9:
        texel["pos"][:3] = texel["pos"][:3] + texel["normal"][:3] * PUSH
    elif pick > 0:
        cx, cy = offset(pick)
        texel["pos"][:3] = texel["pos"][:3] + (texel["dx"][:3] * cx + texel["dy"][:3] * cy)
    ao = ao_tone(r[2], settings.ao_strength, settings.ao_contrast, settings.ao_floor)
    return F(r[0]), bits & 0x00FFFFFF, state, ao


def gather_set(tw, size, valid, chart, ds):
    """bake.rs:286-299 over one tile of width tw: the valid texels on the ds grid, then the first
    valid texel (row-major) of every (ds-block, chart) pair the grid left uncovered. Returns a
    boolean mask over the tile's `size` texels."""
    bw = -(-tw // ds)
    gathered = np.zeros(size, bool)
    seen = set()
    for i in range(size):
        if valid[i] and (i % tw) % ds == 0 and (i // tw) % ds == 0:
            gathered[i] = True
            seen.add((((i // tw) // ds) * bw + (i % tw) // ds, int(chart[i])))
    for i in range(size):
        key = (((i // tw) // ds) * bw + (i % tw) // ds, int(chart[i]))
        if valid[i] and key not in seen:
            seen.add(key)
            gathered[i] = True
    return gathered


def decode_all(res, tex, cov, settings):
    """decode() over every covered texel at once, elementwise in the same FP32 operations:
    `cov` the texels' indices in `tex` (moved in place), `res` their direct-pass results in that
    order. Returns (inside, soft, hard bits, state, occlusion tone) over `cov`."""
    res = np.asarray(res, np.float32)
    inside = res[:, 3] > F(settings.inside_share)
    bits = res[:, 1].view(np.uint32)
    pick = bits >> 24
    state = np.where(pick == 0, VALIDITY_CENTRE,
                     np.where(pick == 9, VALIDITY_PUSHED, VALIDITY_OFFSET)).astype(np.uint8)
    ok = ~inside
    pos = tex["pos"][cov, :3]
    m = ok & (pick == 9)
    pos[m] = pos[m] + tex["normal"][cov[m], :3] * PUSH
    for q in range(1, 9):
        m = ok & (pick == q)
        cx, cy = offset(q)
        pos[m] = pos[m] + (tex["dx"][cov[m], :3] * cx + tex["dy"][cov[m], :3] * cy)
    tex["pos"][cov, :3] = pos
    v = np.maximum(F(1.0) - F(settings.ao_strength) * res[:, 2], F(0.0))
    if settings.ao_contrast != 1.0:
        v = powf(v, settings.ao_contrast)
    ao = np.maximum(v, F(settings.ao_floor))
    return inside, res[:, 0].copy(), bits & 0x00FFFFFF, state, ao


def gather_set_all(tw, size, valid, chart, ds):
    """gather_set() over a tile with array operations: the grid texels, then of the other valid
    texels whose (block, chart) pair no grid texel holds, the first of each pair."""
    i = np.arange(size)
    x, y = i % tw, i // tw
    valid = np.asarray(valid, bool)
    key = ((y // ds) * (-(-tw // ds)) + x // ds).astype(np.int64) << 32 | np.asarray(chart, np.int64)
    grid = valid & (x % ds == 0) & (y % ds == 0)
    cand = valid & ~grid & ~np.isin(key, key[grid])
    _, first = np.unique(key[cand], return_index=True)
    out = grid.copy()
    out[np.nonzero(cand)[0][first]] = True
    return out


def bake(inp: dict) -> dict:
    """inp: texels (T,) TEXEL of the dirty tiles, tile after tile; tiles (nt, 6) of the dirty
    tiles; light0 (aw * ah, 4) the clean tiles' light or None; aw, ah; nodes, tris, mats, charts,
    params; settings (lightmap/bake fields); sun (3,), irr (3,); probes (np,) TEXEL of the dirty
    probes or None, volume_params; quant (optional): the 8-bit node records (quant_records), whose
    boxes the passes then traverse as the device's quantized form does; sky_sh (optional): the
    sky's L1 term (4 x 3)."""
    set_sky_sh(inp.get("sky_sh"))
    try:
        return _bake(inp)
    finally:
        set_sky_sh(None)


def _bake(inp: dict) -> dict:
    st = inp["settings"]
    q = inp.get("quant")

    def run(pass_, nodes, tris, mats, charts, P, items, key_base, light):
        if q is None:
            return run_fp32(pass_, nodes, tris, mats, charts, P, items, key_base, light)
        return run_quant(pass_, nodes, q, tris, mats, charts, P, items, key_base, light)
    tex = inp["texels"].copy()
    tiles = inp["tiles"]
    aw, ah = inp["aw"], inp["ah"]
    P = inp["params"]
    nodes, tris, mats, charts = inp["nodes"], inp["tris"], inp["mats"], inp["charts"]
    g = int(P["counts"][0][0])
    sun = inp["sun"].astype(np.float32)
    irr = np.asarray(inp["irr"], np.float32)
    offs = np.concatenate([[0], np.cumsum(tiles[:, 2] * tiles[:, 3])]).astype(np.int64)
    T = int(offs[-1])
    covered = tex["pos"][:, 3] > 0.5
    chart = np.where(covered, tex["normal"][:, 3].view(np.uint32), np.uint32(0xFFFFFFFF)).astype(np.uint32)
    normal = np.ascontiguousarray(tex["normal"][:, :3], np.float32)

    # the direct pass over the covered texels (bake.rs:149-151)
    cov = np.nonzero(covered)[0]
    res = run(0, nodes, tris, mats, charts, P, tex[cov], 0, None)
    valid = covered.copy()
    state = np.zeros(T, np.uint8)
    soft = np.zeros(T, np.float32)
    ao = np.ones(T, np.float32)
    hard = np.zeros(T, np.uint32)
    inside, sf, hb, stt, tone = decode_all(res, tex, cov, st)
    ok = cov[~inside]
    valid[cov[inside]] = False
    soft[ok], hard[ok], state[ok], ao[ok] = sf[~inside], hb[~inside], stt[~inside], tone[~inside]
    known = np.zeros(T, np.uint8)
    field = np.zeros(T, np.float32)
    direct = np.zeros((T, 3), np.float32)
    indirect = np.zeros((T, 3), np.float32)
    rings = st.padding + 4
    for t in range(len(tiles)):
        w, h = int(tiles[t, 2]), int(tiles[t, 3])
        sl = slice(offs[t], offs[t + 1])
        ch, nr, vl = chart[sl].copy(), normal[sl].copy(), valid[sl].astype(np.uint8)
        a = smooth(w, h, ch, nr, ao[sl], vl, 1)
        s = smooth(w, h, ch, nr, soft[sl], vl, 1)
        f = sun_field(w, h, ch, nr, hard[sl].copy(), vl, g, F(st.sdf_radius))
        c = np.maximum(nr[:, 0] * sun[0] + nr[:, 1] * sun[1] + nr[:, 2] * sun[2], F(0))
        d = np.stack([irr[0] * c * s, irr[1] * c * s, irr[2] * c * s], axis=1).astype(np.float32)
        for arr in (a, f):
            kk = vl.copy()
            fill(w, h, ch, arr, kk, rings)
        kn = vl.copy()
        d = np.ascontiguousarray(d)
        fill(w, h, ch, d, kn, rings)
        ao[sl], field[sl], direct[sl] = a, f, d
        stt = state[sl]
        stt[covered[sl] & (stt == VALIDITY_NONE) & (kn > 0)] = VALIDITY_FILLED
        state[sl] = stt
        known[sl] = np.where(vl > 0, 2, kn)

    def write_light(light):
        for t in range(len(tiles)):
            w, h, x0, y0 = (int(v) for v in tiles[t, 2:6])
            sl = slice(offs[t], offs[t + 1])
            kn = known[sl].reshape(h, w)
            val = np.zeros((h, w, 4), np.float32)
            dd = (direct[sl] + indirect[sl]).reshape(h, w, 3)
            val[..., :3] = np.where(kn[..., None] >= 1, dd, 0)
            val[..., 3] = np.where(kn >= 1, 1.0, 0.0)
            light[y0:y0 + h, x0:x0 + w] = val

    # the clean tiles' light from the cache (bake.rs:276-278 writes every tile; the tiles given
    # here are the dirty ones)
    l0 = inp.get("light0")
    if l0 is not None and len(l0):
        light = np.array(l0, np.float32).reshape(ah, aw, 4)
    else:
        light = np.zeros((ah, aw, 4), np.float32)
    write_light(light)

    # the gather set (bake.rs:282-299)
    ds = st.indirect_downscale
    gathered = np.zeros(T, bool)
    for t in range(len(tiles)):
        o, e = int(offs[t]), int(offs[t + 1])
        gathered[o:e] = gather_set_all(int(tiles[t, 2]), e - o, valid[o:e], chart[o:e], ds)
    gat = np.nonzero(gathered)[0]
    grings = st.padding + 2 * ds + 2
    for _ in range(st.gi_bounces):
        if len(gat) == 0:
            break
        gres = run(1, nodes, tris, mats, charts, P, tex[gat], 0, light.reshape(-1, 4))
        e = np.zeros((T, 3), np.float32)
        e[gat] = gres[:, :3]
        for t in range(len(tiles)):
            w, h = int(tiles[t, 2]), int(tiles[t, 3])
            sl = slice(offs[t], offs[t + 1])
            gm = gathered[sl].astype(np.uint8)
            ee = smooth(w, h, chart[sl].copy(), normal[sl].copy(), e[sl], gm,
                        st.gi_filter * ds)
            gk = gm.copy()
            fill(w, h, chart[sl].copy(), ee, gk, grings)
            indirect[sl] = ee
        write_light(light)

    probes = None
    if inp.get("probes") is not None:
        probes = run(2, nodes, tris, mats, charts, inp["volume_params"], inp["probes"], 0,
                     light.reshape(-1, 4)).reshape(-1, 16)
    return {"light": light, "known": known, "state": state, "field": field, "ao": ao,
            "direct": direct, "indirect": indirect, "probes": probes, "gathered": gat,
            "texels": tex, "valid": valid, "soft": soft}


def clean_light(tiles_all, tile_dirty, cache, aw, ah):
    """write_light (bake.rs:435-444) of the clean tiles from their cached TileLight: the atlas a
    rebake's first bounce starts from. cache: known, direct, indirect over every tile's texels,
    tiles in layout order."""
    light = np.zeros((ah, aw, 4), np.float32)
    o = 0
    for t in range(len(tiles_all)):
        w, h, x0, y0 = (int(v) for v in tiles_all[t, 2:6])
        sl = slice(o, o + w * h)
        o += w * h
        if tile_dirty[t]:
            continue
        kn = cache.known[sl].reshape(h, w)
        dd = (cache.direct[sl] + cache.indirect[sl]).reshape(h, w, 3)
        val = np.zeros((h, w, 4), np.float32)
        val[..., :3] = np.where(kn[..., None] >= 1, dd, 0)
        val[..., 3] = np.where(kn >= 1, 1.0, 0.0)
        light[y0:y0 + h, x0:x0 + w] = val
    return light
