"""ORACLE — test infrastructure, not product code.

The adjoint's checker: the bounces of the reference's bake (crates/lightbaker/src/bake.rs:243-330,
with the gather of trace.wgsl:245-270) replayed in FP64 from the gather rays' records, carrying
the derivative by every parameter as extra channels (forward mode). The smoothing and fill
(filter.rs:20-94) and write_light (bake.rs:435-444) are linear in the values, with weights and
masks fixed by the geometry, so the tangents ride through them as more channels. This is
independent of the package's reverse-mode kernels: a wrong transpose, index or sign there gives
another gradient than the contraction of these tangents.

Parameters, in the order the package's gradient uses: albedo rgb per material, zenith, horizon,
ground, the sky's L1 coefficients c0..c3 (the term 0.282095 c0 + 0.488603 (c1, c2, c3) . d added
where a ray escapes, no counterpart in the reference), sun irradiance (rgb each).
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from . import gather_dirs, lib

PI = math.pi


def _bind():
    L = lib()
    if not hasattr(L, "_adjoint_bound"):
        VP, I, SZ, U32 = C.c_void_p, C.c_int, C.c_size_t, C.c_uint32
        L.orc_gather_records.argtypes = [VP, SZ, VP, VP, VP, VP, VP, SZ, U32, VP]
        L.orc_gather_records.restype = I
        L.orc_smooth64.argtypes = [I, I, VP, VP, VP, I, VP, I, VP]
        L.orc_smooth64.restype = None
        L.orc_fill64.argtypes = [I, I, VP, VP, I, VP, I]
        L.orc_fill64.restype = None
        L._adjoint_bound = True
    return L


def gather_records(nodes, tris, mats, charts, params, items) -> np.ndarray:
    """(items, rays, 2) u32: word, tag per gather ray (orc_gather_records)."""
    L = _bind()
    items = np.ascontiguousarray(items)
    rays = int(params["counts"][0][2])
    out = np.zeros((len(items), rays, 2), np.uint32)
    nodes, tris = np.ascontiguousarray(nodes), np.ascontiguousarray(tris)
    mats = np.ascontiguousarray(mats, np.float32)
    ch = None if charts is None or len(charts) == 0 else np.ascontiguousarray(charts, np.float32)
    params = np.ascontiguousarray(params)
    assert L.orc_gather_records(nodes.ctypes.data, len(nodes), tris.ctypes.data, mats.ctypes.data,
                                None if ch is None else ch.ctypes.data, params.ctypes.data,
                                items.ctypes.data, len(items), 0, out.ctypes.data) == 0
    return out


def smooth64(w, h, chart, normal, v, valid, r):
    L = _bind()
    v = np.ascontiguousarray(v, np.float64).reshape(w * h, -1)
    out = np.empty_like(v)
    L.orc_smooth64(w, h, chart.ctypes.data, normal.ctypes.data, v.ctypes.data, v.shape[1],
                   np.ascontiguousarray(valid, np.uint8).ctypes.data, r, out.ctypes.data)
    return out


def fill64(w, h, chart, v, valid, rings):
    """In place on v ((w * h, C) float64) and valid."""
    L = _bind()
    L.orc_fill64(w, h, chart.ctypes.data, v.ctypes.data, v.shape[1], valid.ctypes.data, rings)


def setup(inp: dict, out: dict) -> dict:
    """What the FP64 bounces need, from oracle/bake_ref.bake's inputs and outputs: the records of
    the gather items (the texels at their final sample points), the filled cos x sun share G
    (bake.rs:243-256 at unit irradiance), the masks, and where each texel sits in the atlas."""
    st = inp["settings"]
    tiles = inp["tiles"]
    aw, ah = int(inp["aw"]), int(inp["ah"])
    tex = inp["texels"]
    offs = np.concatenate([[0], np.cumsum(tiles[:, 2] * tiles[:, 3])]).astype(np.int64)
    T = int(offs[-1])
    covered = tex["pos"][:, 3] > 0.5
    chart = np.where(covered, tex["normal"][:, 3].view(np.uint32),
                     np.uint32(0xFFFFFFFF)).astype(np.uint32)
    normal = np.ascontiguousarray(tex["normal"][:, :3], np.float32)
    sun = np.asarray(inp["sun"], np.float32)
    valid = out["valid"].astype(np.uint8)
    gat = out["gathered"]
    gmask = np.zeros(T, np.uint8)
    gmask[gat] = 1
    G = np.zeros(T)
    atlas = np.zeros(T, np.int64)
    rings = st.padding + 4
    for t in range(len(tiles)):
        w, h, x0, y0 = (int(v) for v in tiles[t, 2:6])
        sl = slice(offs[t], offs[t + 1])
        ch, nr = chart[sl].copy(), normal[sl].copy()
        s1 = smooth64(w, h, ch, nr, out["soft"][sl].astype(np.float64), valid[sl], 1)[:, 0]
        c = np.maximum(nr[:, 0] * sun[0] + nr[:, 1] * sun[1] + nr[:, 2] * sun[2], np.float32(0))
        g = np.ascontiguousarray((c.astype(np.float64) * s1)[:, None])
        fill64(w, h, ch, g, valid[sl].copy(), rings)
        G[sl] = g[:, 0]
        k = np.arange(w * h)
        atlas[sl] = (y0 + k // w) * aw + x0 + k % w
    inv = np.full(aw * ah, -1, np.int64)
    inv[atlas] = np.arange(T)
    rec = gather_records(inp["nodes"], inp["tris"], inp["mats"], inp["charts"], inp["params"],
                         out["texels"][gat])
    l0 = inp.get("light0")
    light0 = (np.zeros((aw * ah, 4)) if l0 is None or len(l0) == 0
              else np.asarray(l0, np.float64).reshape(-1, 4))
    dirs = gather_dirs(inp["params"], out["texels"][gat]).astype(np.float64)
    return {"tiles": tiles, "offs": offs, "T": T, "chart": chart, "normal": normal, "dirs": dirs,
            "known": out["known"], "gat": gat, "gmask": gmask, "G": G, "atlas": atlas, "inv": inv,
            "rec": rec, "light0": light0, "aw": aw, "ah": ah, "settings": st,
            "materials": len(inp["mats"])}


SH0, SH1 = 0.282095, 0.488603
FB0, FB2 = PI * SH0, 2.0 * PI / 3.0 * SH1  # the L1 term's irradiance on an upward face


def forward(S: dict, albedo, zenith, horizon, ground, sun, sky_sh=None,
            tangents: bool = True) -> np.ndarray:
    """The dirty texels' light after the bounces, (T, 3, K): channel 0 the value, channel 1 + p
    its derivative by parameter p (K = 1 without tangents)."""
    st = S["settings"]
    M = S["materials"]
    albedo = np.asarray(albedo, np.float64).reshape(M, 3)
    z, hz, gr, sn = (np.asarray(v, np.float64) for v in (zenith, horizon, ground, sun))
    sh = (np.zeros((4, 3)) if sky_sh is None
          else np.asarray(sky_sh, np.float64).reshape(4, 3))
    P = 3 * M + 24
    K = 1 + P if tangents else 1
    iz, ih, ig, ish, isun = 1 + 3 * M, 4 + 3 * M, 7 + 3 * M, 10 + 3 * M, 22 + 3 * M
    T, known, G = S["T"], S["known"], S["G"]
    lit_t = known >= 1
    D = np.zeros((T, 3, K))
    for c in range(3):
        D[:, c, 0] = sn[c] * G
        if tangents:
            D[:, c, isun + c] = G
    Lt = np.where(lit_t[:, None, None], D, 0.0)
    rec, gat, inv, light0 = S["rec"], S["gat"], S["inv"], S["light0"]
    n, R = rec.shape[0], rec.shape[1]
    word, tag = rec[:, :, 0], rec[:, :, 1]
    kind, mat = tag >> 30, (tag & 0x3FFFFFFF).astype(np.int64)
    fb = (z + hz) * 0.5 * PI + FB0 * sh[0] + FB2 * sh[2]
    dirs = S["dirs"]
    ds = st.indirect_downscale
    rg, grings = st.gi_filter * ds, st.padding + 2 * ds + 2
    tiles, offs = S["tiles"], S["offs"]
    rows = np.arange(n)
    for _ in range(st.gi_bounces if n else 0):
        E = np.zeros((n, 3, K))
        for r in range(R):
            kd, wd, m = kind[:, r], word[:, r], mat[:, r]
            y = wd.view(np.float32).astype(np.float64)
            up = (kd == 3) & (y >= 0)
            dn = (kd == 3) & (y < 0)
            sk = kd == 3
            d = dirs[:, r]
            basis = [np.full(n, SH0), SH1 * d[:, 0], SH1 * d[:, 1], SH1 * d[:, 2]]
            for c in range(3):
                E[up, c, 0] += hz[c] * (1.0 - y[up]) + z[c] * y[up]
                E[dn, c, 0] += gr[c]
                E[sk, c, 0] += sum(basis[j][sk] * sh[j, c] for j in range(4))
                if tangents:
                    E[up, c, iz + c] += y[up]
                    E[up, c, ih + c] += 1.0 - y[up]
                    E[dn, c, ig + c] += 1.0
                    for j in range(4):
                        E[sk, c, ish + 3 * j + c] += basis[j][sk]
            hit = kd <= 1
            ev = np.zeros((n, 3, K))
            lit = np.zeros(n, bool)
            ti = np.where(kd == 0, inv[np.where(kd == 0, wd, 0).astype(np.int64)], -1)
            own = (kd == 0) & (ti >= 0)
            own &= lit_t[np.maximum(ti, 0)]
            ev[own] = Lt[ti[own]]
            lit |= own
            clean = (kd == 0) & (ti < 0)
            clean &= light0[np.where(clean, wd, 0).astype(np.int64), 3] > 0.5
            ev[clean, :, 0] = light0[wd[clean].astype(np.int64), :3]
            lit |= clean
            fbr = hit & ~lit
            for c in range(3):
                ev[fbr, c, 0] = fb[c]
                if tangents:
                    ev[fbr, c, iz + c] = PI / 2
                    ev[fbr, c, ih + c] = PI / 2
                    ev[fbr, c, ish + c] = FB0
                    ev[fbr, c, ish + 6 + c] = FB2
            a = albedo[m]
            E[hit] += (a[hit] / PI)[:, :, None] * ev[hit]
            if tangents:
                hr = rows[hit]
                for c in range(3):
                    E[hr, c, 1 + m[hit] * 3 + c] += ev[hit, c, 0] / PI
        E *= PI / R
        e = np.zeros((T, 3, K))
        e[gat] = E
        I = np.zeros((T, 3, K))
        for t in range(len(tiles)):
            w, h = int(tiles[t, 2]), int(tiles[t, 3])
            sl = slice(offs[t], offs[t + 1])
            ch, nr = S["chart"][sl].copy(), S["normal"][sl].copy()
            gm = S["gmask"][sl]
            v = smooth64(w, h, ch, nr, e[sl].reshape(w * h, -1), gm, rg)
            fill64(w, h, ch, v, gm.copy(), grings)
            I[sl] = v.reshape(-1, 3, K)
        Lt = np.where(lit_t[:, None, None], D + I, 0.0)
    return Lt


def atlas_light(S: dict, Lt: np.ndarray) -> np.ndarray:
    """The final atlas (aw * ah, 4) from the texels' values: the clean tiles' light, then every
    dirty texel's (bake.rs:435-444)."""
    light = S["light0"].copy()
    light[S["atlas"], :3] = Lt[:, :, 0]
    light[S["atlas"], 3] = (S["known"] >= 1).astype(np.float64)
    return light


def gradient(S: dict, Lt: np.ndarray, dlight: np.ndarray) -> np.ndarray:
    """The loss's gradient by every parameter, from its derivative by the atlas's rgb."""
    d = np.asarray(dlight, np.float64).reshape(-1, 4)[S["atlas"], :3]
    d = np.where((S["known"] >= 1)[:, None], d, 0.0)
    return np.einsum("tc,tcp->p", d, Lt[:, :, 1:])
