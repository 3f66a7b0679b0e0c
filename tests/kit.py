"""Shared test inputs: a scene's BVH, sampled items, charts and a light atlas.

Items are surface points sampled on the scene's triangles (area-weighted, seeded), each with its
face normal, a 0.25 m texel frame and its mesh's chart, so every code path of the passes runs:
open texels, texels under or inside boxes (the inside-solid retries), chart lookups and the
fallback. Charts map world (x, z) to a 96 x 96 atlas.
"""
from __future__ import annotations

import numpy as np

from lightbaker_amd import _native as N
from lightbaker_amd.gpu import Bvh
from lightbaker_amd.scenes import Scene, make_params

ATLAS = 96


def scene_bvh(scene: Scene, with_charts: bool = True):
    pos, mesh, mat = scene.flat()
    chart = mesh.astype(np.uint32) if with_charts else np.full(len(mesh), 0xffffffff, np.uint32)
    bvh = Bvh.build(pos, mat, chart, scene.casts())
    return bvh, pos, mesh


def charts_for(scene: Scene, lo: float = -10.0, scale: float = 4.0) -> np.ndarray:
    rows = []
    for _ in scene.meshes:
        rows.append([[scale, 0.0, 0.0, -lo * scale + 0.5], [0.0, 0.0, scale, -lo * scale + 0.5]])
    return np.array(rows, dtype=np.float32).reshape(-1, 8)


def random_light(rng, n=ATLAS * ATLAS) -> np.ndarray:
    light = rng.uniform(0.0, 2.0, (n, 4)).astype(np.float32)
    light[:, 3] = (rng.uniform(size=n) < 0.8).astype(np.float32)
    return light


def sample_items(pos: np.ndarray, mesh: np.ndarray, n: int, rng, texel: float = 0.25):
    tri = pos.reshape(-1, 3, 3).astype(np.float64)
    e1 = tri[:, 1] - tri[:, 0]
    e2 = tri[:, 2] - tri[:, 0]
    cr = np.cross(e1, e2)
    area = np.linalg.norm(cr, axis=1)
    keep = area > 1e-9
    p = np.where(keep, area, 0.0)
    p /= p.sum()
    t = rng.choice(len(tri), size=n, p=p)
    u, v = rng.uniform(size=n), rng.uniform(size=n)
    flip = u + v > 1
    u[flip], v[flip] = 1 - u[flip], 1 - v[flip]
    pts = tri[t, 0] + e1[t] * u[:, None] + e2[t] * v[:, None]
    nrm = cr[t] / area[t, None]
    dx = e1[t] / np.linalg.norm(e1[t], axis=1)[:, None] * texel
    dy = np.cross(nrm, dx)
    items = np.zeros(n, dtype=N.TEXEL)
    items["pos"][:, :3] = pts
    items["pos"][:, 3] = 1.0
    items["normal"][:, :3] = nrm
    items["normal"][:, 3] = mesh[t].astype(np.uint32).view(np.float32)
    items["dx"][:, :3] = dx
    items["dy"][:, :3] = dy
    return items


def probe_items(scene: Scene, n: int | None = None) -> np.ndarray:
    """volume.rs:69-75: probes as texels (pos, normal.w = probe index)."""
    v = scene.volume
    nx = int(round((v.max[0] - v.min[0]) / v.step)) + 1
    ny = int(round((v.max[1] - v.min[1]) / v.step)) + 1
    nz = int(round((v.max[2] - v.min[2]) / v.step)) + 1
    k = np.arange(nx * ny * nz)
    items = np.zeros(len(k), dtype=N.TEXEL)
    items["pos"][:, 0] = v.min[0] + (k % nx) * v.step
    items["pos"][:, 1] = v.min[1] + (k // (nx * nz)) * v.step
    items["pos"][:, 2] = v.min[2] + ((k // nx) % nz) * v.step
    items["pos"][:, 3] = 1.0
    items["normal"][:, 1] = 1.0
    items["normal"][:, 3] = k.astype(np.uint32).view(np.float32)
    if n is not None:
        items = items[:: max(1, len(items) // n)][:n].copy()
    return items


def params_for(scene: Scene) -> np.ndarray:
    return make_params(scene, ATLAS, ATLAS)


def bits(a: np.ndarray) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float32).view(np.uint32)


def host_bake_inputs(scene: Scene) -> dict:
    """oracle/bake_ref.bake's inputs built on the host alone (no GPU): the package's layout,
    rasterizer and BVH (restatements of layout.rs, raster.rs and bvh.rs), every tile dirty."""
    from lightbaker_amd.bake import layout
    L = layout(scene)
    texels = np.concatenate([L.rasterize(t).reshape(-1) for t in range(len(L.tiles))])
    pos, _, mat = scene.flat()
    bvh = Bvh.build(pos, mat, L.tri_charts, scene.casts())
    mats = np.array([[*m.albedo, m.opacity] for m in scene.materials], np.float32)
    P = make_params(scene, L.width, L.height)
    return {"settings": scene.settings, "texels": texels, "tiles": L.tiles, "aw": L.width,
            "ah": L.height, "nodes": bvh.nodes, "tris": bvh.tris, "mats": mats,
            "charts": L.affines.reshape(-1, 8).astype(np.float32), "params": P,
            "sun": P["sun_dir"][0][:3].copy(), "irr": np.asarray(scene.sun_irradiance, np.float32),
            "probes": None, "volume_params": None, "sky_sh": scene.sky_sh}
