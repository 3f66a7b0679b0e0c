"""Scenes in the reference's model (crates/lightbaker/src/scene.rs: SceneHeader and per-mesh
non-indexed positions, Y up, metres) and seeded generators for the workloads this package is
measured on.

Generators (numpy PCG64 streams, so a seed always gives the same scene):
  small_scene(seed)      20x20 m ground quad + 50 yawed boxes (~600 tris), grey albedo 0.5,
                         sun 45 deg elevation, sky 0.5 everywhere, 4 texels/m, 64 AO / 64 GI rays,
                         1 bounce, volume 1 m up to 8 m
  map_scene(seed)        100x100 m: 256x256 fBm heightfield (4 octaves, 6 m) + 400 boxes and
                         32-sided cylinders, 16 materials, 8 texels/m, 256 GI rays, 3 bounces,
                         volume 2 m up to 20 m, 16 tiles of 25 m
  reference_test_scene(bx)  the scene of crates/lightbaker/tests/bake.rs:29-72
"""
from __future__ import annotations

import copy
import math
from dataclasses import dataclass, field

import numpy as np

from . import _native as N


@dataclass
class Material:
    name: str
    albedo: tuple
    receives: bool = True
    casts: bool = True
    opacity: float = 1.0


@dataclass
class Mesh:
    name: str
    material: int
    positions: np.ndarray  # (triangles * 3, 3) float32


@dataclass
class Settings:
    """lightmap (scene.rs:85-97) and bake (scene.rs:99-118) settings."""
    texels_per_meter: float = 4.0
    tile_size: float = 16.0
    chart_angle_deg: float = 40.0
    padding: int = 2
    atlas_width: int = 1024
    indirect_downscale: int = 2
    sun_samples: int = 16
    ao_rays: int = 64
    ao_reach: float = 0.8
    ao_strength: float = 1.0
    ao_floor: float = 0.25
    ao_contrast: float = 1.0
    sdf_radius: float = 2.0
    gi_rays: int = 256
    gi_bounces: int = 2
    gi_reach: float = 12.0
    inside_share: float = 0.25
    gi_filter: int = 1


@dataclass
class Volume:
    min: tuple
    max: tuple
    step: float
    rays: int


@dataclass
class Scene:
    sun_direction: tuple
    sun_radius_deg: float
    sun_irradiance: tuple
    zenith: tuple
    horizon: tuple
    ground: tuple
    materials: list
    meshes: list
    settings: Settings = field(default_factory=Settings)
    volume: Volume | None = None
    # per-tile texels per metre, {(key x, key z): density}; tiles not listed take
    # settings.texels_per_meter (no counterpart in the reference's one density, scene.rs:84)
    tile_density: dict | None = None
    # the sky's L1 spherical-harmonic term, (4, 3): c0, c1, c2, c3 rgb, added to the gradient sky
    # where a ray escapes (no counterpart in the reference; None: the reference's sky)
    sky_sh: np.ndarray | None = None

    def triangle_count(self) -> int:
        return sum(len(m.positions) // 3 for m in self.meshes)

    def flat(self):
        """All triangles (n, 9), each one's mesh and material."""
        pos = np.concatenate([m.positions.reshape(-1, 9) for m in self.meshes]).astype(np.float32)
        mesh = np.concatenate([np.full(len(m.positions) // 3, i, np.uint32)
                               for i, m in enumerate(self.meshes)])
        mat = np.array([self.meshes[i].material for i in mesh], dtype=np.uint32)
        return pos, mesh, mat

    def materials_array(self) -> np.ndarray:
        """bake.rs:103: albedo rgb, opacity."""
        return np.array([[*m.albedo, m.opacity] for m in self.materials], dtype=np.float32)

    def casts(self) -> np.ndarray:
        return np.array([m.casts for m in self.materials], dtype=np.uint8)

    def copy(self) -> "Scene":
        return copy.deepcopy(self)


# the sky's L1 term's irradiance on an upward face per unit c0 and c2, FP32
SH_FALLBACK_C0 = np.float32(math.pi * 0.282095)
SH_FALLBACK_C2 = np.float32(2.0 * math.pi / 3.0 * 0.488603)


def make_params(scene: Scene, atlas_w: int, atlas_h: int) -> np.ndarray:
    """Params as bake.rs:105-120 assembles them (LIFT 0.01, INSIDE_REACH 0.3, seed 0)."""
    s = scene.settings
    sun = np.asarray(scene.sun_direction, dtype=np.float32)
    sun = sun / np.float32(math.sqrt(float(np.dot(sun, sun))))
    g = int(math.isqrt(s.sun_samples))
    z = np.asarray(scene.zenith, np.float32)
    h = np.asarray(scene.horizon, np.float32)
    sky_avg = (z + h) * np.float32(0.5) * np.float32(math.pi)
    if scene.sky_sh is not None:
        # the sky's L1 term's irradiance on an upward face: pi 0.282095 c0 + 2 pi / 3 0.488603 c2
        sh = np.asarray(scene.sky_sh, np.float32).reshape(4, 3)
        sky_avg = sky_avg + (SH_FALLBACK_C0 * sh[0] + SH_FALLBACK_C2 * sh[2])
    p = np.zeros(1, dtype=N.PARAMS)
    p["sun_dir"] = [sun[0], sun[1], sun[2], np.float32(math.tan(math.radians(scene.sun_radius_deg)))]
    p["sun_irr"] = [*scene.sun_irradiance, 0.0]
    p["zenith"] = [*scene.zenith, 0.0]
    p["horizon"] = [*scene.horizon, 0.0]
    p["ground"] = [*scene.ground, min(np.float32(0.3), np.float32(s.ao_reach))]
    p["fallback"] = [*sky_avg, 0.0]
    p["counts"] = [g, s.ao_rays, s.gi_rays, 0]
    p["reach"] = [s.ao_reach, s.gi_reach, 0.01, 0.0]
    p["atlas"] = [atlas_w, atlas_h, 0, np.float32(s.inside_share).view(np.uint32)]
    return p


# ---- geometry helpers (tests/bake.rs:10-26) -------------------------------------------------

def quad(a, b, c, d) -> list:
    return [a, b, c, a, c, d]


def cube(mn, mx) -> np.ndarray:
    """A closed box, faces outward, counter-clockwise seen from outside (tests/bake.rs:14-26)."""
    def p(x, y, z):
        return (mx[0] if x > 0 else mn[0], mx[1] if y > 0 else mn[1], mx[2] if z > 0 else mn[2])
    faces = [
        quad(p(0, 1, 0), p(0, 1, 1), p(1, 1, 1), p(1, 1, 0)),
        quad(p(0, 0, 0), p(1, 0, 0), p(1, 0, 1), p(0, 0, 1)),
        quad(p(0, 0, 1), p(1, 0, 1), p(1, 1, 1), p(0, 1, 1)),
        quad(p(1, 0, 0), p(0, 0, 0), p(0, 1, 0), p(1, 1, 0)),
        quad(p(1, 0, 1), p(1, 0, 0), p(1, 1, 0), p(1, 1, 1)),
        quad(p(0, 0, 0), p(0, 0, 1), p(0, 1, 1), p(0, 1, 0)),
    ]
    return np.array([v for f in faces for v in f], dtype=np.float32)


def yawed_box(cx, cz, y0, sx, sy, sz, yaw) -> np.ndarray:
    """A box of size (sx, sy, sz) standing at (cx, y0, cz), turned by `yaw` about Y."""
    b = cube((-sx / 2, 0.0, -sz / 2), (sx / 2, sy, sz / 2)).astype(np.float64)
    c, s = math.cos(yaw), math.sin(yaw)
    x = b[:, 0] * c + b[:, 2] * s
    z = -b[:, 0] * s + b[:, 2] * c
    return np.stack([x + cx, b[:, 1] + y0, z + cz], axis=1).astype(np.float32)


def cylinder(cx, cz, y0, r, hgt, sides=32, rings=4) -> np.ndarray:
    """A capped cylinder, outward-facing, `sides` facets round and `rings` bands high."""
    tris = []
    ang = [2 * math.pi * k / sides for k in range(sides)]
    ring = [(cx + r * math.cos(a), cz + r * math.sin(a)) for a in ang]
    for j in range(rings):
        ya, yb = y0 + hgt * j / rings, y0 + hgt * (j + 1) / rings
        for k in range(sides):
            (x0, z0), (x1, z1) = ring[k], ring[(k + 1) % sides]
            # seen from outside counter-clockwise: bottom k, top k, top k+1, bottom k+1
            tris += quad((x0, ya, z0), (x0, yb, z0), (x1, yb, z1), (x1, ya, z1))
    top, bot = y0 + hgt, y0
    for k in range(sides):
        (x0, z0), (x1, z1) = ring[k], ring[(k + 1) % sides]
        tris += [(cx, top, cz), (x1, top, z1), (x0, top, z0)]
        tris += [(cx, bot, cz), (x0, bot, z0), (x1, bot, z1)]
    return np.array(tris, dtype=np.float32)


# ---- generators --------------------------------------------------------------------------------

def small_scene(seed: int = 1) -> Scene:
    """20x20 m ground quad plus 50 random yawed boxes (edges U[0.5, 4] m), grey albedo 0.5."""
    rng = np.random.default_rng(seed)
    ground = np.array(quad((-10, 0, -10), (-10, 0, 10), (10, 0, 10), (10, 0, -10)), np.float32)
    boxes = []
    for _ in range(50):
        sx, sy, sz = rng.uniform(0.5, 4.0, 3)
        cx, cz = rng.uniform(-9.0, 9.0, 2)
        boxes.append(yawed_box(cx, cz, 0.0, sx, sy, sz, rng.uniform(0, 2 * math.pi)))
    az = rng.uniform(0, 2 * math.pi)
    el = math.radians(45.0)
    sun = (math.cos(el) * math.cos(az), math.sin(el), math.cos(el) * math.sin(az))
    st = Settings(texels_per_meter=4.0, tile_size=16.0, atlas_width=1024, sun_samples=16,
                  ao_rays=64, gi_rays=64, gi_bounces=1)
    return Scene(
        sun_direction=sun, sun_radius_deg=0.5, sun_irradiance=(3.0, 3.0, 3.0),
        zenith=(0.5, 0.5, 0.5), horizon=(0.5, 0.5, 0.5), ground=(0.5, 0.5, 0.5),
        materials=[Material("grey", (0.5, 0.5, 0.5))],
        meshes=[Mesh("ground", 0, ground), Mesh("boxes", 0, np.concatenate(boxes))],
        settings=st, volume=Volume((-10.0, 0.0, -10.0), (10.0, 8.0, 10.0), 1.0, 64))


def fit_scene(seed: int = 5) -> Scene:
    """small_scene's layout (a 20 x 20 m ground and 50 yawed boxes) with 16 materials (albedo
    U[0.2, 0.8] per channel; the ground is material 0, each box one of the 16; material 15
    receives no light map, so rays that meet it take the fallback irradiance), a sky graded from
    horizon to zenith over a darker ground plus an L1 spherical-harmonic term (brighter above and
    towards +x), and 3 bounces: the scene light is fitted on."""
    rng = np.random.default_rng(seed)
    ground = np.array(quad((-10, 0, -10), (-10, 0, 10), (10, 0, 10), (10, 0, -10)), np.float32)
    meshes = [Mesh("ground", 0, ground)]
    for k in range(50):
        sx, sy, sz = rng.uniform(0.5, 4.0, 3)
        cx, cz = rng.uniform(-9.0, 9.0, 2)
        geo = yawed_box(cx, cz, 0.0, sx, sy, sz, rng.uniform(0, 2 * math.pi))
        meshes.append(Mesh(f"box{k}", int(rng.integers(0, 16)), geo))
    mats = [Material(f"m{i}", tuple(float(v) for v in rng.uniform(0.2, 0.8, 3)),
                     receives=i != 15)
            for i in range(16)]
    az = rng.uniform(0, 2 * math.pi)
    el = math.radians(45.0)
    sun = (math.cos(el) * math.cos(az), math.sin(el), math.cos(el) * math.sin(az))
    st = Settings(texels_per_meter=4.0, tile_size=16.0, atlas_width=1024, sun_samples=16,
                  ao_rays=64, gi_rays=64, gi_bounces=3)
    return Scene(
        sun_direction=sun, sun_radius_deg=0.5, sun_irradiance=(3.0, 3.0, 3.0),
        zenith=(0.35, 0.45, 0.6), horizon=(0.5, 0.5, 0.5), ground=(0.25, 0.22, 0.2),
        materials=mats, meshes=meshes, settings=st, volume=None,
        sky_sh=np.array([[0.40, 0.42, 0.45], [0.05, 0.04, 0.03], [0.12, 0.10, 0.08],
                         [-0.03, -0.02, -0.01]], np.float32))


def rebake_scene(seed: int = 3, size: float = 200.0, densities: tuple | None = (4, 16)) -> Scene:
    """A size x size m map from small_scene's generator at its density (yawed boxes, edges
    U[0.5, 4] m, 50 per 400 m^2 on average, on a ground of 5 m quads, grey albedo 0.5,
    constant sky 0.5, sun at 45 degrees), in 25 m tiles, each laid out at its own texel density
    drawn uniformly from the integers in `densities` (texels per metre; None: 4 everywhere), so
    the tiles' texel counts differ up to (16 / 4)^2 = 16-fold through chart density."""
    rng = np.random.default_rng(seed)
    cells = int(size // 5)
    ground = []
    for i in range(cells):
        for j in range(cells):
            x, z = i * 5.0, j * 5.0
            ground += quad((x, 0, z), (x, 0, z + 5), (x + 5, 0, z + 5), (x + 5, 0, z))
    tiles = int(math.ceil(size / 25.0))
    boxes = []
    for tx in range(tiles):
        for tz in range(tiles):
            # 50 boxes per 400 m^2 on average, 3 m from the tile's edges: every face of a box
            # falls in its tile, so the map has exactly tiles^2 tiles
            for _ in range(int(rng.poisson(625.0 / 400.0 * 50))):
                sx, sy, sz = rng.uniform(0.5, 4.0, 3)
                cx = tx * 25.0 + rng.uniform(3.0, 22.0)
                cz = tz * 25.0 + rng.uniform(3.0, 22.0)
                boxes.append(yawed_box(cx, cz, 0.0, sx, sy, sz, rng.uniform(0, 2 * math.pi)))
    az = rng.uniform(0, 2 * math.pi)
    el = math.radians(45.0)
    sun = (math.cos(el) * math.cos(az), math.sin(el), math.cos(el) * math.sin(az))
    st = Settings(texels_per_meter=4.0, tile_size=25.0, atlas_width=4096, sun_samples=16,
                  ao_rays=64, gi_rays=64, gi_bounces=1)
    dens = None
    if densities is not None:
        drng = np.random.default_rng(seed + 1000)
        lo, hi = densities
        dens = {(tx, tz): float(drng.integers(lo, hi + 1))
                for tx in range(tiles) for tz in range(tiles)}
    return Scene(
        sun_direction=sun, sun_radius_deg=0.5, sun_irradiance=(3.0, 3.0, 3.0),
        zenith=(0.5, 0.5, 0.5), horizon=(0.5, 0.5, 0.5), ground=(0.5, 0.5, 0.5),
        materials=[Material("grey", (0.5, 0.5, 0.5))],
        meshes=[Mesh("ground", 0, np.array(ground, np.float32)),
                Mesh("boxes", 0, np.concatenate(boxes))],
        settings=st, volume=Volume((0.0, 0.0, 0.0), (size, 8.0, size), 1.0, 64),
        tile_density=dens)


def building(cx, cz, radius, height, sides, yaw) -> np.ndarray:
    """A prism over a regular `sides`-gon (walls, a fan roof, no floor): 3 * sides - 2
    triangles, counter-clockwise seen from outside."""
    a = yaw + np.arange(sides) * (2 * math.pi / sides)
    ring = np.stack([cx + radius * np.cos(a), cz + radius * np.sin(a)], axis=1)
    tris = []
    for k in range(sides):
        (x0, z0), (x1, z1) = ring[k], ring[(k + 1) % sides]
        tris += [(x0, 0.0, z0), (x0, height, z0), (x1, height, z1)]
        tris += [(x0, 0.0, z0), (x1, height, z1), (x1, 0.0, z1)]
    for k in range(1, sides - 1):
        tris += [(ring[0][0], height, ring[0][1]), (ring[k + 1][0], height, ring[k + 1][1]),
                 (ring[k][0], height, ring[k][1])]
    return np.array(tris, dtype=np.float32)


def city_scene(seed: int = 4, buildings: int = 2000, size: float = 400.0,
               texels_per_meter: float = 8.0) -> Scene:
    """A size x size m plane (10 m quads) with `buildings` prisms of 40 to 1 200 triangles
    (14 to 400 sides), footprint radius U[3, 9] m, heights lognormal (mu 2.5, sigma 0.6) m,
    one of 16 materials each: 1.24 M triangles at the defaults, flattened (the reference has no
    instancing). 3 bounces, 8 texels/m."""
    rng = np.random.default_rng(seed)
    cells = int(size // 10)
    ground = []
    for i in range(cells):
        for j in range(cells):
            x, z = i * 10.0, j * 10.0
            ground += quad((x, 0, z), (x, 0, z + 10), (x + 10, 0, z + 10), (x + 10, 0, z))
    meshes = [Mesh("ground", 0, np.array(ground, np.float32))]
    for k in range(buildings):
        sides = int(rng.integers(14, 401))
        r = rng.uniform(3.0, 9.0)
        h = float(np.exp(rng.normal(2.5, 0.6)))
        cx, cz = rng.uniform(r, size - r, 2)
        meshes.append(Mesh(f"b{k}", int(rng.integers(0, 16)),
                           building(cx, cz, r, h, sides, rng.uniform(0, 2 * math.pi))))
    mats = [Material(f"m{i}", tuple(float(v) for v in rng.uniform(0.15, 0.85, 3)))
            for i in range(16)]
    el = math.radians(rng.uniform(20.0, 70.0))
    az = rng.uniform(0, 2 * math.pi)
    sun = (math.cos(el) * math.cos(az), math.sin(el), math.cos(el) * math.sin(az))
    st = Settings(texels_per_meter=texels_per_meter, tile_size=25.0, atlas_width=16384,
                  sun_samples=16, ao_rays=64, gi_rays=64, gi_bounces=3)
    return Scene(
        sun_direction=sun, sun_radius_deg=0.5, sun_irradiance=(3.0, 3.0, 3.0),
        zenith=(0.2, 0.25, 0.35), horizon=(0.3, 0.3, 0.32), ground=(0.1, 0.1, 0.1),
        materials=mats, meshes=meshes, settings=st, volume=None)


def fbm(n: int, octaves: int, rng) -> np.ndarray:
    """(n+1)^2 heights in [-1, 1]: octaves of bilinearly upsampled value noise, halving amplitude."""
    h = np.zeros((n + 1, n + 1))
    amp, total = 1.0, 0.0
    for o in range(octaves):
        cells = 4 * 2 ** o
        g = rng.uniform(-1.0, 1.0, (cells + 1, cells + 1))
        t = np.linspace(0.0, cells, n + 1)
        i = np.minimum(t.astype(int), cells - 1)
        f = t - i
        f = f * f * (3 - 2 * f)
        a = g[i][:, i] * (1 - f)[None, :] + g[i][:, i + 1] * f[None, :]
        b = g[i + 1][:, i] * (1 - f)[None, :] + g[i + 1][:, i + 1] * f[None, :]
        h += amp * (a * (1 - f)[:, None] + b * f[:, None])
        total += amp
        amp *= 0.5
    return h / total


def map_scene(seed: int = 2) -> Scene:
    """The 100x100 m map: fBm terrain, 400 boxes and cylinders, 16 materials, 3 bounces."""
    rng = np.random.default_rng(seed)
    n, size, amp = 256, 100.0, 6.0
    hgt = fbm(n, 4, rng) * amp
    xs = np.linspace(0.0, size, n + 1)
    X, Z = np.meshgrid(xs, xs, indexing="ij")
    P = np.stack([X, hgt, Z], axis=-1)
    a, b = P[:-1, :-1], P[1:, :-1]
    c, d = P[1:, 1:], P[:-1, 1:]
    # each cell: (a, d, c) and (a, c, b): counter-clockwise seen from above (normal +y)
    t1 = np.stack([a, d, c], axis=2).reshape(-1, 3, 3)
    t2 = np.stack([a, c, b], axis=2).reshape(-1, 3, 3)
    terrain = np.stack([t1, t2], axis=1).reshape(-1, 3).astype(np.float32)

    def height_at(x, z):
        i = min(int(x / size * n), n - 1)
        j = min(int(z / size * n), n - 1)
        return float(hgt[i:i + 2, j:j + 2].min())

    meshes = [Mesh("terrain", 0, terrain)]
    for k in range(400):
        x, z = rng.uniform(3.0, size - 3.0, 2)
        mat = int(rng.integers(1, 16))
        y0 = height_at(x, z) - 0.2
        if rng.uniform() < 0.5:
            sx, sy, sz = rng.uniform(0.8, 5.0, 3)
            geo = yawed_box(x, z, y0, sx, sy + 0.2, sz, rng.uniform(0, 2 * math.pi))
        else:
            r, hh = rng.uniform(0.3, 1.5), rng.uniform(1.0, 6.0)
            geo = cylinder(x, z, y0, r, hh + 0.2)
        meshes.append(Mesh(f"object{k}", mat, geo))
    mats = [Material(f"m{i}", tuple(float(v) for v in rng.uniform(0.15, 0.85, 3)))
            for i in range(16)]
    el = math.radians(rng.uniform(20.0, 70.0))
    az = rng.uniform(0, 2 * math.pi)
    sun = (math.cos(el) * math.cos(az), math.sin(el), math.cos(el) * math.sin(az))
    st = Settings(texels_per_meter=8.0, tile_size=25.0, atlas_width=4096, sun_samples=16,
                  ao_rays=64, gi_rays=256, gi_bounces=3)
    return Scene(
        sun_direction=sun, sun_radius_deg=0.5, sun_irradiance=(3.0, 3.0, 3.0),
        zenith=(0.2, 0.25, 0.35), horizon=(0.3, 0.3, 0.32), ground=(0.1, 0.1, 0.1),
        materials=mats, meshes=meshes, settings=st,
        volume=Volume((0.0, 0.0, 0.0), (100.0, 20.0, 100.0), 2.0, 256))


def reference_test_scene(bx: float) -> Scene:
    """crates/lightbaker/tests/bake.rs:29-72: ground 64x64 m of 4 m quads, a grey box at
    (bx, 0, 10), a red wall along x = 40, a strip narrower than a texel."""
    ground = []
    for i in range(16):
        for j in range(16):
            x, z = i * 4.0, j * 4.0
            ground += quad((x, 0, z), (x, 0, z + 4), (x + 4, 0, z + 4), (x + 4, 0, z))
    boxes = cube((bx, 0.0, 10.0), (bx + 2.0, 2.0, 12.0))
    wall = cube((40.0, 0.0, 30.0), (40.5, 4.0, 50.0))
    strip = np.array(quad((30, 0.2, 20), (30, 2.6, 20), (30, 2.6, 20.05), (30, 0.2, 20.05)),
                     np.float32)
    st = Settings(texels_per_meter=4.0, tile_size=16.0, chart_angle_deg=40.0, padding=2,
                  atlas_width=1024, indirect_downscale=2, sun_samples=16, ao_rays=64,
                  ao_reach=0.8, ao_strength=1.0, ao_floor=0.25, ao_contrast=1.0, sdf_radius=2.0,
                  gi_rays=256, gi_bounces=2, gi_reach=12.0, inside_share=0.25, gi_filter=1)
    return Scene(
        sun_direction=(-1.0, 0.8, 0.0), sun_radius_deg=0.5, sun_irradiance=(3.0, 3.0, 3.0),
        zenith=(0.2, 0.25, 0.35), horizon=(0.3, 0.3, 0.32), ground=(0.1, 0.1, 0.1),
        materials=[Material("ground", (0.5, 0.5, 0.5)), Material("grey", (0.5, 0.5, 0.5)),
                   Material("red", (0.9, 0.05, 0.05))],
        meshes=[Mesh("ground", 0, np.array(ground, np.float32)), Mesh("box", 1, boxes),
                Mesh("wall", 2, wall), Mesh("strip", 1, strip)],
        settings=st, volume=Volume((0.0, 0.5, 0.0), (64.0, 4.5, 64.0), 2.0, 128))
