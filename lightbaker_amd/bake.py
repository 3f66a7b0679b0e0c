"""bake() (crates/lightbaker/src/bake.rs:94) over the C-ABI: the scene in, the lightmap atlas,
the indirect atlas, the validity, the irradiance volume and the layout's chart maps out."""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _native as N
from .gpu import Gpu
from .scenes import Scene


@dataclass
class Bake:
    """bake.rs:36-47 `Bake` plus the layout facts the tests and tools need."""
    info: dict
    direct: np.ndarray      # (ah, aw, 2): sun field, occlusion
    indirect: np.ndarray    # (ih, iw, 3)
    validity: np.ndarray    # (ah, aw) u8
    volume: np.ndarray | None  # (ny, nz, nx, 4, 4)
    charts: np.ndarray      # (charts, 2, 4) affine rows
    tri_charts: np.ndarray  # (triangles,) u32
    light: np.ndarray       # (ah, aw, 4)
    chart_tiles: np.ndarray
    tiles: np.ndarray       # (tiles, 6): key x, key z, w, h, atlas x, atlas y
    cache: "TileCache"      # what the reference's cache would hold after this bake
    # the atlases as they ship (output.rs:99-110), encoded on the device
    direct_bc5: np.ndarray | None = None        # (aw * ah,) bytes of BC5 blocks
    indirect_rgb9e5: np.ndarray | None = None   # (ih, iw) u32
    volume_bytes: np.ndarray | None = None      # (probes, 16) u8; scale in info["volume_scale"]


@dataclass
class TileCache:
    """bake.rs:61-69 `TileLight` of every tile, texels tile after tile (layout order, row-major
    in a tile), and volume.rs:81-84's probes as traced (before the inside fill)."""
    known: np.ndarray       # u8: 2 own light, 1 filled, 0 none
    state: np.ndarray       # u8 validity
    field: np.ndarray       # f32
    ao: np.ndarray          # f32
    direct: np.ndarray      # (n, 3) f32
    indirect: np.ndarray    # (n, 3) f32
    probes: np.ndarray | None  # (probes, 16) f32


def _cache_desc(cache: TileCache | None, tile_dirty, probe_dirty):
    """The lb_bake_cache of a rebake, and the arrays it borrows (None: bake everything)."""
    if tile_dirty is None and probe_dirty is None:
        return None, []
    keep = []

    def ptr(a, dtype):
        a = np.ascontiguousarray(a, dtype=dtype)
        keep.append(a)
        return a.ctypes.data

    c = N.BakeCache()
    if tile_dirty is not None:
        c.tile_dirty = ptr(np.asarray(tile_dirty).astype(bool), np.uint8)
    if probe_dirty is not None:
        c.probe_dirty = ptr(np.asarray(probe_dirty).astype(bool), np.uint8)
    if cache is not None:
        c.known = ptr(cache.known, np.uint8)
        c.state = ptr(cache.state, np.uint8)
        c.field = ptr(cache.field, np.float32)
        c.ao = ptr(cache.ao, np.float32)
        c.direct = ptr(cache.direct, np.float32)
        c.indirect = ptr(cache.indirect, np.float32)
        c.n_texels = len(cache.known)
        if cache.probes is not None:
            c.probes = ptr(cache.probes, np.float32)
            c.n_probes = len(cache.probes)
    keep.append(c)
    return c, keep


def _tile_cache(lib, h) -> "TileCache":
    pr = _array(lib, h, N.BAKE_PROBES, np.float32)
    return TileCache(
        known=_array(lib, h, N.BAKE_TILE_KNOWN, np.uint8),
        state=_array(lib, h, N.BAKE_TILE_STATE, np.uint8),
        field=_array(lib, h, N.BAKE_TILE_FIELD, np.float32),
        ao=_array(lib, h, N.BAKE_TILE_AO, np.float32),
        direct=_array(lib, h, N.BAKE_TILE_DIRECT, np.float32).reshape(-1, 3),
        indirect=_array(lib, h, N.BAKE_TILE_INDIRECT, np.float32).reshape(-1, 3),
        probes=pr.reshape(-1, 16) if len(pr) else None,
    )


def scene_desc(scene: Scene):
    """The C descriptor of a scene, and the arrays it borrows (keep them alive while it is used)."""
    keep = []
    mats = (N.MaterialDesc * len(scene.materials))()
    for i, m in enumerate(scene.materials):
        mats[i].albedo[:] = list(m.albedo)
        mats[i].opacity = m.opacity
        mats[i].receives = int(m.receives)
        mats[i].casts = int(m.casts)
    meshes = (N.MeshDesc * len(scene.meshes))()
    for i, m in enumerate(scene.meshes):
        p = np.ascontiguousarray(m.positions, dtype=np.float32).reshape(-1)
        keep.append(p)
        meshes[i].material = m.material
        meshes[i].triangles = len(p) // 9
        meshes[i].positions = p.ctypes.data_as(C.POINTER(C.c_float))
    st = scene.settings
    d = N.SceneDesc()
    d.sun_direction[:] = list(scene.sun_direction)
    d.sun_radius_deg = scene.sun_radius_deg
    d.sun_irradiance[:] = list(scene.sun_irradiance)
    d.zenith[:] = list(scene.zenith)
    d.horizon[:] = list(scene.horizon)
    d.ground[:] = list(scene.ground)
    d.materials = mats
    d.n_materials = len(scene.materials)
    d.meshes = meshes
    d.n_meshes = len(scene.meshes)
    for k in ("texels_per_meter", "tile_size", "chart_angle_deg", "padding", "atlas_width",
              "indirect_downscale", "sun_samples", "ao_rays", "ao_reach", "ao_strength",
              "ao_floor", "ao_contrast", "sdf_radius", "gi_rays", "gi_bounces", "gi_reach",
              "inside_share", "gi_filter"):
        setattr(d, k, getattr(st, k))
    if scene.volume is not None:
        d.has_volume = 1
        d.volume_min[:] = list(scene.volume.min)
        d.volume_max[:] = list(scene.volume.max)
        d.volume_step = scene.volume.step
        d.volume_rays = scene.volume.rays
    if scene.sky_sh is not None:
        d.sky_sh[:] = [float(v) for v in np.asarray(scene.sky_sh, np.float32).reshape(12)]
    if scene.tile_density:
        dens = (N.TileDensity * len(scene.tile_density))()
        for i, ((x, z), v) in enumerate(sorted(scene.tile_density.items())):
            dens[i].x, dens[i].z, dens[i].texels_per_meter = int(x), int(z), float(v)
        d.tile_density = dens
        d.n_tile_density = len(scene.tile_density)
        keep.append(dens)
    keep += [mats, meshes]
    return d, keep


def _array(lib, h, which, dtype):
    p = C.c_void_p()
    n = lib.lb_bake_array(h, which, C.byref(p))
    if n == 0:
        return np.zeros(0, dtype=dtype)
    # a ctypes array over the bake's bytes (string_at stops at 2 GiB)
    buf = (C.c_ubyte * n).from_address(p.value)
    return np.frombuffer(buf, dtype=np.uint8).view(dtype).copy()


def bake(scene: Scene, gpu: Gpu, form: int = N.FORM_PLAIN, graph: bool = False,
         threads: int = 0, cache: TileCache | None = None, tile_dirty=None,
         probe_dirty=None, replay: bool = False) -> Bake:
    """A bake; with `tile_dirty` / `probe_dirty` (a flag per layout tile / probe) a rebake that
    traces only the dirty ones and takes the others from `cache` (bake.rs:123-136). With
    `replay`, the bounces after the first replay the first one's ray records."""
    lib = N.lib()
    d, keep = scene_desc(scene)
    cd, ckeep = _cache_desc(cache, tile_dirty, probe_dirty)
    opt = N.BakeOptions(int(form), int(graph), int(threads),
                        C.pointer(cd) if cd is not None else None, int(replay))
    h = C.c_void_p()
    N.check(lib.lb_bake_run(gpu.handle, C.byref(d), C.byref(opt), C.byref(h)))
    try:
        inf = N.BakeInfo()
        N.check(lib.lb_bake_info_get(h, C.byref(inf)))
        info = {}
        for k, _ in N.BakeInfo._fields_:
            v = getattr(inf, k)
            info[k] = list(v) if hasattr(v, "__len__") else v
        aw, ah = inf.atlas_width, inf.atlas_height
        iw, ih = inf.indirect_width, inf.indirect_height
        vol = _array(lib, h, N.BAKE_VOLUME, np.float32)
        if len(vol):
            nx, ny, nz = info["volume_size"]
            vol = vol.reshape(ny, nz, nx, 4, 4)
        else:
            vol = None
        return Bake(
            info=info,
            direct=_array(lib, h, N.BAKE_DIRECT, np.float32).reshape(ah, aw, 2),
            indirect=_array(lib, h, N.BAKE_INDIRECT, np.float32).reshape(ih, iw, 3),
            validity=_array(lib, h, N.BAKE_VALIDITY, np.uint8).reshape(ah, aw),
            volume=vol,
            charts=_array(lib, h, N.BAKE_CHARTS, np.float32).reshape(-1, 2, 4),
            tri_charts=_array(lib, h, N.BAKE_TRI_CHARTS, np.uint32),
            light=_array(lib, h, N.BAKE_LIGHT, np.float32).reshape(ah, aw, 4),
            chart_tiles=_array(lib, h, N.BAKE_CHART_TILES, np.uint32),
            tiles=_array(lib, h, N.BAKE_TILES, np.int32).reshape(-1, 6),
            cache=_tile_cache(lib, h),
            direct_bc5=_array(lib, h, N.BAKE_DIRECT_BC5, np.uint8),
            indirect_rgb9e5=_array(lib, h, N.BAKE_INDIRECT_RGB9E5, np.uint32).reshape(ih, iw),
            volume_bytes=(_array(lib, h, N.BAKE_VOLUME_BYTES, np.uint8).reshape(-1, 16)
                          if vol is not None else None),
        )
    finally:
        lib.lb_bake_destroy(h)
        del keep, ckeep


class BakeSession:
    """lb_bake_prepare / execute / finish: a bake whose device part can be repeated on resident
    data (the benchmark's step) and timed pass by pass."""

    def __init__(self, scene: Scene, gpu: Gpu, form: int = N.FORM_TUNED, graph: bool = True,
                 threads: int = 0, cache: TileCache | None = None, tile_dirty=None,
                 probe_dirty=None, replay: bool = False):
        self.lib = N.lib()
        self.gpu = gpu
        self.scene = scene
        self._desc, self._keep = scene_desc(scene)
        cd, ckeep = _cache_desc(cache, tile_dirty, probe_dirty)
        self._keep += ckeep
        opt = N.BakeOptions(int(form), int(graph), int(threads),
                            C.pointer(cd) if cd is not None else None, int(replay))
        h = C.c_void_p()
        N.check(self.lib.lb_bake_prepare(gpu.handle, C.byref(self._desc), C.byref(opt),
                                         C.byref(h)))
        self.handle = h
        N.own(self, N.BAKE)

    def info(self) -> dict:
        inf = N.BakeInfo()
        N.check(self.lib.lb_bake_info_get(self.handle, C.byref(inf)))
        out = {}
        for k, _ in N.BakeInfo._fields_:
            v = getattr(inf, k)
            out[k] = list(v) if hasattr(v, "__len__") else v
        return out

    def execute(self, stats: bool = False):
        if not stats:
            N.check(self.lib.lb_bake_execute(self.handle, None))
            return None
        st = (N.Stats * 3)()
        N.check(self.lib.lb_bake_execute(self.handle, C.cast(st, C.c_void_p)))
        return [{k: getattr(s, k) for k, _ in N.Stats._fields_} for s in st]

    def finish(self) -> None:
        N.check(self.lib.lb_bake_finish(self.handle))

    def time_pass(self, pass_: int, reps: int = 5) -> float:
        ms = C.c_float()
        N.check(self.lib.lb_bake_time_pass(self.handle, int(pass_), int(reps), C.byref(ms)))
        return ms.value

    def array(self, which: int, dtype) -> np.ndarray:
        return _array(self.lib, self.handle, which, dtype)

    def tile_cache(self) -> TileCache:
        """After finish: every tile's TileLight and the traced probes."""
        return _tile_cache(self.lib, self.handle)

    def close(self) -> None:
        if self.handle:
            self.lib.lb_bake_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


@dataclass
class Layout:
    """Layout::build (layout.rs:89-115) as arrays; `rasterize(t)` gives tile t's texels."""
    width: int
    height: int
    affines: np.ndarray     # (charts, 2, 4)
    chart_tiles: np.ndarray
    tiles: np.ndarray       # (tiles, 6): key x, key z, w, h, atlas x, atlas y
    tri_charts: np.ndarray
    _scene: Scene = None

    def rasterize(self, t: int) -> np.ndarray:
        lib = N.lib()
        d, keep = scene_desc(self._scene)
        h = C.c_void_p()
        N.check(lib.lb_layout_build(C.byref(d), C.byref(h)))
        try:
            w, hh = int(self.tiles[t, 2]), int(self.tiles[t, 3])
            out = np.zeros(w * hh, dtype=N.TEXEL)
            N.check(lib.lb_layout_rasterize(C.byref(d), h, t, out.ctypes.data, len(out)))
            return out.reshape(hh, w)
        finally:
            lib.lb_layout_destroy(h)


def layout(scene: Scene) -> Layout:
    lib = N.lib()
    d, keep = scene_desc(scene)
    h = C.c_void_p()
    N.check(lib.lb_layout_build(C.byref(d), C.byref(h)))
    try:
        w, hh = C.c_uint32(), C.c_uint32()
        nc, nt = C.c_size_t(), C.c_size_t()
        N.check(lib.lb_layout_size(h, C.byref(w), C.byref(hh), C.byref(nc), C.byref(nt)))
        aff = np.zeros((nc.value, 8), np.float32)
        ct = np.zeros(nc.value, np.uint32)
        tiles = np.zeros((nt.value, 6), np.int32)
        tc = np.zeros(scene.triangle_count(), np.uint32)
        N.check(lib.lb_layout_get(h, aff.ctypes.data, ct.ctypes.data, tiles.ctypes.data,
                                  tc.ctypes.data))
        return Layout(w.value, hh.value, aff.reshape(-1, 2, 4), ct, tiles, tc, scene)
    finally:
        lib.lb_layout_destroy(h)
