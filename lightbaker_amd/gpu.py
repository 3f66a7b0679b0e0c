"""The reference's GPU boundary (crates/lightbaker/src/gpu/mod.rs) over the C-ABI.

`Gpu()` is `Gpu::new` (gpu/mod.rs:74), `Gpu.scene` is `Gpu::scene` (:129), `Gpu.run` is
`Gpu::run` (:140); `Pass` is gpu/mod.rs:25-33. Errors raise `NativeError` with the library's
message (the reference's anyhow strings: "no GPU adapter", "no GPU device", "GPU: ...").
"""
from __future__ import annotations

import ctypes as C
import enum

import numpy as np

from . import _native as N


class Pass(enum.IntEnum):
    DIRECT = N.PASS_DIRECT
    GATHER = N.PASS_GATHER
    VOLUME = N.PASS_VOLUME

    @property
    def width(self) -> int:
        """Results an item writes (gpu/mod.rs:44-46)."""
        return 4 if self is Pass.VOLUME else 1


class Bvh:
    """Bvh::build (bvh.rs:50-106) run by the native host code."""

    def __init__(self, nodes: np.ndarray, tris: np.ndarray):
        self.nodes = nodes
        self.tris = tris

    @staticmethod
    def build(positions: np.ndarray, material: np.ndarray, chart: np.ndarray,
              casts: np.ndarray) -> "Bvh":
        lib = N.lib()
        pos = np.ascontiguousarray(positions, dtype=np.float32).reshape(-1, 9)
        mat = np.ascontiguousarray(material, dtype=np.uint32)
        ch = np.ascontiguousarray(chart, dtype=np.uint32)
        cs = np.ascontiguousarray(casts, dtype=np.uint8)
        h = C.c_void_p()
        N.check(lib.lb_bvh_build(N.ptr(pos), len(pos), N.ptr(mat), N.ptr(ch), N.ptr(cs), len(cs),
                                 C.byref(h)))
        try:
            p = C.c_void_p()
            nn = lib.lb_bvh_nodes(h, C.byref(p))
            nodes = np.frombuffer(C.string_at(p, nn * 32), dtype=N.NODE).copy()
            nt = lib.lb_bvh_tris(h, C.byref(p))
            tris = np.frombuffer(C.string_at(p, nt * 48), dtype=N.TRI).copy()
        finally:
            lib.lb_bvh_destroy(h)
        return Bvh(nodes, tris)


class SceneBuffers:
    """The scene as the passes read it (gpu/mod.rs:65-71), resident on the device."""

    def __init__(self, gpu: "Gpu", handle: C.c_void_p):
        self.gpu = gpu
        self.handle = handle
        N.own(self, N.SCENE)

    def close(self) -> None:
        if self.handle:
            N.lib().lb_scene_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceArray:
    """A library-owned device buffer holding a numpy array's bytes."""

    def __init__(self, gpu: "Gpu", nbytes: int):
        self.gpu = gpu
        self.nbytes = nbytes
        p = C.c_void_p()
        N.check(N.lib().lb_buffer_alloc(gpu.handle, max(nbytes, 16), C.byref(p)))
        self.ptr = p
        N.own(self, N.ARRAY)

    @staticmethod
    def of(gpu: "Gpu", a: np.ndarray) -> "DeviceArray":
        a = np.ascontiguousarray(a)
        d = DeviceArray(gpu, a.nbytes)
        N.check(N.lib().lb_upload(gpu.handle, d.ptr, N.ptr(a), a.nbytes))
        return d

    def get(self, dtype, shape) -> np.ndarray:
        out = np.empty(shape, dtype=dtype)
        assert out.nbytes <= self.nbytes
        N.check(N.lib().lb_download(self.gpu.handle, N.ptr(out), self.ptr, out.nbytes))
        return out

    def free(self) -> None:
        if self.ptr and self.gpu.handle:
            N.lib().lb_buffer_free(self.gpu.handle, self.ptr)
        self.ptr = None

    close = free

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Gpu:
    def __init__(self):
        self._lib = N.lib()
        h = C.c_void_p()
        N.check(self._lib.lb_open(C.byref(h)))
        self.handle = h
        self.adapter = self._lib.lb_adapter_name(h).decode()
        self.gpu = self
        N.own(self, N.GPU)

    def close(self) -> None:
        """Closes this Gpu's adjoints, bake sessions, scenes and buffers, then the Gpu."""
        if self.handle:
            N.close_owned(self)
            self._lib.lb_close(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def scene(self, bvh: Bvh, materials: np.ndarray, charts: np.ndarray,
              sky_sh: np.ndarray | None = None) -> SceneBuffers:
        """Gpu::scene; `sky_sh` (4 x 3: c0..c3 rgb) the sky's L1 term (lb_scene_set_sky)."""
        mats = np.ascontiguousarray(materials, dtype=np.float32).reshape(-1, 4)
        ch = np.ascontiguousarray(charts, dtype=np.float32).reshape(-1, 8)
        nodes = np.ascontiguousarray(bvh.nodes)
        tris = np.ascontiguousarray(bvh.tris)
        h = C.c_void_p()
        N.check(self._lib.lb_scene_create(self.handle, N.ptr(nodes), len(nodes), N.ptr(tris),
                                          len(tris), N.ptr(mats), len(mats),
                                          N.ptr(ch) if len(ch) else None, len(ch), C.byref(h)))
        sb = SceneBuffers(self, h)
        if sky_sh is not None:
            sh = np.ascontiguousarray(np.asarray(sky_sh, np.float32).reshape(12))
            N.check(self._lib.lb_scene_set_sky(h, N.ptr(sh)))
        return sb

    def set_form(self, form: int) -> None:
        N.check(self._lib.lb_set_form(self.handle, int(form)))

    TUNING_KEYS = ("workgroup", "rays_per_lane", "lds_stack", "nodes", "refill", "schedule",
                   "chunk", "leaf_batch", "waves", "xcd_split", "gather_order")

    def set_tuning(self, workgroup: int = None, rays_per_lane: int = None, lds_stack: int = None,
                   nodes: int = None, refill: int = None, schedule: int = None,
                   chunk: int = None, leaf_batch: int = None, waves: int = None,
                   xcd_split: int = None, gather_order: int = None,
                   pass_: Pass = None) -> None:
        """Sets the TUNED form's knobs of one pass, or of all three (starting from the gather
        pass's) when `pass_` is None; the knobs not given keep their value."""
        cur = self.tuning(pass_)
        given = dict(workgroup=workgroup, rays_per_lane=rays_per_lane, lds_stack=lds_stack,
                     nodes=nodes, refill=refill, schedule=schedule, chunk=chunk,
                     leaf_batch=leaf_batch, waves=waves, xcd_split=xcd_split,
                     gather_order=gather_order)
        cur.update({k: v for k, v in given.items() if v is not None})
        t = N.Tuning(**{k: int(cur[k]) for k in self.TUNING_KEYS})
        if pass_ is None:
            N.check(self._lib.lb_set_tuning(self.handle, C.byref(t)))
        else:
            N.check(self._lib.lb_set_pass_tuning(self.handle, int(pass_), C.byref(t)))

    def tuning(self, pass_: Pass = None) -> dict:
        """The knobs of one pass (the gather pass's when `pass_` is None)."""
        t = N.Tuning()
        if pass_ is None:
            N.check(self._lib.lb_get_tuning(self.handle, C.byref(t)))
        else:
            N.check(self._lib.lb_get_pass_tuning(self.handle, int(pass_), C.byref(t)))
        return {k: getattr(t, k) for k in self.TUNING_KEYS}

    def run(self, pass_: Pass, scene: SceneBuffers, params: np.ndarray, texels: np.ndarray,
            light: np.ndarray | None) -> np.ndarray:
        """Gpu::run: each item's results, (items * width, 4) float32 in item order."""
        pass_ = Pass(pass_)
        tx = np.ascontiguousarray(texels, dtype=N.TEXEL)
        n = len(tx)
        out = np.zeros((n * pass_.width, 4), dtype=np.float32)
        if n == 0:
            return out
        lt = None if light is None or len(light) == 0 else np.ascontiguousarray(
            light, dtype=np.float32).reshape(-1, 4)
        pr = np.ascontiguousarray(params, dtype=N.PARAMS).reshape(1)
        N.check(self._lib.lb_run(self.handle, int(pass_), scene.handle, N.ptr(pr), N.ptr(tx), n,
                                 N.ptr(lt), 0 if lt is None else len(lt), N.ptr(out), len(out)))
        return out

    # ---- device-resident path ----

    def run_device(self, pass_: Pass, form: int, scene: SceneBuffers, params: np.ndarray,
                   items: DeviceArray, n: int, light: DeviceArray | None, n_light: int,
                   out: DeviceArray, key_base: int = 0) -> None:
        pr = np.ascontiguousarray(params, dtype=N.PARAMS).reshape(1)
        N.check(self._lib.lb_run_device(self.handle, int(pass_), int(form), scene.handle,
                                        N.ptr(pr), items.ptr, n, key_base,
                                        None if light is None else light.ptr, n_light, out.ptr))

    def run_stats(self, pass_: Pass, form: int, scene: SceneBuffers, params: np.ndarray,
                  items: DeviceArray, n: int, light: DeviceArray | None, n_light: int,
                  out: DeviceArray, key_base: int = 0) -> dict:
        pr = np.ascontiguousarray(params, dtype=N.PARAMS).reshape(1)
        st = N.Stats()
        N.check(self._lib.lb_run_stats(self.handle, int(pass_), int(form), scene.handle,
                                       N.ptr(pr), items.ptr, n, key_base,
                                       None if light is None else light.ptr, n_light, out.ptr,
                                       C.byref(st)))
        return {k: getattr(st, k) for k, _ in N.Stats._fields_}

    def sync(self) -> None:
        N.check(self._lib.lb_sync(self.handle))

    def timer_begin(self) -> None:
        N.check(self._lib.lb_timer_begin(self.handle))

    def timer_end(self) -> float:
        ms = C.c_float()
        N.check(self._lib.lb_timer_end(self.handle, C.byref(ms)))
        return ms.value
