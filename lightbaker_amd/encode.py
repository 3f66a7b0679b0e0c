"""The output codecs (crates/lightbaker/src/encode.rs:7-74, output.rs:101,108,165-181) on the
device, over the C-ABI: BC5 blocks of two 8-bit channels, RGB9E5 words, the volume's probe
bytes. Each takes numpy arrays, runs the HIP kernels on the gpu's stream and returns numpy.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _native as N
from .gpu import DeviceArray, Gpu


def bc5(gpu: Gpu, rg: np.ndarray, w: int, h: int) -> np.ndarray:
    """encode.rs:37-50 `bc5`: (h, w, 2) bytes -> w * h bytes, each 4 x 4 block's red BC4 block
    then its green one, blocks row-major. w and h must be multiples of 4."""
    a = np.ascontiguousarray(rg, np.uint8)
    if a.size != w * h * 2:
        raise ValueError(f"bc5: {a.size} bytes for {w} x {h} pairs")
    src, dst = DeviceArray.of(gpu, a), DeviceArray(gpu, w * h)
    N.check(N.lib().lb_encode_bc5(gpu.handle, src.ptr, w, h, dst.ptr))
    return dst.get(np.uint8, (w * h,))


def rgb9e5(gpu: Gpu, rgb: np.ndarray) -> np.ndarray:
    """encode.rs:53-69 over (n, 3) f32 -> n u32."""
    a = np.ascontiguousarray(rgb, np.float32).reshape(-1, 3)
    n = len(a)
    src, dst = DeviceArray.of(gpu, a), DeviceArray(gpu, n * 4)
    N.check(N.lib().lb_encode_rgb9e5(gpu.handle, src.ptr, n, dst.ptr))
    return dst.get(np.uint32, (n,))


def rgb9e5_decode(p: np.ndarray) -> np.ndarray:
    """encode.rs:71-74: (n,) u32 -> (n, 3) f32."""
    p = np.asarray(p, np.uint32)
    s = np.ldexp(np.float32(1), (p >> 27).astype(np.int32) - 24).astype(np.float32)
    q = np.stack([p & 511, (p >> 9) & 511, (p >> 18) & 511], -1).astype(np.float32)
    return q * s[:, None]


def volume_bytes(gpu: Gpu, probes: np.ndarray) -> tuple[np.ndarray, float]:
    """output.rs:108,165-181: (n, 16) f32 probes -> ((n, 16) bytes, scale)."""
    a = np.ascontiguousarray(probes, np.float32).reshape(-1, 16)
    n = len(a)
    src, dst = DeviceArray.of(gpu, a), DeviceArray(gpu, n * 16)
    scale = C.c_float()
    N.check(N.lib().lb_encode_volume(gpu.handle, src.ptr, n, dst.ptr, C.byref(scale)))
    return dst.get(np.uint8, (n, 16)), scale.value
