"""lb_run_device_noalloc against lb_run_device: the same kernels over the same inputs, with the
TUNED lock-step launch's queue heads and spill area taken from the library's static per-device
scratch instead of buffers it grows. Bar: bit-exact (the same kernel, the same item-to-lane
mapping and the same summation order; only where the stack spills differs), on two calls in a
row, the second over fewer items, so the scratch holds the first call's contents when the
second starts.
"""
from __future__ import annotations

import numpy as np
import pytest

from lightbaker_amd import _native as N
from lightbaker_amd.gpu import DeviceArray, Pass
from lightbaker_amd.scenes import small_scene

from . import kit

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def setup(gpu):
    scene = small_scene(1)
    bvh, pos, mesh = kit.scene_bvh(scene)
    rng = np.random.default_rng(23)
    items = kit.sample_items(pos, mesh, 3000, rng)
    probes = kit.probe_items(scene, 800)
    charts = kit.charts_for(scene)
    light = kit.random_light(rng)
    sb = gpu.scene(bvh, scene.materials_array(), charts)
    yield dict(items=items, probes=probes, light=light, params=kit.params_for(scene), sb=sb)
    sb.close()


def _run(gpu, s, fn, pass_, form, items, key_base=0):
    lib = N.lib()
    n = len(items)
    di = DeviceArray.of(gpu, items)
    dl = DeviceArray.of(gpu, s["light"])
    do = DeviceArray(gpu, n * 16 * pass_.width)
    pr = np.ascontiguousarray(s["params"], dtype=N.PARAMS).reshape(1)
    N.check(getattr(lib, fn)(gpu.handle, int(pass_), int(form), s["sb"].handle, N.ptr(pr), di.ptr,
                             n, key_base, dl.ptr, len(s["light"]), do.ptr))
    gpu.sync()
    out = do.get(np.float32, (n * pass_.width, 4))
    for d in (di, dl, do):
        d.free()
    return out


@pytest.mark.parametrize("form", [N.FORM_TUNED, N.FORM_PLAIN], ids=["tuned", "plain"])
@pytest.mark.parametrize("pass_", [Pass.DIRECT, Pass.GATHER, Pass.VOLUME])
def test_noalloc_equals_run_device(gpu, setup, pass_, form):
    src = setup["probes"] if pass_ == Pass.VOLUME else setup["items"]
    for n, key_base in ((len(src), 0), (len(src) // 3, 17)):
        print(f"no-alloc {pass_.name} form {form}: {n} items, key base {key_base}", flush=True)
        items = src[:n]
        want = _run(gpu, setup, "lb_run_device", pass_, form, items, key_base)
        got = _run(gpu, setup, "lb_run_device_noalloc", pass_, form, items, key_base)
        assert np.array_equal(kit.bits(got), kit.bits(want)), (pass_.name, n)


def test_noalloc_full_residency(gpu, setup):
    """The benchmark's gather settings (8 waves a SIMD, 32 lanes an item of 256 rays, 16 LDS
    entries) over enough items to fill every CU: 524 288 resident lanes spilling 32 entries each
    is the whole 64 MiB spill area, which the launch must take, not refuse."""
    gpu.set_tuning(pass_=Pass.GATHER, workgroup=256, rays_per_lane=8, lds_stack=16, waves=8,
                   schedule=0)
    params = np.ascontiguousarray(setup["params"], dtype=N.PARAMS).copy()
    params["counts"][0][2] = 256
    rng = np.random.default_rng(29)
    idx = rng.integers(0, len(setup["items"]), 20000)
    items = np.ascontiguousarray(setup["items"][idx])
    s = dict(setup, params=params)
    print(f"no-alloc full residency: {len(items)} items x 256 rays", flush=True)
    want = _run(gpu, s, "lb_run_device", Pass.GATHER, N.FORM_TUNED, items)
    got = _run(gpu, s, "lb_run_device_noalloc", Pass.GATHER, N.FORM_TUNED, items)
    assert np.array_equal(kit.bits(got), kit.bits(want))


def test_noalloc_shallow_lds_stack(gpu, setup):
    """Eight entries in LDS: 40 spill a lane. The launch either fits the scratch and equals
    lb_run_device, or is refused with the scratch's size named, never run over it."""
    gpu.set_tuning(pass_=Pass.GATHER, lds_stack=8)
    items = setup["items"][:1000]
    want = _run(gpu, setup, "lb_run_device", Pass.GATHER, N.FORM_TUNED, items)
    try:
        got = _run(gpu, setup, "lb_run_device_noalloc", Pass.GATHER, N.FORM_TUNED, items)
    except N.NativeError as e:
        assert "static scratch holds" in str(e)
        return
    assert np.array_equal(kit.bits(got), kit.bits(want))


def test_noalloc_refuses_the_pooled_schedule(gpu, setup):
    gpu.set_tuning(pass_=Pass.GATHER, schedule=1)
    with pytest.raises(N.NativeError, match="pooled schedule"):
        _run(gpu, setup, "lb_run_device_noalloc", Pass.GATHER, N.FORM_TUNED, setup["items"][:64])
    lib = N.lib()
    assert lib.lb_run_device_noalloc(None, 1, 1, None, None, None, 0, 0, None, 0, None) != 0
    assert b"null" in lib.lb_last_error()
