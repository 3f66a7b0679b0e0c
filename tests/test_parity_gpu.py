"""Parity of the HIP passes with the oracle (the FP32 C restatement of trace.wgsl) on the same
items, keys and scene, through the C-ABI.

Bar: the PLAIN form reproduces every output bit for bit (hit-derived bits, counts and FP32 sums in
the reference's order). The TUNED form reproduces every hit-derived quantity exactly (shadow
bits, pick, sun/soft/back counts) and the ray sums within 2e-6 relative + 1e-6 absolute (its
sums are Kahan-compensated in a different order); with 8-bit quantized nodes the same bar holds
against the oracle traversing the same 8-bit records (oracle.run_quant), no item excepted.
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from lightbaker_amd import _native as N
from lightbaker_amd.gpu import Pass
from lightbaker_amd.scenes import small_scene

from . import kit

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def setup(gpu):
    scene = small_scene(1)
    bvh, pos, mesh = kit.scene_bvh(scene)
    rng = np.random.default_rng(7)
    items = kit.sample_items(pos, mesh, 1500, rng)
    charts = kit.charts_for(scene)
    light = kit.random_light(rng)
    mats = scene.materials_array()
    params = kit.params_for(scene)
    sb = gpu.scene(bvh, mats, charts)
    probes = kit.probe_items(scene, 600)
    yield dict(scene=scene, bvh=bvh, items=items, charts=charts, light=light, mats=mats,
               params=params, sb=sb, probes=probes)
    sb.close()


def _oracle(s, pass_, items, quant=False):
    """The oracle's pass; with `quant`, over the 8-bit node records the TUNED form traverses."""
    if quant:
        return oracle.run_quant(int(pass_), s["bvh"].nodes, oracle.quant_records(s["bvh"].nodes),
                                s["bvh"].tris, s["mats"], s["charts"], s["params"], items, 0,
                                s["light"])
    return oracle.run(int(pass_), s["bvh"].nodes, s["bvh"].tris, s["mats"], s["charts"],
                      s["params"], items, 0, s["light"])


@pytest.mark.parametrize("pass_", [Pass.DIRECT, Pass.GATHER, Pass.VOLUME])
def test_plain_bit_exact(gpu, setup, pass_):
    items = setup["probes"] if pass_ == Pass.VOLUME else setup["items"]
    print(f"plain {pass_.name}: {len(items)} items", flush=True)
    gpu.set_form(N.FORM_PLAIN)
    got = gpu.run(pass_, setup["sb"], setup["params"], items, setup["light"])
    want = _oracle(setup, pass_, items)
    gb, wb = kit.bits(got), kit.bits(want)
    bad = np.nonzero((gb != wb).any(axis=1))[0]
    assert len(bad) == 0, f"{len(bad)} of {len(gb)} rows differ, first {bad[:5]}: " \
                          f"{got[bad[:3]]} vs {want[bad[:3]]}"


def _compare_tuned(pass_, got, want, quant):
    n = len(want)
    if pass_ == Pass.DIRECT:
        exact_cols = [0, 1, 3]
        sum_cols = [2]
    elif pass_ == Pass.GATHER:
        exact_cols = [3]
        sum_cols = [0, 1, 2]
    else:
        got = got.reshape(-1, 16)
        want = want.reshape(-1, 16)
        exact_cols = [3, 7, 11, 15]
        sum_cols = [0, 1, 2, 4, 5, 6, 8, 9, 10, 12, 13, 14]
    gb, wb = kit.bits(got), kit.bits(want)
    exact_bad = (gb[:, exact_cols] != wb[:, exact_cols]).any(axis=1)
    tol = 2e-6 * np.abs(want[:, sum_cols]) + 1e-6
    sum_bad = (np.abs(got[:, sum_cols] - want[:, sum_cols]) > tol).any(axis=1)
    bad = exact_bad | sum_bad
    # `want` is the oracle over the node records the kernel traverses (the 8-bit ones for QUANT):
    # the same hits in the same order, so one bar for both layouts
    idx = np.nonzero(bad)[0]
    assert len(idx) == 0, f"{len(idx)} of {n} differ (quant {quant}), first {idx[:5]}: " \
                          f"{got[idx[:3]]} vs {want[idx[:3]]}"


@pytest.mark.parametrize("nodes", [N.NODES_WIDE, N.NODES_QUANT])
@pytest.mark.parametrize("pass_", [Pass.DIRECT, Pass.GATHER, Pass.VOLUME])
def test_tuned_matches(gpu, setup, pass_, nodes):
    items = setup["probes"] if pass_ == Pass.VOLUME else setup["items"]
    print(f"tuned {pass_.name} nodes={nodes}: {len(items)} items", flush=True)
    gpu.set_form(N.FORM_TUNED)
    gpu.set_tuning(256, 2, 16, nodes, 32, 0, leaf_batch=64)
    try:
        got = gpu.run(pass_, setup["sb"], setup["params"], items, setup["light"])
    finally:
        gpu.set_form(N.FORM_PLAIN)
    want = _oracle(setup, pass_, items, quant=nodes == N.NODES_QUANT)
    _compare_tuned(pass_, got, want, nodes == N.NODES_QUANT)


@pytest.mark.parametrize("cfg", [(64, 1, 8, 0, 64), (128, 4, 24, 6, 16), (512, 8, 24, 8, 1),
                                 (256, 2, 32, 0, 32), (256, 8, 48, 6, 8), (256, 4, 16, 8, 24),
                                 (192, 4, 16, 8, 16)])
def test_tuned_configs(gpu, setup, cfg):
    """Every autotuner setting gives the same results (workgroup, rays/lane, LDS stack, waves
    a SIMD, leaf batch), on every pass."""
    wg, rpl, depth, waves, leaf_batch = cfg
    for pass_ in (Pass.DIRECT, Pass.GATHER, Pass.VOLUME):
        print(f"tuned {pass_.name} cfg={cfg}", flush=True)
        items = setup["probes"] if pass_ == Pass.VOLUME else setup["items"]
        gpu.set_form(N.FORM_TUNED)
        gpu.set_tuning(wg, rpl, depth, N.NODES_WIDE, 32, 0, waves=waves, leaf_batch=leaf_batch)
        try:
            got = gpu.run(pass_, setup["sb"], setup["params"], items, setup["light"])
        finally:
            gpu.set_form(N.FORM_PLAIN)
            gpu.set_tuning(**DEFAULT)
        _compare_tuned(pass_, got, _oracle(setup, pass_, items), False)


def test_key_base_slices(gpu, setup):
    """A slice run with key_base = its offset gives the rows of the whole run (gpu/mod.rs:153-154)."""
    items = setup["items"]
    d_items = __import__("lightbaker_amd.gpu", fromlist=["DeviceArray"]).DeviceArray
    whole = gpu.run(Pass.GATHER, setup["sb"], setup["params"], items, setup["light"])
    part = items[500:900].copy()
    di = d_items.of(gpu, part)
    dl = d_items.of(gpu, setup["light"])
    do = d_items(gpu, len(part) * 16)
    gpu.run_device(Pass.GATHER, N.FORM_PLAIN, setup["sb"], setup["params"], di, len(part), dl,
                   len(setup["light"]), do, key_base=500)
    got = do.get(np.float32, (len(part), 4))
    assert np.array_equal(kit.bits(got), kit.bits(whole[500:900]))


def test_counters_match_oracle(gpu, setup):
    """The kernels' traversal counters equal the oracle's on the same rays (PLAIN form)."""
    from lightbaker_amd.gpu import DeviceArray
    items = setup["items"][:300]
    di = DeviceArray.of(gpu, items)
    dl = DeviceArray.of(gpu, setup["light"])
    do = DeviceArray(gpu, len(items) * 16)
    st = gpu.run_stats(Pass.GATHER, N.FORM_PLAIN, setup["sb"], setup["params"], di, len(items),
                       dl, len(setup["light"]), do)
    _, want = oracle.run(1, setup["bvh"].nodes, setup["bvh"].tris, setup["mats"],
                         setup["charts"], setup["params"], items, 0, setup["light"],
                         counters=True)
    assert st == want


# ---- the pooled schedule -----------------------------------------------------------------

STREAM = dict(workgroup=256, rays_per_lane=2, lds_stack=16, nodes=N.NODES_WIDE, refill=32,
              schedule=1, chunk=0, leaf_batch=24)
DEFAULT = dict(workgroup=256, rays_per_lane=2, lds_stack=16, nodes=N.NODES_WIDE, refill=32,
               schedule=0, chunk=0, leaf_batch=64, waves=0)


def _run_tuned(gpu, setup, pass_, **tuning):
    items = setup["probes"] if pass_ == Pass.VOLUME else setup["items"]
    gpu.set_form(N.FORM_TUNED)
    gpu.set_tuning(**dict(STREAM, **tuning))
    try:
        return gpu.run(pass_, setup["sb"], setup["params"], items, setup["light"])
    finally:
        gpu.set_form(N.FORM_PLAIN)
        gpu.set_tuning(**DEFAULT)


def _assert_bits(got, want):
    gb, wb = kit.bits(got), kit.bits(want)
    bad = np.nonzero((gb != wb).any(axis=1))[0]
    assert len(bad) == 0, f"{len(bad)} of {len(gb)} rows differ, first {bad[:5]}: " \
                          f"{got[bad[:3]]} vs {want[bad[:3]]}"


@pytest.mark.parametrize("pass_", [Pass.DIRECT, Pass.GATHER, Pass.VOLUME])
def test_pooled_bit_exact(gpu, setup, pass_):
    """The pooled schedule traces each ray in the reference's order and folds an item's hits in
    the reference's ray order with its arithmetic: every output bit equals the oracle's."""
    print(f"pooled {pass_.name}", flush=True)
    items = setup["probes"] if pass_ == Pass.VOLUME else setup["items"]
    _assert_bits(_run_tuned(gpu, setup, pass_), _oracle(setup, pass_, items))


@pytest.mark.parametrize("cfg", [(64, 8, 1, 0, 1), (512, 48, 64, 0, 64), (256, 16, 32, 5000, 8),
                                 (128, 24, 8, 777, 32), (192, 16, 32, 0, 16)])
def test_pooled_configs(gpu, setup, cfg):
    """Workgroup, LDS stack, refill threshold, leaf batch and chunking (777 rays: a few items a
    chunk) leave every output bit as it is."""
    wg, depth, refill, chunk, leaf_batch = cfg
    for pass_ in (Pass.DIRECT, Pass.GATHER, Pass.VOLUME):
        print(f"pooled {pass_.name} cfg={cfg}", flush=True)
        items = setup["probes"] if pass_ == Pass.VOLUME else setup["items"]
        got = _run_tuned(gpu, setup, pass_, workgroup=wg, lds_stack=depth, refill=refill,
                         chunk=chunk, leaf_batch=leaf_batch)
        _assert_bits(got, _oracle(setup, pass_, items))


@pytest.mark.parametrize("pass_", [Pass.DIRECT, Pass.GATHER, Pass.VOLUME])
def test_pooled_quant(gpu, setup, pass_):
    items = setup["probes"] if pass_ == Pass.VOLUME else setup["items"]
    print(f"pooled quant {pass_.name}", flush=True)
    got = _run_tuned(gpu, setup, pass_, nodes=N.NODES_QUANT)
    _compare_tuned(pass_, got, _oracle(setup, pass_, items), True)


@pytest.mark.parametrize("pass_", [Pass.DIRECT, Pass.GATHER, Pass.VOLUME])
def test_pooled_counters(gpu, setup, pass_):
    """The pooled kernels trace exactly the rays the reference traces (the inside test's
    rounds included) and visit exactly its nodes: their counters equal the oracle's."""
    from lightbaker_amd.gpu import DeviceArray
    items = (setup["probes"] if pass_ == Pass.VOLUME else setup["items"])[:300].copy()
    print(f"pooled counters {pass_.name}", flush=True)
    di = DeviceArray.of(gpu, items)
    dl = DeviceArray.of(gpu, setup["light"])
    do = DeviceArray(gpu, len(items) * 16 * pass_.width)
    gpu.set_tuning(**STREAM)
    try:
        st = gpu.run_stats(pass_, N.FORM_TUNED, setup["sb"], setup["params"], di, len(items),
                           dl, len(setup["light"]), do)
    finally:
        gpu.set_tuning(**DEFAULT)
    _, want = oracle.run(int(pass_), setup["bvh"].nodes, setup["bvh"].tris, setup["mats"],
                         setup["charts"], setup["params"], items, 0, setup["light"],
                         counters=True)
    assert st == want


@pytest.mark.parametrize("pass_", [Pass.DIRECT, Pass.GATHER, Pass.VOLUME])
def test_tuned_xcd_split_same_results(gpu, setup, pass_):
    """The chiplet-split queue hands out the same items in another order; every item's rays and
    sums are its own, so the results are bit for bit those of one queue, and match the oracle
    to the TUNED bar."""
    items = setup["probes"] if pass_ == Pass.VOLUME else setup["items"]
    gpu.set_form(N.FORM_TUNED)
    out = {}
    for x in (0, 1):
        gpu.set_tuning(pass_=pass_, xcd_split=x)
        out[x] = gpu.run(pass_, setup["sb"], setup["params"], items, setup["light"])
    assert np.array_equal(kit.bits(out[0]), kit.bits(out[1]))
    _compare_tuned(pass_, out[1], _oracle(setup, pass_, items), gpu.tuning(pass_)["nodes"] == 1)


@pytest.mark.parametrize("order", [1, 2])
def test_tuned_gather_elevation_bands(gpu, setup, order):
    """The gather's rays in elevation bands go to other lanes in another order: the same rays,
    so the same hits, and sums within the TUNED bar."""
    gpu.set_form(N.FORM_TUNED)
    gpu.set_tuning(pass_=Pass.GATHER, gather_order=order)
    got = gpu.run(Pass.GATHER, setup["sb"], setup["params"], setup["items"], setup["light"])
    _compare_tuned(Pass.GATHER, got, _oracle(setup, Pass.GATHER, setup["items"]),
                   gpu.tuning(Pass.GATHER)["nodes"] == 1)
