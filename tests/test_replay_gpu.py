"""Bounces replayed from the first bounce's ray records (lb_bake_options.replay): every bounce
traces the same gather rays (the same keys and sample points, bake.rs:300-330 with
trace.wgsl:248-262), so the hits of the first give the later ones' light with the new atlas.

Bar: a replayed PLAIN bake equals the traced PLAIN bake bit for bit (the replay adds each ray's
light in the reference's order with its arithmetic); the TUNED form's replayed bake is within
1e-4 relative of it (its first bounce sums in another order), and the pooled schedule's (with
FP32 node records) replayed bake equals the traced PLAIN bake bit for bit.
"""
from __future__ import annotations

import numpy as np
import pytest

from lightbaker_amd import _native as N
from lightbaker_amd.bake import bake
from lightbaker_amd.scenes import fit_scene, reference_test_scene

pytestmark = pytest.mark.gpu

OUTS = ("light", "direct", "indirect", "validity")


def same(a, b):
    return np.array_equal(np.asarray(a).view(np.uint8), np.asarray(b).view(np.uint8))


@pytest.fixture(scope="module", params=["reference-test-scene", "fit-scene"])
def traced(request, gpu):
    scene = reference_test_scene(8.0) if request.param == "reference-test-scene" else fit_scene(5)
    print(f"replay: traced PLAIN bake of the {request.param}", flush=True)
    return scene, bake(scene, gpu, form=N.FORM_PLAIN)


def test_plain_replay_is_bit_exact(gpu, traced):
    scene, want = traced
    got = bake(scene, gpu, form=N.FORM_PLAIN, replay=True)
    for k in OUTS + ("volume",):
        if getattr(want, k) is not None:
            assert same(getattr(got, k), getattr(want, k)), k


def test_pooled_replay_is_bit_exact(gpu, traced):
    scene, want = traced
    gpu.set_tuning(schedule=1, nodes=N.NODES_WIDE)
    got = bake(scene, gpu, form=N.FORM_TUNED, graph=True, replay=True)
    for k in OUTS:
        assert same(getattr(got, k), getattr(want, k)), k


def test_tuned_replay_within_tolerance(gpu, traced):
    scene, want = traced
    got = bake(scene, gpu, form=N.FORM_TUNED, graph=True, replay=True)
    a, b = got.light.reshape(-1, 4), want.light.reshape(-1, 4)
    err = np.abs(a - b) / (np.abs(b) + 1e-3)
    print(f"tuned replay: max relative difference {err.max():.2e}", flush=True)
    assert err.max() < 1e-4
    assert same(got.validity, want.validity)


@pytest.mark.parametrize("order", [1, 2])
def test_tuned_gather_orders_within_tolerance(gpu, traced, order):
    """Elevation-band order (then ranked, or in every bounce): the bake within 1e-4 of PLAIN."""
    scene, want = traced
    from lightbaker_amd.gpu import Pass
    gpu.set_tuning(pass_=Pass.GATHER, gather_order=order)
    got = bake(scene, gpu, form=N.FORM_TUNED, graph=True)
    a, b = got.light.reshape(-1, 4), want.light.reshape(-1, 4)
    err = np.abs(a - b) / (np.abs(b) + 1e-3)
    assert err.max() < 1e-4, err.max()
    assert same(got.validity, want.validity)
