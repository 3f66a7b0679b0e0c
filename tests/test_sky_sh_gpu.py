"""The sky's L1 spherical-harmonic term on the GPU (lb_scene_set_sky): the small scene under the
gradient sky plus an L1 term, against the oracle with the same term (oracle.set_sky_sh).

Bar: PLAIN and the pooled schedule bit for bit on the gather and volume passes; the TUNED
lock-step form within the FP32 bar of the other parity tests (hit-derived bits exact, sums within
2e-6 relative + 1e-6); all-zero coefficients leave every output bit as without the term.
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from lightbaker_amd import _native as N
from lightbaker_amd.gpu import Pass
from lightbaker_amd.scenes import small_scene

from . import kit
from .test_parity_gpu import DEFAULT, STREAM, _assert_bits, _compare_tuned
from .test_sky_sh import SH

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def setup(gpu):
    scene = small_scene(1)
    scene.sky_sh = SH
    bvh, pos, mesh = kit.scene_bvh(scene)
    rng = np.random.default_rng(71)
    s = dict(scene=scene, bvh=bvh, items=kit.sample_items(pos, mesh, 800, rng),
             probes=kit.probe_items(scene, 300), charts=kit.charts_for(scene),
             light=kit.random_light(rng), mats=scene.materials_array(),
             params=kit.params_for(scene))
    s["sb"] = gpu.scene(bvh, s["mats"], s["charts"], sky_sh=SH)
    yield s
    s["sb"].close()


def _items(s, pass_):
    return s["probes"] if pass_ == Pass.VOLUME else s["items"]


def _oracle(s, pass_):
    oracle.set_sky_sh(SH)
    try:
        return oracle.run(int(pass_), s["bvh"].nodes, s["bvh"].tris, s["mats"], s["charts"],
                          s["params"], _items(s, pass_), 0, s["light"])
    finally:
        oracle.set_sky_sh(None)


@pytest.mark.parametrize("pass_", [Pass.GATHER, Pass.VOLUME])
def test_plain_with_l1_sky_bit_exact(gpu, setup, pass_):
    print(f"L1 sky PLAIN {pass_.name}", flush=True)
    gpu.set_form(N.FORM_PLAIN)
    got = gpu.run(pass_, setup["sb"], setup["params"], _items(setup, pass_), setup["light"])
    _assert_bits(got, _oracle(setup, pass_))


@pytest.mark.parametrize("pass_", [Pass.GATHER, Pass.VOLUME])
def test_tuned_with_l1_sky(gpu, setup, pass_):
    print(f"L1 sky TUNED {pass_.name}", flush=True)
    gpu.set_form(N.FORM_TUNED)
    gpu.set_tuning(**DEFAULT)
    try:
        got = gpu.run(pass_, setup["sb"], setup["params"], _items(setup, pass_), setup["light"])
    finally:
        gpu.set_form(N.FORM_PLAIN)
    _compare_tuned(pass_, got, _oracle(setup, pass_), False)


@pytest.mark.parametrize("pass_", [Pass.GATHER, Pass.VOLUME])
def test_pooled_with_l1_sky_bit_exact(gpu, setup, pass_):
    print(f"L1 sky pooled {pass_.name}", flush=True)
    gpu.set_form(N.FORM_TUNED)
    gpu.set_tuning(**STREAM)
    try:
        got = gpu.run(pass_, setup["sb"], setup["params"], _items(setup, pass_), setup["light"])
    finally:
        gpu.set_form(N.FORM_PLAIN)
        gpu.set_tuning(**DEFAULT)
    _assert_bits(got, _oracle(setup, pass_))


def test_zero_coefficients_change_nothing(gpu, setup):
    s = setup
    plain = small_scene(1)
    P = kit.params_for(plain)
    gpu.set_form(N.FORM_PLAIN)
    without = gpu.scene(s["bvh"], s["mats"], s["charts"])
    zero = gpu.scene(s["bvh"], s["mats"], s["charts"], sky_sh=np.zeros((4, 3), np.float32))
    try:
        for pass_ in (Pass.GATHER, Pass.VOLUME):
            print(f"L1 sky zero {pass_.name}", flush=True)
            a = gpu.run(pass_, without, P, _items(s, pass_), s["light"])
            b = gpu.run(pass_, zero, P, _items(s, pass_), s["light"])
            _assert_bits(b, a)
    finally:
        zero.close()
        without.close()
