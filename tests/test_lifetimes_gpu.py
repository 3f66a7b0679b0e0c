"""Bookkeeping of the bake's captured bounce graph, its ray ranking, and handle lifetimes.

The bounce graph holds its kernels' arguments by value: the gather item count, the gather's
tuning, and the addresses of the buffers they read, among them the TUNED form's GPU-wide spill
area, which a launch with another tuning regrows. A graph is replayed only while all of these
still hold, and captured again otherwise (lb_bake_info.graph_captures counts the captures).
These tests check that bookkeeping; none of them replays a stale graph.
"""
from __future__ import annotations

import numpy as np
import pytest

from lightbaker_amd import _native as N
from lightbaker_amd.bake import BakeSession
from lightbaker_amd.gpu import Gpu, Pass
from lightbaker_amd.scenes import small_scene

pytestmark = pytest.mark.gpu


def scene2():
    """The small scene with two bounces: the second takes the rays ranked by the first."""
    s = small_scene(1)
    s.settings.gi_bounces = 2
    return s


def _light(sess):
    return sess.array(N.BAKE_LIGHT, np.float32).reshape(-1, 4).copy()


def _close(a, b, tol=1e-4):
    err = np.abs(a - b) / (np.abs(b) + 1e-3)
    return float(err.max())


def test_graph_recaptured_when_spill_area_regrows():
    """A direct-pass tuning with a shallower LDS stack regrows the GPU-wide spill area in the
    next execute's direct launch; the bounce graph, which spills into that area, is captured
    again instead of replayed, and the light stays within the lock-step tolerance. The GPU is
    opened here, so its spill area holds only what this bake's launches asked for (a shared
    one may have grown past the new need in earlier launches)."""
    print("graph key: spill area regrown between executes", flush=True)
    g = Gpu()
    g.set_tuning(256, 4, 16, N.NODES_QUANT, 32, 0, leaf_batch=16)
    sess = BakeSession(scene2(), g, form=N.FORM_TUNED, graph=True)
    sess.execute()
    first = _light(sess)
    assert sess.info()["graph_captures"] == 1
    sess.execute()
    assert sess.info()["graph_captures"] == 1, "same key: the graph is replayed"
    assert np.array_equal(_light(sess), first)
    g.set_tuning(pass_=Pass.DIRECT, lds_stack=8)
    sess.execute()
    assert sess.info()["graph_captures"] == 2, "a regrown spill area captures again"
    assert _close(_light(sess), first) < 1e-4
    sess.close()
    g.close()


def test_graph_recaptured_on_every_gather_knob(gpu):
    """Each gather knob a TUNED bake's graph depends on (the chiplet split and the ray order
    included) captures the graph again when it changes."""
    print("graph key: gather knobs", flush=True)
    sess = BakeSession(scene2(), gpu, form=N.FORM_TUNED, graph=True)
    sess.execute()
    first = _light(sess)
    n = sess.info()["graph_captures"]
    for knob, value in (("xcd_split", 1), ("gather_order", 1), ("leaf_batch", 8)):
        gpu.set_tuning(pass_=Pass.GATHER, **{knob: value})
        sess.execute()
        got = sess.info()["graph_captures"]
        assert got == n + 1, f"{knob}: {got} captures, {n + 1} expected"
        n = got
        assert _close(_light(sess), first) < 1e-4, knob
    sess.close()


def test_ranking_survives_graph_replays(gpu):
    """The gather's step counts and their order are read back after the second execute of a
    graph-replayed bake too (the replay re-runs the ranking), equal to the first's."""
    print("ranking after a graph replay", flush=True)
    sess = BakeSession(scene2(), gpu, form=N.FORM_TUNED, graph=True)
    sess.execute()
    cost = sess.array(N.BAKE_GATHER_COST, np.uint8).copy()
    order = sess.array(N.BAKE_GATHER_ORDER, np.uint8).copy()
    assert len(cost) > 0 and len(order) == len(cost)
    sess.execute()
    assert sess.info()["graph_captures"] == 1
    assert np.array_equal(sess.array(N.BAKE_GATHER_COST, np.uint8), cost)
    assert np.array_equal(sess.array(N.BAKE_GATHER_ORDER, np.uint8), order)
    sess.close()


def test_time_pass_after_replayed_bounces(gpu):
    """A bake whose later bounces replay ray records ranks nothing, so timing its gather takes
    the rays in azimuth order (no ranking is read) and leaves the bake's results as they were."""
    print("time_pass after a replayed bake", flush=True)
    sess = BakeSession(scene2(), gpu, form=N.FORM_TUNED, graph=True, replay=True)
    sess.execute()
    assert len(sess.array(N.BAKE_GATHER_ORDER, np.uint8)) == 0
    light = _light(sess)
    ms = sess.time_pass(Pass.GATHER, 2)
    assert ms > 0.0
    sess.execute()
    assert _close(_light(sess), light) < 1e-4
    sess.close()


def test_stats_execute_keeps_ranking_valid(gpu):
    """An execute with the traversal counters traces every bounce unranked; a later time_pass
    then still reads the ranking of the last ranked execute, a permutation of each item's rays."""
    print("stats execute between ranked executes", flush=True)
    sess = BakeSession(scene2(), gpu, form=N.FORM_TUNED, graph=True)
    sess.execute()
    order = sess.array(N.BAKE_GATHER_ORDER, np.uint8).copy()
    rays = small_scene(1).settings.gi_rays
    perm = order.reshape(-1, rays)
    assert (np.sort(perm, axis=1) == np.arange(rays, dtype=np.uint8)).all()
    stats = sess.execute(stats=True)
    assert stats[1]["rays"] > 0
    assert sess.time_pass(Pass.GATHER, 1) > 0.0
    sess.close()


def test_gpu_closed_before_its_bake(gpu):
    """Closing a Gpu closes its bake sessions first; a bake kept by another reference after its
    Gpu closed holds no dangling handle (the library keeps the device handle until the last bake
    over it goes)."""
    print("Gpu closed before its bake session", flush=True)
    g = Gpu()
    sess = BakeSession(small_scene(1), g, form=N.FORM_PLAIN, graph=False)
    sess.execute()
    g.close()
    assert sess.handle is None
    sess.close()  # a second close does nothing

    # the library alone: a bake outlives lb_close of its gpu
    lib = N.lib()
    import ctypes as C
    from lightbaker_amd.bake import scene_desc
    h = C.c_void_p()
    N.check(lib.lb_open(C.byref(h)))
    d, keep = scene_desc(small_scene(1))
    opt = N.BakeOptions(N.FORM_PLAIN, 0, 0, None, 0)
    b = C.c_void_p()
    N.check(lib.lb_bake_prepare(h, C.byref(d), C.byref(opt), C.byref(b)))
    N.check(lib.lb_bake_execute(b, None))
    lib.lb_close(h)
    N.check(lib.lb_bake_execute(b, None))  # the gpu handle is still alive inside the bake
    lib.lb_bake_destroy(b)
    del keep
