"""The TUNED form's autotuner: picks, for each pass on its own, the schedule, workgroup size, rays
per lane, LDS stack depth, refill threshold, leaf batch, node layout, residency and the queue's
split over the chiplets with the best time on a
prepared bake's own launches, then checks that the bake still gives the starting setting's light
atlas.

Search: coordinate descent from the current setting, one knob at a time over its range, at most
`rounds` rounds, each candidate timed as the median of `reps` launches of its pass. A knob that
the candidate's schedule ignores is skipped (rays per lane, waves and the chiplet split for the
pooled schedule, refill for the lock-step one); a candidate the device cannot launch (an LDS stack larger than a workgroup's
LDS) is skipped too. The node layout stays as set unless `tune_nodes`: the 8-bit and FP32 boxes
may keep different ones of two triangles a few ulps apart along a ray, which moves a texel's
light by more than rounding, so a change of layout cannot pass the final check.
"""
from __future__ import annotations

import numpy as np

from . import _native as N
from .gpu import Pass

RANGES = {
    "schedule": [0, 1],
    "rays_per_lane": [1, 2, 4, 8, 16],
    "refill": [16, 32, 48, 56],
    "leaf_batch": [8, 16, 24, 32, 48, 64],
    "workgroup": [64, 128, 192, 256, 512],
    "lds_stack": [8, 16, 24, 32, 48],
    "nodes": [N.NODES_WIDE, N.NODES_QUANT],
    "waves": [0, 6, 8],
    "xcd_split": [0, 1],
}


def _skip(dim: str, cand: dict) -> bool:
    if dim == "refill" and cand["schedule"] == 0:
        return True
    if dim in ("waves", "xcd_split") and cand["schedule"] == 1:
        return True
    return dim == "rays_per_lane" and cand["schedule"] == 1


def _tune_pass(sess, pass_: Pass, reps: int, rounds: int, log,
               tune_nodes: bool = False) -> tuple[dict, float, int]:
    gpu = sess.gpu

    def measure(c):
        try:
            gpu.set_tuning(pass_=pass_, **c)
            return sorted(sess.time_pass(pass_, 1) for _ in range(reps))[reps // 2]
        except N.NativeError:
            return float("inf")

    cur = gpu.tuning(pass_)
    best = measure(cur)
    tried = 1
    for _ in range(rounds):
        changed = False
        for dim, values in RANGES.items():
            if dim == "nodes" and not tune_nodes:
                continue
            for v in values:
                cand = dict(cur, **{dim: v})
                if v == cur[dim] or _skip(dim, cand):
                    continue
                t = measure(cand)
                tried += 1
                if log:
                    log(f"autotune {pass_.name} {cand}: {t:.2f} ms")
                if t < best * 0.995:
                    best, cur, changed = t, cand, True
        if not changed:
            break
    gpu.set_tuning(pass_=pass_, **cur)
    return cur, best, tried


def autotune(sess, reps: int = 3, rounds: int = 2, log=None, tune_nodes: bool = False) -> dict:
    """Tunes the session's GPU in place, pass by pass (the session must be TUNED and executed
    once). Returns each pass's pick and time; reverts every pass if the picks change the light
    atlas by more than the compensated sums' rounding (1e-4 relative)."""
    gpu = sess.gpu
    start = {p: gpu.tuning(p) for p in Pass}
    sess.execute()
    sess.finish()
    ref_light = sess.array(N.BAKE_LIGHT, np.float32).copy()
    picks, ms, tried = {}, {}, 0
    for p in (Pass.GATHER, Pass.DIRECT, Pass.VOLUME):
        picks[p.name.lower()], ms[p.name.lower()], n = _tune_pass(sess, p, reps, rounds, log,
                                                                  tune_nodes)
        tried += n
    sess.execute()
    sess.finish()
    light = sess.array(N.BAKE_LIGHT, np.float32)
    err = float(np.max(np.abs(light - ref_light) / (np.abs(ref_light) + 1e-3)))
    if not err < 1e-4:
        for p, t in start.items():
            gpu.set_tuning(pass_=p, **t)
        return {"config": {p.name.lower(): t for p, t in start.items()}, "tried": tried,
                "rejected": picks, "max_rel_diff": err}
    return {"config": picks, "ms": ms, "tried": tried, "max_rel_diff": err}
