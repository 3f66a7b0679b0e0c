"""The TUNED form's autotuner: picks the workgroup size, rays per lane, LDS stack depth, refill
threshold and node layout with the best time on a prepared bake's own passes, checking that
every candidate gives the same results as the starting configuration.

Search: coordinate descent from the current setting, one dimension at a time over its range
(schedule lock-step/pooled, workgroup 64..512, rays/lane 1..8 for lock-step, LDS stack 8..48,
refill and leaf batch for pooled, WIDE/QUANT nodes), at most two rounds; each
candidate timed on the gather pass (weighted by the bounces) and the direct pass,
the median of `reps` launches. A candidate the device cannot launch (an LDS stack larger than a
workgroup's LDS) is skipped.
"""
from __future__ import annotations

import numpy as np

from . import _native as N
from .gpu import Pass

RANGES = {
    "workgroup": [64, 128, 256, 512],
    "rays_per_lane": [1, 2, 4, 8],
    "lds_stack": [8, 16, 24, 32, 48],
    "refill": [16, 32, 48, 56],
    "nodes": [N.NODES_WIDE, N.NODES_QUANT],
    "schedule": [0, 1],
    "leaf_batch": [8, 16, 24, 32, 64],
}
ORDER = ["schedule", "rays_per_lane", "refill", "leaf_batch", "workgroup", "lds_stack", "nodes"]


def _cost(sess, bounces: int, reps: int) -> float:
    g = sorted(sess.time_pass(Pass.GATHER, 1) for _ in range(reps))[reps // 2]
    d = sorted(sess.time_pass(Pass.DIRECT, 1) for _ in range(reps))[reps // 2]
    return g * max(bounces, 1) + d


def autotune(sess, bounces: int, reps: int = 3, rounds: int = 2, log=None) -> dict:
    """Tunes the session's GPU in place (the session must be TUNED and executed once)."""
    gpu = sess.gpu
    cur = gpu.tuning()

    def apply(c):
        gpu.set_tuning(**c)

    def measure(c):
        try:
            apply(c)
            return _cost(sess, bounces, reps)
        except N.NativeError:
            return float("inf")

    start = dict(cur)
    sess.execute()
    sess.finish()
    ref_light = sess.array(N.BAKE_LIGHT, np.float32).copy()
    best = measure(cur)
    history = [(dict(cur), best)]
    for _ in range(rounds):
        changed = False
        for dim in ORDER:
            for v in RANGES[dim]:
                if v == cur[dim]:
                    continue
                if dim in ("refill", "leaf_batch") and cur["schedule"] == 0:
                    continue
                if dim == "rays_per_lane" and cur["schedule"] == 1:
                    continue
                cand = dict(cur, **{dim: v})
                t = measure(cand)
                history.append((cand, t))
                if log:
                    log(f"autotune {cand}: {t:.2f} ms")
                if t < best * 0.995:
                    best, cur, changed = t, cand, True
        if not changed:
            break
    apply(cur)
    # the pick must bake what the starting setting bakes (ray sums to compensated rounding)
    sess.execute()
    sess.finish()
    light = sess.array(N.BAKE_LIGHT, np.float32)
    err = float(np.max(np.abs(light - ref_light) / (np.abs(ref_light) + 1e-3)))
    if not err < 1e-4:
        apply(start)
        return {"config": start, "ms": history[0][1], "tried": len(history),
                "rejected": cur, "max_rel_diff": err}
    return {"config": cur, "ms": best, "tried": len(history), "max_rel_diff": err}
