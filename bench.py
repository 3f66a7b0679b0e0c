"""Benchmark: Grays/s of the 100 x 100 m, 3-bounce bake on one MI355X.

A step is the device part of one bake of the map (lightbaker_amd.scenes.map_scene(2): a 256 x 256
fBm heightfield plus 400 boxes and cylinders, 16 materials, 8 texels/m, 64 AO rays, 16 sun rays,
256 gather rays, 3 bounces, volume every 2 m to 20 m): the direct pass, its texel passes, the
gather set, three bounces (gather + texel passes, replayed from a hipGraph) and the volume pass,
all on inputs already resident in HBM. `value` is the rays those passes trace per second
(counted by the kernels' own counters in a separate run), over the wall time of K such steps.

Also reported: the roofline of the dominant kernel (the gather pass), the PLAIN form's time per
step on the same map, the gather pass through the reference's host-buffer boundary (lb_run, the
PCIe copies inside the call) beside the same pass on resident buffers (`config.host_boundary`),
and the oracle (the C restatement of the passes) timed on the host's cores on a bounded sample of
the gather pass.

`--workload rebake` measures a rebake instead: lightbaker_amd.scenes.rebake_scene(3), a 200 x
200 m map in 64 tiles of 25 m, each laid out at its own density of 4-16 texels/m (texel counts
more than tenfold apart), baked once whole (its
tiles' light is the cache), then a seeded subset of its tiles and their probe chunks rebaked; a
step is the device part of that rebake, every dirty tile's items in one list per pass.

`--workload city` measures the 400 x 400 m city (lightbaker_amd.scenes.city_scene(4): 2 000
buildings, 1.24 M triangles, 8 texels/m, 3 bounces) the same way, with the autotuner's earlier
picks on the city (`--autotune` runs it again); its CPU
leg compares the PLAIN form and the TUNED form with quantized and FP32 nodes against the FP64
tracer.

    python bench.py [--gpus 1] [--steps 5] [--warmup 2] [--form tuned|plain] [--no-cpu]
                    [--workload map|rebake|city]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E


def node_bytes(stats: dict, form: int, nodes_layout: int) -> int:
    """Algorithmic bytes of one pass from its traversal counters, with each form's record sizes:
    PLAIN reads the node (32 B) at every visit and both children (2 x 32 B) at an inner visit
    (trace.wgsl:90,111-112); TUNED reads one record with both children's boxes and references at
    an inner visit (WIDE 2 x 32 B, QUANT 32 B) and nothing at a leaf. Every triangle test 48 B,
    every material load 16 B, every bounced() 32 B of chart rows and 16 B of light texel."""
    if form == 0:
        nb = (stats["inner_visits"] + stats["leaf_visits"]) * 32 + stats["inner_visits"] * 64
    else:
        nb = stats["inner_visits"] * (64 if nodes_layout == 0 else 32)
    return nb + stats["tri_tests"] * 48 + stats["material_loads"] * 16 + stats["light_gathers"] * 48


def kernel_record(kernel: str) -> dict | None:
    """What rocprofv3 measured for a launch of `kernel` on the map's gather pass, recorded with
    the commands that measured it in profiles/gather_map_scene_roofline.json: HBM bytes
    (FETCH_SIZE and WRITE_SIZE, scaled by what the counters read for a copy of known size), VALU
    and SALU wave instructions, the launch's GPU cycles (GRBM_GUI_ACTIVE per XCD) and SIMDs; None
    when no record names the kernel."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                        "gather_map_scene_roofline.json")
    try:
        with open(path) as f:
            recs = json.load(f)["kernels"]
    except (OSError, ValueError, KeyError):
        return None
    return recs.get(kernel)


def roofline(kernel: str, ms: float, algorithmic_bytes: float, record: dict | None) -> dict:
    """The roofline position of one launch of the dominant kernel.

    HBM: the bytes rocprofv3 measured a launch moving (FETCH + WRITE, calibrated), over the launch
    time measured here with HIP events, against the 8 TB/s peak. Instruction issue: the VALU wave
    instructions the counters measured, over the launch's cycles times the SIMDs (each SIMD
    issues at most one VALU wave instruction a cycle), from the profiled run alone. The
    traversal's algorithmic bytes (every node, triangle, material and light read a ray makes, by
    the kernel's own counters) are reported beside them: most are served by L2 and the MALL, so
    their rate is no HBM figure."""
    out = {"bound": "hbm", "achieved": None, "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "frac": None, "traffic": None, "kernel": kernel, "ms_per_launch": ms,
           "algorithmic": {"bytes_per_launch": algorithmic_bytes,
                           "gbs": algorithmic_bytes / (ms / 1e3) / 1e9,
                           "note": "demand on the memory system, mostly L2 and MALL hits"},
           "issue": None, "binding": None, "source": None}
    if record:
        hbm = record["hbm_bytes"]
        out["traffic"] = hbm
        out["achieved"] = hbm / (ms / 1e3) / 1e9
        out["frac"] = out["achieved"] / HBM_PEAK_GBS
        cycles = record["gui_active_cycles"]
        issue = record["valu_insts"] / (cycles * record["simds"])
        out["issue"] = {"unit": "VALU wave instructions a SIMD cycle", "achieved": issue,
                        "peak": 1.0, "frac": issue,
                        "salu_per_valu": record["salu_insts"] / record["valu_insts"],
                        "lanes_active": record.get("lanes_active"),
                        "ms_profiled": record.get("ms_profiled")}
        out["binding"] = "instruction issue" if issue > out["frac"] else "hbm"
        out["source"] = "profiles/gather_map_scene_roofline.md"
    return out


def cpu_baseline(sess, gpu, volume_rays: int, target_s: float = 5.0,
                 max_items: int = 500000) -> dict:
    """The CPU leg: the oracle (oracle/trace_ref.c, FP32, OpenMP over items) timed on a sample of
    the gather pass (sized to about `target_s` seconds, then the median of 3 runs after that
    warm-up), and the error of the PLAIN form and of the TUNED form with each node layout
    against the oracle's FP64 tracer on identical rays (gather: the gather items, a random
    `max_items` of them where there are more; direct: 200 000 random covered texels; volume:
    20 000 random probes), with the final light atlas."""
    import oracle
    from lightbaker_amd import _native as N
    from lightbaker_amd.gpu import Bvh, Pass
    texels = sess.array(N.BAKE_TEXELS, N.TEXEL)
    gat = sess.array(N.BAKE_GATHER_INDEX, np.uint32)
    nodes = sess.array(N.BAKE_NODES, N.NODE)
    tris = sess.array(N.BAKE_TRIS, N.TRI)
    mats = sess.array(N.BAKE_MATERIALS, np.float32).reshape(-1, 4)
    charts = sess.array(N.BAKE_CHARTS, np.float32).reshape(-1, 8)
    params = sess.array(N.BAKE_PARAMS, N.PARAMS)
    light = sess.array(N.BAKE_LIGHT, np.float32).reshape(-1, 4)
    rays_per_item = int(params["counts"][0][2])
    items = texels[gat]
    rng = np.random.default_rng(0)
    order = rng.permutation(len(items))

    def run(n):
        sample = np.ascontiguousarray(items[order[:n]])
        t0 = time.perf_counter()
        oracle.run(1, nodes, tris, mats, charts, params, sample, 0, light)
        return time.perf_counter() - t0

    n = 256
    dt = run(n)
    while dt < 1.0 and n < len(items):
        n = min(len(items), n * 4)
        dt = run(n)
    n = int(min(len(items), max(256, n * target_s / max(dt, 1e-3))))
    runs = sorted(run(n) for _ in range(3))
    dt = runs[1]
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    base = {"value": n * rays_per_item / dt / 1e9, "unit": "Grays/s", "cores": threads,
            "kind": "port",
            "sample": f"{n} gather items x {rays_per_item} rays of the last bounce "
                      f"(random subset, final light atlas), median of 3 runs "
                      f"({', '.join(f'{t:.2f}' for t in runs)} s)"}

    # error against FP64 on identical rays; the key of an item is its index in the pass
    sb = gpu.scene(Bvh(nodes, tris), mats, charts)

    def forms(pass_, its, lt, prm=params):
        out = {}
        gpu.set_form(N.FORM_PLAIN)
        out["plain"] = gpu.run(pass_, sb, prm, its, lt)
        gpu.set_form(N.FORM_WATERTIGHT)
        out["plain_watertight"] = gpu.run(pass_, sb, prm, its, lt)
        gpu.set_form(N.FORM_TUNED)
        keep = gpu.tuning(pass_)["nodes"]
        for name, layout in (("tuned_quant", N.NODES_QUANT), ("tuned_wide", N.NODES_WIDE)):
            gpu.set_tuning(pass_=pass_, nodes=layout)
            out[name] = gpu.run(pass_, sb, prm, its, lt)
        gpu.set_tuning(pass_=pass_, nodes=keep)
        gpu.set_form(N.FORM_PLAIN)
        return out

    def errors(got, ref, cols, direct=False):
        r = ref[:, cols]
        scale = float(np.max(np.abs(r))) or 1.0
        res = {}
        for name, g in got.items():
            d = g[:, cols].astype(np.float64) - r
            res[name] = {"rmse": float(np.sqrt(np.mean(d * d))),
                         "max_abs_of_full_scale": float(np.max(np.abs(d))) / scale}
        # items the watertight test changes in a hit-derived count (back share; for the direct
        # pass also shadow bits, pick and sun share) or beyond its rounding of t (1e-4 relative):
        # rays that met another triangle or none
        a, b = got["plain"], got["plain_watertight"]
        c = [0, 2] if direct else [0, 1, 2]
        differ = ~np.isclose(a[:, c], b[:, c], rtol=1e-4, atol=1e-5).all(axis=1)
        differ |= a[:, 3] != b[:, 3]
        if direct:
            differ |= a[:, 1].view(np.uint32) != b[:, 1].view(np.uint32)
            differ |= a[:, 0] != b[:, 0]
        res["watertight_items_differing"] = float(np.mean(differ))
        return res

    acc = {}
    # a sample keeps its items' keys: each run gets the same items in the same order
    pick = np.sort(rng.permutation(len(gat))[:max_items]) if len(gat) > max_items else slice(None)
    g_items = np.ascontiguousarray(texels[gat][pick])
    got = forms(Pass.GATHER, g_items, light)
    ref = oracle.run64(1, nodes, tris, mats, charts, params, g_items, 0, light)
    acc["gather_irradiance"] = dict(errors(got, ref, [0, 1, 2]), items=len(g_items))
    cov = np.nonzero(texels["pos"][:, 3] > 0.5)[0]
    d_items = np.ascontiguousarray(texels[cov[rng.permutation(len(cov))[:200000]]])
    got = forms(Pass.DIRECT, d_items, None)
    ref = oracle.run64(0, nodes, tris, mats, charts, params, d_items, 0, None)
    acc["direct_occlusion"] = dict(errors(got, ref, [2], True), items=len(d_items))
    acc["direct_soft_sun"] = dict(errors(got, ref, [0], True), items=len(d_items))
    # the volume (trace.wgsl:272-324): a sample of probes, each keyed by its index in the grid
    # (x fastest, then z, then y: volume.rs:39), its L0 and L1 radiance sums and its sun share
    inf = sess.info()
    nx, ny, nz = inf["volume_size"]
    if nx * ny * nz:
        k = np.sort(rng.permutation(nx * ny * nz)[:20000])
        f32 = np.float32
        o, step = np.asarray(inf["volume_origin"], f32), f32(inf["volume_step"])
        p_items = np.zeros(len(k), N.TEXEL)
        p_items["pos"][:, 0] = o[0] + (k % nx).astype(f32) * step
        p_items["pos"][:, 1] = o[1] + (k // (nx * nz)).astype(f32) * step
        p_items["pos"][:, 2] = o[2] + ((k // nx) % nz).astype(f32) * step
        p_items["pos"][:, 3] = 1.0
        p_items["normal"][:, 1] = 1.0
        p_items["normal"][:, 3] = k.astype(np.uint32).view(f32)
        pv = params.copy()
        pv["counts"][0][2] = volume_rays
        got = {n: v.reshape(-1, 16) for n, v in forms(Pass.VOLUME, p_items, light, pv).items()}
        ref = oracle.run64(2, nodes, tris, mats, charts, pv, p_items, 0, light).reshape(-1, 16)
        acc["volume_l0"] = dict(errors(got, ref, [0, 1, 2]), items=len(p_items))
        acc["volume_l1"] = dict(errors(got, ref, [4, 5, 6, 8, 9, 10, 12, 13, 14]),
                                items=len(p_items))
        acc["volume_sun"] = dict(errors(got, ref, [3]), items=len(p_items))
    sb.close()
    base["error_vs_fp64"] = acc

    # the tile dependencies' filter (deps.rs:104-121) in the oracle, one core, 4 zones
    from oracle import deps_ref
    pos = tris_of(sess)
    zones = np.tile(np.array([[0, 0, 0, 30, 10, 30]], np.float32), (4, 1))
    zones[:, [0, 2, 3, 5]] += np.arange(4, dtype=np.float32)[:, None] * 20
    sun = params["sun_dir"][0][:3].astype(np.float32)
    t0 = time.perf_counter()
    deps_ref.bits(pos, None, zones, sun, np.float32(pos[:, 1::3].max()))
    dt = time.perf_counter() - t0
    base["deps_filter"] = {"value": len(zones) * len(pos) / dt / 1e9, "unit": "Gpairs/s",
                           "cores": 1, "sample": f"4 zones x {len(pos)} triangles"}
    return base


def host_boundary(sess, gpu, form: int, reps: int = 3) -> dict:
    """The gather pass through the reference's own boundary, lb_run (gpu/mod.rs:140-196: host
    texels and light in, host results out, so the copies over PCIe are inside the call), beside
    lb_run_device on the same items and light already in HBM. The bake's gather items and final
    light atlas, the bake's tuning, unranked (lb_run has no first bounce to rank by); the median
    of `reps` calls after one warm-up each way."""
    from lightbaker_amd import _native as N
    from lightbaker_amd.gpu import Bvh, DeviceArray, Pass
    texels = sess.array(N.BAKE_TEXELS, N.TEXEL)
    items = np.ascontiguousarray(texels[sess.array(N.BAKE_GATHER_INDEX, np.uint32)])
    light = np.ascontiguousarray(sess.array(N.BAKE_LIGHT, np.float32).reshape(-1, 4))
    params = sess.array(N.BAKE_PARAMS, N.PARAMS)
    sb = gpu.scene(Bvh(sess.array(N.BAKE_NODES, N.NODE), sess.array(N.BAKE_TRIS, N.TRI)),
                   sess.array(N.BAKE_MATERIALS, np.float32).reshape(-1, 4),
                   sess.array(N.BAKE_CHARTS, np.float32).reshape(-1, 8))
    gpu.set_form(form)
    n = len(items)

    def timed(fn):
        fn()
        ts = []
        for _ in range(reps):
            gpu.sync()
            t0 = time.perf_counter()
            fn()
            gpu.sync()
            ts.append((time.perf_counter() - t0) * 1000.0)
        return float(np.median(ts))

    ms_host = timed(lambda: gpu.run(Pass.GATHER, sb, params, items, light))
    d_items, d_light = DeviceArray.of(gpu, items), DeviceArray.of(gpu, light)
    d_out = DeviceArray(gpu, n * 16)
    ms_dev = timed(lambda: gpu.run_device(Pass.GATHER, form, sb, params, d_items, n, d_light,
                                          len(light), d_out))
    for a in (d_items, d_light, d_out):
        a.free()
    sb.close()
    gpu.set_form(N.FORM_PLAIN)
    rays = n * int(params["counts"][0][2])
    return {"pass": "gather", "items": n, "rays": rays,
            "bytes_to_device": items.nbytes + light.nbytes, "bytes_to_host": n * 16,
            "ms_host_buffers": ms_host, "grays_host_buffers": rays / ms_host / 1e6,
            "ms_device_resident": ms_dev, "grays_device_resident": rays / ms_dev / 1e6}


def tris_of(sess) -> np.ndarray:
    """The scene's triangles' corners (n, 9), from the bake's triangle records."""
    from lightbaker_amd import _native as N
    t = sess.array(N.BAKE_TRIS, N.TRI)
    v0 = t["v0"][:, :3]
    return np.concatenate([v0, v0 + t["e1"][:, :3], v0 + t["e2"][:, :3]], 1).astype(np.float32)


def deps_timing(gpu, scene) -> dict:
    """The tile dependencies (deps.rs:72-121, bake.rs:124-128): every tile's and probe chunk's
    hash with the triangle filter on the device, end to end (host streams and BLAKE3 included)."""
    from lightbaker_amd import deps
    g = deps.global_hash(scene)
    deps.tile_hashes(gpu, scene, g)  # warm
    t0 = time.perf_counter()
    h = deps.tile_hashes(gpu, scene, g)
    s_tiles = time.perf_counter() - t0
    t0 = time.perf_counter()
    keys, _ = deps.probe_hashes(gpu, scene, g)
    s_probes = time.perf_counter() - t0
    n_tris = scene.triangle_count()
    # the filter alone (lb_deps_swept: upload, one launch, read-back) over as many zones, each a
    # tile-sized box somewhere on the scene, as the bake has tiles and chunks
    import ctypes as C
    from lightbaker_amd import _native as N
    pos = np.concatenate([m.positions.reshape(-1, 9) for m in scene.meshes]).astype(np.float32)
    nz = len(h) + len(keys)
    rng = np.random.default_rng(0)
    lo, hi = pos.reshape(-1, 3).min(0), pos.reshape(-1, 3).max(0)
    c = rng.uniform(lo, hi, (nz, 3)).astype(np.float32)
    half = np.float32(scene.settings.tile_size / 2 + scene.settings.gi_reach)
    zones = np.ascontiguousarray(np.concatenate([c - half, c + half], 1), np.float32)
    sun = np.asarray(scene.sun_direction, np.float32)
    sun = np.ascontiguousarray(sun / np.linalg.norm(sun), np.float32)
    bits = np.zeros((nz, (len(pos) + 31) // 32), np.uint32)
    top = float(pos[:, 1::3].max())
    for _ in range(2):
        t0 = time.perf_counter()
        N.check(N.lib().lb_deps_swept(gpu.handle, N.ptr(pos), len(pos), None, N.ptr(zones), nz,
                                      N.ptr(sun), C.c_float(top), N.ptr(bits)))
        s_filter = time.perf_counter() - t0
    return {"tiles": len(h), "probe_chunks": len(keys), "triangles": n_tris,
            "pairs": nz * n_tris, "s_tile_hashes": s_tiles, "s_probe_hashes": s_probes,
            "s_filter": s_filter, "filter_gpairs_per_s": nz * n_tris / s_filter / 1e9}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--form", choices=["tuned", "plain"], default="tuned")
    ap.add_argument("--nodes", choices=["wide", "quant"], default="quant")
    ap.add_argument("--schedule", type=int, choices=[0, 1], default=0,
                    help="TUNED schedule before autotuning: 0 lock-step, 1 pooled")
    ap.add_argument("--workload", choices=["map", "rebake", "city"], default="map")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--gather-order", type=int, choices=[0, 1, 2], default=None,
                    help="lock-step gather: 0 azimuth then ranked, 1 elevation bands then "
                         "ranked, 2 elevation bands every bounce; default: the bench's")
    ap.add_argument("--xcd-split", type=int, choices=[0, 1], default=None,
                    help="lock-step passes: items split over the 8 chiplets' queues (1) or one "
                         "queue (0); default: the library's")
    ap.add_argument("--autotune", action="store_true",
                    help="autotune also where the workload skips it by default (city)")
    ap.add_argument("--profile-pass", choices=["direct", "gather", "volume"], default="gather",
                    help="the pass --profile-gather launches")
    ap.add_argument("--profile-gather", type=int, default=0, metavar="N",
                    help="prepare, execute once, then N gather launches only (for counters)")
    ap.add_argument("--copy-kernel", type=int, default=0, metavar="MIB",
                    help="only copy MIB MiB device-to-device with torch (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path (asked for "
              f"{args.gpus}); run with --gpus 1", file=sys.stderr)
        return 2

    if args.copy_kernel:
        import torch
        n = args.copy_kernel * (1 << 20) // 4
        a = torch.ones(n, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(f"copied {args.copy_kernel} MiB three times", file=sys.stderr)
        return 0

    from lightbaker_amd import _native as N
    from lightbaker_amd.bake import BakeSession, bake
    from lightbaker_amd.gpu import Gpu, Pass
    from lightbaker_amd.scenes import city_scene, map_scene, rebake_scene

    form = N.FORM_TUNED if args.form == "tuned" else N.FORM_PLAIN
    layout = N.NODES_WIDE if args.nodes == "wide" else N.NODES_QUANT
    t_scene = time.perf_counter()
    scene = {"map": lambda: map_scene(2), "rebake": lambda: rebake_scene(3),
             "city": lambda: city_scene(4)}[args.workload]()
    if args.workload == "city" and not args.autotune:
        # one city bake is 15 G rays: the autotuner's hundred trial launches take minutes, so by
        # default the picks it made on the city before are used
        args.no_autotune = True
    t_scene = time.perf_counter() - t_scene
    gpu = Gpu()
    gpu.set_tuning(256, 4, 16, layout, 32, args.schedule, leaf_batch=16, waves=6)
    # the autotuner's usual picks on this map, so a run without it uses the same kernels
    gpu.set_tuning(pass_=Pass.GATHER, rays_per_lane=8, waves=8, xcd_split=1)
    gpu.set_tuning(pass_=Pass.DIRECT, rays_per_lane=8, waves=8, leaf_batch=8)
    if args.workload == "city":
        # the autotuner's picks on the city, where they differ from the map's: a shallower LDS
        # stack for the gather (its tree is taller, the spill rarer than the LDS it frees), a
        # deeper one for the direct pass
        gpu.set_tuning(pass_=Pass.GATHER, lds_stack=8, leaf_batch=16)
        gpu.set_tuning(pass_=Pass.DIRECT, lds_stack=16)
    if args.xcd_split is not None:
        for p in Pass:
            gpu.set_tuning(pass_=p, xcd_split=args.xcd_split)
    if args.gather_order is not None:
        gpu.set_tuning(pass_=Pass.GATHER, gather_order=args.gather_order)
    print(f"# {gpu.adapter}; scene {scene.triangle_count()} triangles ({t_scene:.1f} s)",
          file=sys.stderr, flush=True)
    rebake = {}
    if args.workload == "rebake":
        # the cache: one whole bake; then a seeded share of the tiles and their probe chunks
        first = bake(scene, gpu, form=N.FORM_TUNED, graph=True)
        rng = np.random.default_rng(33)
        nt = len(first.tiles)
        dirty = np.zeros(nt, bool)
        dirty[rng.choice(nt, int(rng.integers(1, nt + 1)), replace=False)] = True
        keys = {(int(first.tiles[t, 0]), int(first.tiles[t, 1])) for t in np.nonzero(dirty)[0]}
        v = scene.volume
        nx, ny, nz = first.info["volume_size"]
        k = np.arange(nx * ny * nz)
        f = np.float32
        px = f(v.min[0]) + (k % nx).astype(f) * f(v.step)
        pz = f(v.min[2]) + ((k // nx) % nz).astype(f) * f(v.step)
        size = f(scene.settings.tile_size)
        pk = zip(np.floor(px / size).astype(int), np.floor(pz / size).astype(int))
        pdirty = np.array([(int(a), int(b)) in keys for a, b in pk], bool)
        rebake = {"cache": first.cache, "tile_dirty": dirty, "probe_dirty": pdirty}
        del first
    t0 = time.perf_counter()
    sess = BakeSession(scene, gpu, form=form, graph=True, **rebake)
    prep_s = time.perf_counter() - t0
    if args.profile_gather:
        sess.execute()
        pp = {"direct": Pass.DIRECT, "gather": Pass.GATHER, "volume": Pass.VOLUME}[args.profile_pass]
        ms = sess.time_pass(pp, reps=args.profile_gather)
        print(f"# {args.profile_gather} {args.profile_pass} launches, {ms:.2f} ms each",
              file=sys.stderr)
        return 0
    tuned = None
    if form == N.FORM_TUNED and not args.no_autotune:
        from lightbaker_amd.autotune import autotune
        t0 = time.perf_counter()
        sess.execute()
        tuned = autotune(sess, log=lambda m: print("# " + m, file=sys.stderr, flush=True))
        tuned["seconds"] = time.perf_counter() - t0
        print(f"# autotune: {tuned}", file=sys.stderr, flush=True)
        layout = gpu.tuning(Pass.GATHER)["nodes"]
    stats = sess.execute(stats=True)
    rays = sum(s["rays"] for s in stats)
    info = sess.info()
    print(f"# prepared in {prep_s:.1f} s: {info['texels_baked']} texels, "
          f"{info['gather_items']} gather items, {info['probes']} probes, "
          f"{info['bvh_nodes']} nodes; rays/step {rays}", file=sys.stderr, flush=True)
    for _ in range(args.warmup):
        sess.execute()
    gpu.sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        sess.execute()
    gpu.sync()
    dt = time.perf_counter() - t0
    ms_step = dt / args.steps * 1000.0
    info = sess.info()
    t0 = time.perf_counter()
    sess.finish()
    finish_s = time.perf_counter() - t0

    # the tail on the device (csrc/atlas.hip): algorithmic bytes of its kernels, per baked texel
    # (TileLight: 47 B read, 34 B written), per atlas texel (the atlas: 22 B read where a tile
    # lies, 9 B written; BC5: 8 B read, 1 B written), per indirect texel (3 x 13 B for the sums'
    # output and two rings, RGB9E5: 12 B read, 4 B written) and per probe (64 B read, 16 written)
    fin = sess.info()
    n_atlas = fin["atlas_width"] * fin["atlas_height"]
    n_ind = fin["indirect_width"] * fin["indirect_height"]
    fin_bytes = (fin["texels"] * 81 + fin["texels"] * 22 + n_atlas * (9 + 9)
                 + n_ind * (3 * 13 + 16) + fin["probes"] * 80)

    bounces = scene.settings.gi_bounces
    ms_gather = sess.time_pass(Pass.GATHER, reps=5)
    g = stats[1]
    per_launch = {k: v / bounces for k, v in g.items()}
    bytes_launch = node_bytes(per_launch, form, layout) + info["gather_items"] * (64 + 16)
    tuning = gpu.tuning()
    if form == N.FORM_PLAIN:
        kernel = "k_gather_plain<false, false>"
    elif tuning["schedule"] == 0:
        kernel = f"k_gather_tuned<{layout}, 0, {tuning['waves']}>"
    else:
        kernel = f"k_pool<0, {layout}, false>"
    # the recorded counters are the map's gather launch
    roof = roofline(kernel, ms_gather, bytes_launch,
                    kernel_record(kernel) if args.workload == "map" else None)
    print(f"# gather launch {ms_gather:.2f} ms, {bytes_launch / 1e9:.3f} GB algorithmic, "
          f"{per_launch['rays'] / ms_gather / 1e6:.2f} Grays/s", file=sys.stderr, flush=True)

    # the PLAIN form on the same bake: the median of 5 steps after 2 warm-ups, and its gather
    # launch's roofline
    plain_ms = plain_roof = plain_steps = None
    if not args.no_plain and form == N.FORM_TUNED:
        ps = BakeSession(scene, gpu, form=N.FORM_PLAIN, graph=True, **rebake)
        pstats = ps.execute(stats=True)
        for _ in range(2):
            ps.execute()
        plain_steps = []
        for _ in range(5):
            gpu.sync()
            t0 = time.perf_counter()
            ps.execute()
            gpu.sync()
            plain_steps.append((time.perf_counter() - t0) * 1000.0)
        plain_ms = float(np.median(plain_steps))
        pms = ps.time_pass(Pass.GATHER, reps=5)
        pl = {k: v / bounces for k, v in pstats[1].items()}
        pbytes = node_bytes(pl, N.FORM_PLAIN, layout) + info["gather_items"] * (64 + 16)
        pk = "k_gather_plain<false, false>"
        plain_roof = roofline(pk, pms, pbytes,
                              kernel_record(pk) if args.workload == "map" else None)
        ps.close()

    # the same bake with the later bounces replaying the first bounce's ray records (the same
    # rays, so the same results; the traced rays a step drop to direct + one bounce + volume)
    replay_ms = None
    if form == N.FORM_TUNED and scene.settings.gi_bounces > 1:
        rs = BakeSession(scene, gpu, form=form, graph=True, replay=True, **rebake)
        rs.execute()
        gpu.sync()
        t0 = time.perf_counter()
        for _ in range(max(args.steps, 1)):
            rs.execute()
        gpu.sync()
        replay_ms = (time.perf_counter() - t0) * 1000.0 / max(args.steps, 1)
        rs.close()
        print(f"# replayed bounces: {replay_ms:.2f} ms a step", file=sys.stderr, flush=True)

    cpu = None if args.no_cpu else cpu_baseline(
        sess, gpu, scene.volume.rays if scene.volume is not None else 1)
    dep = deps_timing(gpu, scene)
    hostb = host_boundary(sess, gpu, form) if args.workload == "map" else None
    if args.workload == "map":
        metric = "Grays/s, 100x100 m 3-bounce bake (direct + 3 gather bounces + volume)"
        workload = "map_scene(2): 100x100 m, 8 texels/m, 256 GI rays, 3 bounces, 16 tiles"
    elif args.workload == "city":
        metric = "Grays/s, 400x400 m 1.2M-triangle city 3-bounce bake"
        workload = ("city_scene(4): 400x400 m, 2 000 buildings, 1.24 M triangles, 8 texels/m, "
                    "64 GI rays, 3 bounces, 256 tiles")
    else:
        metric = "Grays/s, rebake of a seeded share of a 200x200 m map's 64 tiles"
        workload = (f"rebake_scene(3): 200x200 m, 4-16 texels/m by tile, 64 GI rays, 1 bounce, "
                    f"{info['tiles_baked']} of {info['tiles']} tiles dirty, "
                    f"{info['probes_baked']} of {info['probes']} probes")
    out = {
        "metric": metric,
        "value": rays / (ms_step / 1e3) / 1e9,
        "unit": "Grays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": {"map": "synthetic: seeded map generator (fBm terrain + boxes + cylinders), "
                        "no dataset",
                 "rebake": "synthetic: seeded generator (yawed boxes on a ground of 5 m quads), "
                           "no dataset",
                 "city": "synthetic: seeded city generator (prisms of 14-400 sides on a plane), "
                         "no dataset"}[args.workload],
        "config": {
            "workload": workload,
            "form": args.form, "nodes": "wide" if layout == N.NODES_WIDE else "quant",
            "graph": True, "tuning": tuned,
            "triangles": scene.triangle_count(), "bvh_nodes": info["bvh_nodes"],
            "texels": info["texels_baked"], "gather_items": info["gather_items"],
            "probes": info["probes"], "rays_per_step": rays,
            "rays_direct": stats[0]["rays"], "rays_gather": stats[1]["rays"],
            "rays_volume": stats[2]["rays"],
            "ms_direct": info["ms_direct"], "ms_direct_texel": info["ms_direct_texel"],
            "ms_bounces": info["ms_bounces"], "ms_volume": info["ms_volume"],
            "s_prepare_host": prep_s, "s_finish_host": finish_s,
            "s_prepare_parts": {k: info[k] for k in ("s_layout", "s_bvh", "s_raster", "s_total")},
            "ms_finish_device": fin["ms_finish"],
            "finish_bytes": fin_bytes, "finish_gbps": fin_bytes / (fin["ms_finish"] / 1e3) / 1e9
            if fin["ms_finish"] > 0 else None,
            "s_per_bake_end_to_end": prep_s + ms_step / 1e3 + finish_s,
            "tiles_baked": info["tiles_baked"],
            "ms_per_tile": ms_step / max(info["tiles_baked"], 1),
            "plain_ms_per_step": plain_ms,
            "plain_ms_steps": plain_steps,
            "plain_roofline": plain_roof,
            "tuned_vs_plain": (plain_ms / ms_step) if plain_ms else None,
            "replay_ms_per_step": replay_ms,
            "replay_rays_traced_per_step": (stats[0]["rays"] + stats[1]["rays"] // bounces
                                            + stats[2]["rays"]) if replay_ms else None,
            "replay_vs_plain": (plain_ms / replay_ms) if plain_ms and replay_ms else None,
            "deps": dep,
            "host_boundary": hostb,
        },
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))
    sess.close()
    gpu.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
