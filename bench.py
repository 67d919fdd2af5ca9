#!/usr/bin/env python3
"""Train-step throughput of the MI355X tiled radiance-mesh pipeline.

Workload (default): 200,000 vertices in 8 Gaussian clusters (sigma 0.15) inside
[-1, 1]^3, Delaunay (~1.35M tets), SH degree 3 fp32, 16 cameras on a Fibonacci
sphere (radius 3, 50 deg) at 800x800; one step = forward + backward of all 16
views + one Adam update (lr_sh 2.5e-3, lr_verts 1e-4, lr_dens 1e-2,
lr_grads 1e-3), L1 loss against targets rendered from a perturbed held-out
parameter set.  The step is captured once into a hipGraph and replayed.

Prints ONE JSON line.  ``value`` = views/sec over the timed steps (inputs
resident in HBM); ``roofline`` prices the dominant kernel's algorithmic bytes
(per tile-tet pair record fetch, image and gradient traffic) against 8 TB/s,
timed with HIP events recorded on the renderer's stream inside the replayed
graph; ``cpu_baseline`` times the CPU oracle (a C restatement of the
reference kernels, single thread) on a bounded sample of the same workload.
"""

from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "train-step views/sec (fwd+bwd+Adam) at ~1.3M tets 800x800; % HBM peak"
HBM_PEAK_GBS = 8000.0
LR = dict(lr_verts=1e-4, lr_dens=1e-2, lr_grads=1e-3, lr_sh=2.5e-3)


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def build_scene(seed, n_points, views, width):
    from rmeshvk_amd.camera import fibonacci_cameras
    from rmeshvk_amd.scenes import cluster_points, make_scene

    rng = np.random.default_rng(seed)
    scene = make_scene(cluster_points(n_points, rng, 8, 0.15), rng, sh_degree=3)
    cams = fibonacci_cameras(views, 3.0, 50.0, width, width)
    # held-out perturbed parameters for the targets
    held = dict(scene)
    held["densities"] = (scene["densities"] * np.exp(rng.normal(0, 0.2, scene["densities"].shape))).astype(np.float32)
    held["sh"] = (scene["sh"] + rng.normal(0, 0.05, scene["sh"].shape)).astype(np.float32)
    held["color_grads"] = (scene["color_grads"] + rng.normal(0, 0.02, scene["color_grads"].shape)).astype(np.float32)
    return scene, held, cams


def make_renderer(scene, width, views, tile_size):
    from rmeshvk_amd import RMeshRenderer

    return RMeshRenderer(scene["vertices"], scene["indices"], scene["base_colors"], scene["densities"],
                         scene["color_grads"], scene["circumdata"], width, width, device_id=0,
                         sh_coeffs=scene["sh"], sh_degree=3, max_views=views, tile_size=tile_size)


def algorithmic_bytes(stage, P, V, W, H, T, N, M):
    """Per-launch algorithmic bytes (96 B per tile-tet pair record + image/gradient terms)."""
    if stage == "forward":
        return 96 * P + V * (16 * W * H + 8 * T)
    if stage == "backward":
        return 96 * P + V * (32 * W * H + 8 * T) + 12 * N + 28 * M
    raise ValueError(stage)


def profiled_traffic(kernel, tile_size):
    """HBM bytes per launch of `kernel` at this tile size from the committed
    rocprofv3 FETCH_SIZE/WRITE_SIZE summary (profiles/*_hbm_traffic.md, counters
    already scaled by the copy-kernel calibration); None when not profiled."""
    import glob
    import re

    best = None
    for path in sorted(glob.glob(os.path.join(ROOT, "profiles", "*_hbm_traffic.md"))):
        for line in open(path):
            m = re.match(r"\| `(.+?)` \| (\d+) \| [\d.]+ \| [\d.]+ \| ([\d.]+) \| ([\d.]+) \|", line)
            if not m or kernel not in m.group(1) or f"<{tile_size}," not in m.group(1).replace(" ", ""):
                continue
            n = int(m.group(2))
            if best is None or n > best[0]:
                best = (n, m.group(1), float(m.group(3)) * 1e9, float(m.group(4)) * 1e9, os.path.basename(path))
    if best is None:
        return None
    return {"bytes": best[2] + best[3], "read": best[2], "written": best[3], "kernel": best[1],
            "source": "profiles/" + best[4]}


def profiled_valu(kernel, tile_size):
    """VALU issue rate and lane utilisation of `kernel` at this tile size from
    the committed rocprofv3 SQ-counter summary (profiles/*_valu_lds.md); None
    when not profiled."""
    import glob
    import re

    best = None
    for path in sorted(glob.glob(os.path.join(ROOT, "profiles", "*_valu_lds.md"))):
        for line in open(path):
            m = re.match(r"\| `(.+?)` \| (\d+) \| ([\d.]+) \| [\d.]+ \| [\d.]+ \| (\d+)% \| (\d+)% \|", line)
            if not m or kernel not in m.group(1) or f"<{tile_size}," not in m.group(1).replace(" ", ""):
                continue
            n = int(m.group(2))
            if best is None or n > best[0]:
                best = (n, m.group(1), float(m.group(3)), int(m.group(4)), int(m.group(5)), os.path.basename(path))
    if best is None:
        return None
    return {"issue": min(best[3], 100) / 100.0, "lane_util": best[4] / 100.0, "valu_instr_G": best[2],
            "kernel": best[1], "source": "profiles/" + best[5]}


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def profiled_atomics(kernel, tile_size):
    """Float atomic requests per launch of `kernel` (TCP->TCC atomics without
    return) from the committed rocprofv3 summary (profiles/*_atomics.md);
    None when not profiled."""
    import glob
    import re

    best = None
    for path in sorted(glob.glob(os.path.join(ROOT, "profiles", "*_atomics.md"))):
        for line in open(path):
            m = re.match(r"\| `(.+?)` \| (\d+) \| ([\d.eE+]+) \|", line)
            if not m or kernel not in m.group(1) or f"<{tile_size}," not in m.group(1).replace(" ", ""):
                continue
            if best is None or int(m.group(2)) > best[0]:
                best = (int(m.group(2)), m.group(1), float(m.group(3)), os.path.basename(path))
    if best is None:
        return None
    return {"requests_per_launch": best[2], "kernel": best[1], "source": "profiles/" + best[3]}


def atomic_ceiling():
    """Measured float-atomic rate of the L2 for scattered lanes (k_atomic_probe:
    2^27 adds, each lane on its own 64-byte line, lines spread over 2^16)."""
    import ctypes

    from rmeshvk_amd import _lib

    ms = ctypes.c_float()
    n = 1 << 27
    _lib.check(_lib.load().rmesh_atomic_probe(0, 1 << 16, n, ctypes.byref(ms)))
    return n / (ms.value * 1e-3)


def cpu_baseline(scene, cams, width, tile_size, warmup=3, timed=10):
    """The CPU oracle (a C restatement of the reference kernels, compositing
    tiles on every host thread OpenMP gives it) on whole views of the
    workload: per view SH colours, projection, pair generation, sort, ranges,
    forward, L1 loss gradient, adjoint and SH adjoint, plus 1/V of one Adam
    update over every parameter (one update per step of V views).  `warmup`
    untimed views, then the median of `timed` views."""
    from oracle import oracle as orc

    W = H = width
    _, _, T = orc.tile_grid(W, H, tile_size)
    n_params = len(scene["vertices"]) + len(scene["densities"]) + len(scene["color_grads"]) + len(scene["sh"])
    z = np.zeros(n_params, np.float32)
    t0 = time.perf_counter()
    orc.adam(z, z, z, z, 1, 1e-3)
    t_adam = time.perf_counter() - t0
    gt = np.zeros((H, W, 3), np.float32)
    times = []
    P = 0
    for i in range(warmup + timed):
        cam = cams[i % len(cams)]
        t0 = time.perf_counter()
        colors = orc.sh_colors(scene, cam[0])
        tt, _ = orc.project(scene, cam, W, H, tile_size)
        keys, vals, P = orc.pairs(scene, cam, W, H, tt, tile_size)
        sk, sv = orc.sort(keys, vals, orc.sorting_bits(T))
        rg = orc.ranges(sk, T)
        fw = orc.forward(scene, cam, W, H, sv, rg, tile_size, colors=colors)
        dl, _ = orc.loss_grad(fw["rgba"], gt, 0)
        g = orc.backward(scene, cam, W, H, sv, rg, fw["rgba"], dl, tile_size, colors=colors)
        orc.sh_backward(scene, cam[0], g["d_base_colors"])
        if i >= warmup:
            times.append(time.perf_counter() - t0 + t_adam / len(cams))
    per_view = float(np.median(times))
    threads = orc.get_threads()
    return {
        "value": 1.0 / per_view,
        "unit": "views/s",
        "cores": threads,
        "kind": "port",
        "cpu_model": _cpu_model(),
        "sample": (f"whole views of the workload ({P} pairs in the last): SH colours, projection, pairs, sort, "
                   f"ranges, forward, L1 loss, adjoint, SH adjoint and 1/{len(cams)} of an Adam update over all "
                   f"{n_params} parameters; {warmup} warm-up views, median of {timed} views "
                   f"(min {min(times):.2f} s, max {max(times):.2f} s), {threads} OpenMP threads"),
        "seconds_measured": float(np.sum(times)),
    }


def mode_summary(r, ms_per_step, stages):
    return {"ms_per_step": round(ms_per_step, 3), "forward_ms": round(stages["forward"], 4),
            "backward_ms": round(stages["backward"], 4), "device_bytes": r.device_bytes()}


def time_steps(r, step, steps, V, warmup=2):
    """Warm-up + timed graph-replayed steps in the renderer's current mode."""
    for _ in range(warmup):
        step()
    acc = {k: 0.0 for k in r.STAGES}
    r.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
        for k, v in r.stage_times().items():
            acc[k] += v
    r.synchronize()
    ms = 1000.0 * (time.perf_counter() - t0) / steps
    st = {k: v / steps for k, v in acc.items()}
    out = mode_summary(r, ms, st)
    out["views_per_sec"] = round(V * 1000.0 / ms, 2)
    out["binning_sort_ms"] = round(st["binning"] + st["pairs"] + st["sort"] + st["ranges"], 4)
    out["sort_ms"] = round(st["sort"], 4)
    return out


def config_e(seed, steps):
    """The large-scene leg: 2M points (16 clusters + 20 % uniform, ~13M tets),
    SH degree 3, 4 views at 1920x1080, forward + backward + Adam per step in
    both adjoint modes (graph-replayed), with device memory held."""
    from rmeshvk_amd import RMeshRenderer
    from rmeshvk_amd.scenes import config_scene

    t0 = time.perf_counter()
    scene, cams = config_scene("E", seed)
    M = len(scene["indices"]) // 4
    log(f"config E scene: {M} tets ({time.perf_counter() - t0:.1f} s)")
    Wd, Hd = 1920, 1080
    args = (scene["vertices"], scene["indices"], scene["base_colors"], scene["densities"], scene["color_grads"],
            scene["circumdata"], Wd, Hd)
    held = dict(scene)
    rng = np.random.default_rng(seed + 1)
    held["densities"] = (scene["densities"] * np.exp(rng.normal(0, 0.2, M))).astype(np.float32)
    tgt = RMeshRenderer(*args[:3], held["densities"], *args[4:], device_id=0, sh_coeffs=scene["sh"], sh_degree=3,
                        max_views=len(cams), tile_size=16)
    targets = np.ascontiguousarray(tgt.forward_batched(cams)[..., :3])
    del tgt
    r = RMeshRenderer(*args, device_id=0, sh_coeffs=scene["sh"], sh_degree=3, max_views=len(cams), tile_size=16)
    r.set_targets(targets)
    r.set_graph_mode(True)
    r.set_timing(True)
    step = lambda: r.train_step_batched(cams, 0, LR["lr_verts"], LR["lr_dens"], LR["lr_grads"], LR["lr_sh"])  # noqa
    out = {"workload": f"2M-point 16-cluster + 20% uniform Delaunay mesh ({M} tets), SH degree 3, "
                       f"{len(cams)} views at {Wd}x{Hd}", "tets": M}
    for mode in ("undo", "checkpoint"):
        r.set_adjoint_mode(mode)
        r.set_graph_mode(True)
        out[mode] = time_steps(r, step, steps, len(cams))
        log(f"config E {mode}: {out[mode]}")
    out["pairs_per_step"] = r.pair_count()
    return out


def config_d(seed, steps):
    """The uneven multi-scene leg: 32 scenes (vertex counts log-uniform in
    [5k, 100k], uniform or clustered; SH degree from {0, 1, 3}; 256^2, 512^2 or
    1024^2), 4 views each, forward + backward + Adam of all 128 views per step
    in one renderer (one launch per stage), graph-replayed, autotuned.  Against
    it: the same scenes stepped one after another, each in its own renderer
    (also graph-replayed and autotuned on its views)."""
    from rmeshvk_amd import MultiSceneRenderer, RMeshRenderer
    from rmeshvk_amd.scenes import config_d_scenes

    t0 = time.perf_counter()
    items = config_d_scenes(seed)
    scenes, cams = [], []
    for sc, cc, res in items:
        sc.update(width=res, height=res, max_views=4)
        scenes.append(sc)
        cams.append(cc)
    log(f"config D: {len(scenes)} scenes, {sum(len(s['indices']) // 4 for s in scenes)} tets "
        f"({time.perf_counter() - t0:.1f} s)")
    nv = [len(c) for c in cams]
    flat = [c for cc in cams for c in cc]
    held = []
    rng = np.random.default_rng(seed + 7)
    for s in scenes:
        h = dict(s)
        h["densities"] = (s["densities"] * np.exp(rng.normal(0, 0.2, s["densities"].shape))).astype(np.float32)
        held.append(h)
    tr = MultiSceneRenderer(held, tile_size=16, device_id=0)
    targets = [im[..., :3].copy() for im in tr.forward(nv, flat)]
    del tr
    mr = MultiSceneRenderer(scenes, tile_size=16, device_id=0)
    mr.set_targets(nv, targets)
    tune = mr.autotune(nv, flat, loss_type=0, reps=2)
    log(f"config D autotune: {tune}")
    mr.set_scene_timing(True)
    lr = (LR["lr_verts"], LR["lr_dens"], LR["lr_grads"], LR["lr_sh"])
    V = sum(nv)
    out = {"workload": f"{len(scenes)} scenes ({sum(len(s['indices']) // 4 for s in scenes)} tets), SH degrees "
                       f"{sorted(set(int(s['sh_degree']) for s in scenes))}, resolutions "
                       f"{sorted(set(int(s['width']) for s in scenes))}, {V} views per step"}
    ms_batch = None
    for order in ("interleaved", "by_scene"):
        # the compositing grids' tile order across scenes: all scenes' tiles
        # heaviest first, or scene after scene (fewest pairs first)
        mr.set_scene_order(order)
        mr.set_graph_mode(True)
        for _ in range(2):
            mr.train_step(nv, flat, 0, *lr)
        lat = []
        mr.synchronize()
        tb = time.perf_counter()
        for _ in range(steps):
            mr.train_step(nv, flat, 0, *lr)
            lat.append(mr.scene_latency())
        mr.synchronize()
        ms = 1000.0 * (time.perf_counter() - tb) / steps
        lat = np.concatenate(lat)
        key = "batched" if order == "interleaved" else "batched_by_scene"
        out[key] = {"ms_per_step": round(ms, 3), "views_per_sec": round(V * 1000.0 / ms, 2),
                    "tuning": tune["tuning"], "sort_digit_bits": tune["sort_digit_bits"], "scene_order": order,
                    "scene_latency_ms_p50": round(float(np.percentile(lat, 50)), 3),
                    "scene_latency_ms_p99": round(float(np.percentile(lat, 99)), 3),
                    "latency_note": "per scene: step start to its last adjoint tile (device wall clock); the "
                                    "shared SH adjoint + Adam stage follows"}
        if order == "interleaved":
            ms_batch = ms
        log(f"config D {order}: {json.dumps(out[key])}")
    del mr
    # the same scenes one renderer at a time
    per = []
    for s, c, t in zip(scenes, cams, targets):
        r = RMeshRenderer(s["vertices"], s["indices"], s["base_colors"], s["densities"], s["color_grads"],
                          s["circumdata"], s["width"], s["height"], device_id=0, sh_coeffs=s["sh"],
                          sh_degree=int(s["sh_degree"]), max_views=len(c), tile_size=16)
        r.set_targets(np.ascontiguousarray(t))
        r.autotune(c, loss_type=0, reps=1)
        r.set_graph_mode(True)
        for _ in range(2):
            r.train_step_batched(c, 0, *lr)
        r.synchronize()
        ts = time.perf_counter()
        for _ in range(steps):
            r.train_step_batched(c, 0, *lr)
        r.synchronize()
        per.append(1000.0 * (time.perf_counter() - ts) / steps)
        del r
    per = np.array(per)
    out["per_scene_renderers"] = {"ms_per_step_all_scenes": round(float(per.sum()), 3),
                                  "views_per_sec": round(V * 1000.0 / float(per.sum()), 2),
                                  "scene_step_ms_p50": round(float(np.percentile(per, 50)), 3),
                                  "scene_step_ms_p99": round(float(np.percentile(per, 99)), 3)}
    out["speedup_batched"] = round(float(per.sum()) / ms_batch, 3)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--views", type=int, default=16)
    ap.add_argument("--width", type=int, default=800)
    ap.add_argument("--points", type=int, default=200000)
    ap.add_argument("--tile-size", type=int, default=12)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--no-autotune", action="store_true", help="keep --tile-size and the default kernel variants")
    ap.add_argument("--tuning", default="",
                    help="TS,F,B[,D]: tile size, kernel variants and sort digit bits, no autotune")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-error", action="store_true", help="skip the forward_error (densification statistics) pass")
    ap.add_argument("--no-interval", action="store_true", help="skip the forward_interval_tiled pass")
    ap.add_argument("--no-load", action="store_true", help="skip timing the .rmesh scene-file parser")
    ap.add_argument("--cpu-warmup", type=int, default=3, help="untimed CPU-oracle views before the timed ones")
    ap.add_argument("--cpu-views", type=int, default=10, help="timed CPU-oracle views (the median is reported)")
    ap.add_argument("--adjoint", choices=("undo", "checkpoint"), default="undo",
                    help="adjoint mode of the timed steps (the other mode is timed after them)")
    ap.add_argument("--no-modes", action="store_true", help="skip timing the other adjoint mode")
    ap.add_argument("--no-digits", action="store_true", help="skip timing the sort at the other digit width")
    ap.add_argument("--config-d", action="store_true",
                    help="also time the 32-scene uneven batch (one renderer) against per-scene renderers")
    ap.add_argument("--config-e", action="store_true",
                    help="also time the 13M-tet 1920x1080 workload (both adjoint modes; the scene build takes minutes)")
    ap.add_argument("--copy-probe-mb", type=int, default=0,
                    help="run a plain device copy of this many MiB (HBM counter calibration)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py: the train step runs on one GPU; --gpus {args.gpus} is not supported "
              "(sharding over GPUs is not implemented)", file=sys.stderr)
        sys.exit(2)

    V, Wd = args.views, args.width
    t_setup = time.perf_counter()
    scene, held, cams = build_scene(args.seed, args.points, V, Wd)
    N, M = len(scene["vertices"]) // 3, len(scene["indices"]) // 4
    log(f"scene: {N} vertices, {M} tets ({time.perf_counter() - t_setup:.1f} s)")

    tgt = make_renderer(held, Wd, V, args.tile_size)
    targets = tgt.forward_batched(cams)[..., :3].copy()
    del tgt
    r = make_renderer(scene, Wd, V, args.tile_size)
    r.set_targets(np.ascontiguousarray(targets))
    log(f"targets rendered; setup {time.perf_counter() - t_setup:.1f} s")

    tuning = None
    if args.tuning:
        tv = [int(x) for x in args.tuning.split(",")]
        r.set_tuning(tv[0], tv[1], tv[2])
        if len(tv) > 3:
            r.set_sort_digit_bits(tv[3])
    elif not args.no_autotune:
        ta = time.perf_counter()
        tuning = r.autotune(cams, loss_type=0, reps=3)
        log(f"autotune: tile/fwd/bwd = {tuning['tuning']}, {tuning['ms']} ({time.perf_counter() - ta:.1f} s)")
    tun = r.get_tuning()
    tile_size = tun[0]

    graph = not args.no_graph
    r.set_adjoint_mode(args.adjoint)
    r.set_graph_mode(graph)
    r.set_timing(True)
    step = lambda: r.train_step_batched(cams, 0, LR["lr_verts"], LR["lr_dens"], LR["lr_grads"], LR["lr_sh"])  # noqa
    loss0 = None
    for i in range(args.warmup):
        loss0 = step()
        log(f"warmup step {i}: loss {loss0:.6f}")
    P = r.pair_count()
    stage_acc = {k: 0.0 for k in r.STAGES}
    r.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        loss = step()
        for k, v in r.stage_times().items():
            stage_acc[k] += v
    r.synchronize()
    elapsed = time.perf_counter() - t0
    ms_per_step = 1000.0 * elapsed / args.steps
    value = V * args.steps / elapsed
    stages = {k: v / args.steps for k, v in stage_acc.items()}
    log(f"timed {args.steps} steps: {ms_per_step:.2f} ms/step, {value:.1f} views/s, loss {loss:.6f}; stages {stages}")

    T = ((Wd + tile_size - 1) // tile_size) ** 2
    dom = max(("forward", "backward"), key=lambda k: stages[k])
    alg = algorithmic_bytes(dom, P, V, Wd, Wd, T, N, M)
    achieved = alg / (stages[dom] * 1e-3) / 1e9
    kname = "k_backward" if dom == "backward" else "k_forward"
    prof = profiled_traffic(kname, tile_size)
    roofline = {"bound": "hbm", "kernel": kname,
                "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": int(prof["bytes"]) if prof else None,
                "algorithmic_bytes": int(alg), "kernel_ms": round(stages[dom], 4), "pairs": P}
    valu = profiled_valu(kname, tile_size)
    if valu:
        # the binding resource of the compositing kernels: SQ_INSTS_VALU x 4 cycles
        # per SIMD over the GPU cycles (issue), and active lanes per VALU op
        roofline["valu"] = valu
        # below ~95 % issue the VALU does not bind these kernels: cutting the
        # adjoint's VALU instructions by 19 % (12.6 -> 10.2 G per launch, same
        # occupancy) moved its time by 2.6 %; its phases are separated by 7
        # barriers per batch at 4 waves per SIMD, so latency binds
        roofline["binding_resource"] = "valu" if valu["issue"] >= 0.95 else "latency"
    if kname == "k_backward":
        # the adjoint's gradient flush is float atomics into L2: its rate
        # against the probe's (same instruction, one 64-byte line per lane)
        peak = atomic_ceiling()
        roofline["atomics"] = {"peak_per_s": peak,
                               "peak_note": "k_atomic_probe: 2^27 float atomic adds without return, each lane on its "
                                            "own 64-byte line among 2^16 (the flush's pattern), measured in this run"}
        atom = profiled_atomics(kname, tile_size)
        if atom:
            rate = atom["requests_per_launch"] / (stages[dom] * 1e-3)
            roofline["atomics"].update(rate_per_s=rate, frac=round(rate / peak, 4),
                                       requests_per_launch=atom["requests_per_launch"], source=atom["source"])
    if prof:
        roofline["traffic_detail"] = {k: prof[k] for k in ("read", "written", "kernel", "source")}
        roofline["traffic_note"] = ("FETCH_SIZE x2 + WRITE_SIZE x1 per launch (copy-kernel calibration); "
                                    "writes are dominated by the per-(tile, tet) gradient atomics")

    modes = None
    if not args.no_modes:
        # the adjoint's two ways to get each pair's transmittance, on the same
        # renderer and views: the undo from the image alpha (the reference's)
        # and the forward's per-pixel checkpoints every 32 tets
        modes = {args.adjoint: mode_summary(r, ms_per_step, stages)}
        other = "checkpoint" if args.adjoint == "undo" else "undo"
        r.set_adjoint_mode(other)
        r.set_graph_mode(graph)
        r.set_timing(True)
        modes[other] = time_steps(r, step, args.steps, V)
        r.set_adjoint_mode(args.adjoint)
        modes["note"] = ("ms per graph-replayed step, stage times (HIP events) and device memory held by the "
                         "renderer in each adjoint mode; the second mode is timed after the first on the same "
                         "renderer")
        log(f"adjoint modes: {json.dumps(modes)}")

    digits = None
    if not args.no_digits:
        # the sort at the other digit width (8-bit: 256 buckets, 4-bit: 16),
        # same renderer, views and tuning otherwise
        d0 = r.get_sort_digit_bits()
        digits = {str(d0): {"sort_ms": round(stages["sort"], 4), "ms_per_step": round(ms_per_step, 3)}}
        r.set_sort_digit_bits(4 if d0 == 8 else 8)
        r.set_graph_mode(graph)
        r.set_timing(True)
        t = time_steps(r, step, args.steps, V)
        digits[str(r.get_sort_digit_bits())] = {"sort_ms": t["sort_ms"], "ms_per_step": t["ms_per_step"]}
        r.set_sort_digit_bits(d0)
        log(f"sort digit widths: {json.dumps(digits)}")

    plain = None
    if not args.no_plain:
        r.set_graph_mode(False)
        r.set_timing(False)
        # one eager step per view; view v's target is uploaded into slot 0 first
        saved = np.ascontiguousarray(targets)
        r.synchronize()
        tp = time.perf_counter()
        for v in range(V):
            r.set_targets(np.ascontiguousarray(saved[v:v + 1]))
            r.train_step_batched([cams[v]], 0, LR["lr_verts"], LR["lr_dens"], LR["lr_grads"], LR["lr_sh"])
        r.synchronize()
        tpl = time.perf_counter() - tp
        plain = {"views_per_sec": V / tpl, "note": "one eager per-view train step (fwd+bwd+Adam) per view, "
                 "target upload included", "speedup_batched_graph": value / (V / tpl)}
        log(f"plain per-view path: {V / tpl:.1f} views/s")

    err_pass = None
    if not args.no_error:
        # per-tet error statistics of one view (outside the timed step: a
        # densification-time call; the synchronous C ABI copies the two error
        # images in and the [M, 16] + [M, 2] statistics out)
        r.set_graph_mode(False)
        r.set_timing(False)
        r.forward_tiled(*cams[0])
        erng = np.random.default_rng(7)
        pe = erng.random(Wd * Wd, dtype=np.float32)
        se = erng.random(Wd * Wd, dtype=np.float32)
        times = []
        for _ in range(3):
            r.synchronize()
            te = time.perf_counter()
            _, cnt = r.forward_error(pe, se)
            times.append(time.perf_counter() - te)
        err_pass = {"ms": round(1000.0 * min(times), 3), "tets_with_stats": int((cnt[:, 0] > 0).sum()),
                    "note": "forward_error on view 0, host copies included (kernel time: profiles/)"}
        log(f"forward_error: {err_pass['ms']:.2f} ms (view 0, {err_pass['tets_with_stats']} tets with statistics)")

    ivl_pass = None
    if not args.no_interval:
        # interval-tiled forward of view 0 beside the tiled forward of the same
        # view (each: binning, compositing and the synchronous host copies)
        r.set_graph_mode(False)
        r.set_timing(False)
        ti, tb, tf = [], [], []
        irng = np.random.default_rng(11)
        dl_i = irng.normal(0.0, 1.0, (Wd, Wd, 4)).astype(np.float32)
        dl_x = irng.normal(0.0, 1.0, (Wd, Wd, 4)).astype(np.float32)
        dl_d = irng.normal(0.0, 1.0, (Wd, Wd, 5)).astype(np.float32)
        for _ in range(3):
            r.synchronize()
            t0 = time.perf_counter()
            iv = r.forward_interval_tiled(*cams[0])
            ti.append(time.perf_counter() - t0)
            t0 = time.perf_counter()
            r.backward_interval_tiled(dl_i, dl_x, dl_d)
            tb.append(time.perf_counter() - t0)
            t0 = time.perf_counter()
            r.forward_tiled(*cams[0])
            tf.append(time.perf_counter() - t0)
        ivl_pass = {"ms": round(1000.0 * min(ti), 3), "backward_ms": round(1000.0 * min(tb), 3),
                    "forward_tiled_ms": round(1000.0 * min(tf), 3),
                    "mean_alpha": round(float(iv["image"][..., 3].mean()), 4),
                    "note": "forward_interval_tiled on view 0 (image, depth, distortion), its adjoint with image, "
                            "xyzd and distortion gradients, and forward_tiled of the same view; binning and host "
                            "copies included (kernel times: profiles/)"}
        log(f"forward_interval_tiled: {ivl_pass['ms']:.2f} ms, backward_interval_tiled {ivl_pass['backward_ms']:.2f} "
            f"ms (forward_tiled {ivl_pass['forward_tiled_ms']:.2f} ms)")

    load_pass = None
    if not args.no_load:
        # the workload's scene as a .rmesh (SH degree 3, 12 PCA components, as
        # the reference's converter writes it), parsed by the native loader:
        # raw, and gzip level 1; host code, multi-threaded
        import gzip

        from rmeshvk_amd import data as sio

        raw = sio.save_rmesh(None, scene["vertices"], scene["indices"], scene["densities"], scene["color_grads"],
                             scene["sh"].reshape(M, 3, 16), 3, k_components=12, compress=False)
        gz = gzip.compress(raw, compresslevel=1)
        tr, tg = [], []
        for _ in range(3):
            t0 = time.perf_counter()
            sio.load_rmesh(raw)
            tr.append(time.perf_counter() - t0)
        for _ in range(2):
            t0 = time.perf_counter()
            sio.load_rmesh(gz)
            tg.append(time.perf_counter() - t0)
        load_pass = {"raw_ms": round(1000.0 * min(tr), 1), "gzip_ms": round(1000.0 * min(tg), 1),
                     "raw_MB": round(len(raw) / 1e6, 1), "gzip_MB": round(len(gz) / 1e6, 1),
                     "tets_per_s_raw": round(M / min(tr)),
                     "note": "load_rmesh of this workload's scene (f16 decode, PCA decompression of 48 SH floats "
                             "per tet from 12 components, float64 circumspheres) on the host, up to 16 threads"}
        if not args.no_cpu:
            from oracle import scene_io as sref

            t0 = time.perf_counter()
            sref.parse_rmesh(raw, raw=True)
            load_pass["cpu_baseline_ms"] = round(1000.0 * (time.perf_counter() - t0), 1)
            load_pass["cpu_baseline"] = "numpy restatement of parse_rmesh (oracle), same file"
        log(f"load_rmesh: raw {load_pass['raw_ms']} ms ({load_pass['raw_MB']} MB), gzip {load_pass['gzip_ms']} ms")
        del raw, gz

    cpu = None
    if not args.no_cpu:
        tc = time.perf_counter()
        cpu = cpu_baseline(scene, cams, Wd, tile_size, args.cpu_warmup, args.cpu_views)
        log(f"cpu baseline: {cpu['value']:.4f} views/s on {cpu['cores']} threads ({time.perf_counter() - tc:.1f} s)")

    sort_db = r.get_sort_digit_bits()
    cfg_d = None
    if args.config_d:
        cfg_d = config_d(args.seed, max(3, args.steps // 2))
        log(f"config D: {json.dumps(cfg_d)}")
    cfg_e = None
    if args.config_e:
        del r
        cfg_e = config_e(args.seed, max(3, args.steps // 2))

    copy_probe = None
    if args.copy_probe_mb:
        import ctypes

        from rmeshvk_amd import _lib

        ms = ctypes.c_float()
        nbytes = args.copy_probe_mb << 20
        _lib.check(_lib.load().rmesh_copy_probe(0, nbytes, 3, ctypes.byref(ms)))
        copy_probe = {"bytes_read": nbytes, "bytes_written": nbytes, "ms": ms.value,
                      "GBps": 2 * nbytes / (ms.value * 1e-3) / 1e9}
        log(f"copy probe (k_copy_probe): {args.copy_probe_mb} MiB read + written, {ms.value:.3f} ms, "
            f"{copy_probe['GBps']:.0f} GB/s")

    out = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "views/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": (f"{args.points}-point 8-cluster Delaunay mesh ({M} tets), SH degree 3, {V} views at "
                         f"{Wd}x{Wd}, forward+backward+Adam per step, L1 vs perturbed-parameter targets"),
            "tets": M, "vertices": N, "views_per_step": V, "tile_size": tile_size, "hipgraph": graph,
            "adjoint": args.adjoint,
            "kernel_variants": {"forward": tun[1], "backward": tun[2]},
            "sort_digit_bits": sort_db,
            "autotune_ms": tuning["ms"] if tuning else None,
            "pairs_per_step": P, "loss_first_warmup": loss0, "loss_last": loss,
        },
        "stages_ms": {k: round(v, 4) for k, v in stages.items()},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "plain_per_view": plain,
        "forward_error": err_pass,
        "interval_forward": ivl_pass,
        "scene_load": load_pass,
        "adjoint_modes": modes,
        "sort_digits": digits,
    }
    if cfg_d:
        out["config_d"] = cfg_d
    if cfg_e:
        out["config_e"] = cfg_e
    if copy_probe:
        out["copy_probe"] = copy_probe
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
