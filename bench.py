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


def cpu_baseline(scene, cam, width, stride, tile_size):
    """Oracle (C, one thread) on one view: full binning, SH colours/adjoint and
    Adam over every parameter, forward/backward on every `stride`-th tile."""
    from oracle import oracle as orc

    W = H = width
    _, _, T = orc.tile_grid(W, H, tile_size)
    t0 = time.perf_counter()
    colors = orc.sh_colors(scene, cam[0])
    tt, _ = orc.project(scene, cam, W, H, tile_size)
    keys, vals, P = orc.pairs(scene, cam, W, H, tt, tile_size)
    sk, sv = orc.sort(keys, vals, orc.sorting_bits(T))
    rg = orc.ranges(sk, T)
    t_bin = time.perf_counter() - t0
    t1 = time.perf_counter()
    fw = orc.forward(scene, cam, W, H, sv, rg, tile_size, colors=colors, tile_step=stride)
    t_fwd = time.perf_counter() - t1
    gt = np.zeros((H, W, 3), np.float32)
    dl, _ = orc.loss_grad(fw["rgba"], gt, 0)
    t2 = time.perf_counter()
    g = orc.backward(scene, cam, W, H, sv, rg, fw["rgba"], dl, tile_size, colors=colors, tile_step=stride)
    t_bwd = time.perf_counter() - t2
    t3 = time.perf_counter()
    orc.sh_backward(scene, cam[0], g["d_base_colors"])
    n_params = len(scene["vertices"]) + len(scene["densities"]) + len(scene["color_grads"]) + len(scene["sh"])
    z = np.zeros(n_params, np.float32)
    orc.adam(z, z, z, z, 1, 1e-3)
    t_rest = time.perf_counter() - t3
    per_view = t_bin + stride * (t_fwd + t_bwd) + t_rest
    return {
        "value": 1.0 / per_view,
        "unit": "views/s",
        "cores": 1,
        "kind": "port",
        "sample": (f"1 view of the workload: full projection/pairs/sort/ranges ({P} pairs), SH colours + SH adjoint, "
                   f"Adam over all {n_params} params; forward+backward on every {stride}th tile of {T}, "
                   f"scaled x{stride}; measured {t_bin + t_fwd + t_bwd + t_rest:.1f} s of CPU"),
        "seconds_measured": t_bin + t_fwd + t_bwd + t_rest,
    }


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
    ap.add_argument("--tuning", default="", help="TS,F,B: tile size and kernel variants, no autotune")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-error", action="store_true", help="skip the forward_error (densification statistics) pass")
    ap.add_argument("--no-interval", action="store_true", help="skip the forward_interval_tiled pass")
    ap.add_argument("--no-load", action="store_true", help="skip timing the .rmesh scene-file parser")
    ap.add_argument("--cpu-stride", type=int, default=1)
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
        ts_, fv_, bv_ = (int(x) for x in args.tuning.split(","))
        r.set_tuning(ts_, fv_, bv_)
    elif not args.no_autotune:
        ta = time.perf_counter()
        tuning = r.autotune(cams, loss_type=0, reps=3)
        log(f"autotune: tile/fwd/bwd = {tuning['tuning']}, {tuning['ms']} ({time.perf_counter() - ta:.1f} s)")
    tile_size = r.get_tuning()[0]

    graph = not args.no_graph
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
    if prof:
        roofline["traffic_detail"] = {k: prof[k] for k in ("read", "written", "kernel", "source")}
        roofline["traffic_note"] = ("FETCH_SIZE x2 + WRITE_SIZE x1 per launch (copy-kernel calibration); "
                                    "writes are dominated by the per-(tile, tet) gradient atomics")

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
        cpu = cpu_baseline(scene, cams[0], Wd, args.cpu_stride, tile_size)
        log(f"cpu baseline: {cpu['value']:.4f} views/s ({time.perf_counter() - tc:.1f} s)")

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
            "kernel_variants": {"forward": r.get_tuning()[1], "backward": r.get_tuning()[2]},
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
    }
    if copy_probe:
        out["copy_probe"] = copy_probe
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
