"""Writes tests/golden/config_b_fd.json: float64 central finite differences
of the config-B loss with respect to 64 parameters, at eps 1e-4 (the
configuration's step) and at eps 1e-6.

Config B: 10,000 points uniform in [-1, 1]^3, Delaunay (66,338 tets), SH
degree 3 (DC ~ N(0.5, 0.2), others ~ N(0, 0.3)), densities ~ LogN(0, 1),
colour gradients ~ N(0, 0.1); 8 cameras on a Fibonacci sphere (radius 3,
50 deg) at 512x512.  Loss: the sum over the 8 views of the L2 loss (mean over
pixels of the squared RGB difference, loss_compute.wgsl) against targets
rendered by the float32 oracle from the same scene with densities x 1.2.

16 parameters per group (vertices, densities, colour gradients, SH
coefficients), drawn (seed 0) among those whose float64 analytic gradient is
at least 1e-2 of the group's largest, so that each difference sees its
parameter.  Every loss evaluation runs the float64 oracle over all 8 views
(binning included, so coverage changes are part of the difference).  The
float64 analytic gradient is stored beside each difference.

The loss is only piecewise smooth in a vertex position: a pixel ray crossing
a tet's edge or silhouette as the vertex moves puts a kink in its chord.  At
eps 1e-4 (~0.02 px of motion here) such kinks fall inside the difference
for some vertices and move it by tens of percent; at eps 1e-6 the difference
converges to the derivative (checked: 1e-4, 1e-5, 1e-6, 1e-7 on single
vertices).  Both are stored.

Run from the repository root: python tests/golden/make_config_b_fd.py
(about twenty minutes on 8 cores).
"""

from __future__ import annotations

import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import oracle as orc  # noqa: E402
from rmeshvk_amd.scenes import config_scene  # noqa: E402

EPS = (1e-4, 1e-6)
W = H = 512
GROUPS = (("vertices", "d_vertices"), ("densities", "d_densities"), ("color_grads", "d_color_grads"),
          ("sh", "d_sh"))


def targets(scene, cams):
    s = dict(scene)
    s["densities"] = (scene["densities"] * np.float32(1.2)).astype(np.float32)
    return [orc.render(s, c, W, H)["rgba"][..., :3].copy() for c in cams]


def f64(scene):
    return {k: (np.asarray(v, np.float64) if k in ("vertices", "base_colors", "densities", "color_grads", "sh")
                else v) for k, v in scene.items()}


def total_loss(s64, cams, gts):
    tot = 0.0
    for c, gt in zip(cams, gts):
        img = orc.render(s64, c, W, H, dtype=np.float64)["rgba"]
        tot += orc.loss_grad(img, gt.astype(np.float64), 1, np.float64)[1]
    return tot


def analytic(s64, cams, gts):
    acc = None
    for c, gt in zip(cams, gts):
        _, _, _, g = orc.render_backward(s64, c, W, H, gt.astype(np.float64), 1, dtype=np.float64)
        acc = g if acc is None else {k: acc[k] + g[k] for k in acc}
    return acc


def main():
    t0 = time.time()
    scene, cams = config_scene("B", seed=0)
    gts = targets(scene, cams)
    s64 = f64(scene)
    g = analytic(s64, cams, gts)
    rng = np.random.default_rng(0)
    entries = []
    for name, gk in GROUPS:
        ga = np.abs(g[gk])
        cand = np.flatnonzero(ga >= 1e-2 * ga.max())
        for i in rng.choice(cand, 16, replace=False):
            i = int(i)
            e = {"param": name, "index": i, "analytic_f64": float(g[gk][i])}
            for eps, key in zip(EPS, ("fd_1e-4", "fd_1e-6")):
                sp, sm = dict(s64), dict(s64)
                sp[name] = s64[name].copy()
                sm[name] = s64[name].copy()
                sp[name][i] += eps
                sm[name][i] -= eps
                e[key] = (total_loss(sp, cams, gts) - total_loss(sm, cams, gts)) / (2 * eps)
            entries.append(e)
            print(f"{name}[{i}]: fd(1e-4) {e['fd_1e-4']:.6e} fd(1e-6) {e['fd_1e-6']:.6e} analytic {g[gk][i]:.6e} "
                  f"({time.time() - t0:.0f} s)", flush=True)
    out = {"config": "B", "seed": 0, "width": W, "height": H, "views": len(cams), "eps": list(EPS),
           "loss": "sum over views of the L2 loss vs targets rendered with densities x 1.2",
           "entries": entries}
    with open(os.path.join(ROOT, "tests", "golden", "config_b_fd.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
