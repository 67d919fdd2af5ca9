"""GPU: a scene read back from .rmesh and PLY files renders as the oracle
renders the loaded arrays (the loader's SH layout is the renderer's), at
the images' bar (1e-4 max-abs)."""

import math

import numpy as np
import pytest

from oracle import oracle as orc
from rmeshvk_amd import data as io
from rmeshvk_amd.camera import make_camera
from rmeshvk_amd.scenes import make_scene, uniform_points

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("fmt", ["rmesh", "ply"])
def test_loaded_scene_renders(fmt):
    rng = np.random.default_rng(21)
    s = make_scene(uniform_points(1500, rng), rng, sh_degree=2)
    M = len(s["indices"]) // 4
    sh = s["sh"].reshape(M, 3, 9)
    if fmt == "rmesh":
        blob = io.save_rmesh(None, s["vertices"], s["indices"], s["densities"], s["color_grads"], sh, 2,
                             k_components=12)
        scene, shd, _ = io.load_rmesh(blob)
    else:
        blob = io.save_ply(None, s["vertices"], s["indices"], s["densities"], s["color_grads"], sh, 2)
        scene, shd, _ = io.load_ply(blob)
    w, h = 160, 120
    cam = make_camera((0.3, -2.7, 0.7), (0.0, 0.0, 0.0), math.radians(60), w, h)
    r = io.renderer_from_scene(scene, shd, w, h, device_id=0)
    img = r.forward_tiled(*cam)
    ref_scene = dict(scene, base_colors=np.zeros(3 * M, np.float32), sh=shd["coeffs"], sh_degree=shd["degree"])
    ref = orc.render(ref_scene, cam, w, h)
    assert ref["rgba"][..., 3].max() > 0.5
    assert np.abs(img - ref["rgba"]).max() <= 1e-4
