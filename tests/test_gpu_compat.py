"""GPU: API parity of the viewer-style entry points.

- forward() + backward_hw() (lib.rs:932-1025, 2255-2378): forward() is served
  by the tiled forward, so backward_hw is the tiled adjoint over its state;
  checked against forward_tiled + backward (1e-5 relative L2: float atomics
  in another order).
- f16 SH (sh_half, project_compute.wgsl:267-318): colours from the
  coefficients rounded to f16 as the reference's viewer packs them; the image
  equals the oracle's render with the f16-rounded coefficients (1e-4), and
  differs from the f32 one; training is refused.
"""

import numpy as np
import pytest

from oracle import oracle as orc

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def config_b():
    """An SH degree-3 mesh (2000 points) seen from two cameras."""
    import math

    from rmeshvk_amd.camera import make_camera
    from rmeshvk_amd.scenes import make_scene, uniform_points

    rng = np.random.default_rng(31)
    scene = make_scene(uniform_points(2000, rng), rng, sh_degree=3)
    w, h = 192, 160
    cams = [make_camera((0.3, -2.7, 0.7), (0.0, 0.0, 0.0), math.radians(60), w, h),
            make_camera((-2.2, 1.0, 1.1), (0.0, 0.1, 0.0), math.radians(55), w, h)]
    return scene, cams, w, h


def _renderer(scene, w, h, **kw):
    from rmeshvk_amd import RMeshRenderer

    return RMeshRenderer(scene["vertices"], scene["indices"], scene["base_colors"], scene["densities"],
                         scene["color_grads"], scene["circumdata"], w, h, device_id=0, sh_coeffs=scene["sh"],
                         sh_degree=int(scene["sh_degree"]), **kw)


def _rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def test_forward_backward_hw(config_b):
    scene, cams, w, h = config_b
    r = _renderer(scene, w, h)
    dl = np.random.default_rng(4).normal(0.0, 1.0, (h, w, 4)).astype(np.float32)
    img = r.forward(*cams[0])
    g_hw = r.backward_hw(dl)
    img_t = r.forward_tiled(*cams[0])
    g = r.backward(dl)
    assert np.array_equal(img, img_t)
    for k in ("d_vertices", "d_densities", "d_base_colors", "d_color_grads"):
        assert _rel(g_hw[k], g[k]) <= 1e-5, k


def test_sh_half_forward(config_b):
    scene, cams, w, h = config_b
    r = _renderer(scene, w, h, sh_half=True)
    img = r.forward_tiled(*cams[0])
    half = dict(scene, sh=scene["sh"].astype(np.float16).astype(np.float32))
    ref = orc.render(half, cams[0], w, h)
    full = orc.render(scene, cams[0], w, h)
    assert ref["rgba"][..., 3].max() > 0.5
    assert np.abs(img - ref["rgba"]).max() <= 1e-4
    assert np.abs(ref["rgba"] - full["rgba"]).max() > 1e-4  # the rounding is visible
    with pytest.raises(NotImplementedError, match="forward only"):
        r.backward(np.zeros((h, w, 4), np.float32))
    with pytest.raises(NotImplementedError, match="forward only"):
        r.train_step(*cams[0], np.zeros((h, w, 3), np.float32), 1e-3, 1e-3, 1e-3, 0)
