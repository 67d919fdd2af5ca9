"""GPU: per-tet error statistics (RMeshRenderer.forward_error, lib.rs:1621-1711;
error_tiled_compute.wgsl) against the oracle on config A at each tile size.

Bars: each statistic column within 1e-4 relative L2 (float32 sums in another
order; the GPU slab divides by a 1-ulp reciprocal).  Pixel counts: equal for
all but 1 % of the tets that receive any, and then within 3.  The slab's
t_max - t_min cancels, so a thin chord's optical depth carries the
reciprocal's rounding at ~1e-4 relative, and a pixel whose log T crosses the
-5.5 cut-off within that margin counts one tet more or less (the reference's
own WGSL division is no closer to the oracle).  Max-weight column within 2
where the counts agree.
"""

import math

import numpy as np
import pytest

from oracle import oracle as orc
from rmeshvk_amd.camera import make_camera

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def config_a():
    from rmeshvk_amd.scenes import config_scene

    scene, cams = config_scene("A", seed=0)
    return scene, cams[0], 256, 256


def _renderer(scene, w, h, **kw):
    from rmeshvk_amd import RMeshRenderer

    return RMeshRenderer(scene["vertices"], scene["indices"], scene["base_colors"], scene["densities"],
                         scene["color_grads"], scene["circumdata"], w, h, device_id=0, **kw)


@pytest.mark.parametrize("ts", [8, 12, 16])
def test_forward_error_matches_oracle(ts, config_a):
    scene, cam, w, h = config_a
    r = _renderer(scene, w, h, tile_size=ts)
    r.forward_tiled(*cam)
    rng = np.random.default_rng(ts)
    pe = rng.uniform(0.0, 1.0, h * w).astype(np.float32)
    se = rng.uniform(-0.5, 0.5, h * w).astype(np.float32)
    err, cnt = r.forward_error(pe, se)
    ref = orc.render(scene, cam, w, h, ts=ts)
    e_ref, c_ref = orc.error_stats(scene, cam, w, h, ref["sorted_vals"], ref["ranges"], pe, se, ts=ts)
    assert c_ref[:, 0].sum() > 10000
    for k in range(16):
        nb = np.linalg.norm(e_ref[:, k].astype(np.float64))
        if nb == 0:
            assert not np.any(err[:, k]), k
            continue
        rel = np.linalg.norm(err[:, k].astype(np.float64) - e_ref[:, k]) / nb
        assert rel < 1e-4, (k, rel)
    same = cnt[:, 0] == c_ref[:, 0]
    assert np.count_nonzero(~same) <= 0.01 * np.count_nonzero(c_ref[:, 0])
    assert np.abs(cnt[:, 0] - c_ref[:, 0]).max() <= 3
    assert np.abs(cnt[same, 1] - c_ref[same, 1]).max() <= 2


def test_forward_error_argument_checks(config_a):
    scene, cam, w, h = config_a
    r = _renderer(scene, w, h)
    with pytest.raises(ValueError):
        r.forward_error(np.zeros(w * h, np.float32), np.zeros(w * h, np.float32))  # no forward yet
    r.forward_tiled(*cam)
    with pytest.raises(ValueError, match="pixel_err must have"):
        r.forward_error(np.zeros(7, np.float32), np.zeros(w * h, np.float32))
    with pytest.raises(TypeError):
        r.forward_error(np.zeros(w * h, np.float64), np.zeros(w * h, np.float32))


def test_forward_error_min_t(config_a):
    """min_t shifts the ray origin: entry/exit moments follow it."""
    scene, cam, w, h = config_a
    r = _renderer(scene, w, h)
    r.set_min_t(0.4)
    r.forward_tiled(*cam)
    pe = np.full(h * w, 0.5, np.float32)
    err, cnt = r.forward_error(pe, pe)
    ref = orc.render(scene, cam, w, h, min_t=0.4)
    e_ref, c_ref = orc.error_stats(scene, cam, w, h, ref["sorted_vals"], ref["ranges"], pe, pe, min_t=0.4)
    for k in (0, 6, 9, 13):
        nb = np.linalg.norm(e_ref[:, k].astype(np.float64))
        assert np.linalg.norm(err[:, k].astype(np.float64) - e_ref[:, k]) / nb < 1e-4, k
