"""GPU: many scenes of uneven tet count, resolution and SH degree in one
renderer (one launch per stage over all scenes' views).

Each scene of the batch must give what a renderer holding that scene alone
gives: images bit for bit (same lists, same arithmetic), gradients of a
zero-learning-rate step within 1e-5 relative L2 (float atomics in another
order), parameters after two steps with learning rates within 1e-5; images
of one scene per SH degree against the oracle (1e-4); graph replay equal to
eager; per-scene latencies reported; the scene-major tile order gives the
same images and gradients (its effect on latency needs a batch larger than
the GPU's resident workgroups: bench.py --config-d).
"""

import math

import numpy as np
import pytest

from oracle import oracle as orc
from rmeshvk_amd.camera import fibonacci_cameras, make_camera
from rmeshvk_amd.scenes import cluster_points, make_scene, uniform_points

pytestmark = pytest.mark.gpu

# (points, SH degree (-1: no SH), width, height, views)
SPECS = [(900, 3, 96, 80, 3), (300, 0, 64, 64, 2), (2500, 1, 128, 128, 2), (1200, 3, 72, 100, 1), (600, 1, 80, 80, 3)]
TS = 12


def _rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / nb if nb > 0 else np.linalg.norm(a)


@pytest.fixture(scope="module")
def batch():
    rng = np.random.default_rng(2024)
    scenes, cams = [], []
    for i, (npts, deg, w, h, nv) in enumerate(SPECS):
        pts = uniform_points(npts, rng) if i % 2 == 0 else cluster_points(npts, rng, 3, 0.25)
        s = make_scene(pts, rng, sh_degree=deg)
        s.update(width=w, height=h, max_views=4)
        scenes.append(s)
        cams.append(fibonacci_cameras(nv, 3.0, 50.0, w, h))
    return scenes, cams


def _single(s, views):
    from rmeshvk_amd import RMeshRenderer

    return RMeshRenderer(s["vertices"], s["indices"], s["base_colors"], s["densities"], s["color_grads"],
                         s["circumdata"], s["width"], s["height"], device_id=0, sh_coeffs=s["sh"],
                         sh_degree=int(s["sh_degree"]), max_views=views, tile_size=TS)


def _multi(scenes):
    from rmeshvk_amd import MultiSceneRenderer

    return MultiSceneRenderer(scenes, tile_size=TS, device_id=0)


def _flat(cams):
    return [c for cc in cams for c in cc]


def test_multi_images_equal_single_scene_renderers(batch):
    scenes, cams = batch
    mr = _multi(scenes)
    assert mr.num_scenes == len(scenes)
    nv = [len(c) for c in cams]
    imgs = mr.forward(nv, _flat(cams))
    for i, s in enumerate(scenes):
        r = _single(s, len(cams[i]))
        ref = r.forward_batched(cams[i])
        np.testing.assert_array_equal(imgs[i], ref)
    # one scene per SH degree against the oracle
    for i in (0, 1, 2):
        o = orc.render(scenes[i], cams[i][0], scenes[i]["width"], scenes[i]["height"], ts=TS)
        assert np.abs(imgs[i][0] - o["rgba"]).max() < 1e-4, i


def _targets(scenes, cams, seed):
    rng = np.random.default_rng(seed)
    return [np.ascontiguousarray(rng.random((len(c), s["height"], s["width"], 3)).astype(np.float32))
            for s, c in zip(scenes, cams)]


@pytest.mark.parametrize("graph", [False, True])
def test_multi_gradients_equal_single_scene_steps(batch, graph):
    scenes, cams = batch
    nv = [len(c) for c in cams]
    gts = _targets(scenes, cams, 5)
    mr = _multi(scenes)
    mr.set_targets(nv, gts)
    mr.set_graph_mode(graph)
    for _ in range(2):  # (capture and) replay
        loss, lv = mr.train_step(nv, _flat(cams), 1, 0.0, 0.0, 0.0, 0.0)
    assert lv.shape == (sum(nv),)
    o = 0
    for i, s in enumerate(scenes):
        r = _single(s, nv[i])
        r.set_targets(gts[i])
        l1 = r.train_step_batched(cams[i], 1, 0.0, 0.0, 0.0, 0.0)
        assert abs(float(lv[o:o + nv[i]].mean()) - l1) <= 1e-5 * max(abs(l1), 1e-6), i
        o += nv[i]
        g = r.gradients()
        g["d_sh"] = r.sh_gradient()
        for k in ("d_vertices", "d_densities", "d_color_grads", "d_sh"):
            assert _rel(mr.read_scene(i, k), g[k]) <= 1e-5, (i, k)


def test_multi_training_equals_single_scene_training(batch):
    scenes, cams = batch
    nv = [len(c) for c in cams]
    gts = _targets(scenes, cams, 9)
    lrs = (1e-3, 1e-2, 1e-3, 2.5e-3)
    mr = _multi(scenes)
    mr.set_targets(nv, gts)
    mr.set_graph_mode(True)
    mr.set_scene_timing(True)
    for _ in range(2):
        mr.train_step(nv, _flat(cams), 0, *lrs)
    lat = mr.scene_latency()
    assert lat.shape == (len(scenes),) and np.all(lat > 0.0)
    assert mr.step == 2
    for i, s in enumerate(scenes):
        r = _single(s, nv[i])
        r.set_targets(gts[i])
        for _ in range(2):
            r.train_step_batched(cams[i], 0, *lrs)
        p = r.get_params()
        for k in ("vertices", "densities", "color_grads", "sh_coeffs"):
            assert _rel(mr.read_scene(i, k), p[k]) <= 1e-5, (i, k)


def test_multi_uneven_view_counts_and_errors(batch):
    """A scene may have no views in a step; bad layouts are refused."""
    scenes, cams = batch
    mr = _multi(scenes)
    nv = [len(c) for c in cams]
    nv2 = list(nv)
    nv2[1] = 0
    cams2 = [c for i, c in enumerate(cams) if i != 1]
    imgs = mr.forward(nv2, _flat(cams2))
    assert imgs[1].shape[0] == 0
    full = mr.forward(nv, _flat(cams))
    for i in range(len(scenes)):
        if i != 1:
            np.testing.assert_array_equal(imgs[i], full[i])
    with pytest.raises(ValueError):
        mr.forward([5] + nv[1:], _flat(cams) + [cams[0][0]] * 2)  # more views than max_views
    with pytest.raises(NotImplementedError):
        mr.forward_tiled(*cams[0][0])  # single-scene entry point


def test_scene_major_order_same_results_earlier_scenes(batch):
    scenes, cams = batch
    nv = [len(c) for c in cams]
    gts = _targets(scenes, cams, 13)
    out = {}
    for order in ("interleaved", "by_scene"):
        mr = _multi(scenes)
        mr.set_scene_order(order)
        assert mr.get_scene_order() == order
        mr.set_targets(nv, gts)
        mr.set_graph_mode(True)
        mr.set_scene_timing(True)
        imgs = mr.forward(nv, _flat(cams))
        for _ in range(2):
            mr.train_step(nv, _flat(cams), 1, 0.0, 0.0, 0.0, 0.0)
        out[order] = (imgs, {(i, k): mr.read_scene(i, k) for i in range(len(scenes))
                             for k in ("d_vertices", "d_densities", "d_color_grads", "d_sh")}, mr.scene_latency())
    for i in range(len(scenes)):
        np.testing.assert_array_equal(out["by_scene"][0][i], out["interleaved"][0][i])
    for key, g in out["interleaved"][1].items():
        assert _rel(out["by_scene"][1][key], g) <= 1e-5, key
    assert out["by_scene"][2].shape == (len(scenes),) and np.all(out["by_scene"][2] > 0.0)
    with pytest.raises(ValueError):
        _multi(scenes).set_scene_order("random")
