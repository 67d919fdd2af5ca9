"""GPU tests of the training path: aux outputs, SH colours and adjoint, train_step
against the oracle's step, batched vs per-view, hipGraph replay vs eager, and
the reference's autograd finite-difference test (test_autograd.py:138-250).

Bars: images / aux within 1e-4 max-abs; gradients within 1e-3 relative L2;
parameters after Adam within 1e-5 relative (the update divides by sqrt(v), so
gradient rounding reaches the parameters only at this level).
"""

import math

import numpy as np
import pytest

from oracle import oracle as orc
from rmeshvk_amd.camera import colmap_camera, fibonacci_cameras, make_camera
from rmeshvk_amd.scenes import make_scene, single_tet_scene, uniform_points

pytestmark = pytest.mark.gpu

GRAD_TOL = 1e-3


def _rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / nb if nb > 0 else np.linalg.norm(a)


def _renderer(scene, w, h, **kw):
    from rmeshvk_amd import RMeshRenderer

    extra = {}
    if "sh" in scene:
        extra = dict(sh_coeffs=scene["sh"], sh_degree=int(scene["sh_degree"]))
    extra.update(kw)
    return RMeshRenderer(scene["vertices"], scene["indices"], scene["base_colors"], scene["densities"],
                         scene["color_grads"], scene["circumdata"], w, h, device_id=0, **extra)


@pytest.fixture(scope="module")
def small():
    rng = np.random.default_rng(99)
    scene = make_scene(uniform_points(800, rng), rng)
    cams = fibonacci_cameras(3, 3.0, 50.0, 72, 64)
    return scene, cams, 72, 64


@pytest.fixture(scope="module")
def small_sh():
    rng = np.random.default_rng(5)
    scene = make_scene(uniform_points(700, rng), rng, sh_degree=3)
    cams = fibonacci_cameras(3, 3.0, 50.0, 64, 64)
    return scene, cams, 64, 64


def test_render_aux(small):
    scene, cams, w, h = small
    r = _renderer(scene, w, h)
    img, aux = r.forward_tiled(*cams[0], render_aux=True)
    ref = orc.render(scene, cams[0], w, h, want_aux=True)
    assert np.abs(img - ref["rgba"]).max() < 1e-4
    scale = np.maximum(1.0, np.abs(ref["aux"]))
    assert (np.abs(aux - ref["aux"]) / scale).max() < 1e-4


@pytest.mark.parametrize("min_t", [0.0, 0.5])
def test_min_t(small, min_t):
    scene, cams, w, h = small
    r = _renderer(scene, w, h)
    r.set_min_t(min_t)
    img = r.forward_tiled(*cams[1])
    ref = orc.render(scene, cams[1], w, h, min_t=min_t)
    assert np.abs(img - ref["rgba"]).max() < 1e-4


def test_sh_forward_backward(small_sh):
    scene, cams, w, h = small_sh
    cam = cams[0]
    r = _renderer(scene, w, h)
    img = r.forward_tiled(*cam)
    ref = orc.render(scene, cam, w, h)  # SH colours from the oracle
    assert np.abs(img - ref["rgba"]).max() < 1e-4
    rng = np.random.default_rng(0)
    dl = np.ascontiguousarray(rng.normal(0, 1e-3, (h, w, 4)).astype(np.float32))
    g = r.backward(dl)
    gr = orc.backward(scene, cam, w, h, ref["sorted_vals"], ref["ranges"], ref["rgba"], dl, colors=ref["colors"])
    sg = orc.sh_backward(scene, cam[0], gr["d_base_colors"])
    assert _rel(g["d_base_colors"], gr["d_base_colors"]) < GRAD_TOL
    assert _rel(g["d_sh"], sg["d_sh"]) < GRAD_TOL
    assert _rel(g["d_color_grads"], gr["d_color_grads"] + sg["d_color_grads"]) < GRAD_TOL
    assert _rel(g["d_vertices"], gr["d_vertices"] + sg["d_vertices"]) < GRAD_TOL
    assert _rel(g["d_densities"], gr["d_densities"]) < GRAD_TOL


@pytest.mark.parametrize("loss_type", [0, 1])
def test_train_step_matches_oracle(small, loss_type):
    scene, cams, w, h = small
    cam = cams[2]
    r = _renderer(scene, w, h)
    gt = np.ascontiguousarray(np.random.default_rng(1).random((h, w, 3)).astype(np.float32))
    lrs = (1e-3, 1e-2, 1e-3)
    loss = r.train_step(*cam, gt, *lrs, loss_type)
    assert r.step == 1
    fw, dl, ref_loss, g = orc.render_backward(scene, cam, w, h, gt, loss_type)
    assert abs(loss - ref_loss) <= 1e-4 * max(1.0, abs(ref_loss))
    got = r.get_params()
    z = lambda a: np.zeros_like(a)  # noqa: E731
    for name, gk, lr in (("vertices", "d_vertices", lrs[0]), ("densities", "d_densities", lrs[1]),
                         ("color_grads", "d_color_grads", lrs[2])):
        p_ref, _, _ = orc.adam(scene[name], g[gk], z(scene[name]), z(scene[name]), 1, lr)
        # the step moves each parameter by ~lr * sign(g); compare the moves
        move, move_ref = got[name] - scene[name], p_ref - scene[name]
        agree = np.mean(np.abs(move - move_ref) <= 1e-3 * lr + 1e-6 * np.abs(scene[name]))
        assert agree > 0.995, (name, agree)
    np.testing.assert_array_equal(got["base_colors"], scene["base_colors"])  # never optimised (lib.rs:405-433)


def test_batched_matches_per_view(small_sh):
    scene, cams, w, h = small_sh
    r1 = _renderer(scene, w, h)
    singles = [r1.forward_tiled(*c) for c in cams]
    rb = _renderer(scene, w, h, max_views=len(cams))
    imgs = rb.forward_batched(cams)
    for v in range(len(cams)):
        np.testing.assert_array_equal(imgs[v], singles[v])


def test_batched_gradients_sum_views(small_sh):
    scene, cams, w, h = small_sh
    V = len(cams)
    rb = _renderer(scene, w, h, max_views=V)
    gts = np.ascontiguousarray(np.random.default_rng(2).random((V, h, w, 3)).astype(np.float32))
    rb.set_targets(gts)
    rb.train_step_batched(cams, 1, 0.0, 0.0, 0.0, 0.0)  # zero learning rates: gradients only
    got = rb.gradients()
    acc = {k: 0.0 for k in got}
    for v in range(V):
        _, _, _, g = orc.render_backward(scene, cams[v], w, h, gts[v], 1)
        for k in acc:
            acc[k] = acc[k] + g[k].astype(np.float64)
    for k in got:
        assert _rel(got[k], acc[k]) < GRAD_TOL, k


def test_batched_sh_adam_matches_oracle(small_sh):
    """The SH adjoint's fused Adam moves every coefficient like the oracle's
    Adam step on the view-summed SH gradient (one step: each move is about
    lr * sign(g), so the moves are compared as in test_train_step_matches_oracle)."""
    scene, cams, w, h = small_sh
    V = len(cams)
    rb = _renderer(scene, w, h, max_views=V)
    gts = np.ascontiguousarray(np.random.default_rng(4).random((V, h, w, 3)).astype(np.float32))
    rb.set_targets(gts)
    lr_sh = 2.5e-3
    rb.train_step_batched(cams, 1, 0.0, 0.0, 0.0, lr_sh)
    d_sh = np.zeros(len(scene["sh"]), np.float64)
    for v in range(V):
        d_sh += orc.render_backward(scene, cams[v], w, h, gts[v], 1)[3]["d_sh"]
    sh0 = scene["sh"]
    p_ref, _, _ = orc.adam(sh0, d_sh.astype(np.float32), np.zeros_like(sh0), np.zeros_like(sh0), 1, lr_sh)
    got = rb.get_params()["sh_coeffs"]
    move, move_ref = got - sh0, p_ref - sh0
    assert np.count_nonzero(move_ref) > 0.5 * len(sh0)
    agree = np.mean(np.abs(move - move_ref) <= 1e-3 * lr_sh + 1e-6 * np.abs(sh0))
    assert agree > 0.995, agree


def test_graph_replay_matches_eager(small_sh):
    scene, cams, w, h = small_sh
    V = len(cams)
    gts = np.ascontiguousarray(np.random.default_rng(3).random((V, h, w, 3)).astype(np.float32))
    runs = []
    for graph in (False, True):
        r = _renderer(scene, w, h, max_views=V)
        r.set_targets(gts)
        r.set_graph_mode(graph)
        losses = [r.train_step_batched(cams, 0, 1e-4, 1e-2, 1e-3, 2.5e-3) for _ in range(4)]
        assert r.step == 4
        runs.append((losses, r.get_params()))
    (l0, p0), (l1, p1) = runs
    np.testing.assert_allclose(l0, l1, rtol=1e-4)
    for k in ("vertices", "densities", "color_grads", "sh_coeffs"):
        assert _rel(p1[k], p0[k]) < 1e-5, k


def test_autograd_finite_diff():
    """test_autograd.py:138-250 through RMeshForward: single tet, 32x32, eps 1e-3, rel_err < 0.1."""
    import torch

    from rmesh_wgpu import RMeshRenderer
    from rmesh_wgpu.autograd import RMeshForward

    W = H = 32
    s = single_tet_scene(42)
    cam_np, vp_np, ci_np = colmap_camera(W, H, 1.0)
    r = RMeshRenderer(s["vertices"], s["indices"], s["base_colors"], s["densities"], s["color_grads"],
                      s["circumdata"], W, H)
    params = {k: torch.tensor(s[k], dtype=torch.float32, requires_grad=True)
              for k in ("vertices", "base_colors", "densities", "color_grads")}
    img = RMeshForward.apply(r, torch.tensor(cam_np), torch.tensor(vp_np), torch.tensor(ci_np),
                             params["vertices"], params["base_colors"], params["densities"], params["color_grads"])
    loss = 0.5 * (img[:, :, :3] ** 2).sum()
    loss.backward()

    def f(vals):
        r.update_params(vals["vertices"], vals["base_colors"], vals["densities"], vals["color_grads"])
        im = r.forward_tiled(cam_np, vp_np, ci_np)[:, :, :3].astype(np.float64)
        return 0.5 * np.sum(im * im)

    eps = 1e-3
    base = {k: s[k].copy() for k in params}
    for k in params:
        an = params[k].grad.numpy()
        for i in range(len(base[k])):
            plus = {kk: v.copy() for kk, v in base.items()}
            minus = {kk: v.copy() for kk, v in base.items()}
            plus[k][i] += eps
            minus[k][i] -= eps
            num = (f(plus) - f(minus)) / (2 * eps)
            a = float(an[i])
            rel = abs(a - num) / max(abs(a), abs(num), 1e-7)
            assert rel < 0.1 or abs(a - num) < 5e-5, (k, i, a, num)


def test_stage_timing_graph_matches_untimed(small_sh):
    """Timing mode replays one graph per stage with events between the launches;
    results equal the single-graph replay and every stage reports a duration."""
    scene, cams, w, h = small_sh
    V = len(cams)
    gts = np.ascontiguousarray(np.random.default_rng(5).random((V, h, w, 3)).astype(np.float32))
    runs = []
    for timing in (False, True):
        r = _renderer(scene, w, h, max_views=V)
        r.set_targets(gts)
        r.set_graph_mode(True)
        r.set_timing(timing)
        losses = [r.train_step_batched(cams, 0, 1e-4, 1e-2, 1e-3, 2.5e-3) for _ in range(3)]
        if timing:
            st = r.stage_times()
            assert set(st) == set(r.STAGES)
            assert all(v >= 0.0 for v in st.values()) and st["forward"] > 0.0 and st["backward"] > 0.0
        runs.append((losses, r.get_params()))
    (l0, p0), (l1, p1) = runs
    np.testing.assert_allclose(l0, l1, rtol=1e-4)
    for k in ("vertices", "densities", "color_grads", "sh_coeffs"):
        assert _rel(p1[k], p0[k]) < 1e-5, k
