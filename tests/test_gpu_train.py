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


def _adam_exact(name, got, p0, grad, lr):
    """Every parameter equals the oracle's Adam step (step 1, zero moments)
    applied to the GPU's own gradient, to two ulps of the parameter plus
    1e-6 lr: the update itself is compared, with no tolerance for sign flips."""
    z = np.zeros_like(p0)
    p_ref, _, _ = orc.adam(p0, np.asarray(grad, np.float32), z, z, 1, lr)
    tol = 2.0 * np.spacing(np.abs(p_ref).astype(np.float32)) + 1e-6 * lr
    bad = np.abs(got.astype(np.float64) - p_ref) > tol
    assert not bad.any(), (name, int(bad.sum()), np.flatnonzero(bad)[:8])
    assert np.count_nonzero(got != p0) > 0.5 * len(p0), name  # the step moved the group


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
    gg = r.gradients()  # the step's gradients, left in the device buffers
    for name, gk, lr in (("vertices", "d_vertices", lrs[0]), ("densities", "d_densities", lrs[1]),
                         ("color_grads", "d_color_grads", lrs[2])):
        assert _rel(gg[gk], g[gk]) < GRAD_TOL, gk
        _adam_exact(name, got[name], scene[name], gg[gk], lr)
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
    """The SH adjoint's fused Adam: a step with an SH learning rate moves the
    coefficients as the oracle's Adam step on the view-summed SH gradient that
    the same step computes without one (checked against the oracle's SH
    adjoint at 1e-3).

    The fused step cannot hand out its own gradient, and two runs of the
    adjoint differ in the order of their float atomics, so the reference
    gradient carries run-to-run noise.  Two gradient-only runs bound it per
    coefficient (their difference, floored at 1e-6 of the gradient); the
    tolerance of each move adds the Adam step's sensitivity to that noise,
    lr eps 4 noise / (|g| - 4 noise + eps)^2 (step 1, zero moments: the move is
    lr g / (|g| + eps)).  Coefficients whose gradient lies within 4 noise of 0
    (sign not determined) only have to move by at most lr; they are asserted
    to be under 1 % of the group."""
    scene, cams, w, h = small_sh
    V = len(cams)
    gts = np.ascontiguousarray(np.random.default_rng(4).random((V, h, w, 3)).astype(np.float32))
    lr_sh, eps = 2.5e-3, 1e-8
    runs = []
    for _ in range(2):
        ra = _renderer(scene, w, h, max_views=V)
        ra.set_targets(gts)
        ra.train_step_batched(cams, 1, 0.0, 0.0, 0.0, 0.0)  # gradients only: d_sh written
        runs.append(ra.sh_gradient().astype(np.float64))
    d_sh_gpu = runs[0]
    d_sh = np.zeros(len(scene["sh"]), np.float64)
    for v in range(V):
        d_sh += orc.render_backward(scene, cams[v], w, h, gts[v], 1)[3]["d_sh"]
    assert _rel(d_sh_gpu, d_sh) < GRAD_TOL
    rb = _renderer(scene, w, h, max_views=V)
    rb.set_targets(gts)
    rb.train_step_batched(cams, 1, 0.0, 0.0, 0.0, lr_sh)
    got = rb.get_params()["sh_coeffs"].astype(np.float64)
    p0 = scene["sh"]
    z = np.zeros_like(p0)
    p_ref, _, _ = orc.adam(p0, d_sh_gpu.astype(np.float32), z, z, 1, lr_sh)
    ag = np.abs(d_sh_gpu)
    noise = np.abs(runs[0] - runs[1]) + 1e-6 * ag
    open_sign = ag <= 4.0 * noise
    assert open_sign.mean() < 0.01, open_sign.mean()
    lo = np.maximum(ag - 4.0 * noise, 0.0)
    tol = (2.0 * np.spacing(np.abs(p_ref).astype(np.float32)) + 1e-6 * lr_sh
           + lr_sh * eps * 4.0 * noise / (lo + eps) ** 2)
    bad = ~open_sign & (np.abs(got - p_ref) > tol)
    assert not bad.any(), (int(bad.sum()), np.flatnonzero(bad)[:8])
    assert np.all(np.abs(got - p0)[open_sign] <= lr_sh * (1 + 1e-6) + 2.0 * np.spacing(np.abs(p0[open_sign])))
    assert np.count_nonzero(got != p0) > 0.5 * len(p0)  # the step moved the group


def test_batched_base_color_lr_graph(small):
    """lr_base (scenes without SH) is part of the captured step: changing it
    between graph replays gives the eager result, and 0 leaves the group out."""
    scene, cams, w, h = small
    V = len(cams)
    gts = np.ascontiguousarray(np.random.default_rng(6).random((V, h, w, 3)).astype(np.float32))
    runs = []
    for graph in (False, True):
        r = _renderer(scene, w, h, max_views=V)
        r.set_targets(gts)
        r.set_graph_mode(graph)
        for lr_b in (1e-2, 0.0, 3e-3):
            r.train_step_batched(cams, 0, 1e-4, 1e-2, 1e-3, lr_base=lr_b)
        runs.append(r.get_params())
    p0, p1 = runs
    assert _rel(p0["base_colors"], scene["base_colors"]) > 1e-4  # the group moved
    for k in ("vertices", "densities", "color_grads", "base_colors"):
        assert _rel(p1[k], p0[k]) < 1e-5, k
    # a step with lr_base = 0 after the moves leaves the base colours where they are
    r = _renderer(scene, w, h, max_views=V)
    r.set_targets(gts)
    r.set_graph_mode(True)
    r.train_step_batched(cams, 0, 1e-4, 1e-2, 1e-3, lr_base=1e-2)
    b1 = r.get_params()["base_colors"]
    r.train_step_batched(cams, 0, 1e-4, 1e-2, 1e-3, lr_base=0.0)
    np.testing.assert_array_equal(r.get_params()["base_colors"], b1)


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


def test_graph_capacity_overflow_regrows():
    """A captured step whose pair count outgrows the pair buffers: the replay
    flags the overflow, leaves the parameters untouched, grows the buffers and
    redoes the step eagerly (step advances by one); the next step captures a
    new graph.  Both steps equal an eager renderer's."""
    from rmeshvk_amd.camera import fibonacci_cameras

    rng = np.random.default_rng(21)
    scene = make_scene(uniform_points(12000, rng), rng)
    cams = fibonacci_cameras(3, 3.0, 50.0, 384, 384)
    V, w, h = len(cams), 384, 384
    dim = dict(scene)
    dim["densities"] = (scene["densities"] * 1e-7).astype(np.float32)  # every tet culled: no pairs
    gts = np.ascontiguousarray(np.random.default_rng(8).random((V, h, w, 3)).astype(np.float32))
    rs = []
    for graph in (True, False):
        r = _renderer(dim, w, h, max_views=V, tile_size=8)
        r.set_targets(gts)
        r.set_graph_mode(graph)
        r.train_step_batched(cams, 0, 1e-4, 1e-2, 1e-3)
        assert r.pair_count() == 0 and r.step == 1
        rs.append(r)
    ra, rb = rs
    # densities back to the scene's: ~1.57M pairs, past the 2^20-pair initial buffers
    for r in rs:
        p = r.get_params()
        r.update_params(p["vertices"], p["base_colors"], scene["densities"], p["color_grads"])
    for step in (2, 3):
        la = ra.train_step_batched(cams, 0, 1e-4, 1e-2, 1e-3)
        lb = rb.train_step_batched(cams, 0, 1e-4, 1e-2, 1e-3)
        assert ra.step == step and rb.step == step
        assert ra.pair_count() == rb.pair_count() > (1 << 20)
        assert abs(la - lb) <= 1e-5 * abs(lb)
        pa, pb = ra.get_params(), rb.get_params()
        for k in ("vertices", "densities", "color_grads"):
            assert _rel(pa[k], pb[k]) < 1e-5, (step, k)


def test_autograd_device_path_equals_host_path():
    """RMeshModule with its parameters on the GPU runs zero-copy (torch's
    stream, device-to-device copies): the same image and gradients as with
    CPU parameters (host arrays)."""
    import torch

    from rmesh_wgpu.autograd import RMeshModule

    rng = np.random.default_rng(17)
    s = make_scene(uniform_points(500, rng), rng)
    W = H = 64
    cam = fibonacci_cameras(1, 3.0, 50.0, W, H)[0]
    args = (s["vertices"], s["indices"], s["base_colors"], s["densities"], s["color_grads"], s["circumdata"], W, H)
    target = torch.rand((H, W, 3), generator=torch.Generator().manual_seed(0))
    out = {}
    for dev in ("cpu", "cuda"):
        m = RMeshModule(*args, device_id=0).to(dev)
        img = m(*(torch.tensor(c) for c in cam))
        assert img.device.type == dev
        loss = ((img[..., :3] - target.to(dev)) ** 2).mean() + 0.1 * img[..., 3].mean()
        loss.backward()
        out[dev] = (img.detach().cpu().numpy(),
                    {k: getattr(m, k).grad.detach().cpu().numpy() for k in
                     ("vertices", "base_colors", "densities", "color_grads")})
    np.testing.assert_array_equal(out["cpu"][0], out["cuda"][0])
    for k, g in out["cpu"][1].items():
        assert np.linalg.norm(g) > 0, k
        assert _rel(out["cuda"][1][k], g) <= 1e-6, k
