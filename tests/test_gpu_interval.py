"""GPU: interval-tiled forward (RMeshRenderer.forward_interval_tiled,
lib.rs:1218-1390; interval_generate.wgsl, interval_tiled_rasterize.wgsl)
against the oracle's restatement over the same tile lists.

Bars: image within 1e-4 max-abs (the images' bar).  Depth (xyzd.w) and the
first four distortion channels within 1e-4 relative to each channel's
largest magnitude: both sides evaluate the same float32 expressions, so the
differences are expf roundings (device vs libm) carried through the NDC-z
linearization, whose 1 - z denominator amplifies them.  The self term
(channel 4) sums (1 - 2u e^-u - e^-2u) / sigma per pair, which is ~u^3 / 3
after the cancellation of three O(1) terms: one ulp of e^-u in either
evaluation moves it by ~2e-7 / sigma, ~1e-4 of the channel's maximum over a
pixel's pairs, hence 1e-3 there.  xyz of xyzd is exactly zero (no vertex
normals on this path).
"""

import numpy as np
import pytest

from oracle import oracle as orc

pytestmark = pytest.mark.gpu

IMG_TOL = 1e-4
DEPTH_TOL = np.array([1e-4])
DIST_TOL = np.array([1e-4, 1e-4, 1e-4, 1e-4, 1e-3])


def _renderer(scene, w, h, **kw):
    from rmeshvk_amd import RMeshRenderer

    extra = {}
    if "sh" in scene:
        extra = dict(sh_coeffs=scene["sh"], sh_degree=int(scene["sh_degree"]))
    extra.update(kw)
    return RMeshRenderer(scene["vertices"], scene["indices"], scene["base_colors"], scene["densities"],
                         scene["color_grads"], scene["circumdata"], w, h, device_id=0, **extra)


@pytest.fixture(scope="module", params=["A", "B"])
def config(request):
    from rmeshvk_amd.scenes import config_scene

    scene, cams = config_scene(request.param, seed=0)
    return request.param, scene, cams, 256, 256


def _check(got, ref):
    img = np.abs(got["image"] - ref["image"]).max()
    assert img <= IMG_TOL, ("image", img)
    assert not np.any(got["xyzd"][..., :3])
    for name, a, b, tol in (("depth", got["xyzd"][..., 3:], ref["xyzd"][..., 3:], DEPTH_TOL),
                            ("distortion", got["distortion"], ref["distortion"], DIST_TOL)):
        scale = np.abs(b).reshape(-1, b.shape[-1]).max(axis=0)
        assert np.all(scale > 0), name
        err = np.abs(a - b).reshape(-1, b.shape[-1]).max(axis=0) / scale
        assert np.all(err <= tol), (name, err)


@pytest.mark.parametrize("ts", [8, 12, 16])
def test_interval_forward_matches_oracle(ts, config):
    name, scene, cams, w, h = config
    r = _renderer(scene, w, h, tile_size=ts)
    for cam in cams[:2]:
        got = r.forward_interval_tiled(*cam)
        ref = orc.render(scene, cam, w, h, ts=ts)
        iv = orc.interval_forward(scene, cam, w, h, ref["sorted_vals"], ref["ranges"], ts=ts, colors=ref["colors"])
        assert iv["image"][..., 3].max() > 0.5
        _check(got, iv)


def test_interval_forward_ragged_image(config):
    """An image size that is no multiple of the tile size (partial edge tiles)."""
    name, scene, cams, _, _ = config
    w, h = 203, 117
    r = _renderer(scene, w, h, tile_size=16)
    from rmeshvk_amd.camera import make_camera

    cam = make_camera((0.2, -2.6, 0.5), (0.0, 0.0, 0.0), 0.9, w, h)
    got = r.forward_interval_tiled(*cam)
    ref = orc.render(scene, cam, w, h, ts=16)
    iv = orc.interval_forward(scene, cam, w, h, ref["sorted_vals"], ref["ranges"], ts=16, colors=ref["colors"])
    _check(got, iv)


def test_interval_forward_outputs_and_state(config):
    name, scene, cams, w, h = config
    r = _renderer(scene, w, h)
    out = r.forward_interval_tiled(*cams[0])
    assert set(out) == {"image", "xyzd", "distortion"}
    assert out["image"].shape == (h, w, 4) and out["xyzd"].shape == (h, w, 4)
    assert out["distortion"].shape == (h, w, 5)
    # the interval forward leaves no state for the tiled adjoint
    with pytest.raises(ValueError):
        r.backward(np.zeros((h, w, 4), np.float32))
    # a tiled forward afterwards renders as usual
    img = r.forward_tiled(*cams[0])
    ref = orc.render(scene, cams[0], w, h, ts=12)
    assert np.abs(img - ref["rgba"]).max() <= IMG_TOL
    with pytest.raises(TypeError):
        r.forward_interval_tiled(cams[0][0].astype(np.float64), cams[0][1], cams[0][2])


# ---------------------------------------------------------------------------
# backward_interval_tiled (lib.rs:2049-2240): GPU adjoint + chain back against
# the oracle's restatement.  Bar: each output within 1e-3 relative L2 (the
# gradients' bar): float32 sums in another order, device vs libm exp.
# ---------------------------------------------------------------------------
GRAD_TOL = 1e-3
_OUTS = ("d_vertices", "d_densities", "d_base_colors", "d_color_grads", "d_vertex_normals")


def _rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / nb if nb > 0 else np.linalg.norm(a)


@pytest.mark.parametrize("ts", [8, 12, 16])
@pytest.mark.parametrize("extra", [False, True])
def test_interval_backward_matches_oracle(ts, extra, config):
    name, scene, cams, w, h = config
    r = _renderer(scene, w, h, tile_size=ts)
    cam = cams[0]
    r.forward_interval_tiled(*cam)
    rng = np.random.default_rng(ts + 7 * extra)
    dl = rng.normal(0.0, 1.0, (h, w, 4)).astype(np.float32)
    dx = rng.normal(0.0, 1.0, (h, w, 4)).astype(np.float32) if extra else None
    dd = rng.normal(0.0, 1.0, (h, w, 5)).astype(np.float32) if extra else None
    got = r.backward_interval_tiled(dl, dx, dd)
    ref = orc.render(scene, cam, w, h, ts=ts)
    exp = orc.interval_backward_full(scene, cam, w, h, ref["sorted_vals"], ref["ranges"], dl, dx, dd, ts=ts,
                                     colors=ref["colors"])
    for k in _OUTS:
        if not extra and k == "d_vertex_normals":
            assert not np.any(got[k]) and not np.any(exp[k])
            continue
        assert np.linalg.norm(exp[k]) > 0, k
        assert _rel(got[k], exp[k]) <= GRAD_TOL, (k, _rel(got[k], exp[k]))


def test_interval_backward_state_and_arguments(config):
    name, scene, cams, w, h = config
    r = _renderer(scene, w, h)
    dl = np.zeros((h, w, 4), np.float32)
    with pytest.raises(ValueError, match="requires a preceding forward_interval_tiled"):
        r.backward_interval_tiled(dl)
    r.forward_interval_tiled(*cams[0])
    with pytest.raises(ValueError):
        r.backward_interval_tiled(np.zeros(7, np.float32))
    with pytest.raises(ValueError):
        r.backward_interval_tiled(dl, dl_d_distortion=np.zeros((h, w, 4), np.float32))
    g = r.backward_interval_tiled(dl)
    assert all(not np.any(g[k]) for k in _OUTS)
    # a tiled forward in between invalidates the interval state
    r.forward_tiled(*cams[0])
    with pytest.raises(ValueError):
        r.backward_interval_tiled(dl)


def test_autograd_interval_mode(config):
    """RMeshModule(render_mode="interval_tiled"): the image of
    forward_interval_tiled and the gradients of backward_interval_tiled."""
    import torch

    from rmeshvk_amd.autograd import RMeshModule

    name, scene, cams, w, h = config
    # the module renders the base colours: config B's SH stay out of both
    # sides, and its zero base colours are replaced by random ones
    if "sh" in scene:
        scene = {k: v for k, v in scene.items() if k != "sh"}
        m = len(scene["indices"]) // 4
        scene["base_colors"] = np.random.default_rng(5).uniform(0.2, 0.8, 3 * m).astype(np.float32)
    mod = RMeshModule(scene["vertices"], scene["indices"], scene["base_colors"], scene["densities"],
                      scene["color_grads"], scene["circumdata"], w, h, device_id=0)
    cam = [torch.from_numpy(np.asarray(c)) for c in cams[0]]
    img = mod(*cam, render_mode="interval_tiled")
    dl = torch.from_numpy(np.random.default_rng(3).normal(0.0, 1.0, (h, w, 4)).astype(np.float32))
    (img * dl).sum().backward()
    r = _renderer(scene, w, h)
    exp_img = r.forward_interval_tiled(*cams[0])["image"]
    g = r.backward_interval_tiled(dl.numpy())
    np.testing.assert_array_equal(img.detach().numpy(), exp_img)
    for p, k in ((mod.vertices, "d_vertices"), (mod.base_colors, "d_base_colors"), (mod.densities, "d_densities"),
                 (mod.color_grads, "d_color_grads")):
        assert _rel(p.grad.numpy(), g[k]) <= 1e-5, k
