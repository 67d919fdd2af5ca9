"""GPU: the checkpointed adjoint (RMESH_ADJOINT_CHECKPOINT) through the C ABI.

- forward_tiled + backward in checkpoint mode against the oracle's
  checkpointed adjoint (float32): images within 1e-4 max-abs, every gradient
  within 1e-3 relative L2, at tile sizes 8/12/16, on a plain and on a
  saturating scene (dense tets: the forward stops compositing most pixels
  inside the mesh, where the two adjoints differ);
- on the saturating scene the checkpointed GPU adjoint is also within 1e-3
  of the float64 checkpointed oracle, which float64 finite differences pin
  (tests/test_oracle_grad.py), while the undo mode is not;
- the batched, graph-replayed train step in checkpoint mode equals the sum of
  per-view checkpointed steps;
- mode switches: a backward needs a forward in its own mode;
- the adjoint variants with more than 32 tets per batch, which run this mode
  with 32.
"""

import math

import numpy as np
import pytest

from oracle import oracle as orc
from rmeshvk_amd.camera import fibonacci_cameras
from rmeshvk_amd.scenes import make_scene, uniform_points

pytestmark = pytest.mark.gpu

GRAD_TOL = 1e-3
KEYS = ("d_vertices", "d_densities", "d_color_grads", "d_base_colors")


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


def _plain():
    rng = np.random.default_rng(31)
    s = make_scene(uniform_points(900, rng), rng)
    return s, fibonacci_cameras(2, 3.0, 50.0, 80, 72)[1], 80, 72


def _saturating():
    rng = np.random.default_rng(32)
    s = make_scene(uniform_points(900, rng), rng)
    s["densities"] = (s["densities"] * 30.0).astype(np.float32)
    return s, fibonacci_cameras(2, 3.0, 50.0, 80, 72)[0], 80, 72


def _f64(scene):
    return {k: (np.asarray(v, np.float64) if k in ("vertices", "base_colors", "densities", "color_grads", "sh")
                else v) for k, v in scene.items()}


def _oracle_ck(scene, cam, w, h, ts, dl, dtype=np.float32):
    s = scene if dtype == np.float32 else _f64(scene)
    ref = orc.render(scene, cam, w, h, ts)  # float32 binning: the lists the GPU uses
    fw = orc.forward(s, cam, w, h, ref["sorted_vals"], ref["ranges"], ts, dtype=dtype, want_checkpoints=True)
    g = orc.backward(s, cam, w, h, ref["sorted_vals"], ref["ranges"], fw["rgba"], np.asarray(dl, dtype), ts,
                     dtype=dtype, checkpoints=fw)
    return ref, fw, g


@pytest.mark.parametrize("make", [_plain, _saturating], ids=["plain", "saturating"])
@pytest.mark.parametrize("ts", [8, 12, 16])
def test_checkpointed_backward_matches_oracle(make, ts):
    scene, cam, w, h = make()
    r = _renderer(scene, w, h, tile_size=ts)
    r.set_adjoint_mode("checkpoint")
    assert r.get_adjoint_mode() == "checkpoint"
    img = r.forward_tiled(*cam)
    dl = np.ascontiguousarray(np.random.default_rng(ts).normal(0, 1e-2, (h, w, 4)).astype(np.float32))
    g = r.backward(dl)
    ref, fw, go = _oracle_ck(scene, cam, w, h, ts, dl)
    assert np.abs(img - ref["rgba"]).max() <= 1e-4
    for k in KEYS:
        assert _rel(g[k], go[k]) <= GRAD_TOL, (k, _rel(g[k], go[k]))


@pytest.mark.parametrize("bv", [7, 8])
def test_checkpointed_with_wide_batch_variants(bv):
    """The adjoint variants with more than 32 tets per batch run the
    checkpointed mode with 32 (a batch spans at most two checkpoint groups)."""
    scene, cam, w, h = _saturating()
    r = _renderer(scene, w, h, tile_size=16)
    r.set_tuning(16, 0, bv)
    r.set_adjoint_mode("checkpoint")
    img = r.forward_tiled(*cam)
    dl = np.ascontiguousarray(np.random.default_rng(bv).normal(0, 1e-2, (h, w, 4)).astype(np.float32))
    g = r.backward(dl)
    ref, fw, go = _oracle_ck(scene, cam, w, h, 16, dl)
    assert np.abs(img - ref["rgba"]).max() <= 1e-4
    for k in KEYS:
        assert _rel(g[k], go[k]) <= GRAD_TOL, (k, _rel(g[k], go[k]))


def test_checkpointed_is_exact_where_undo_is_not():
    """Saturating scene: the checkpointed GPU adjoint is within 1e-3 of the
    float64 checkpointed oracle (the forward's exact adjoint); the undo mode's
    gradients, which give tets behind saturated pixels a share, are not."""
    scene, cam, w, h = _saturating()
    dl = np.ascontiguousarray(np.random.default_rng(5).normal(0, 1e-2, (h, w, 4)).astype(np.float32))
    _, _, g64 = _oracle_ck(scene, cam, w, h, 12, dl, np.float64)
    r = _renderer(scene, w, h)
    r.forward_tiled(*cam)
    g_undo = r.backward(dl)
    r.set_adjoint_mode("checkpoint")
    r.forward_tiled(*cam)
    g_ck = r.backward(dl)
    for k in KEYS:
        assert _rel(g_ck[k], g64[k]) <= GRAD_TOL, (k, _rel(g_ck[k], g64[k]))
    assert max(_rel(g_undo[k], g64[k]) for k in KEYS) > 1e-2


def test_checkpointed_batched_graph_step_sums_views():
    rng = np.random.default_rng(41)
    scene = make_scene(uniform_points(700, rng), rng, sh_degree=3)
    w = h = 64
    cams = fibonacci_cameras(3, 3.0, 50.0, w, h)
    V = len(cams)
    gts = np.ascontiguousarray(np.random.default_rng(2).random((V, h, w, 3)).astype(np.float32))
    rb = _renderer(scene, w, h, max_views=V)
    rb.set_adjoint_mode("checkpoint")
    rb.set_targets(gts)
    rb.set_graph_mode(True)
    for _ in range(2):  # capture, then a replay
        rb.train_step_batched(cams, 1, 0.0, 0.0, 0.0, 0.0)
    gb = rb.gradients()
    gb["d_sh"] = rb.sh_gradient()
    r1 = _renderer(scene, w, h)
    r1.set_adjoint_mode("checkpoint")
    acc = {k: np.zeros(len(v)) for k, v in gb.items()}
    for v in range(V):
        r1.train_step(*cams[v], gts[v], 0.0, 0.0, 0.0, 1)
        g = r1.gradients()
        g["d_sh"] = r1.sh_gradient()
        for k in acc:
            acc[k] += g[k]
    for k in acc:
        assert _rel(gb[k], acc[k]) <= 1e-5, (k, _rel(gb[k], acc[k]))
    # and a batched step in checkpoint mode differs from one in undo mode only
    # through the transmittance it starts from (same images)
    ru = _renderer(scene, w, h, max_views=V)
    ru.set_targets(gts)
    ru.train_step_batched(cams, 1, 0.0, 0.0, 0.0, 0.0)
    np.testing.assert_array_equal(ru.read_buffer(0, V * w * h * 4), rb.read_buffer(0, V * w * h * 4))


def test_mode_switch_needs_a_new_forward():
    scene, cam, w, h = _plain()
    r = _renderer(scene, w, h)
    r.forward_tiled(*cam)
    r.set_adjoint_mode("checkpoint")
    dl = np.zeros((h, w, 4), np.float32)
    with pytest.raises(ValueError):
        r.backward(dl)
    r.forward_tiled(*cam, render_aux=True)  # the aux forward keeps no checkpoints
    with pytest.raises(ValueError):
        r.backward(dl)
    r.forward_tiled(*cam)
    r.backward(dl)
    with pytest.raises(ValueError):
        r.set_adjoint_mode("other")
    assert r.device_bytes() > 0
