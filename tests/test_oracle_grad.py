"""CPU: the oracle's adjoints against float64 central finite differences of its
own float64 forward, and its SH evaluation against scipy's spherical harmonics.

The reference checks its adjoint the same way in float32
(kernel_tests.rs:733-1180, rel 0.10, eps 1e-3; test_autograd.py:138-250);
in float64 with eps 1e-6 the agreement is ~1e-9, so the bar here is 1e-6
relative L2 over the sampled entries — a dropped term, a wrong sign or a
swapped vertex slot fails it by orders of magnitude.
"""

from __future__ import annotations

import math

import numpy as np
import pytest

from oracle import oracle as orc
from rmeshvk_amd.camera import colmap_camera, fibonacci_cameras, make_camera
from rmeshvk_amd.scenes import make_scene, regular_tet_scene, single_tet_scene, two_tet_scene, uniform_points

FD_TOL = 1e-6
PARAMS = (("vertices", "d_vertices"), ("densities", "d_densities"), ("color_grads", "d_color_grads"),
          ("base_colors", "d_base_colors"))


def _f64(scene):
    return {k: (np.asarray(v, np.float64) if k in ("vertices", "densities", "color_grads", "base_colors", "sh")
                else v) for k, v in scene.items()}


def _fd_backward(scene, cam, w, h, seed=0, n_per=16, eps=1e-6, ts=12, checkpointed=False):
    rng = np.random.default_rng(seed)
    s = _f64(scene)
    fw = orc.render(s, cam, w, h, ts, dtype=np.float64)
    dl = rng.normal(0.0, 1.0, (h, w, 4))  # includes the alpha channel's path
    ck = None
    if checkpointed:
        ck = orc.forward(s, cam, w, h, fw["sorted_vals"], fw["ranges"], ts, dtype=np.float64, want_checkpoints=True)
    g = orc.backward(s, cam, w, h, fw["sorted_vals"], fw["ranges"], fw["rgba"], dl, ts, dtype=np.float64,
                     checkpoints=ck)

    def loss(sc):
        return float(np.sum(orc.render(sc, cam, w, h, ts, dtype=np.float64)["rgba"] * dl))

    res = {}
    for name, gk in PARAMS:
        n = len(s[name])
        idx = rng.choice(n, min(n, n_per), replace=False)
        an, num = [], []
        for i in idx:
            e = eps * max(1.0, abs(float(s[name][i])))  # relative step for large values (dense tets)
            sp = dict(s)
            sp[name] = s[name].copy()
            sp[name][i] += e
            sm = dict(s)
            sm[name] = s[name].copy()
            sm[name][i] -= e
            num.append((loss(sp) - loss(sm)) / (2 * e))
            an.append(g[gk][i])
        an, num = np.array(an), np.array(num)
        assert np.linalg.norm(num) > 1e-3, name  # the sample must see the parameter
        res[name] = np.linalg.norm(an - num) / np.linalg.norm(num)
    return res


def _scene_regular():
    s = regular_tet_scene(3.0)
    s["color_grads"] = np.array([0.3, -0.2, 0.25], np.float32)
    s["base_colors"] = np.array([0.9, 0.6, 0.3], np.float32)
    return s, make_camera((0.3, -2.0, 0.7), (0.0, 0.0, 0.0), math.pi / 2, 48, 48), 48, 48


def _scene_two():
    return (two_tet_scene(np.random.default_rng(1)),
            make_camera((0.5, -2.0, 1.5), (0.5, 0.3, 0.0), math.radians(60), 48, 48), 48, 48)


def _scene_single():
    return single_tet_scene(42), colmap_camera(32, 32, 1.0), 32, 32


def _scene_delaunay():
    rng = np.random.default_rng(7)
    s = make_scene(uniform_points(40, rng), rng)  # mixed-sign colours: exercises the clamps
    return s, fibonacci_cameras(2, 3.0, 50.0, 48, 48)[1], 48, 48


@pytest.mark.parametrize("make", [_scene_regular, _scene_two, _scene_single, _scene_delaunay],
                         ids=["regular", "two_tet", "single_tet", "delaunay40"])
def test_backward_matches_finite_differences(make):
    s, cam, w, h = make()
    res = _fd_backward(s, cam, w, h)
    for name, rel in res.items():
        assert rel < FD_TOL, (name, rel)


def _scene_saturating():
    """Dense tets: most covered pixels pass log T < -13.8 inside the mesh, so
    the forward stops compositing them and the tets behind never reach the
    image."""
    rng = np.random.default_rng(3)
    s = make_scene(uniform_points(60, rng), rng)
    s["densities"] = (s["densities"] * 40.0).astype(np.float32)
    return s, fibonacci_cameras(2, 3.0, 50.0, 40, 40)[0], 40, 40


@pytest.mark.parametrize("make", [_scene_regular, _scene_two, _scene_delaunay, _scene_saturating],
                         ids=["regular", "two_tet", "delaunay40", "saturating"])
def test_checkpointed_backward_matches_finite_differences(make):
    """The checkpointed adjoint (undo restarted every 32 tets from the
    forward's stored log T, tets after a pixel's stop skipped) is the exact
    adjoint of the forward, saturated pixels included."""
    s, cam, w, h = make()
    res = _fd_backward(s, cam, w, h, checkpointed=True, seed=2)
    for name, rel in res.items():
        assert rel < FD_TOL, (name, rel)


def _saturation_fraction(s, cam, w, h):
    fw = orc.render(_f64(s), cam, w, h, dtype=np.float64)
    a = fw["rgba"][..., 3]
    return float(np.mean(a > 1.0 - np.exp(-13.8)))


def test_undo_adjoint_differs_where_pixels_saturate():
    """The reference's undo walks back through tets the forward never
    composited (the pixel had saturated), and gives them gradients: on the
    saturating scene it misses the finite differences by far more than the
    bar, while the checkpointed adjoint above meets them — the FD check
    separates the two."""
    s, cam, w, h = _scene_saturating()
    assert _saturation_fraction(s, cam, w, h) > 0.3
    res = _fd_backward(s, cam, w, h, checkpointed=False, seed=2)
    assert max(res.values()) > 1e-2, res


def test_checkpointed_equals_undo_without_saturation():
    """Where no pixel saturates the two adjoints agree to float64 rounding:
    both reconstruct the same transmittances."""
    s, cam, w, h = _scene_delaunay()
    s = _f64(s)
    assert _saturation_fraction(s, cam, w, h) == 0.0
    fw = orc.render(s, cam, w, h, dtype=np.float64)
    ck = orc.forward(s, cam, w, h, fw["sorted_vals"], fw["ranges"], dtype=np.float64, want_checkpoints=True)
    dl = np.random.default_rng(3).normal(0.0, 1.0, (h, w, 4))
    a = orc.backward(s, cam, w, h, fw["sorted_vals"], fw["ranges"], fw["rgba"], dl, dtype=np.float64)
    b = orc.backward(s, cam, w, h, fw["sorted_vals"], fw["ranges"], fw["rgba"], dl, dtype=np.float64, checkpoints=ck)
    for k in a:
        assert np.linalg.norm(a[k] - b[k]) <= 1e-10 * np.linalg.norm(a[k]), k


def test_checkpoints_hold_the_forward_transmittance():
    """Every stored checkpoint is the pixel's log T before tile-list position
    32 j, recomputed here by compositing only the first 32 j tets of each
    tile (float64); the final log T is the image's -log(1 - alpha)."""
    rng = np.random.default_rng(12)
    s = _f64(make_scene(uniform_points(900, rng), rng))
    w = h = 48
    cam = fibonacci_cameras(1, 3.0, 50.0, w, h)[0]
    ts = 8
    fw = orc.render(s, cam, w, h, ts, dtype=np.float64)
    ck = orc.forward(s, cam, w, h, fw["sorted_vals"], fw["ranges"], ts, dtype=np.float64, want_checkpoints=True)
    np.testing.assert_allclose(np.exp(ck["fin_logt"]), 1.0 - fw["rgba"][..., 3], rtol=1e-12, atol=1e-15)
    tx, ty, T = orc.tile_grid(w, h, ts)
    rg = fw["ranges"]
    checked = 0
    for tile in range(T):
        rs, re = int(rg[2 * tile]), int(rg[2 * tile + 1])
        n = re - rs if re > rs else 0
        for j in range(1, (n + 31) // 32):
            q = 32 * j
            part = rg.copy()
            part[2 * tile] = re - q  # the first q tets in compositing order
            sub = orc.forward(s, cam, w, h, fw["sorted_vals"], part, ts, dtype=np.float64, tiles=(tile, tile + 1))
            alpha = sub["rgba"][..., 3]
            slot = (rs >> 5) + tile + j - 1
            got = ck["ck"][slot * ts * ts:(slot + 1) * ts * ts].reshape(ts, ts)
            x0, y0 = (tile % tx) * ts, (tile // tx) * ts
            want = np.log(np.maximum(1.0 - alpha[y0:y0 + ts, x0:x0 + ts], 1e-300))
            stop = ck["fin_stop"][y0:y0 + ts, x0:x0 + ts]
            live = stop[: want.shape[0], : want.shape[1]] > q - 1  # the pixel was still compositing at q
            gw = got[: want.shape[0], : want.shape[1]]
            np.testing.assert_allclose(gw[live], want[live], rtol=1e-12, atol=1e-12)
            checked += int(live.sum())
    assert checked > 100


def test_oracle_adjoint_independent_of_thread_count():
    """Tiles run in parallel; per-tile sums are added in tile order, so one
    thread and many give the same bits."""
    rng = np.random.default_rng(5)
    s = make_scene(uniform_points(1200, rng), rng)
    w = h = 96
    cam = fibonacci_cameras(1, 3.0, 50.0, w, h)[0]
    fw = orc.render(s, cam, w, h, 8)
    dl = np.random.default_rng(6).normal(0.0, 1e-3, (h, w, 4)).astype(np.float32)
    outs = []
    n0 = orc.get_threads()
    try:
        for n in (1, 7):
            orc.set_threads(n)
            outs.append(orc.backward(s, cam, w, h, fw["sorted_vals"], fw["ranges"], fw["rgba"], dl, 8))
    finally:
        orc.set_threads(n0)
    for k in outs[0]:
        np.testing.assert_array_equal(outs[0][k], outs[1][k])


def test_loss_gradient_chain():
    """L1/L2 loss gradient composed with the adjoint equals the FD of the loss."""
    s, cam, w, h = _scene_two()
    s = _f64(s)
    gt = np.random.default_rng(4).random((h, w, 3))
    for lt in (0, 1):
        _, _, loss, g = orc.render_backward(s, cam, w, h, gt, lt, dtype=np.float64)
        eps = 1e-6
        for name, gk in PARAMS[:2]:
            i = 1
            sp = dict(s)
            sp[name] = s[name].copy()
            sp[name][i] += eps
            sm = dict(s)
            sm[name] = s[name].copy()
            sm[name][i] -= eps
            lp = orc.loss_grad(orc.render(sp, cam, w, h, dtype=np.float64)["rgba"], gt, lt, np.float64)[1]
            lm = orc.loss_grad(orc.render(sm, cam, w, h, dtype=np.float64)["rgba"], gt, lt, np.float64)[1]
            num = (lp - lm) / (2 * eps)
            assert abs(g[gk][i] - num) <= 1e-6 * max(abs(num), 1e-3), (lt, name, g[gk][i], num)


# --------------------------------------------------------------------------
# SH (sh_eval_compute.wgsl)
# --------------------------------------------------------------------------
def _direction_grid(nth=8, nph=16):
    xg, wg = np.polynomial.legendre.leggauss(nth)
    dirs, wts = [], []
    for ct, wt in zip(xg, wg):
        st = math.sqrt(1 - ct * ct)
        for k in range(nph):
            ph = 2 * math.pi * k / nph
            dirs.append([st * math.cos(ph), st * math.sin(ph), ct])
            wts.append(wt * 2 * math.pi / nph)
    return np.array(dirs), np.array(wts)


def _basis_from_oracle(dirs, nb=16, a=0.05):
    """Y_j(d) recovered from the oracle's colours: one tiny tet per (direction,
    basis function) at 5 d from a camera at the origin, SH one-hot with amplitude
    a, zero colour gradient; colour = softplus(a Y_j(d) + 0.5)."""
    base = np.array([[1, 1, 1], [-1, -1, 1], [-1, 1, -1], [1, -1, -1]], np.float64) * 0.01
    D = len(dirs)
    verts = np.concatenate([base + 5 * d for d in dirs for _ in range(nb)])
    sh = np.zeros((D * nb, 3, nb))
    for t in range(D * nb):
        sh[t, :, t % nb] = a
    scene = {"vertices": verts.ravel(), "indices": np.arange(4 * D * nb, dtype=np.uint32), "sh": sh.ravel(),
             "sh_degree": 3, "color_grads": np.zeros(3 * D * nb)}
    col = orc.sh_colors(scene, np.zeros(3, np.float32), dtype=np.float64).reshape(-1, 3)
    x = 0.1 * np.log(np.expm1(10.0 * col))  # softplus (beta 10) inverse
    assert np.allclose(x[:, 0], x[:, 2]) and np.allclose(x[:, 0], x[:, 1])
    return ((x[:, 0] - 0.5) / a).reshape(D, nb)


def test_sh_basis_is_scipy_real_sh():
    """The 16 basis functions equal scipy's real spherical harmonics with the
    Condon-Shortley phase (m < 0: sqrt2 Im Y_l^|m|, m > 0: sqrt2 Re Y_l^m),
    index l^2 + l + m; hence orthonormal on the sphere."""
    from scipy.special import sph_harm_y

    dirs, wts = _direction_grid()
    Y = _basis_from_oracle(dirs)
    th = np.arccos(np.clip(dirs[:, 2], -1, 1))
    ph = np.arctan2(dirs[:, 1], dirs[:, 0])
    R = np.zeros_like(Y)
    for l in range(4):  # noqa: E741
        for m in range(-l, l + 1):
            i = l * l + l + m
            if m < 0:
                R[:, i] = math.sqrt(2) * np.imag(sph_harm_y(l, -m, th, ph))
            elif m == 0:
                R[:, i] = np.real(sph_harm_y(l, 0, th, ph))
            else:
                R[:, i] = math.sqrt(2) * np.real(sph_harm_y(l, m, th, ph))
    np.testing.assert_allclose(Y, R, atol=1e-9)
    G = (Y * wts[:, None]).T @ Y
    np.testing.assert_allclose(G, np.eye(16), atol=1e-9)


def test_sh_degree0_reduction():
    """Degree 0: colour = softplus(sh0 / (2 sqrt(pi)) + 0.5 + g . (v0 - centroid))."""
    rng = np.random.default_rng(3)
    s = make_scene(uniform_points(50, rng), rng, sh_degree=0)
    s = _f64(s)
    cam = np.array([0.3, -2.5, 1.0], np.float32)
    got = orc.sh_colors(s, cam, dtype=np.float64).reshape(-1, 3)
    V = s["vertices"].reshape(-1, 3)
    tv = V[s["indices"].reshape(-1, 4).astype(np.int64)]
    off = np.einsum("mk,mk->m", s["color_grads"].reshape(-1, 3), tv[:, 0] - tv.mean(1))
    x = s["sh"].reshape(-1, 3) / (2 * math.sqrt(math.pi)) + 0.5 + off[:, None]
    sp = np.where(10 * x > 20, x, 0.1 * np.log1p(np.exp(10 * x)))
    np.testing.assert_allclose(got, sp, rtol=1e-12, atol=1e-14)


@pytest.mark.parametrize("deg", [0, 1, 2, 3])
def test_sh_backward_matches_finite_differences(deg):
    rng = np.random.default_rng(deg)
    s = _f64(make_scene(uniform_points(30, rng), rng, sh_degree=deg))
    cam = np.array([0.4, -2.2, 0.9], np.float32)
    M = len(s["densities"])
    w = rng.normal(size=3 * M)
    g = orc.sh_backward(s, cam, w, dtype=np.float64)

    def f(sc):
        return float(np.dot(orc.sh_colors(sc, cam, dtype=np.float64), w))

    eps = 1e-6
    for name, gk in (("sh", "d_sh"), ("color_grads", "d_color_grads"), ("vertices", "d_vertices")):
        idx = rng.choice(len(s[name]), min(len(s[name]), 20), replace=False)
        an, num = [], []
        for i in idx:
            sp = dict(s)
            sp[name] = s[name].copy()
            sp[name][i] += eps
            sm = dict(s)
            sm[name] = s[name].copy()
            sm[name][i] -= eps
            num.append((f(sp) - f(sm)) / (2 * eps))
            an.append(g[gk][i])
        an, num = np.array(an), np.array(num)
        assert np.linalg.norm(an - num) <= FD_TOL * max(np.linalg.norm(num), 1e-6), (name, an, num)
