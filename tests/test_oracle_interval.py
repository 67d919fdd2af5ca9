"""CPU: the oracle's interval-tiled forward (forward_interval_tiled,
interval_generate.wgsl + interval_tiled_rasterize.wgsl) against independent
float64 references.

The interval path interpolates NDC z over screen-space triangles and
linearizes it with the reference's near/far (0.01 / 1000); with a camera
built with the same near/far and planar faces, the linearized chord is the
ray chord, so with a constant colour per tet (g = 0) the image equals the float64
quadrature renderer (tests/npref.py) and depth x ray scale equals the
chord-weighted depth.  With g != 0 the reference interpolates the colour
offsets with screen-space (not perspective-correct) barycentrics; as the
projection becomes affine that rule tends to the linear colour field the
quadrature integrates, which pins it (test_colour_offsets_in_the_affine_limit).
Distortion (Mip-NeRF 360 state: x, y, z = sum w, w = sum w t, self term)
is checked against float64 quadrature of the weight T sigma e^-sigma (t - t0)
over the chords.  Bars: float32 NDC z at a 1e5 near/far ratio loses ~1e-4,
hence 5e-4 on the image.
"""

from __future__ import annotations

import math

import numpy as np

import npref
from oracle import oracle as orc
from rmeshvk_amd.camera import make_camera
from rmeshvk_amd.scenes import regular_tet_scene, two_tet_scene

TOL = 5e-4


def _cam(eye, target, fov, w, h):
    return make_camera(eye, target, fov, w, h, near=0.01, far=1000.0)


def _run(scene, cam, w, h):
    ref = orc.render(scene, cam, w, h)
    return ref, orc.interval_forward(scene, cam, w, h, ref["sorted_vals"], ref["ranges"])


def test_single_tet_image_and_depth():
    s = regular_tet_scene(2.0)
    s["base_colors"] = np.array([1.0, 0.8, 0.9], np.float32)
    s["color_grads"] = np.zeros(3, np.float32)
    w, h = 48, 40
    cam = _cam((0.7, -2.0, 0.9), (0.0, 0.1, 0.0), math.radians(70), w, h)
    assert npref.endpoint_colours_positive(s, cam, w, h)
    ref, iv = _run(s, cam, w, h)
    q = npref.render(s, cam, w, h)
    assert q[..., 3].max() > 0.5
    assert np.abs(iv["image"] - q).max() < TOL
    # depth: view z x ray scale is the ray parameter; compare with the
    # quadrature's sum over the chord of T w(t) t, which for one tet is the
    # closed form m1 - t0 (1 - e^-u) ... here checked through the weights
    o, d = npref.pixel_rays(cam, w, h)
    t0, t1, hit = npref.tet_chords(np.asarray(s["vertices"], np.float64).reshape(1, 4, 3), o, d)
    t0, t1, hit = t0[:, 0], t1[:, 0], hit[:, 0]
    ci = np.asarray(cam[2], np.float64)
    px, py = np.meshgrid(np.arange(w) + 0.5, np.arange(h) + 0.5)
    rsc = np.sqrt(((px - ci[14]) / ci[12]) ** 2 + ((py - ci[15]) / ci[13]) ** 2 + 1.0).ravel()
    od = 2.0 * (t1 - t0)
    phi = np.where(od > 0, (1 - np.exp(-od)) / np.where(od > 0, od, 1), 1.0)
    w0, w1 = phi - np.exp(-od), 1 - phi
    depth = np.where(hit, w0 * t1 + w1 * t0, 0.0)  # in ray parameter units
    got = iv["xyzd"][..., 3].ravel() * rsc
    assert np.abs(got - depth).max() < 2e-3 * max(1.0, np.abs(depth).max())
    assert np.all(iv["xyzd"][..., :3] == 0.0)  # no vertex normals on this path


def test_two_tets_image():
    s = two_tet_scene(np.random.default_rng(1))
    s["color_grads"] = np.zeros_like(s["color_grads"])
    w, h = 64, 64
    cam = _cam((0.5, -2.0, 1.5), (0.5, 0.3, 0.0), math.radians(60), w, h)
    assert npref.endpoint_colours_positive(s, cam, w, h)
    ref, iv = _run(s, cam, w, h)
    q = npref.render(s, cam, w, h)
    assert (q[..., 3] > 0.01).sum() > 100
    assert np.abs(iv["image"] - q).max() < TOL
    assert np.abs(iv["image"] - ref["rgba"]).max() < TOL


def _gl(a, b, n=48):
    x, wt = np.polynomial.legendre.leggauss(n)
    return 0.5 * (b - a)[..., None] * x + 0.5 * (b + a)[..., None], 0.5 * (b - a)[..., None] * wt


def test_distortion_moments():
    """One tet (T = 1): z = 1 - e^-u, w = int w t, v = int int w w |s - t|;
    two tets: x - y = 2 sum_{i<j} int int w_i(s) w_j(t) (t - s)."""
    s = two_tet_scene(np.random.default_rng(1))
    w, h = 48, 48
    cam = _cam((0.5, -2.0, 1.5), (0.5, 0.3, 0.0), math.radians(60), w, h)
    ref, iv = _run(s, cam, w, h)
    o, d = npref.pixel_rays(cam, w, h)
    V = np.asarray(s["vertices"], np.float64).reshape(-1, 3)
    tv = V[np.asarray(s["indices"]).reshape(-1, 4)]
    t0, t1, hit = npref.tet_chords(tv, o, d)
    sig = np.asarray(s["densities"], np.float64)
    order = npref.depth_order(s, cam[0])
    P = len(d)
    logT = np.zeros(P)
    mom0, mom1, selfd, cross = [], [], np.zeros(P), np.zeros(P)
    for m in order:
        a, b, ok = t0[:, m], t1[:, m], hit[:, m]
        ts_, wts = _gl(np.where(ok, a, 0.0), np.where(ok, b, 0.0))
        T = np.exp(logT)[:, None]
        wfun = T * sig[m] * np.exp(-sig[m] * (ts_ - np.where(ok, a, 0.0)[:, None]))
        m0 = np.where(ok, (wfun * wts).sum(1), 0.0)
        m1 = np.where(ok, (wfun * ts_ * wts).sum(1), 0.0)
        # self term: int int w(s) w(t) |s - t| = 2 int_{s<t}; nested quadrature
        ss, wss = _gl(np.repeat(np.where(ok, a, 0.0)[:, None], ts_.shape[1], 1), ts_, 24)
        ws_ = T[..., None] * sig[m] * np.exp(-sig[m] * (ss - np.where(ok, a, 0.0)[:, None, None]))
        inner = (ws_ * (ts_[..., None] - ss) * wss).sum(-1)
        sd = np.where(ok, 2.0 * (wfun * inner * wts).sum(1), 0.0)
        for pm0, pm1 in zip(mom0, mom1):
            cross += 2.0 * (pm0 * m1 - pm1 * m0)
        mom0.append(m0)
        mom1.append(m1)
        selfd += sd
        logT = np.where(ok, logT - sig[m] * (b - a), logT)
    ds = iv["distortion"].reshape(-1, 5)
    z = sum(mom0)
    wsum = sum(mom1)
    assert (z > 0.05).sum() > 100
    np.testing.assert_allclose(ds[:, 2], z, atol=2e-3)
    np.testing.assert_allclose(ds[:, 3], wsum, atol=2e-3 * max(1.0, np.abs(wsum).max()))
    np.testing.assert_allclose(ds[:, 4], selfd, atol=2e-3 * max(1.0, np.abs(selfd).max()))
    np.testing.assert_allclose(ds[:, 0] - ds[:, 1], cross, atol=2e-3 * max(1.0, np.abs(cross).max()))


# ---------------------------------------------------------------------------
# Adjoint (backward_interval_tiled.wgsl + interval_chain_back.wgsl).  The
# reference differentiates the image with the screen-space geometry (pixel
# positions of the silhouette, barycentric weights, crossing parameters)
# held fixed: its per-tile stage yields the derivative with respect to each
# slot's front/back NDC z and colour offsets, the base colour and the
# density, and its chain back maps those through the NDC z and offsets of
# the vertices.  Both stages are pinned by float64 central differences at
# exactly that level: the compositing over frozen decompositions, and the
# attributes of moved vertices under frozen choices and weights.
# ---------------------------------------------------------------------------

_ATTR = np.r_[10:30, 60:64]  # slot zf, zb, of, ob; base colour; density


def _mesh_scene(seed, n=60):
    from rmeshvk_amd.scenes import make_scene, uniform_points

    rng = np.random.default_rng(seed)
    s = make_scene(uniform_points(n, rng), rng)
    s["densities"] = (0.5 * s["densities"]).astype(np.float32)
    s["base_colors"] = np.abs(s["base_colors"]) + 0.2
    return {k: (np.asarray(v, np.float64) if k in ("vertices", "base_colors", "densities", "color_grads")
                else v) for k, v in s.items()}


def test_interval_adjoint_colour_matches_finite_differences():
    s = _mesh_scene(5)
    w, h = 40, 32
    cam = _cam((0.3, -2.8, 0.6), (0.0, 0.0, 0.0), math.radians(60), w, h)
    ref = orc.render(s, cam, w, h, ts=8)
    ivl = orc.interval_generate(s, cam, w, h, dtype=np.float64)
    fwd = orc.interval_composite(ivl, cam, w, h, ref["sorted_vals"], ref["ranges"], ts=8, dtype=np.float64)
    assert fwd["image"][..., 3].max() > 0.3
    rng = np.random.default_rng(1)
    dl = rng.normal(0.0, 1.0, (h, w, 4))
    d = orc.interval_backward(ivl, cam, w, h, ref["sorted_vals"], ref["ranges"], fwd, dl, ts=8, dtype=np.float64)

    def loss(v):
        f = orc.interval_composite(v, cam, w, h, ref["sorted_vals"], ref["ranges"], ts=8, dtype=np.float64)
        return float(np.sum(f["image"] * dl))

    live = np.flatnonzero(np.abs(d[:, _ATTR]).sum(axis=1) > 0)
    assert len(live) > 20
    got, fd = [], []
    eps = 1e-6
    for t in rng.choice(live, 24, replace=False):
        for k in _ATTR:
            p = ivl.copy()
            p[t, k] += eps
            lp = loss(p)
            p[t, k] -= 2 * eps
            fd.append((lp - loss(p)) / (2 * eps))
            got.append(d[t, k])
    got, fd = np.array(got), np.array(fd)
    assert np.linalg.norm(got - fd) / np.linalg.norm(fd) < 1e-5


def test_interval_chain_back_matches_finite_differences():
    s = _mesh_scene(6, n=30)
    w, h = 40, 32
    cam = _cam((0.4, -2.6, 0.8), (0.0, 0.0, 0.0), math.radians(60), w, h)
    N, M = len(s["vertices"]) // 3, len(s["indices"]) // 4
    vn = np.random.default_rng(2).normal(0.0, 1.0, N * 3)
    ivl = orc.interval_generate(s, cam, w, h, vertex_normals=vn, dtype=np.float64)
    assert np.count_nonzero(ivl[:, 64] == 3) > 0 and np.count_nonzero(ivl[:, 64] == 4) > 0
    rng = np.random.default_rng(3)
    d_ivl = np.zeros_like(ivl)
    d_ivl[:, 10:64] = rng.normal(0.0, 1.0, (M, 54))
    d_ivl[ivl[:, 64] == 0] = 0.0
    g = orc.interval_chain(s, cam, w, h, ivl, d_ivl, dtype=np.float64)

    def S(sc, normals):
        a = orc.interval_attrs(sc, cam, w, h, ivl, vertex_normals=normals, dtype=np.float64)
        return float(np.sum(a[:, 10:64] * d_ivl[:, 10:64]))

    eps = 1e-6
    for key, out in (("vertices", "d_vertices"), ("color_grads", "d_color_grads"), ("base_colors", "d_base_colors")):
        fd = np.zeros_like(s[key])
        for i in range(len(fd)):
            sp = dict(s)
            sp[key] = s[key].copy()
            sp[key][i] += eps
            lp = S(sp, vn)
            sp[key][i] -= 2 * eps
            fd[i] = (lp - S(sp, vn)) / (2 * eps)
        assert np.linalg.norm(g[out] - fd) / np.linalg.norm(fd) < 1e-6, key
    fdn = np.zeros(N * 3)
    for i in range(N * 3):
        p = vn.copy()
        p[i] += eps
        lp = S(s, p)
        p[i] -= 2 * eps
        fdn[i] = (lp - S(s, p)) / (2 * eps)
    assert np.linalg.norm(g["d_vertex_normals"] - fdn) / np.linalg.norm(fdn) < 1e-6
    np.testing.assert_array_equal(g["d_densities"], d_ivl[:, 63])


def test_colour_offsets_in_the_affine_limit():
    """g != 0: the reference interpolates the colour offsets with screen-space
    barycentrics; they equal the perspective-correct (linear-in-space) colour
    as the projection becomes affine.  A 0.1-sized tet seen from 100 units
    (depth spread / depth ~ 1e-3) in float64: the image equals the float64
    quadrature of the linear colour field to ~|g| size 1e-3, while g moves
    the image by ~5e-2."""
    from rmeshvk_amd.scenes import regular_tet_scene

    s = regular_tet_scene(10.0)
    s["vertices"] = (s["vertices"] * np.float32(0.1)).astype(np.float32)
    s["base_colors"] = np.array([0.6, 0.5, 0.7], np.float32)
    s["color_grads"] = np.array([0.4, -0.3, 0.5], np.float32)
    from rmeshvk_amd.scenes import circumspheres

    s["circumdata"] = circumspheres(s["vertices"], s["indices"])
    w = h = 40
    fov = 2.0 * math.atan(0.1 * w / (2.0 * 30.0 * 100.0))  # the tet spans ~30 px
    eye = np.array([35.0, -90.0, 30.0])
    eye = eye / np.linalg.norm(eye) * 100.0
    cam = _cam(tuple(eye), (0.0, 0.0, 0.0), fov, w, h)
    assert npref.endpoint_colours_positive(s, cam, w, h)
    ref = orc.render(s, cam, w, h)
    f64 = {k: (np.asarray(v, np.float64) if k in ("vertices", "base_colors", "densities", "color_grads") else v)
           for k, v in s.items()}
    iv = orc.interval_forward(f64, cam, w, h, ref["sorted_vals"], ref["ranges"], dtype=np.float64)
    q = npref.render(s, cam, w, h)
    assert q[..., 3].max() > 0.3
    err = np.abs(iv["image"] - q).max()
    assert err < 5e-4, err
    g0 = dict(f64, color_grads=np.zeros(3))
    iv0 = orc.interval_forward(g0, cam, w, h, ref["sorted_vals"], ref["ranges"], dtype=np.float64)
    assert np.abs(iv0["image"] - iv["image"]).max() > 2e-2
