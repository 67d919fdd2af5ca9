"""CPU: the oracle's interval-tiled forward (forward_interval_tiled,
interval_generate.wgsl + interval_tiled_rasterize.wgsl) against independent
float64 references.

The interval path interpolates NDC z over screen-space triangles and
linearizes it with the reference's near/far (0.01 / 1000); with a camera
built with the same near/far and planar faces, the linearized chord is the
ray chord, so with a constant colour per tet (g = 0) the image equals the float64
quadrature renderer (tests/npref.py) and depth x ray scale equals the
chord-weighted depth.  With g != 0 the reference interpolates the colour
offsets with screen-space (not perspective-correct) barycentrics; that rule
has no independent reference here (parity unpinned beyond GPU vs oracle).
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
