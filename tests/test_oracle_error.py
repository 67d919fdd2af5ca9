"""CPU: the oracle's per-tet error statistics (forward_error,
error_tiled_compute.wgsl:88-373) against closed forms and an independent
float64 brute force over pixel rays (tests/npref.py chords, no tiling, no
spans, no sort).

Pinned: the weight column is the sum of the pixels' optical depths (one tet:
exp(log T) = 1), which must agree with the oracle forward's alpha = 1 - e^-od;
the error-weighted columns are closed forms of the weight for constant error
images; entry/exit/centroid moments against float64 chord endpoints; the
pixel count and the max-weight column by brute force; a dense front tet
shields a back tet exactly as the -5.5 log-transmittance rule says.
"""

from __future__ import annotations

import math

import numpy as np

import npref
from oracle import oracle as orc
from rmeshvk_amd.camera import make_camera
from rmeshvk_amd.scenes import regular_tet_scene, two_tet_scene

REL = 2e-4  # float32 oracle vs float64 brute force


def _binned(scene, cam, w, h, ts=12):
    ref = orc.render(scene, cam, w, h, ts=ts)
    return ref


def test_single_tet_closed_forms():
    s = regular_tet_scene(2.5)
    w, h = 64, 56
    cam = make_camera((0.3, -2.2, 0.6), (0.0, 0.05, 0.0), math.radians(75), w, h)
    ref = _binned(s, cam, w, h)
    pe = np.ones((h, w), np.float32)
    se = np.full((h, w), 2.0, np.float32)
    err, cnt = orc.error_stats(s, cam, w, h, ref["sorted_vals"], ref["ranges"], pe, se)
    e = err[0]
    # brute force over pixel rays (float64 chords)
    o, d = npref.pixel_rays(cam, w, h)
    tv_ = np.asarray(s["vertices"], np.float64).reshape(1, 4, 3)
    t0, t1, hit = npref.tet_chords(tv_, o, d)
    hit, t0, t1 = hit[:, 0], t0[:, 0], t1[:, 0]
    od = np.clip(2.5 * (t1 - t0), 1e-8, 88.0)
    od = np.where(hit, od, 0.0)
    assert hit.sum() > 200
    np.testing.assert_allclose(e[0], od.sum(), rtol=REL)
    # the oracle forward's alpha carries the same optical depths
    alpha = ref["rgba"][..., 3].ravel().astype(np.float64)
    np.testing.assert_allclose(e[0], (-np.log1p(-alpha[alpha > 0])).sum(), rtol=REL)
    # closed forms of the weighted columns
    assert e[12] == e[0]
    assert e[1] == e[0] and e[2] == e[0]
    np.testing.assert_allclose(e[5], 2.0 * e[0], rtol=1e-6)
    assert e[3] == 0.0 and e[4] == 0.0
    cen = tv_[0].mean(axis=0)
    np.testing.assert_allclose(e[13:16] / e[12], cen, rtol=1e-5, atol=1e-6)
    # entry / exit moments
    ent = o[None, :] + d * t0[:, None]
    ext = o[None, :] + d * t1[:, None]
    np.testing.assert_allclose(e[6:9], (od[:, None] * np.where(hit[:, None], ent, 0)).sum(0), rtol=REL, atol=1e-4)
    np.testing.assert_allclose(e[9:12], (od[:, None] * np.where(hit[:, None], ext, 0)).sum(0), rtol=REL, atol=1e-4)
    # pixel count and max weight
    assert abs(int(cnt[0, 0]) - int(hit.sum())) <= 2
    assert abs(int(cnt[0, 1]) - int(65535.0 * od.max())) <= 8


def test_pixel_error_weighting():
    s = regular_tet_scene(1.5)
    w, h = 48, 48
    cam = make_camera((0.0, -2.0, 0.4), (0.0, 0.0, 0.0), math.radians(70), w, h)
    ref = _binned(s, cam, w, h)
    rng = np.random.default_rng(3)
    pe = rng.uniform(0.0, 2.0, (h, w)).astype(np.float32)
    se = rng.uniform(-1.0, 1.0, (h, w)).astype(np.float32)
    err, _ = orc.error_stats(s, cam, w, h, ref["sorted_vals"], ref["ranges"], pe, se)
    o, d = npref.pixel_rays(cam, w, h)
    t0, t1, hit = npref.tet_chords(np.asarray(s["vertices"], np.float64).reshape(1, 4, 3), o, d)
    od = np.where(hit[:, 0], np.clip(1.5 * (t1[:, 0] - t0[:, 0]), 1e-8, 88.0), 0.0)
    p, q = pe.ravel().astype(np.float64), se.ravel().astype(np.float64)
    np.testing.assert_allclose(err[0, 1], (od * p).sum(), rtol=REL)
    np.testing.assert_allclose(err[0, 2], (od * p * p).sum(), rtol=REL)
    np.testing.assert_allclose(err[0, 5], (od * q).sum(), rtol=REL, atol=1e-5)


def test_front_tet_shields_back_tet():
    """Depth order, exp(log T) weighting and the -5.5 cut-off over two tets."""
    s = two_tet_scene(np.random.default_rng(1))
    s["densities"] = np.array([9.0, 0.7], np.float32)  # a dense tet shields whatever is behind it
    w, h = 64, 64
    cam = make_camera((0.5, -2.0, 1.5), (0.5, 0.3, 0.0), math.radians(60), w, h)
    ref = _binned(s, cam, w, h)
    pe = np.ones((h, w), np.float32)
    err, cnt = orc.error_stats(s, cam, w, h, ref["sorted_vals"], ref["ranges"], pe, pe)
    o, d = npref.pixel_rays(cam, w, h)
    tv = np.asarray(s["vertices"], np.float64).reshape(-1, 3)[np.asarray(s["indices"]).reshape(-1, 4)]
    t0, t1, hit = npref.tet_chords(tv, o, d)
    order = npref.depth_order(s, cam[0])
    log_t = np.zeros(len(d))
    wsum = np.zeros(2)
    count = np.zeros(2, int)
    for t in order:  # front to back
        od = np.clip(s["densities"][t] * (t1[:, t] - t0[:, t]), 1e-8, 88.0)
        live = hit[:, t] & (log_t >= -5.5) & (np.exp(log_t) >= 1e-6)
        wsum[t] += (np.exp(log_t) * od)[live].sum()
        count[t] += live.sum()
        log_t = np.where(live, log_t - od, log_t)
    assert count.min() > 20  # both tets receive statistics somewhere
    np.testing.assert_allclose(err[:, 0], wsum, rtol=5e-4)
    assert np.all(np.abs(cnt[:, 0] - count) <= 2)
