"""CPU: the oracle's forward path (projection, pairs, sort, ranges, compositing)
against an independent float64 quadrature renderer (tests/npref.py), closed
forms of the regular tet, and brute-force tile coverage.

Scenes here keep every chord's colour positive, since the reference clamps
endpoint colours at 0 (a kink the quadrature does not model); the clamp
itself is exercised by the GPU-vs-oracle parity tests.
"""

from __future__ import annotations

import itertools
import math

import numpy as np
import pytest

import npref
from oracle import oracle as orc
from rmeshvk_amd.camera import fibonacci_cameras, make_camera
from rmeshvk_amd.scenes import make_scene, regular_tet_scene, two_tet_scene, uniform_points

IMG_TOL = 5e-5  # float32 oracle vs float64 quadrature


def _positive_colour_scene(n_points, seed):
    rng = np.random.default_rng(seed)
    s = make_scene(uniform_points(n_points, rng), rng)
    M = len(s["densities"])
    s["base_colors"] = rng.uniform(0.3, 1.0, 3 * M).astype(np.float32)
    s["color_grads"] = rng.uniform(-0.08, 0.08, 3 * M).astype(np.float32)
    return s


def test_regular_tet_closed_form():
    """g = 0: RGB = c * alpha with alpha = 1 - exp(-sigma * chord)."""
    s = regular_tet_scene(3.0)
    s["base_colors"] = np.array([0.9, 0.5, 0.2], np.float32)
    cam = make_camera((0.0, -2.0, 0.5), (0.0, 0.0, 0.0), math.pi / 2, 64, 64)
    o, d = npref.pixel_rays(cam, 64, 64)
    t0, t1, hit = npref.tet_chords(np.asarray(s["vertices"], np.float64).reshape(1, 4, 3), o, d)
    alpha = np.where(hit[:, 0], 1 - np.exp(-3.0 * (t1[:, 0] - t0[:, 0])), 0.0).reshape(64, 64)
    img = orc.render(s, cam, 64, 64)["rgba"]
    assert alpha.max() > 0.5
    np.testing.assert_allclose(img[..., 3], alpha, atol=IMG_TOL)
    for c in range(3):
        np.testing.assert_allclose(img[..., c], s["base_colors"][c] * alpha, atol=IMG_TOL)


def test_regular_tet_linear_colour_quadrature():
    s = regular_tet_scene(2.0)
    s["base_colors"] = np.array([1.0, 0.8, 0.9], np.float32)
    s["color_grads"] = np.array([0.3, -0.2, 0.25], np.float32)
    cam = make_camera((0.7, -2.0, 0.9), (0.0, 0.1, 0.0), math.radians(70), 48, 40)
    assert npref.endpoint_colours_positive(s, cam, 48, 40)
    ref = npref.render(s, cam, 48, 40)
    img = orc.render(s, cam, 48, 40)["rgba"]
    assert np.abs(img - ref).max() < IMG_TOL


def test_two_tet_quadrature():
    s = two_tet_scene(np.random.default_rng(1))
    cam = make_camera((0.5, -2.0, 1.5), (0.5, 0.3, 0.0), math.radians(60), 64, 64)
    assert npref.endpoint_colours_positive(s, cam, 64, 64)
    ref = npref.render(s, cam, 64, 64)
    img = orc.render(s, cam, 64, 64)["rgba"]
    assert (ref[..., 3] > 0.01).sum() > 100
    assert np.abs(img - ref).max() < IMG_TOL


@pytest.mark.parametrize("view", [0, 1, 2])
def test_delaunay_quadrature(view):
    """~280 tets: depth order, tile binning, spans and the closed-form integral."""
    s = _positive_colour_scene(60, 7)
    cam = fibonacci_cameras(3, 3.0, 50.0, 64, 64)[view]
    assert npref.endpoint_colours_positive(s, cam, 64, 64)
    ref = npref.render(s, cam, 64, 64)
    img = orc.render(s, cam, 64, 64)["rgba"]
    assert (ref[..., 3] > 0.1).sum() > 500
    assert np.abs(img - ref).max() < IMG_TOL


def test_min_t_reparametrises_without_clipping():
    """min_t moves the ray origin to cam + min_t * dir (rasterize_compute.wgsl:351-352)
    but the slab test still takes the whole line, so the image is the min_t = 0
    image (up to rounding)."""
    s = _positive_colour_scene(60, 7)
    cam = fibonacci_cameras(3, 3.0, 50.0, 64, 64)[1]
    ref = npref.render(s, cam, 64, 64)
    img = orc.render(s, cam, 64, 64, min_t=2.5)["rgba"]
    assert np.abs(img - ref).max() < IMG_TOL


@pytest.mark.parametrize("ts", [8, 12, 16])
def test_tile_size_invariance(ts):
    s = _positive_colour_scene(60, 11)
    cam = fibonacci_cameras(2, 3.0, 50.0, 70, 52)[0]
    a = orc.render(s, cam, 70, 52, ts)["rgba"]
    b = orc.render(s, cam, 70, 52, 12)["rgba"]
    assert np.abs(a - b).max() < 1e-5


def test_aux_normals_depth_single_tet():
    """Single tet: normals = alpha * outward unit normal of the entry face;
    depth = the transmittance-weighted mean ray parameter, by quadrature."""
    s = regular_tet_scene(2.0)
    cam = make_camera((0.6, -2.0, 0.8), (0.0, 0.0, 0.0), math.radians(70), 40, 40)
    out = orc.render(s, cam, 40, 40, want_aux=True)
    aux, img = out["aux"], out["rgba"]
    o, d = npref.pixel_rays(cam, 40, 40)
    tv = np.asarray(s["vertices"], np.float64).reshape(4, 3)
    t0, t1, hit = npref.tet_chords(tv[None], o, d)
    t0, t1, hit = t0[:, 0], t1[:, 0], hit[:, 0]
    # entry face: the face plane the ray crosses at t0, found by distance to each plane
    centroid = tv.mean(0)
    p_in = o[None, :] + d * t0[:, None]
    normals = np.zeros((len(d), 3))
    best = np.full(len(d), np.inf)
    for opp in range(4):
        tri = [i for i in range(4) if i != opp]
        n = np.cross(tv[tri[1]] - tv[tri[0]], tv[tri[2]] - tv[tri[0]])
        n /= np.linalg.norm(n)
        if np.dot(n, centroid - tv[tri[0]]) > 0:
            n = -n  # outward
        dist = np.abs((p_in - tv[tri[0]]) @ n)
        upd = dist < best
        best = np.where(upd, dist, best)
        normals = np.where(upd[:, None], n[None, :], normals)
    alpha = np.where(hit, 1 - np.exp(-2.0 * (t1 - t0)), 0.0)
    L = np.where(hit, t1 - t0, 0.0)
    x, wq = np.polynomial.legendre.leggauss(24)
    depth = np.zeros(len(d))
    for xi, wi in zip(x, wq):
        ss = 0.5 * L * (xi + 1)
        depth += 0.5 * L * wi * 2.0 * np.exp(-2.0 * ss) * (t0 + ss)
    depth = np.where(hit, depth, 0.0)
    sel = (alpha > 1e-3).reshape(40, 40)
    assert sel.sum() > 100
    np.testing.assert_allclose(aux[..., :3][sel], (alpha[:, None] * normals).reshape(40, 40, 3)[sel], atol=1e-4)
    np.testing.assert_allclose(aux[..., 3][sel], depth.reshape(40, 40)[sel], atol=1e-4)
    np.testing.assert_allclose(img[..., 3], alpha.reshape(40, 40), atol=IMG_TOL)


# --------------------------------------------------------------------------
# binning
# --------------------------------------------------------------------------
def test_binning_covers_every_hit_pixel():
    """Every pixel whose ray has a non-grazing chord through a tet lies in a tile
    the tet was binned to (the scanline is conservative)."""
    s = _positive_colour_scene(80, 3)
    w, h, ts = 72, 60, 12
    for cam in fibonacci_cameras(3, 3.0, 50.0, w, h):
        tt, _ = orc.project(s, cam, w, h, ts)
        keys, vals, P = orc.pairs(s, cam, w, h, tt, ts)
        assert P == int(tt.sum())
        tx = (w + ts - 1) // ts
        binned = set(zip((keys >> np.uint64(32)).astype(np.int64).tolist(), vals.astype(np.int64).tolist()))
        o, d = npref.pixel_rays(cam, w, h)
        V = np.asarray(s["vertices"], np.float64).reshape(-1, 3)
        tv = V[np.asarray(s["indices"], np.int64).reshape(-1, 4)]
        t0, t1, hit = npref.tet_chords(tv, o, d)
        solid = hit & (t1 - t0 > 1e-3) & (t0 > 0)
        pix, tet = np.nonzero(solid)
        tiles = (pix // w // ts) * tx + (pix % w) // ts
        missing = [(a, b) for a, b in zip(tiles.tolist(), tet.tolist()) if (a, b) not in binned]
        assert not missing, missing[:10]


def _hull_meets_rect(pts, x0, y0, x1, y1, eps):
    """Separating-axis test between the convex hull of 2-D points and the
    rectangle [x0, x1] x [y0, y1] grown by eps (float64)."""
    if pts[:, 0].max() < x0 - eps or pts[:, 0].min() > x1 + eps:
        return False
    if pts[:, 1].max() < y0 - eps or pts[:, 1].min() > y1 + eps:
        return False
    corners = np.array([[x0 - eps, y0 - eps], [x1 + eps, y0 - eps], [x1 + eps, y1 + eps], [x0 - eps, y1 + eps]])
    for i, j in itertools.combinations(range(len(pts)), 2):
        e = pts[j] - pts[i]
        n = np.array([-e[1], e[0]])
        if not n.any():
            continue
        side = (pts - pts[i]) @ n
        if side.max() <= 0.0 or side.min() >= 0.0:  # a hull edge: every point on one side
            s_out = 1.0 if side.max() <= 0.0 else -1.0
            if np.all(s_out * ((corners - pts[i]) @ n) > 0.0):
                return False
    return True


def test_binning_is_tight():
    """Every binned (tile, tet) pair's tile meets the tet's projected hull: the
    hull of its four vertices' float64 projections, tested against the tile's
    rectangle by separating axes (independent of the scanline), within 1e-3
    px for float32 rounding.  Over-binning (a tile past the hull's extent in
    its row band) fails it.  The one exception is the reference's own: a row
    band whose hull extent lies left of the image (or right of it) is clamped
    into the edge column (tile_gen_scan_compute.wgsl:184-185,
    clamp(max(floor(x / T), 0), tile_min_x, tile_max_x)), so the edge
    columns' rectangles reach past the image on that side."""
    for seed, (w, h, ts) in ((3, (72, 60, 12)), (4, (64, 64, 8)), (5, (50, 70, 16))):
        s = _positive_colour_scene(80, seed)
        V = np.asarray(s["vertices"], np.float64).reshape(-1, 3)
        tets = np.asarray(s["indices"], np.int64).reshape(-1, 4)
        tx = (w + ts - 1) // ts
        for cam in fibonacci_cameras(3, 3.0, 50.0, w, h):
            tt, _ = orc.project(s, cam, w, h, ts)
            keys, vals, P = orc.pairs(s, cam, w, h, tt, ts)
            vp = np.asarray(cam[1], np.float64).reshape(4, 4).T  # column-major
            clip = np.c_[V, np.ones(len(V))] @ vp.T
            ndc = clip[:, :3] / (clip[:, 3:4] + 1e-6)
            proj = np.c_[(ndc[:, 0] + 1.0) * 0.5 * w, (1.0 - ndc[:, 1]) * 0.5 * h]
            tiles = (keys >> np.uint64(32)).astype(np.int64)
            real = keys != np.uint64(0xFFFFFFFFFFFFFFFF)
            assert real.sum() > 50
            for tile, t in zip(tiles[real].tolist(), vals[real].astype(np.int64).tolist()):
                c, y0 = tile % tx, (tile // tx) * ts
                x0 = -np.inf if c == 0 else c * ts
                x1 = np.inf if c == tx - 1 else (c + 1) * ts
                assert _hull_meets_rect(proj[tets[t]], x0, y0, x1, y0 + ts, 1e-3), (seed, tile, t)


def test_hull_rect_test_detects_misses():
    """The separating-axis check above rejects a rectangle beside a triangle's
    hypotenuse (inside its bounding box) and accepts one across it."""
    tri = np.array([[0.0, 0.0], [10.0, 0.0], [0.0, 10.0], [0.0, 0.0]])
    assert not _hull_meets_rect(tri, 7.0, 7.0, 9.0, 9.0, 1e-3)
    assert _hull_meets_rect(tri, 4.0, 4.0, 6.0, 6.0, 1e-3)
    assert _hull_meets_rect(tri, 5.0, 5.0 - 1e-4, 6.0, 6.0, 1e-3)  # touches (5, 5) on the hypotenuse


def test_keys_decode_to_tile_and_depth():
    s = _positive_colour_scene(80, 5)
    w = h = 64
    cam = fibonacci_cameras(2, 3.0, 50.0, w, h)[1]
    tt, _ = orc.project(s, cam, w, h, 12)
    keys, vals, P = orc.pairs(s, cam, w, h, tt, 12)
    cd = np.asarray(s["circumdata"], np.float32).reshape(-1, 4)[vals]
    c = np.asarray(cam[0], np.float32)
    dx, dy, dz = cd[:, 0] - c[0], cd[:, 1] - c[1], cd[:, 2] - c[2]
    depth32 = ((dx * dx + dy * dy) + dz * dz) - cd[:, 3]
    low = (keys & np.uint64(0xFFFFFFFF)).astype(np.uint32)
    np.testing.assert_array_equal(low, ~depth32.astype(np.float32).view(np.uint32))
    d2 = ((cd[:, :3].astype(np.float64) - c) ** 2).sum(1)
    depth64 = d2 - cd[:, 3]
    err = np.abs((~low).view(np.float32).astype(np.float64) - depth64)
    assert np.all(err <= 1e-6 * (d2 + cd[:, 3]))  # float32 rounding of the operands
    assert np.all((keys >> np.uint64(32)) < np.uint64(((w + 11) // 12) ** 2))


def test_f32_f64_oracle_agree():
    s = _positive_colour_scene(60, 2)
    cam = fibonacci_cameras(2, 3.0, 50.0, 48, 48)[0]
    a = orc.render(s, cam, 48, 48)["rgba"]
    b = orc.render(s, cam, 48, 48, dtype=np.float64)["rgba"]
    assert np.abs(a - b).max() < IMG_TOL


def test_aux_entropy_channels_two_tets():
    """aux[4..8] (rasterize_compute.wgsl:447-456): per covered pair
    w = T alpha and m = (t_min + t_max) / 2, summed as (w, w log w, w m,
    (w m) log(w m)); checked against float64 chords of both tets in depth
    order (the front tet's alpha gives the back tet's T)."""
    s = two_tet_scene(np.random.default_rng(3))
    w = h = 48
    cam = make_camera((0.9, -1.6, 1.9), (0.5, 0.35, 0.0), math.radians(45), w, h)
    out = orc.render(s, cam, w, h, want_aux=True)
    aux = out["aux"].reshape(-1, 8)
    o, d = npref.pixel_rays(cam, w, h)
    tv = np.asarray(s["vertices"], np.float64).reshape(-1, 3)[np.asarray(s["indices"]).reshape(-1, 4)]
    t0, t1, hit = npref.tet_chords(tv, o, d)
    t0, t1 = np.where(hit, t0, 0.0), np.where(hit, t1, 0.0)
    order = npref.depth_order(s, cam[0])
    dens = np.asarray(s["densities"], np.float64)
    e = np.zeros((len(d), 4))
    T = np.ones(len(d))
    for k in order:
        L = np.where(hit[:, k], t1[:, k] - t0[:, k], 0.0)
        a = 1.0 - np.exp(-dens[k] * L)
        wt = T * a
        m = 0.5 * (t0[:, k] + t1[:, k])
        wc = wt * m
        e[:, 0] += np.where(hit[:, k], wt, 0.0)
        e[:, 1] += np.where(hit[:, k] & (wt > 1e-20), wt * np.log(np.maximum(wt, 1e-300)), 0.0)
        e[:, 2] += np.where(hit[:, k], wc, 0.0)
        e[:, 3] += np.where(hit[:, k] & (wc > 1e-20), wc * np.log(np.maximum(wc, 1e-300)), 0.0)
        T = T * np.where(hit[:, k], 1.0 - a, 1.0)
    # both tets behind one another somewhere, and pixels away from grazing chords
    assert np.count_nonzero(hit.all(axis=1)) > 50
    sel = np.abs(e[:, 0]) > 1e-3
    np.testing.assert_allclose(aux[sel, 4:8], e[sel], atol=2e-4, rtol=1e-4)
