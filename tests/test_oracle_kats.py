"""CPU: the oracle against the reference's own known-answer tests and closed forms.

Fixture: tests/golden/reference_kats.json (literals of the reference tests,
written by tests/golden/make_reference_kats.py).  Each check is chosen so that
a plausible slip in the oracle (wrong sign, dropped term, swapped index, an
order that is not stable) fails it.
"""

from __future__ import annotations

import json
import math
import os

import numpy as np
import pytest

from oracle import oracle as orc
from rmeshvk_amd.camera import look_at, make_camera, perspective_matrix
from rmeshvk_amd.scenes import circumspheres, regular_tet_scene

HERE = os.path.dirname(os.path.abspath(__file__))
with open(os.path.join(HERE, "golden", "reference_kats.json")) as _f:
    KATS = json.load(_f)


# --------------------------------------------------------------------------
# loss (kernel_tests.rs:176-305)
# --------------------------------------------------------------------------
def test_loss_l1_kat():
    k = KATS["loss_l1"]
    w, h = k["width"], k["height"]
    rgba = np.tile(np.array(k["rendered_rgba"], np.float32), (h, w, 1))
    gt = np.tile(np.array(k["ground_truth_rgb"], np.float32), (h, w, 1))
    dl, loss = orc.loss_grad(rgba, gt, k["loss_type"])
    exp = np.array(k["expected_dl"], np.float32)
    assert np.abs(dl - exp).max() < k["tol"]
    # loss = mean over pixels of the summed |diff| = 0.3 + 0.1 + 0.1
    assert abs(loss - 0.5) < 1e-6


def test_loss_l2_closed_form():
    rng = np.random.default_rng(0)
    h, w = 20, 30
    rgba = rng.random((h, w, 4)).astype(np.float32)
    gt = rng.random((h, w, 3)).astype(np.float32)
    dl, loss = orc.loss_grad(rgba, gt, 1)
    d = rgba[..., :3].astype(np.float64) - gt
    np.testing.assert_allclose(dl[..., :3], 2.0 * d / (w * h), rtol=1e-6, atol=1e-12)
    assert np.all(dl[..., 3] == 0.0)
    assert abs(loss - (d * d).sum() / (w * h)) < 1e-6 * loss  # differences rounded to float32


# --------------------------------------------------------------------------
# Adam (kernel_tests.rs:313-420; adam_compute.wgsl)
# --------------------------------------------------------------------------
def test_adam_kat():
    k = KATS["adam"]
    p0 = np.array(k["params"], np.float32)
    g = np.array(k["grads"], np.float32)
    z = np.zeros_like(p0)
    p, m, v = orc.adam(p0, g, z, z, k["step"], k["lr"], k["beta1"], k["beta2"], k["eps"])
    assert np.all(np.abs(p - p0) > k["min_move"])
    assert np.all(np.sign(p - p0) == -np.sign(g))
    assert np.abs(m - np.array(k["expected_m"], np.float32)).max() < k["m_tol"]
    one_minus_b2 = float(np.float32(1.0) - np.float32(k["beta2"]))  # the constant as float32 forms it
    np.testing.assert_allclose(v, one_minus_b2 * g.astype(np.float64) ** 2, rtol=1e-6)
    # bias-corrected first step moves by lr * g / (|g| + eps) = lr * sign(g)
    np.testing.assert_allclose(p - p0, -k["lr"] * np.sign(g), rtol=1e-4)


def test_adam_constant_gradient_steps():
    """With a constant gradient, m_k = (1 - b1^k) g and v_k = (1 - b2^k) g^2, so
    every bias-corrected step moves by lr * sign(g)."""
    g = np.array([0.5, -2.0, 1e-3, 0.0], np.float32)
    p = np.array([1.0, 1.0, 1.0, 1.0], np.float32)
    m = np.zeros(4, np.float32)
    v = np.zeros(4, np.float32)
    lr = 1e-2
    for step in range(1, 6):
        p, m, v = orc.adam(p, g, m, v, step, lr)
        np.testing.assert_allclose(m, (1 - 0.9 ** step) * g, rtol=2e-5, atol=1e-12)
        np.testing.assert_allclose(v, (1 - 0.999 ** step) * g.astype(np.float64) ** 2, rtol=2e-4, atol=1e-18)
    np.testing.assert_allclose(p[:3], 1.0 - 5 * lr * np.sign(g[:3]), rtol=1e-4)
    assert p[3] == 1.0  # zero gradient: no move


def test_adam_nan_gradient_is_zero():
    p, m, v = orc.adam(np.ones(2, np.float32), np.array([np.nan, 1.0], np.float32), np.zeros(2, np.float32),
                       np.zeros(2, np.float32), 1, 0.1)
    assert p[0] == 1.0 and m[0] == 0.0 and v[0] == 0.0
    assert p[1] < 1.0


# --------------------------------------------------------------------------
# sort (kernel_tests.rs:96-167; drs.rs:23-26, stable LSD over sorting_bits)
# --------------------------------------------------------------------------
def test_sort_kat_64_keys():
    k = KATS["radix_sort"]
    rng = np.random.default_rng(42424242)
    keys32 = rng.integers(0, 2 ** 32, k["n"], dtype=np.uint64) % k["key_modulus"]
    vals = np.arange(k["n"], dtype=np.uint32)
    sk, sv = orc.sort(keys32, vals, k["sorting_bits"])
    assert np.all(np.diff(sk.astype(np.int64)) >= 0)
    np.testing.assert_array_equal(sk, keys32[sv])


@pytest.mark.parametrize("n,bits", [(1, 32), (7, 40), (1000, 48), (65537, 56), (20000, 64)])
def test_sort_is_stable_argsort_of_masked_key(n, bits):
    rng = np.random.default_rng(n + bits)
    # few distinct masked keys -> many ties; random bits above sorting_bits must be ignored
    low = rng.integers(0, 50, n, dtype=np.uint64) << np.uint64(bits - 8)
    high = rng.integers(0, 2 ** 16, n, dtype=np.uint64) << np.uint64(48) if bits < 64 else np.zeros(n, np.uint64)
    mask = np.uint64((1 << bits) - 1) if bits < 64 else np.uint64(0xFFFFFFFFFFFFFFFF)
    keys = (low | (high & ~mask)) if bits < 64 else low
    vals = rng.permutation(n).astype(np.uint32)
    sk, sv = orc.sort(keys, vals, bits)
    order = np.argsort(keys & mask, kind="stable")
    np.testing.assert_array_equal(sv, vals[order])
    np.testing.assert_array_equal(sk, keys[order])


def test_sorting_bits_table():
    for num_tiles, bits in KATS["sorting_bits"]["table"]:
        assert orc.sorting_bits(num_tiles) == bits, num_tiles


def test_sentinel_fill_kat():
    """Padding pairs beyond the generated count hold key 0xFFFFFFFF:0xFFFFFFFF and value 0
    (kernel_tests.rs:428-545)."""
    k = KATS["tile_fill"]
    s = regular_tet_scene()
    cam = make_camera((0.0, -2.0, 0.5), (0.0, 0.0, 0.0), math.pi / 2, 64, 64)
    tt, _ = orc.project(s, cam, 64, 64, 12)
    P = int(tt.sum())
    keys, vals, n = orc.pairs(s, cam, 64, 64, tt, 12, cap=P + 5)
    assert n == P
    sentinel = np.uint64((k["sentinel_key"][0] << 32) | k["sentinel_key"][1])
    assert np.all(keys[P:] == sentinel)
    assert np.all(vals[P:] == k["sentinel_value"])


# --------------------------------------------------------------------------
# ranges (tile_ranges_compute.wgsl): closed form from sorted tile ids
# --------------------------------------------------------------------------
def _ranges_closed_form(tiles, num_tiles, n_real):
    r = np.zeros(2 * num_tiles, np.uint32)
    t = tiles[:n_real]
    for tile in np.unique(t):
        lo = np.searchsorted(t, tile, "left")
        hi = np.searchsorted(t, tile, "right")
        r[2 * tile] = lo
        r[2 * tile + 1] = hi
    return r


def test_ranges_closed_form():
    rng = np.random.default_rng(3)
    T = 97
    tiles = np.sort(rng.choice(T, 500, replace=True)).astype(np.uint64)
    keys = (tiles << np.uint64(32)) | rng.integers(0, 2 ** 32, 500, dtype=np.uint64)
    np.testing.assert_array_equal(orc.ranges(keys, T), _ranges_closed_form(tiles, T, 500))


def test_ranges_trailing_sentinel_quirk():
    """When sentinel pairs lie inside the pair count (the projection counted more
    tiles than the scanline generated) they sort last and the last real tile's
    end is never written (it stays 0): its tile-change boundary is a sentinel,
    which tile_ranges_compute.wgsl:21-36 skips; every other tile matches the
    closed form."""
    tiles = np.array([0, 0, 2, 2, 2, 5], np.uint64)
    sentinel = np.uint64(0xFFFFFFFFFFFFFFFF)
    keys = np.concatenate([(tiles << np.uint64(32)) | np.uint64(7), np.full(3, sentinel)])
    got = orc.ranges(keys, 8)
    exp = _ranges_closed_form(tiles, 8, 6)
    exp[2 * 5 + 1] = 0
    np.testing.assert_array_equal(got, exp)


# --------------------------------------------------------------------------
# circumspheres (test_util.rs:133-163)
# --------------------------------------------------------------------------
def test_circumsphere_kat_and_equidistance():
    k = KATS["circumsphere"]
    got = orc.circumspheres(np.array(k["vertices"], np.float32), np.array(k["indices"], np.uint32))
    np.testing.assert_allclose(got, k["expected"], atol=1e-6)
    rng = np.random.default_rng(1)
    v = rng.normal(size=(40, 3)).astype(np.float32)
    idx = np.stack([rng.choice(40, 4, replace=False) for _ in range(30)]).astype(np.uint32)
    cs = orc.circumspheres(v.ravel(), idx.ravel()).reshape(-1, 4)
    np.testing.assert_allclose(cs, circumspheres(v.ravel(), idx.ravel()).reshape(-1, 4), rtol=1e-4, atol=1e-5)
    for t in range(30):
        d2 = ((v[idx[t]].astype(np.float64) - cs[t, :3]) ** 2).sum(1)
        np.testing.assert_allclose(d2, cs[t, 3], rtol=2e-3)


# --------------------------------------------------------------------------
# known tet: camera, projection and visibility (kernel_tests.rs:1547-1671, 1839-1914)
# --------------------------------------------------------------------------
def _known_cam():
    k = KATS["known_tet"]
    return k, make_camera(k["eye"], k["target"], math.radians(k["fov_y_deg"]), k["width"], k["height"],
                          k["near"], k["far"])


def test_known_tet_cpu_projection():
    k, _ = _known_cam()
    vp = perspective_matrix(math.radians(k["fov_y_deg"]), 1.0, k["near"], k["far"]) @ look_at(k["eye"], k["target"])
    np.testing.assert_allclose(vp @ np.linalg.inv(vp), np.eye(4), atol=1e-4)
    v = np.array(k["vertices"], np.float64).reshape(4, 3)
    clip = (vp @ np.c_[v, np.ones(4)].T).T
    assert np.all(clip[:, 3] > 0)
    ndc = clip[:, :3] / clip[:, 3:4]
    assert np.all(np.abs(ndc[:, :2]) <= 1) and np.all((ndc[:, 2] >= 0) & (ndc[:, 2] <= 1))
    px = (ndc[:, 0] + 1) * 0.5 * k["width"]
    py = (1 - ndc[:, 1]) * 0.5 * k["height"]
    assert (px.max() - px.min()) * (py.max() - py.min()) >= k["expected_pixel_area_min"]


def test_known_tet_visibility():
    k, cam = _known_cam()
    s = regular_tet_scene(k["density"])
    tt, vis = orc.project(s, cam, k["width"], k["height"], k["tile_size"])
    assert int(vis.sum()) == k["expected_visible"]
    assert int(tt[0]) >= k["expected_tiles_touched_min"]
    tt64, vis64 = orc.project(s, cam, k["width"], k["height"], k["tile_size"], dtype=np.float64)
    np.testing.assert_array_equal(tt64, tt)


def test_known_tet_center_ray_alpha():
    """Centre pixel: hit with t_enter > 0 and alpha = 1 - exp(-3 chord) > 0.01,
    chord from an independent slab test; the oracle's pixel alpha equals it."""
    k, cam = _known_cam()
    s = regular_tet_scene(k["density"])
    w, h = k["width"], k["height"]
    from npref import pixel_rays, tet_chords

    o, d = pixel_rays(cam, w, h)
    t0, t1, hit = tet_chords(np.array(k["vertices"], np.float64).reshape(1, 4, 3), o, d)
    c = (h // 2) * w + w // 2
    assert hit[c, 0] and t0[c, 0] > 0 and t1[c, 0] > t0[c, 0]
    alpha = 1 - math.exp(-k["density"] * (t1[c, 0] - t0[c, 0]))
    assert alpha > k["expected_center_alpha_min"]
    img = orc.render(s, cam, w, h, k["tile_size"])["rgba"]
    assert abs(img[h // 2, w // 2, 3] - alpha) < 1e-5
