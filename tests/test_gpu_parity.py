"""GPU parity of the HIP pipeline against the CPU oracle, through the C ABI.

Bars (stated per check): tile counts, pre-sort pairs, sorted pairs and
per-tile ranges bit-exact; images within 1e-4 max-abs (FP32); gradients
within 1e-3 relative L2.
"""

import math

import numpy as np
import pytest

from oracle import oracle as orc
from rmeshvk_amd.camera import colmap_camera, fibonacci_cameras, make_camera
from rmeshvk_amd.scenes import (cluster_points, make_scene, regular_tet_scene, single_tet_scene,
                                two_tet_scene, uniform_points)

pytestmark = pytest.mark.gpu

IMG_TOL = 1e-4
GRAD_TOL = 1e-3


def _renderer(scene, w, h, **kw):
    from rmeshvk_amd import RMeshRenderer

    return RMeshRenderer(scene["vertices"], scene["indices"], scene["base_colors"], scene["densities"],
                         scene["color_grads"], scene["circumdata"], w, h, device_id=0, **kw)


def _cases():
    rng = np.random.default_rng(1234)
    fov = math.radians(90.0)
    cases = {
        "regular_tet": (regular_tet_scene(), make_camera((0.0, -2.0, 0.5), (0, 0, 0), fov, 64, 64), 64, 64),
        "two_tet": (two_tet_scene(rng), make_camera((3.0, 0.4, 1.0), (0.5, 0.4, 0.0), fov, 64, 64), 64, 64),
        "single_tet": (single_tet_scene(42), colmap_camera(32, 32, 1.0), 32, 32),
    }
    s = make_scene(uniform_points(600, rng), rng)
    cases["uniform_600_odd"] = (s, make_camera((2.2, -1.1, 0.9), (0, 0, 0), math.radians(55.0), 100, 76), 100, 76)
    s = make_scene(cluster_points(1500, rng, 3, 0.2), rng)
    cases["cluster_inside"] = (s, make_camera((0.1, 0.05, 0.02), (1.0, 0.3, 0.2), math.radians(70.0), 96, 96), 96, 96)
    return cases


CASES = _cases()


@pytest.fixture(scope="module")
def config_a():
    from rmeshvk_amd.scenes import config_scene

    scene, cams = config_scene("A", seed=0)
    return scene, cams[0], 256, 256


def _check_binning(r, scene, cam, w, h, ts):
    ref = orc.render(scene, cam, w, h, ts=ts)
    # pre-sort pairs (K1 counts + K7 keys/values), bit-exact
    keys, vals, P = r.debug_pairs(*cam)
    assert P == ref["P"]
    np.testing.assert_array_equal(keys, ref["keys"])
    np.testing.assert_array_equal(vals, ref["vals"])
    img = r.forward_tiled(*cam)
    tt = r.debug_tiles_touched(0)
    np.testing.assert_array_equal(tt, ref["tiles_touched"])
    assert r.read_instance_count() == int(ref["visible"].sum())
    sk, sv, rg = r.debug_sorted()
    np.testing.assert_array_equal(sk, ref["sorted_keys"])
    np.testing.assert_array_equal(sv, ref["sorted_vals"])
    np.testing.assert_array_equal(rg, ref["ranges"])
    return ref, img


@pytest.mark.parametrize("name", list(CASES))
def test_binning_and_image(name):
    scene, cam, w, h = CASES[name]
    r = _renderer(scene, w, h)
    ref, img = _check_binning(r, scene, cam, w, h, 12)
    err = np.abs(img - ref["rgba"]).max()
    print(f"{name}: P={ref['P']} max|img-oracle|={err:.2e}")
    assert err < IMG_TOL


@pytest.mark.parametrize("ts", [8, 16, 32])
def test_tile_sizes(ts, config_a):
    scene, cam, w, h = config_a
    r = _renderer(scene, w, h, tile_size=ts)
    ref, img = _check_binning(r, scene, cam, w, h, ts)
    assert np.abs(img - ref["rgba"]).max() < IMG_TOL


@pytest.mark.parametrize("w", [128, 136])
def test_full_sort_and_depth_first_layouts(w):
    """128x128 at tile 8 is 256 key tiles in one view: the sentinel key masks
    to the last real tile, and the renderer keeps the full (tile, depth) pair
    sort; 136x128 (272 tiles) takes depth-first pair generation (items sorted
    by depth, pairs sorted on the tile bits only).  Both give the oracle's
    lists bit for bit."""
    rng = np.random.default_rng(99)
    s = make_scene(cluster_points(4000, rng, 3, 0.25), rng)
    h = 128
    cam = fibonacci_cameras(2, 3.0, 50.0, w, h)[1]
    r = _renderer(s, w, h, tile_size=8)
    ref, img = _check_binning(r, s, cam, w, h, 8)
    assert np.abs(img - ref["rgba"]).max() < IMG_TOL


def test_config_a_forward_backward(config_a):
    scene, cam, w, h = config_a
    r = _renderer(scene, w, h)
    ref, img = _check_binning(r, scene, cam, w, h, 12)
    assert np.abs(img - ref["rgba"]).max() < IMG_TOL
    rng = np.random.default_rng(3)
    dl = np.ascontiguousarray(rng.normal(0, 1e-3, (h, w, 4)).astype(np.float32))
    g = r.backward(dl)
    gr = orc.backward(scene, cam, w, h, ref["sorted_vals"], ref["ranges"], ref["rgba"], dl)
    for k in ("d_vertices", "d_densities", "d_color_grads", "d_base_colors"):
        a, b = g[k].astype(np.float64), gr[k].astype(np.float64)
        rel = np.linalg.norm(a - b) / np.linalg.norm(b)
        print(f"{k}: rel L2 {rel:.2e}")
        assert rel < GRAD_TOL, (k, rel)


@pytest.mark.parametrize("name", list(CASES))
def test_backward_micro(name):
    scene, cam, w, h = CASES[name]
    r = _renderer(scene, w, h)
    img = r.forward_tiled(*cam)
    ref = orc.render(scene, cam, w, h)
    gt = np.full((h, w, 3), 0.5, dtype=np.float32)
    dl, _ = orc.loss_grad(ref["rgba"], gt, 1)
    dl[..., 3] = np.float32(0.01)  # exercise the alpha-gradient init term
    dl = np.ascontiguousarray(dl, dtype=np.float32)
    g = r.backward(dl)
    gr = orc.backward(scene, cam, w, h, ref["sorted_vals"], ref["ranges"], ref["rgba"], dl)
    for k in ("d_vertices", "d_densities", "d_color_grads", "d_base_colors"):
        a, b = g[k].astype(np.float64), gr[k].astype(np.float64)
        nb = np.linalg.norm(b)
        rel = np.linalg.norm(a - b) / nb if nb > 0 else np.linalg.norm(a)
        assert rel < GRAD_TOL, (name, k, rel)


def test_sort_known_answer():
    """kernel_tests.rs:96-167: 64 keys, ascending order and key<->value correspondence."""
    from rmeshvk_amd import sort_pairs

    rng = np.random.default_rng(42424242)
    keys = (rng.integers(0, 2**32, 64, dtype=np.uint64) % 10000).astype(np.uint64)
    vals = np.arange(64, dtype=np.uint32)
    sk, sv = sort_pairs(keys, vals, 32)
    assert np.all(np.diff(sk.astype(np.int64)) >= 0)
    np.testing.assert_array_equal(sk, keys[sv])


@pytest.mark.parametrize("n,bits", [(1, 40), (4095, 32), (4097, 48), (70001, 48), (1_000_003, 56)])
def test_sort_matches_stable_argsort(n, bits):
    from rmeshvk_amd import sort_pairs

    rng = np.random.default_rng(n)
    tiles = rng.integers(0, 300, n, dtype=np.uint64)
    depth = rng.integers(0, 50, n, dtype=np.uint64)  # many ties: stability matters
    keys = (tiles << np.uint64(32)) | depth
    keys[rng.random(n) < 0.01] = np.uint64(0xFFFFFFFFFFFFFFFF)
    vals = np.arange(n, dtype=np.uint32)
    sk, sv = sort_pairs(keys, vals, bits)
    mask = np.uint64((1 << bits) - 1) if bits < 64 else np.uint64(0xFFFFFFFFFFFFFFFF)
    order = np.argsort(keys & mask, kind="stable")
    np.testing.assert_array_equal(sv, vals[order])
    np.testing.assert_array_equal(sk, keys[order])


def test_empty_view():
    """Camera looking away from the mesh: no pairs, transparent image, zero gradients."""
    scene, _, w, h = CASES["uniform_600_odd"]
    cam = make_camera((0.0, 5.0, 0.0), (0.0, 10.0, 0.0), math.radians(40.0), w, h)
    r = _renderer(scene, w, h)
    img = r.forward_tiled(*cam)
    assert np.all(img == 0.0)
    g = r.backward(np.ones((h, w, 4), dtype=np.float32))
    assert all(np.all(v == 0.0) for v in g.values())


def _variant_grid():
    from rmeshvk_amd import RMeshRenderer

    nf, nb = RMeshRenderer.num_variants()
    grid = [(ts, v % nf, v % nb) for ts in (8, 12, 16) for v in range(max(nf, nb))]
    return grid + [(32, 0, 0), (32, 1, 1)]  # tile 32: two geometries each


@pytest.mark.parametrize("ts,fv,bv", _variant_grid())
def test_kernel_variants_match_oracle(ts, fv, bv, config_a):
    """Every batch geometry the autotuner can pick renders and differentiates
    like the oracle (the geometry changes the schedule, not the arithmetic)."""
    scene, cam, w, h = config_a
    r = _renderer(scene, w, h)
    r.set_tuning(ts, fv, bv)
    assert r.get_tuning() == (ts, fv, bv)
    ref = orc.render(scene, cam, w, h, ts=ts)
    img = r.forward_tiled(*cam)
    assert np.abs(img - ref["rgba"]).max() < IMG_TOL
    rng = np.random.default_rng(ts + fv)
    dl = np.ascontiguousarray(rng.normal(0, 1e-3, (h, w, 4)).astype(np.float32))
    g = r.backward(dl)
    gr = orc.backward(scene, cam, w, h, ref["sorted_vals"], ref["ranges"], ref["rgba"], dl, ts)
    for k in ("d_vertices", "d_densities", "d_color_grads", "d_base_colors"):
        a, b = g[k].astype(np.float64), gr[k].astype(np.float64)
        rel = np.linalg.norm(a - b) / np.linalg.norm(b)
        assert rel < GRAD_TOL, (ts, fv, bv, k, rel)


def test_autotune_keeps_parameters(config_a):
    scene, cam, w, h = config_a
    r = _renderer(scene, w, h, max_views=2)
    cams = [cam, make_camera((0.0, 3.0, 0.5), (0, 0, 0), math.radians(50.0), w, h)]
    r.set_targets(np.zeros((2, h, w, 3), np.float32))
    before = r.get_params()
    res = r.autotune(cams, loss_type=0, reps=1)
    ts, fv, bv = res["tuning"]
    assert ts in (8, 12, 16, 32) and res["ms"]["total"] > 0
    assert res["sort_digit_bits"] in (4, 8)
    after = r.get_params()
    for k in before:
        np.testing.assert_array_equal(before[k], after[k])
    assert r.step == 0
    # the chosen configuration still matches the oracle
    ref = orc.render(scene, cam, w, h, ts=ts)
    assert np.abs(r.forward_tiled(*cam) - ref["rgba"]).max() < IMG_TOL


def test_batched_binning_at_scale():
    """~300K tets x 4 views at 512^2 (> 4096 scan blocks of (tet, view) items):
    the batched sort holds, view by view, exactly the oracle's sorted pairs
    (key tile = view * (T + 1) + tile; tile T of each view is its sentinel
    tile), every view's sentinels sort between its pairs and the next view's,
    and each view's ranges are the oracle's per-view ranges shifted by the
    view's offset — including a tile followed by its view's sentinels, which
    keeps end = 0 as in the reference's per-view sort."""
    from rmeshvk_amd import RMeshRenderer

    rng = np.random.default_rng(21)
    scene = make_scene(cluster_points(45000, rng, 6, 0.2), rng)
    w = h = 512
    cams = fibonacci_cameras(4, 3.0, 50.0, w, h)
    r = RMeshRenderer(scene["vertices"], scene["indices"], scene["base_colors"], scene["densities"],
                      scene["color_grads"], scene["circumdata"], w, h, device_id=0, max_views=len(cams))
    r.forward_batched(cams, copy_out=False)
    sk, sv, rg = r.debug_sorted(views=len(cams))
    _, _, T = orc.tile_grid(w, h, 12)
    KT = T + 1
    M = len(scene["indices"]) // 4
    assert (M * len(cams) + 255) // 256 > 4096
    kt = (sk >> np.uint64(32)).astype(np.int64)
    v_of, t_of = kt // KT, kt % KT
    assert np.all(np.diff(kt) >= 0)
    for v, cam in enumerate(cams):
        tt, _ = orc.project(scene, cam, w, h, 12)
        keys, vals, P = orc.pairs(scene, cam, w, h, tt, 12)
        ok, ov = orc.sort(keys, vals, orc.sorting_bits(T))
        o_real = (ok >> np.uint64(32)) != 0xFFFFFFFF
        mine = v_of == v
        sel = mine & (t_of < T)
        np.testing.assert_array_equal(sk[sel] - (np.uint64(v * KT) << np.uint64(32)), ok[o_real])
        np.testing.assert_array_equal(sv[sel], ov[o_real])
        assert int((mine & (t_of == T)).sum()) == int((~o_real).sum())  # this view's sentinels
        base = int(np.argmax(mine))
        rv = orc.ranges(ok, T).reshape(T, 2).astype(np.int64)
        counts = np.bincount((ok[o_real] >> np.uint64(32)).astype(np.int64), minlength=T)
        exp = np.zeros((T, 2), np.int64)
        ne = counts > 0
        exp[ne, 0] = rv[ne, 0] + base
        exp[ne, 1] = np.where(rv[ne, 1] != 0, rv[ne, 1] + base, 0)
        np.testing.assert_array_equal(rg[2 * v * T:2 * (v + 1) * T].reshape(T, 2), exp)


@pytest.mark.parametrize("ts", [12, 32])
def test_sort_digit_width(ts):
    """4-bit digits (twice the passes) give the 8-bit sort's lists bit for bit,
    one view (the oracle's) and four views in one sort."""
    rng = np.random.default_rng(77)
    scene = make_scene(cluster_points(6000, rng, 4, 0.2), rng)
    w, h = 160, 128
    cams = fibonacci_cameras(4, 3.0, 50.0, w, h)
    out = {}
    for bits in (8, 4):
        r = _renderer(scene, w, h, tile_size=ts, max_views=4)
        r.set_sort_digit_bits(bits)
        assert r.get_sort_digit_bits() == bits
        img = r.forward_tiled(*cams[0])
        one = r.debug_sorted()
        imgs = r.forward_batched(cams)
        out[bits] = (img, one, r.debug_sorted(views=4), imgs)
    ref = orc.render(scene, cams[0], w, h, ts=ts)
    for bits in (8, 4):
        img, (sk, sv, rg), _, _ = out[bits]
        np.testing.assert_array_equal(sk, ref["sorted_keys"])
        np.testing.assert_array_equal(sv, ref["sorted_vals"])
        np.testing.assert_array_equal(rg, ref["ranges"])
        assert np.abs(img - ref["rgba"]).max() < IMG_TOL
    for a, b in zip(out[8][2], out[4][2]):
        np.testing.assert_array_equal(a, b)
    np.testing.assert_array_equal(out[8][3], out[4][3])
    with pytest.raises(ValueError):
        _renderer(scene, w, h).set_sort_digit_bits(5)
