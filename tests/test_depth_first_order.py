"""CPU: the pair order of depth-first pair generation.

The GPU binning sorts the (tet, view) items stably by depth key, emits each
item's pairs in that order, and then sorts the pairs stably on the tile bits
only.  These tests check, in numpy, that this gives exactly the reference's
order (a stable sort of the pairs in item order on the masked (tile, depth)
key, crates/rmesh-sort/src/drs.rs:23-26): on the oracle's own pair lists, on
synthetic items built to tie often (equal depth keys, equal tiles), with
sentinel pairs, and — as the counter-case the renderer excludes — when one
view's sentinel key masks to its last real tile.
"""

from __future__ import annotations

import numpy as np
import pytest

from oracle import oracle as orc
from rmeshvk_amd.camera import fibonacci_cameras
from rmeshvk_amd.scenes import cluster_points, make_scene

SENT = np.uint64(0xFFFFFFFFFFFFFFFF)


def _reference_order(keys, vals, bits):
    mask = np.uint64((1 << bits) - 1)
    o = np.argsort(keys & mask, kind="stable")
    return keys[o], vals[o]


def _depth_first_order(keys, vals, counts, dkey, bits):
    """Items sorted stably by dkey emit their slices of the item-order list;
    then a stable sort on key bits [32, bits)."""
    offs = np.concatenate([[0], np.cumsum(counts.astype(np.int64))])
    items = np.argsort(dkey, kind="stable")
    idx = np.concatenate([np.arange(offs[i], offs[i + 1]) for i in items] + [np.zeros(0, np.int64)])
    k2, v2 = keys[idx], vals[idx]
    tmask = np.uint64((1 << (bits - 32)) - 1)
    o = np.argsort((k2 >> np.uint64(32)) & tmask, kind="stable")
    return k2[o], v2[o]


def _item_depth_keys(scene, cam, counts):
    cd = np.asarray(scene["circumdata"], np.float32).reshape(-1, 4)
    c = np.asarray(cam[0], np.float32)
    dx, dy, dz = cd[:, 0] - c[0], cd[:, 1] - c[1], cd[:, 2] - c[2]
    depth = ((dx * dx + dy * dy) + dz * dz) - cd[:, 3]
    return np.where(counts > 0, ~depth.astype(np.float32).view(np.uint32), np.uint32(0xFFFFFFFF))


@pytest.mark.parametrize("ts", [8, 12, 16])
def test_oracle_pairs_depth_first_equals_reference_sort(ts):
    rng = np.random.default_rng(5)
    s = make_scene(cluster_points(3000, rng, 3, 0.25), rng)
    w, h = 96, 80
    for cam in fibonacci_cameras(3, 3.0, 50.0, w, h):
        tt, _ = orc.project(s, cam, w, h, ts)
        keys, vals, P = orc.pairs(s, cam, w, h, tt, ts)
        _, _, T = orc.tile_grid(w, h, ts)
        bits = orc.sorting_bits(T)
        ok, ov = orc.sort(keys, vals, bits)
        dk, dv = _depth_first_order(keys, vals, tt, _item_depth_keys(s, cam, tt), bits)
        np.testing.assert_array_equal(dk, ok)
        np.testing.assert_array_equal(dv, ov)


def _synthetic(rng, n_items, n_tiles, n_views, depth_levels, sentinel_rate):
    """Item-order pair lists of n_views views (key tiles per view: n_tiles
    real + 1 sentinel), items tet-major, depth keys from a few levels (ties)."""
    kt_per_view = n_tiles + 1
    keys, vals, counts, dkeys = [], [], [], []
    for it in range(n_items):
        t, v = it // n_views, it % n_views
        c = int(rng.integers(0, 5))
        d = np.uint32(rng.integers(0, depth_levels))
        tiles = np.sort(rng.choice(n_tiles, size=c, replace=False)) if c else np.zeros(0, np.int64)
        n_sent = int(rng.random() < sentinel_rate) if c else 0
        ks = [(np.uint64(v * kt_per_view + int(x)) << np.uint64(32)) | np.uint64(d) for x in tiles[:c - n_sent]]
        ks += [(np.uint64(v * kt_per_view + n_tiles) << np.uint64(32)) | np.uint64(0xFFFFFFFF)] * n_sent
        keys += ks
        vals += [t] * (c - n_sent) + [0] * n_sent
        counts.append(c)
        dkeys.append(d if c else np.uint32(0xFFFFFFFF))
    return (np.array(keys, np.uint64), np.array(vals, np.uint32), np.array(counts), np.array(dkeys, np.uint32),
            n_views * kt_per_view)


@pytest.mark.parametrize("seed", range(4))
def test_synthetic_ties_and_sentinels(seed):
    rng = np.random.default_rng(seed)
    keys, vals, counts, dkeys, n_kt = _synthetic(rng, 3000, 37, 4, 6, 0.1)
    bits = orc.sorting_bits(n_kt)
    rk, rv = _reference_order(keys, vals, bits)
    dk, dv = _depth_first_order(keys, vals, counts, dkeys, bits)
    np.testing.assert_array_equal(dk, rk)
    np.testing.assert_array_equal(dv, rv)


def test_one_view_sentinel_on_last_tile_is_the_excluded_case():
    """One view with 256 key tiles (sorting bits 40): the reference's sentinel
    key ~0 masks to tile 255, the last real tile, and sorts behind that
    tile's real pairs by its depth bits (~0).  Depth-first generation emits a
    sentinel at its item's depth position, ahead of deeper items' pairs on
    tile 255, so the renderer keeps the full (tile, depth) sort for this
    layout.  With 257 key tiles (sorting bits 48) the sentinel masks to tile
    0xFFFF, past every real tile, and both orders agree."""
    assert orc.sorting_bits(256) == 40 and orc.sorting_bits(257) == 48
    # item 0: one sentinel (its tet's depth key 0x10); item 1: a real pair on tile 255, depth key 0x20
    keys = np.array([SENT, (np.uint64(255) << np.uint64(32)) | np.uint64(0x20)], np.uint64)
    vals = np.array([0, 7], np.uint32)
    counts = np.array([1, 1])
    dkeys = np.array([0x10, 0x20], np.uint32)
    _, rv = _reference_order(keys, vals, 40)
    _, dv = _depth_first_order(keys, vals, counts, dkeys, 40)
    np.testing.assert_array_equal(rv, [7, 0])
    np.testing.assert_array_equal(dv, [0, 7])
    rk, rv = _reference_order(keys, vals, 48)
    dk, dv = _depth_first_order(keys, vals, counts, dkeys, 48)
    np.testing.assert_array_equal(dk, rk)
    np.testing.assert_array_equal(dv, rv)
