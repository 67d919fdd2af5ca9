#!/usr/bin/env python3
"""Writes reference_kats.json: the inputs and expected outputs that the
reference's own tests state for the train-step path, as data.

Each entry names the test it comes from.  Values are the literals of those
tests (their inputs and the closed-form expectations they assert); nothing
here is computed by this project's code, so the fixture pins the oracle
independently of it.

    python tests/golden/make_reference_kats.py
"""

from __future__ import annotations

import json
import os

HERE = os.path.dirname(os.path.abspath(__file__))


def kats() -> dict:
    n_px = 64 * 64
    return {
        "loss_l1": {
            "source": "crates/rmesh-backward/tests/kernel_tests.rs:176-305",
            "width": 64, "height": 64, "loss_type": 0,
            "rendered_rgba": [0.8, 0.6, 0.4, 1.0],
            "ground_truth_rgb": [0.5, 0.5, 0.5],
            # sign(rendered - gt) / n_pixels, alpha gradient 0
            "expected_dl": [1.0 / n_px, 1.0 / n_px, -1.0 / n_px, 0.0],
            "tol": 1e-5,
        },
        "adam": {
            "source": "crates/rmesh-backward/tests/kernel_tests.rs:313-420",
            "params": [1.0, 2.0, 3.0, 4.0],
            "grads": [0.1, -0.2, 0.3, -0.4],
            "step": 1, "lr": 0.01, "beta1": 0.9, "beta2": 0.999, "eps": 1e-8,
            # m = (1 - beta1) g with m0 = 0; params move against the gradient
            "expected_m": [0.01, -0.02, 0.03, -0.04],
            "m_tol": 1e-6,
            "min_move": 1e-6,
        },
        "radix_sort": {
            "source": "crates/rmesh-backward/tests/kernel_tests.rs:96-167",
            "n": 64, "key_modulus": 10000, "sorting_bits": 32,
            "checks": ["ascending keys", "sorted_keys[i] == keys[values[i]]"],
        },
        "tile_fill": {
            "source": "crates/rmesh-backward/tests/kernel_tests.rs:428-545",
            "sentinel_key": [4294967295, 4294967295],
            "sentinel_value": 0,
        },
        "sorting_bits": {
            "source": "crates/rmesh-sort/src/drs.rs:23-26",
            # num_tiles -> (32 + ceil(log2(num_tiles)) + 7) & ~7, one tile -> 1 tile bit
            "table": [[0, 40], [1, 40], [2, 40], [256, 40], [257, 48], [4489, 48], [65536, 48],
                      [65537, 56], [16777216, 56], [16777217, 64]],
        },
        "known_tet": {
            "source": "crates/rmesh-backward/tests/kernel_tests.rs:1516-1671,1839-1914",
            "vertices": [0.5, 0.5, 0.5, -0.5, -0.5, 0.5, -0.5, 0.5, -0.5, 0.5, -0.5, -0.5],
            "indices": [0, 1, 2, 3],
            "density": 3.0,
            "eye": [0.0, -2.0, 0.5], "target": [0.0, 0.0, 0.0], "up": [0.0, 0.0, 1.0],
            "fov_y_deg": 90.0, "width": 64, "height": 64, "near": 0.01, "far": 100.0,
            "tile_size": 12,
            "expected_visible": 1,
            "expected_tiles_touched_min": 1,
            "expected_pixel_area_min": 1.0,
            "expected_center_alpha_min": 0.01,
        },
        "circumsphere": {
            "source": "crates/rmesh-util/src/test_util.rs:133-163",
            # regular tet inscribed in the unit cube: centre 0, r^2 = 0.75
            "vertices": [0.5, 0.5, 0.5, -0.5, -0.5, 0.5, -0.5, 0.5, -0.5, 0.5, -0.5, -0.5],
            "indices": [0, 1, 2, 3],
            "expected": [0.0, 0.0, 0.0, 0.75],
        },
        "autograd_fd": {
            "source": "crates/rmesh-python/tests/test_autograd.py:138-250",
            "width": 32, "height": 32, "seed": 42, "eps": 1e-3, "rel_tol": 0.1,
        },
        "multi_tet_fd": {
            "source": "crates/rmesh-backward/tests/kernel_tests.rs:733-1180",
            "eps": 1e-3, "rel_tol": 0.10, "abs_tol": 5e-5,
        },
    }


def main() -> None:
    path = os.path.join(HERE, "reference_kats.json")
    with open(path, "w") as f:
        json.dump(kats(), f, indent=1, sort_keys=True)
        f.write("\n")
    print(path)


if __name__ == "__main__":
    main()
