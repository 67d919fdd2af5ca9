"""Seeded synthetic radiance-mesh scenes (host-side data generation).

Points -> ``scipy.spatial.Delaunay`` -> tets with positive orientation
(crates/rmesh-util/src/test_util.rs:175-185 swaps vertices 2 and 3 when the
determinant is negative) -> circumdata (cx, cy, cz, r^2) computed in f64 and
stored f32 (crates/rmesh-data/src/lib.rs:601-710).

Literal micro-scenes reproduce the reference's test geometry:
``regular_tet_scene`` (crates/rmesh-backward/tests/kernel_tests.rs:1516-1540)
and ``two_tet_scene`` geometry (kernel_tests.rs:61-88).
"""

from __future__ import annotations

import numpy as np


def circumspheres(vertices: np.ndarray, indices: np.ndarray) -> np.ndarray:
    """(cx, cy, cz, r^2) per tet in f64, degenerate denominator -> 1 (test_util.rs:133-164)."""
    v = vertices.reshape(-1, 3).astype(np.float64)
    t = indices.reshape(-1, 4).astype(np.int64)
    p0, p1, p2, p3 = v[t[:, 0]], v[t[:, 1]], v[t[:, 2]], v[t[:, 3]]
    a, b, c = p1 - p0, p2 - p0, p3 - p0
    aa = np.einsum("ij,ij->i", a, a)
    bb = np.einsum("ij,ij->i", b, b)
    cc = np.einsum("ij,ij->i", c, c)
    bc, ca, ab = np.cross(b, c), np.cross(c, a), np.cross(a, b)
    denom = 2.0 * np.einsum("ij,ij->i", a, bc)
    denom = np.where(np.abs(denom) < 1e-12, 1.0, denom)
    r = (aa[:, None] * bc + bb[:, None] * ca + cc[:, None] * ab) / denom[:, None]
    out = np.empty((len(t), 4), dtype=np.float64)
    out[:, :3] = p0 + r
    out[:, 3] = np.einsum("ij,ij->i", r, r)
    return out.astype(np.float32).ravel()


def orient_positive(vertices: np.ndarray, tets: np.ndarray) -> np.ndarray:
    v = vertices.reshape(-1, 3).astype(np.float64)
    t = tets.reshape(-1, 4).astype(np.int64).copy()
    p0, p1, p2, p3 = v[t[:, 0]], v[t[:, 1]], v[t[:, 2]], v[t[:, 3]]
    det = np.einsum("ij,ij->i", p1 - p0, np.cross(p2 - p0, p3 - p0))
    neg = det < 0
    t[neg, 2], t[neg, 3] = t[neg, 3].copy(), t[neg, 2].copy()
    return t.astype(np.uint32)


def delaunay_tets(points: np.ndarray) -> np.ndarray:
    from scipy.spatial import Delaunay

    d = Delaunay(points.astype(np.float64))
    return orient_positive(points, d.simplices)


def make_scene(points: np.ndarray, rng: np.random.Generator, sh_degree: int = -1,
               sh_sigma: float = 0.3, dc_mean: float = 0.5, dc_sigma: float = 0.2,
               base_sigma: float = 0.3) -> dict:
    """Per-tet parameters: densities ~ LogN(0, 1), color_grads ~ N(0, 0.1).

    ``sh_degree < 0``: base_colors ~ N(0, base_sigma) (the reference's Python
    path, colours used as given).  ``sh_degree >= 0``: SH coefficients
    [M][3][nc] ~ N(0, sh_sigma) with DC ~ N(dc_mean, dc_sigma).
    """
    verts = points.astype(np.float32)
    tets = delaunay_tets(verts)
    M = len(tets)
    scene = {
        "vertices": verts.ravel().copy(),
        "indices": tets.ravel().copy(),
        "densities": rng.lognormal(0.0, 1.0, M).astype(np.float32),
        "color_grads": rng.normal(0.0, 0.1, M * 3).astype(np.float32),
    }
    scene["circumdata"] = circumspheres(scene["vertices"], scene["indices"])
    if sh_degree >= 0:
        nc = (sh_degree + 1) ** 2
        sh = rng.normal(0.0, sh_sigma, (M, 3, nc)).astype(np.float32)
        sh[:, :, 0] = rng.normal(dc_mean, dc_sigma, (M, 3)).astype(np.float32)
        scene["sh"] = sh.ravel().copy()
        scene["sh_degree"] = sh_degree
        scene["base_colors"] = np.zeros(M * 3, dtype=np.float32)
    else:
        scene["base_colors"] = rng.normal(0.0, base_sigma, M * 3).astype(np.float32)
    return scene


def uniform_points(n: int, rng: np.random.Generator) -> np.ndarray:
    return rng.uniform(-1.0, 1.0, (n, 3))


def cluster_points(n: int, rng: np.random.Generator, n_clusters: int = 8, sigma: float = 0.15,
                   background: float = 0.0) -> np.ndarray:
    """Mixture of Gaussian clusters clipped to [-1, 1]^3, plus a uniform background fraction."""
    n_bg = int(round(n * background))
    n_cl = n - n_bg
    centers = rng.uniform(-0.6, 0.6, (n_clusters, 3))
    lab = rng.integers(0, n_clusters, n_cl)
    pts = np.clip(centers[lab] + rng.normal(0.0, sigma, (n_cl, 3)), -1.0, 1.0)
    if n_bg:
        pts = np.concatenate([pts, rng.uniform(-1.0, 1.0, (n_bg, 3))], axis=0)
    # Clipping can duplicate points on the cube faces; jitter duplicates apart.
    pts = np.unique(pts, axis=0)
    return pts


def regular_tet_scene(density: float = 3.0) -> dict:
    """Regular tet inscribed in the unit cube (kernel_tests.rs:1516-1540)."""
    verts = np.array([0.5, 0.5, 0.5, -0.5, -0.5, 0.5, -0.5, 0.5, -0.5, 0.5, -0.5, -0.5], dtype=np.float32)
    idx = np.array([0, 1, 2, 3], dtype=np.uint32)
    s = {
        "vertices": verts,
        "indices": idx,
        "densities": np.array([density], dtype=np.float32),
        "color_grads": np.zeros(3, dtype=np.float32),
        "base_colors": np.full(3, 0.5, dtype=np.float32),
    }
    s["circumdata"] = circumspheres(verts, idx)
    return s


def two_tet_scene(rng: np.random.Generator) -> dict:
    """Two tets sharing a face, literal geometry of kernel_tests.rs:61-88.

    Densities U(0.5, 3.5) and colour gradients U(-0.05, 0.05) as there, drawn
    from numpy's generator (the reference's ChaCha8 stream is not reproduced).
    """
    verts = np.array([0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.5, 1.0, 0.0, 0.5, 0.3, 0.8, 0.5, 0.3, -0.8],
                     dtype=np.float32)
    idx = np.array([0, 1, 2, 3, 0, 2, 1, 4], dtype=np.uint32)
    s = {
        "vertices": verts,
        "indices": idx,
        "densities": (rng.random(2) * 3.0 + 0.5).astype(np.float32),
        "color_grads": ((rng.random(6) - 0.5) * 0.1).astype(np.float32),
        "base_colors": np.full(6, 0.5, dtype=np.float32),
    }
    s["circumdata"] = circumspheres(verts, idx)
    return s


def single_tet_scene(seed: int = 42) -> dict:
    """Single random tet of the reference's autograd test (test_autograd.py:34-66)."""
    rng = np.random.RandomState(seed)
    radius = 0.6
    verts = (rng.rand(4, 3).astype(np.float32) - 0.5) * 2.0 * radius
    v0, v1, v2, v3 = verts
    det = np.dot(v1 - v0, np.cross(v2 - v0, v3 - v0))
    idx = np.array([0, 1, 3, 2] if det < 0 else [0, 1, 2, 3], dtype=np.uint32)
    base_colors = rng.rand(3).astype(np.float32) * 2.0 - 1.0
    densities = np.array([rng.rand() * 5.0 + 0.5], dtype=np.float32)
    color_grads = (rng.rand(3).astype(np.float32) - 0.5) * 0.2
    s = {
        "vertices": verts.ravel().copy(),
        "indices": idx,
        "base_colors": base_colors,
        "densities": densities,
        "color_grads": color_grads,
    }
    s["circumdata"] = circumspheres(s["vertices"], idx)
    return s


def config_scene(name: str, seed: int = 0) -> tuple[dict, list]:
    """Scenes and cameras of the benchmark configurations A, B, C, E (sizes per name)."""
    from .camera import fibonacci_cameras, make_camera
    import math

    rng = np.random.default_rng(seed)
    if name == "A":
        scene = make_scene(uniform_points(2000, rng), rng, sh_degree=-1, base_sigma=0.3)
        cams = [make_camera((3.0, 0.0, 0.0), (0.0, 0.0, 0.0), math.radians(50.0), 256, 256)]
        return scene, cams
    if name == "B":
        scene = make_scene(uniform_points(10000, rng), rng, sh_degree=3)
        return scene, fibonacci_cameras(8, 3.0, 50.0, 512, 512)
    if name == "C":
        scene = make_scene(cluster_points(200000, rng, 8, 0.15), rng, sh_degree=3)
        return scene, fibonacci_cameras(16, 3.0, 50.0, 800, 800)
    if name == "E":
        # 2,000,000 points: 16 Gaussian clusters + 20 % uniform background; 4 views at 1920x1080
        scene = make_scene(cluster_points(2000000, rng, 16, 0.15, background=0.2), rng, sh_degree=3)
        return scene, fibonacci_cameras(4, 3.0, 50.0, 1920, 1080)
    raise ValueError(f"unknown config {name!r}")


def config_d_scenes(seed: int = 0, n_scenes: int = 32) -> list:
    """The uneven multi-scene batch: per scene a vertex count log-uniform in
    [5k, 100k] (uniform in the cube or 4-8 clusters, by coin), SH degree from
    {0, 1, 3} and a square resolution from {256, 512, 1024}, 4 Fibonacci
    views.  Returns [(scene, cams, resolution)]."""
    from .camera import fibonacci_cameras

    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n_scenes):
        n = int(round(np.exp(rng.uniform(np.log(5000), np.log(100000)))))
        deg = int(rng.choice([0, 1, 3]))
        res = int(rng.choice([256, 512, 1024]))
        if rng.random() < 0.5:
            pts = uniform_points(n, rng)
        else:
            pts = cluster_points(n, rng, int(rng.integers(4, 9)), 0.15)
        scene = make_scene(pts, rng, sh_degree=deg)
        out.append((scene, fibonacci_cameras(4, 3.0, 50.0, res, res), res))
    return out
