"""Camera packing helpers.

Conventions follow the reference's camera module and test setup:

* ``perspective_matrix`` / ``look_at``: wgpu depth [0, 1], Z-up world
  (crates/rmesh-util/src/camera.rs:103-125).
* c2w columns ``[right, -up, forward]`` (pinhole: x right, y down, z forward)
  and intrinsics ``fx = fy = H / (2 tan(fov/2))``, ``cx = W/2``, ``cy = H/2``
  (crates/rmesh-render/tests/common/mod.rs:219-232).
* Flat packing used by every API entry point
  (crates/rmesh-python/src/lib.rs:117-136): ``vp`` is the 4x4 view-projection
  matrix flattened column-major; ``c2w_intrinsics`` is
  ``[c2w_col0.xyz, 0, c2w_col1.xyz, 0, c2w_col2.xyz, 0, fx, fy, cx, cy]``.
"""

from __future__ import annotations

import math

import numpy as np


def perspective_matrix(fov_y: float, aspect: float, near: float, far: float) -> np.ndarray:
    """Row-major 4x4 of camera.rs:108-116 (columns listed there)."""
    f = 1.0 / math.tan(fov_y / 2.0)
    m = np.zeros((4, 4), dtype=np.float64)
    m[0, 0] = f / aspect
    m[1, 1] = f
    m[2, 2] = far / (far - near)
    m[3, 2] = 1.0
    m[2, 3] = -(far * near) / (far - near)
    return m


def _basis(eye, target, up=(0.0, 0.0, 1.0)):
    eye = np.asarray(eye, dtype=np.float64)
    target = np.asarray(target, dtype=np.float64)
    f = target - eye
    f /= np.linalg.norm(f)
    r = np.cross(f, np.asarray(up, dtype=np.float64))
    r /= np.linalg.norm(r)
    u = np.cross(r, f)
    return f, r, u


def look_at(eye, target, up=(0.0, 0.0, 1.0)) -> np.ndarray:
    """Row-major 4x4 of camera.rs:119-129."""
    eye = np.asarray(eye, dtype=np.float64)
    f, r, u = _basis(eye, target, up)
    m = np.eye(4, dtype=np.float64)
    m[0, :3], m[0, 3] = r, -r @ eye
    m[1, :3], m[1, 3] = u, -u @ eye
    m[2, :3], m[2, 3] = f, -f @ eye
    return m


def make_camera(eye, target, fov_y: float, width: int, height: int,
                near: float = 0.01, far: float = 100.0):
    """Return ``(cam_pos[3], vp[16], c2w_intrinsics[16])`` as float32 arrays."""
    eye = np.asarray(eye, dtype=np.float64)
    vp = perspective_matrix(fov_y, width / height, near, far) @ look_at(eye, target)
    f, r, u = _basis(eye, target)
    f_val = 1.0 / math.tan(fov_y / 2.0)
    fx = fy = f_val * height / 2.0
    ci = np.zeros(16, dtype=np.float64)
    ci[0:3] = r
    ci[4:7] = -u
    ci[8:11] = f
    ci[12:16] = [fx, fy, width / 2.0, height / 2.0]
    return (eye.astype(np.float32), vp.flatten(order="F").astype(np.float32), ci.astype(np.float32))


def fibonacci_cameras(n: int, radius: float, fov_deg: float, width: int, height: int,
                      target=(0.0, 0.0, 0.0)):
    """``n`` cameras on a Fibonacci sphere of ``radius`` looking at ``target``."""
    cams = []
    golden = math.pi * (3.0 - math.sqrt(5.0))
    for i in range(n):
        z = 1.0 - (2.0 * i + 1.0) / n
        rr = math.sqrt(max(0.0, 1.0 - z * z))
        th = golden * i
        eye = radius * np.array([rr * math.cos(th), rr * math.sin(th), z])
        cams.append(make_camera(eye + np.asarray(target), target, math.radians(fov_deg), width, height))
    return cams


def colmap_camera(width: int, height: int, distance: float = 1.0):
    """Camera of the reference's autograd test (test_autograd.py:69-121):
    at (0, 0, d) looking at the origin, 45 deg, projection built from intrinsics."""
    cam_pos = np.array([0.0, 0.0, distance], dtype=np.float32)
    R = np.array([[1, 0, 0], [0, -1, 0], [0, 0, -1]], dtype=np.float32)
    t = -R @ cam_pos
    Rt = np.eye(4, dtype=np.float32)
    Rt[:3, :3] = R.T
    Rt[:3, 3] = t
    fov = 45.0 * np.pi / 180.0
    aspect = width / height
    f_val = 1.0 / np.tan(fov / 2.0)
    fx = f_val * width / (2.0 * aspect)
    fy = f_val * height / 2.0
    cx, cy = width / 2.0, height / 2.0
    proj = np.zeros((4, 4), dtype=np.float32)
    proj[0, 0] = 2 * fx / width
    proj[0, 2] = 2 * cx / width - 1
    proj[1, 1] = -2 * fy / height
    proj[1, 2] = 1 - 2 * cy / height
    proj[3, 2] = 1.0
    vp = Rt.T @ proj.T
    R_c2w = R.T
    ci = np.zeros(16, dtype=np.float32)
    ci[0:3] = R_c2w[:, 0]
    ci[4:7] = R_c2w[:, 1]
    ci[8:11] = R_c2w[:, 2]
    ci[12:16] = [fx, fy, cx, cy]
    return cam_pos, vp.ravel().astype(np.float32), ci
