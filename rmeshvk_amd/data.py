"""Scene files: .rmesh and PLY (crates/rmesh-data/src/lib.rs).

Loading goes through the native parser of librmesh_hip.so (``rmesh_scene_*``:
gzip inflate, f16 decode, PCA decompression of the SH coefficients and the
float64 circumspheres, in parallel over tets).  The results are the
renderer's flat arrays:

    scene, sh, tagged = load_rmesh("scene.rmesh")
    r = renderer_from_scene(scene, sh, 800, 800)

``load_rmesh`` / ``load_rmesh_raw`` / ``load_ply`` mirror the reference's
``load_rmesh`` / ``load_rmesh_raw`` / ``load_ply`` (lib.rs:140-148, 711-853),
returning (SceneData, ShCoeffs, tagged) as dicts.  The tagged PBR sections of
an extended .rmesh are counted (``tagged = {"sections": n}``), not decoded:
the PBR viewer they feed is not part of this package; ``None`` when absent.

``save_rmesh`` and ``save_ply`` write the same formats (the layout of the
reference's ``convert.py`` and of the binary PLY its loader reads).
"""

from __future__ import annotations

import ctypes
import gzip
import os

import numpy as np

from . import _lib


def _bytes(src):
    if isinstance(src, (str, os.PathLike)):
        with open(src, "rb") as f:
            return np.frombuffer(f.read(), np.uint8)
    if isinstance(src, np.ndarray):
        return np.ascontiguousarray(src).view(np.uint8).ravel()
    return np.frombuffer(bytes(src), np.uint8)


def _parse(src, fmt):
    lib = _lib.load()
    data = _bytes(src)
    h = ctypes.c_void_p()
    info = _lib.SceneInfo()
    _lib.check(lib.rmesh_scene_parse(data.ctypes.data if data.size else None, data.size, fmt, ctypes.byref(h),
                                     ctypes.byref(info)))
    try:
        N, M, deg = int(info.n_vertices), int(info.n_tets), int(info.sh_degree)
        nc = (deg + 1) ** 2
        scene = {
            "vertices": np.empty(3 * N, np.float32),
            "indices": np.empty(4 * M, np.uint32),
            "densities": np.empty(M, np.float32),
            "color_grads": np.empty(3 * M, np.float32),
            "circumdata": np.empty(4 * M, np.float32),
        }
        coeffs = np.empty(3 * nc * M, np.float32)
        p = lambda a: a.ctypes.data  # noqa: E731
        _lib.check(lib.rmesh_scene_copy(h, p(scene["vertices"]), p(scene["indices"]), p(scene["densities"]),
                                        p(scene["color_grads"]), p(scene["circumdata"]), p(coeffs)))
    finally:
        lib.rmesh_scene_free(h)
    scene["start_pose"] = np.array(info.start_pose[:], np.float32)
    scene["vertex_count"] = N
    scene["tet_count"] = M
    sh = {"coeffs": coeffs, "degree": deg}
    tagged = {"sections": int(info.n_tagged)} if info.n_tagged else None
    return scene, sh, tagged


def load_rmesh(src):
    """.rmesh (gzip-compressed or raw) from a path or bytes (lib.rs:140-144)."""
    return _parse(src, _lib.FILE_RMESH)


def load_rmesh_raw(src):
    """Uncompressed .rmesh from a path or bytes (lib.rs:146-148)."""
    return _parse(src, _lib.FILE_RMESH_RAW)


def load_ply(src):
    """Binary little-endian PLY from a path or bytes (lib.rs:711-853)."""
    return _parse(src, _lib.FILE_PLY)


def renderer_from_scene(scene, sh, width, height, **kw):
    """An RMeshRenderer over a loaded scene, colours from its SH coefficients."""
    from .renderer import RMeshRenderer

    M = int(scene["tet_count"])
    return RMeshRenderer(scene["vertices"], scene["indices"], np.zeros(3 * M, np.float32), scene["densities"],
                         scene["color_grads"], scene["circumdata"], width, height, sh_coeffs=sh["coeffs"],
                         sh_degree=int(sh["degree"]), **kw)


# ---------------------------------------------------------------------------
# writers
# ---------------------------------------------------------------------------

def pca_compress(sh, k):
    """PCA of the per-tet SH rows [M, D] (convert.py compress_matrix): centred
    covariance, eigenvectors by decreasing eigenvalue, the top k.  Returns
    (weights [M, k], basis [k, D], mean [D]) as float16."""
    X = np.asarray(sh, np.float64).reshape(len(sh), -1)
    mean = X.mean(axis=0)
    Xc = X - mean
    cov = Xc.T @ Xc / max(len(X) - 1, 1)
    w, v = np.linalg.eigh(cov)
    top = v[:, np.argsort(w)[::-1]][:, :k]
    return (Xc @ top).astype(np.float16), top.T.astype(np.float16), mean.astype(np.float16)


def encode_densities(s):
    """u8 log encoding of densities (convert.py): log(max(s, 1e-3)) * 20 + 100,
    values <= 1 to 0, clipped to [0, 255] and truncated."""
    d = np.log(np.clip(np.asarray(s, np.float64), 1e-3, None)) * 20 + 100
    d[d <= 1] = 0
    return np.clip(d, 0, 255).astype(np.uint8)


def _pad4(buf):
    buf += b"\x00" * (-len(buf) % 4)


def save_rmesh(path, vertices, indices, densities, color_grads, sh_coeffs, sh_degree, k_components=12,
               start_pose=None, compress=True, tagged=None, level=9):
    """Write a .rmesh (the convert.py layout).  sh_coeffs: [M, 3, (deg+1)^2]
    channel-major; k_components <= 3 (deg+1)^2.  tagged: optional list of
    (tag, array, dtype_code) extension sections (0 f16, 1 f32, 2 u8, 3 i32).
    compress: gzip at `level` (convert.py uses 9).  Returns the bytes (also
    written to path unless it is None)."""
    V = np.ascontiguousarray(vertices, np.float32).reshape(-1, 3)
    I = np.ascontiguousarray(indices, np.uint32).reshape(-1, 4)
    M = len(I)
    nc = (int(sh_degree) + 1) ** 2
    sh = np.asarray(sh_coeffs, np.float32).reshape(M, 3 * nc)
    weights, basis, mean = pca_compress(sh, int(k_components))
    pose = np.zeros(8, np.float32) if start_pose is None else np.asarray(start_pose, np.float32).reshape(8)
    buf = bytearray()
    buf += np.array([len(V), M, int(sh_degree), int(k_components)], np.uint32).tobytes()
    buf += pose.tobytes()
    buf += V.tobytes()
    buf += I.tobytes()
    buf += encode_densities(densities).tobytes()
    _pad4(buf)
    buf += mean.tobytes()
    buf += basis.tobytes()
    buf += weights.tobytes()
    _pad4(buf)
    buf += np.asarray(color_grads, np.float32).reshape(M, 3).astype(np.float16).tobytes()
    if tagged:
        _pad4(buf)
        buf += np.array([0x524D5458, len(tagged)], np.uint32).tobytes()
        for tag, arr, code in tagged:
            a = np.asarray(arr)
            buf += tag.encode("ascii")[:16].ljust(16, b"\x00")
            buf += np.array([code, a.ndim, *a.shape], np.uint32).tobytes()
            payload = a.astype({0: np.float16, 1: np.float32, 2: np.uint8, 3: np.int32}[code]).tobytes()
            buf += np.array([len(payload)], np.uint32).tobytes()
            buf += payload
            _pad4(buf)
    out = gzip.compress(bytes(buf), compresslevel=level) if compress else bytes(buf)
    if path is not None:
        with open(path, "wb") as f:
            f.write(out)
    return out


def save_ply(path, vertices, indices, densities, color_grads, sh_coeffs, sh_degree):
    """Write the binary PLY load_ply reads: vertices (x, y, z) and tetrahedra
    (uchar 4 int indices, s, grd_xyz, sh_i_{r,g,b} interleaved by coefficient).
    sh_coeffs: [M, 3, (deg+1)^2] channel-major.  Returns the bytes."""
    V = np.ascontiguousarray(vertices, np.float32).reshape(-1, 3)
    I = np.ascontiguousarray(indices, np.int32).reshape(-1, 4)
    M = len(I)
    nc = (int(sh_degree) + 1) ** 2
    sh = np.asarray(sh_coeffs, np.float32).reshape(M, 3, nc).transpose(0, 2, 1).reshape(M, 3 * nc)
    head = ["ply", "format binary_little_endian 1.0", f"element vertex {len(V)}", "property float x",
            "property float y", "property float z", f"element tetrahedron {M}",
            "property list uchar int indices", "property float s", "property float grd_x",
            "property float grd_y", "property float grd_z"]
    for c in range(nc):
        head += [f"property float sh_{c}_r", f"property float sh_{c}_g", f"property float sh_{c}_b"]
    head.append("end_header")
    rec = np.dtype([("n", "u1"), ("idx", "<i4", 4), ("s", "<f4"), ("g", "<f4", 3), ("sh", "<f4", 3 * nc)])
    t = np.empty(M, rec)
    t["n"] = 4
    t["idx"] = I
    t["s"] = np.asarray(densities, np.float32)
    t["g"] = np.asarray(color_grads, np.float32).reshape(M, 3)
    t["sh"] = sh
    out = ("\n".join(head) + "\n").encode("ascii") + V.astype("<f4").tobytes() + t.tobytes()
    if path is not None:
        with open(path, "wb") as f:
            f.write(out)
    return out
