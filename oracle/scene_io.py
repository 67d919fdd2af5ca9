"""Numpy restatement of the reference's scene-file parsers — TEST INFRASTRUCTURE ONLY.

Follows crates/rmesh-data/src/lib.rs: parse_rmesh (243-394), the tagged
section walk (396-470), compute_circumspheres_parallel (597-690) and
load_ply (711-853).  Only tests/ and bench.py's cpu_baseline leg use it, as
the checker of the native parser in librmesh_hip.so.

Float conventions: float32 arithmetic where the reference uses f32 (the
density decode (v - 100) / 20 and its expf, taken from the C library as Rust's
f32::exp does on Linux; the PCA sum mean + w_0 b_0 + w_1 b_1 + ... in k
order, no FMA), float64 for the circumspheres with glam's operation order
(dot = x x + y y + z z, v / s per component).

Pinned by: the layout of the reference's writer (convert.py:213-262, the
tagged sections of convert.py:283-300) restated independently in
rmeshvk_amd.data.save_rmesh and checked field by field at fixed offsets in
tests/test_scene_io.py; round trips of known arrays; f16 and PCA by numpy's
own float16 conversion and a float64 reconstruction; the circumspheres by
the regular-tet known answer, equidistance, and orc_circumspheres on
non-degenerate tets.
"""

from __future__ import annotations

import ctypes
import gzip

import numpy as np

TAGGED_MAGIC = 0x524D5458

_libm = ctypes.CDLL("libm.so.6")
_libm.expf.restype = ctypes.c_float
_libm.expf.argtypes = [ctypes.c_float]
# (v - 100) / 20 in float32, then expf: one value per u8 code (lib.rs:287-293)
_DENS_LUT = np.array([_libm.expf(float((np.float32(v) - np.float32(100.0)) / np.float32(20.0))) for v in range(256)],
                     np.float32)


def circumspheres(vertices, indices):
    """compute_circumspheres_parallel: [M, 4] float32 (centre, r^2)."""
    V = np.asarray(vertices, np.float32).reshape(-1, 3).astype(np.float64)
    I = np.asarray(indices, np.int64).reshape(-1, 4)
    p = [V[I[:, k]] for k in range(4)]
    a, b, c = p[1] - p[0], p[2] - p[0], p[3] - p[0]

    def dot(u, v):
        return u[:, 0] * v[:, 0] + u[:, 1] * v[:, 1] + u[:, 2] * v[:, 2]

    def cross(u, v):
        return np.stack([u[:, 1] * v[:, 2] - u[:, 2] * v[:, 1], u[:, 2] * v[:, 0] - u[:, 0] * v[:, 2],
                         u[:, 0] * v[:, 1] - u[:, 1] * v[:, 0]], axis=1)

    aa, bb, cc = dot(a, a), dot(b, b), dot(c, c)
    cbc, cca, cab = cross(b, c), cross(c, a), cross(a, b)
    denom = 2.0 * dot(a, cbc)
    out = np.zeros((len(I), 4), np.float32)
    deg = np.abs(denom) < 1e-10
    safe = np.where(deg, 1.0, denom)
    num = aa[:, None] * cbc + bb[:, None] * cca + cc[:, None] * cab
    rel = num / safe[:, None]
    cen = p[0] + rel
    out[:, :3] = cen.astype(np.float32)
    out[:, 3] = dot(rel, rel).astype(np.float32)
    if deg.any():
        L = lambda u: np.sqrt(dot(u, u))  # noqa: E731
        md = np.maximum.reduce([L(p[1] - p[0]), L(p[2] - p[0]), L(p[3] - p[0]), L(p[2] - p[1]), L(p[3] - p[1]),
                                L(p[3] - p[2])])
        cg = (((p[0] + p[1]) + p[2]) + p[3]) * 0.25
        r = np.minimum.reduce([L(p[k] - cg) for k in range(4)])
        out[deg, :3] = cg[deg].astype(np.float32)
        out[deg, 3] = np.where(md[deg] < 1e-7, (r * r)[deg], 1e20).astype(np.float32)
    return out.ravel()


def _f16(buf, off, n):
    return np.frombuffer(buf, np.float16, n, off).astype(np.float32), off + 2 * n


def parse_rmesh(data, raw=False):
    """(scene, sh, n_tagged) of a .rmesh image (gzip unless raw)."""
    data = bytes(data)
    if not raw and len(data) >= 2 and data[0] == 0x1F and data[1] == 0x8B:
        data = gzip.decompress(data)
    N, M, deg, k = (int(x) for x in np.frombuffer(data, np.uint32, 4, 0))
    pose = np.frombuffer(data, np.float32, 7, 16).copy()
    off = 48
    verts = np.frombuffer(data, np.float32, 3 * N, off).copy()
    off += 12 * N
    idx = np.frombuffer(data, np.uint32, 4 * M, off).copy()
    off += 16 * M
    dens = _DENS_LUT[np.frombuffer(data, np.uint8, M, off)]
    off += M
    off = (off + 3) & ~3
    nc = (deg + 1) ** 2
    dims = 3 * nc
    ek = min(k, dims)
    mean, off = _f16(data, off, dims)
    basis, off = _f16(data, off, ek * dims)
    weights, off = _f16(data, off, M * ek)
    off = (off + 3) & ~3
    grads = np.zeros(3 * M, np.float32)
    tagged_here = off + 4 <= len(data) and int(np.frombuffer(data, np.uint32, 1, off)[0]) == TAGGED_MAGIC
    if not tagged_here:
        n = min(3 * M, (len(data) - off) // 2)
        grads[:n], off = _f16(data, off, n)
    W = weights.reshape(M, ek)
    B = basis.reshape(ek, dims)
    sh = np.broadcast_to(mean, (M, dims)).copy()
    for q in range(ek):
        sh = sh + W[:, q:q + 1] * B[q][None, :]
    off = (off + 3) & ~3
    n_tagged = 0
    if off + 8 <= len(data) and int(np.frombuffer(data, np.uint32, 1, off)[0]) == TAGGED_MAGIC:
        n_tagged = int(np.frombuffer(data, np.uint32, 1, off + 4)[0])
    scene = {"vertices": verts, "indices": idx, "densities": dens, "color_grads": grads,
             "circumdata": circumspheres(verts, idx), "start_pose": pose, "vertex_count": N, "tet_count": M}
    return scene, {"coeffs": sh.astype(np.float32).ravel(), "degree": deg}, n_tagged


def parse_ply(data):
    """(scene, sh) of a binary little-endian PLY image."""
    data = bytes(data)
    end = data.index(b"end_header\n") + len(b"end_header\n")
    N = M = num_sh = 0
    in_tet = False
    for line in data[:end].decode("ascii").splitlines():
        line = line.strip()
        if line.startswith("element vertex "):
            N, in_tet = int(line[15:]), False
        elif line.startswith("element tetrahedron "):
            M, in_tet = int(line[20:]), True
        elif line.startswith("element "):
            in_tet = False
        elif in_tet and line.startswith("property float sh_") and line.endswith("_r"):
            num_sh += 1
    nc = num_sh
    deg = int(np.sqrt(nc)) - 1
    verts = np.frombuffer(data, np.float32, 3 * N, end).copy()
    rec = np.dtype([("n", "u1"), ("idx", "<i4", 4), ("s", "<f4"), ("g", "<f4", 3), ("sh", "<f4", 3 * nc)])
    t = np.frombuffer(data, rec, M, end + 12 * N)
    sh = t["sh"].reshape(M, nc, 3).transpose(0, 2, 1).reshape(-1)
    idx = t["idx"].astype(np.uint32).ravel()
    scene = {"vertices": verts, "indices": idx, "densities": t["s"].copy(), "color_grads": t["g"].ravel().copy(),
             "circumdata": circumspheres(verts, idx), "start_pose": np.zeros(7, np.float32), "vertex_count": N,
             "tet_count": M}
    return scene, {"coeffs": np.ascontiguousarray(sh, np.float32), "degree": deg}
