"""CPU: scene files (crates/rmesh-data/src/lib.rs) through the native parser.

The native parser (rmesh_scene_* in librmesh_hip.so, host code) is checked
bit for bit against the numpy restatement in oracle/scene_io.py, which is
itself pinned here: the byte layout of the reference's writer (convert.py)
field by field at fixed offsets, round trips of known arrays, numpy's own
float16 decode, a float64 reconstruction of the PCA, and the circumspheres'
known answers.  Truncated and inconsistent files must raise, never crash.
"""

from __future__ import annotations

import gzip

import numpy as np
import pytest

from oracle import oracle as orc
from oracle import scene_io as ref
from rmeshvk_amd import data as io
from rmeshvk_amd.scenes import make_scene, regular_tet_scene, uniform_points


def _scene(seed=0, n=400, deg=3):
    rng = np.random.default_rng(seed)
    s = make_scene(uniform_points(n, rng), rng, sh_degree=deg)
    M = len(s["indices"]) // 4
    return s, s["sh"].reshape(M, 3, (deg + 1) ** 2)


def _same(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return a.shape == b.shape and np.array_equal(a.view(np.uint32) if a.dtype == np.float32 else a,
                                                 b.view(np.uint32) if b.dtype == np.float32 else b)


def _check_native_vs_ref(got, exp):
    (gs, gsh, _), (es, esh) = got, exp[:2]
    for k in ("vertices", "indices", "densities", "color_grads", "circumdata", "start_pose"):
        assert _same(gs[k], es[k]), k
    assert gs["vertex_count"] == es["vertex_count"] and gs["tet_count"] == es["tet_count"]
    assert gsh["degree"] == esh["degree"]
    assert _same(gsh["coeffs"], esh["coeffs"])


@pytest.mark.parametrize("deg,k", [(3, 12), (3, 48), (1, 5), (0, 3), (2, 40)])
@pytest.mark.parametrize("compress", [True, False])
def test_rmesh_round_trip_matches_restatement(deg, k, compress):
    s, sh = _scene(1, deg=deg)
    pose = np.arange(8, dtype=np.float32) * 0.5
    blob = io.save_rmesh(None, s["vertices"], s["indices"], s["densities"], s["color_grads"], sh, deg,
                         k_components=k, start_pose=pose, compress=compress)
    got = io.load_rmesh(blob)
    exp = ref.parse_rmesh(blob)
    _check_native_vs_ref(got, exp)
    gs, gsh, tagged = got
    assert tagged is None
    M = len(s["indices"]) // 4
    # the stored fields, from the source arrays
    assert np.array_equal(gs["vertices"], s["vertices"]) and np.array_equal(gs["indices"], s["indices"])
    assert np.array_equal(gs["start_pose"], pose[:7])
    assert np.array_equal(gs["color_grads"], s["color_grads"].astype(np.float16).astype(np.float32))
    codes = io.encode_densities(s["densities"]).astype(np.float64)
    assert np.allclose(gs["densities"], np.exp((codes - 100.0) / 20.0), rtol=1e-6)
    # PCA: the float64 reconstruction of the stored f16 mean / basis / weights
    ek = min(k, 3 * (deg + 1) ** 2)
    raw = gzip.decompress(blob) if compress else blob
    nc3 = 3 * (deg + 1) ** 2
    off = (48 + 12 * len(s["vertices"]) // 3 + 16 * M + M + 3) & ~3
    mean = np.frombuffer(raw, np.float16, nc3, off).astype(np.float64)
    basis = np.frombuffer(raw, np.float16, ek * nc3, off + 2 * nc3).astype(np.float64).reshape(ek, nc3)
    w = np.frombuffer(raw, np.float16, M * ek, off + 2 * nc3 + 2 * ek * nc3).astype(np.float64).reshape(M, ek)
    recon = mean + w @ basis
    assert np.abs(gsh["coeffs"].reshape(M, nc3) - recon).max() < 1e-5 * max(1.0, np.abs(recon).max())
    if ek == nc3:  # full rank: the coefficients come back up to f16 quantisation
        assert np.abs(gsh["coeffs"] - sh.ravel()).max() < 0.02


def test_rmesh_layout_offsets():
    """The convert.py byte layout, read at fixed offsets (convert.py:240-262)."""
    s, sh = _scene(2, n=60, deg=1)
    M, N = len(s["indices"]) // 4, len(s["vertices"]) // 3
    raw = io.save_rmesh(None, s["vertices"], s["indices"], s["densities"], s["color_grads"], sh, 1,
                        k_components=7, compress=False)
    assert list(np.frombuffer(raw, np.uint32, 4, 0)) == [N, M, 1, 7]
    assert np.array_equal(np.frombuffer(raw, np.float32, 3 * N, 48), s["vertices"])
    assert np.array_equal(np.frombuffer(raw, np.uint32, 4 * M, 48 + 12 * N), s["indices"])
    d_off = 48 + 12 * N + 16 * M
    assert np.array_equal(np.frombuffer(raw, np.uint8, M, d_off), io.encode_densities(s["densities"]))
    sh_off = (d_off + M + 3) & ~3
    g_off = (sh_off + 2 * (12 + 7 * 12 + M * 7) + 3) & ~3
    assert len(raw) == g_off + 6 * M
    assert np.array_equal(np.frombuffer(raw, np.float16, 3 * M, g_off), s["color_grads"].astype(np.float16))


def test_density_codes():
    """u8 codes: exp((v - 100) / 20) for every code (lib.rs:287-293)."""
    s = regular_tet_scene(1.0)
    for v in (0, 1, 2, 99, 100, 101, 200, 255):
        dens = np.array([np.exp((v - 99.5) / 20.0)], np.float32) if v else np.array([1e-4], np.float32)
        blob = io.save_rmesh(None, s["vertices"], s["indices"], dens, np.zeros(3, np.float32),
                             np.zeros((1, 3, 1), np.float32), 0, k_components=1, compress=False)
        got = io.load_rmesh(blob)[0]["densities"][0]
        code = int(io.encode_densities(dens)[0])
        assert code == v
        assert got == ref._DENS_LUT[v]
        # float32 (v - 100) / 20 carries up to |x| ulp/2 into the exponent
        assert abs(got - np.exp((v - 100) / 20.0)) <= 1e-6 * np.exp((v - 100) / 20.0)


def test_tagged_sections_and_truncated_gradients():
    s, sh = _scene(3, n=80, deg=0)
    M = len(s["indices"]) // 4
    tag = [("roughness", np.full(M, 0.5), 0), ("tone_curve", np.arange(6.0), 1)]
    blob = io.save_rmesh(None, s["vertices"], s["indices"], s["densities"], s["color_grads"], sh, 0,
                         k_components=3, tagged=tag, compress=False)
    got = io.load_rmesh(blob)
    assert got[2] == {"sections": 2}
    _check_native_vs_ref(got, ref.parse_rmesh(blob))
    # tagged sections where the gradients would start: gradients are zero
    raw = io.save_rmesh(None, s["vertices"], s["indices"], s["densities"], s["color_grads"], sh, 0,
                        k_components=3, compress=False)
    no_grads = raw[:len(raw) - 6 * M]
    tagged = no_grads + np.array([ref.TAGGED_MAGIC, 0], np.uint32).tobytes()
    got = io.load_rmesh(tagged)
    assert not np.any(got[0]["color_grads"]) and got[2] is None
    _check_native_vs_ref(got, ref.parse_rmesh(tagged))
    # truncated gradients: the ones present, then zeros
    part = raw[:len(raw) - 6 * M + 10]
    got = io.load_rmesh(part)
    g = got[0]["color_grads"]
    assert np.array_equal(g[:5], s["color_grads"][:5].astype(np.float16).astype(np.float32)) and not np.any(g[5:])
    _check_native_vs_ref(got, ref.parse_rmesh(part))


def test_rmesh_errors():
    s, sh = _scene(4, n=60, deg=0)
    raw = io.save_rmesh(None, s["vertices"], s["indices"], s["densities"], s["color_grads"], sh, 0,
                        k_components=3, compress=False)
    M, N = len(s["indices"]) // 4, len(s["vertices"]) // 3
    for cut in (0, 10, 47, 48 + 12 * N, 48 + 12 * N + 16 * M + 3):
        with pytest.raises(ValueError):
            io.load_rmesh(raw[:cut])
    bad = bytearray(raw)
    bad[48 + 12 * N:48 + 12 * N + 4] = np.array([N + 5], np.uint32).tobytes()  # index out of range
    with pytest.raises(ValueError, match="references vertex"):
        io.load_rmesh(bytes(bad))
    bad = bytearray(raw)
    bad[8:12] = np.array([7], np.uint32).tobytes()
    with pytest.raises(ValueError, match="SH degree"):
        io.load_rmesh(bytes(bad))
    gz = gzip.compress(raw)
    with pytest.raises(ValueError):
        io.load_rmesh(gz[:len(gz) // 2])
    with pytest.raises(ValueError):
        io.load_rmesh_raw(gz)  # a gzip image is no raw .rmesh


@pytest.mark.parametrize("deg", [0, 1, 3])
def test_ply_round_trip(deg):
    s, sh = _scene(5, deg=deg)
    blob = io.save_ply(None, s["vertices"], s["indices"], s["densities"], s["color_grads"], sh, deg)
    got = io.load_ply(blob)
    _check_native_vs_ref(got, ref.parse_ply(blob) + (None,))
    gs, gsh, tagged = got
    assert tagged is None
    for k in ("vertices", "indices", "densities", "color_grads"):
        assert np.array_equal(gs[k], s[k]), k
    assert gsh["degree"] == deg and np.array_equal(gsh["coeffs"], sh.ravel())


def test_ply_errors():
    s, sh = _scene(6, n=60, deg=1)
    blob = io.save_ply(None, s["vertices"], s["indices"], s["densities"], s["color_grads"], sh, 1)
    with pytest.raises(ValueError, match="binary_little_endian"):
        io.load_ply(blob.replace(b"binary_little_endian", b"binary_big_endian___"))
    with pytest.raises(ValueError, match="perfect square"):
        io.load_ply(blob.replace(b"property float sh_3_r", b"property float sh_3_x"))
    with pytest.raises(ValueError, match="end_header"):
        io.load_ply(blob.replace(b"end_header", b"end_hexder"))
    with pytest.raises(ValueError, match="truncated"):
        io.load_ply(blob[:-7])
    # counts whose byte sizes would wrap 64 bits (12 * 2^62 + 12) are refused, not wrapped
    big = b"4611686018427387905"
    M = len(s["indices"]) // 4
    for old in (b"element vertex 60\n", b"element tetrahedron %d\n" % M):
        assert old in blob
        with pytest.raises(ValueError, match="2\\^32"):
            io.load_ply(blob.replace(old, old.rsplit(b" ", 1)[0] + b" " + big + b"\n"))


def test_circumspheres():
    # the regular tet's known answer (centre 0, r^2 = 0.75)
    s = regular_tet_scene(1.0)
    c = ref.circumspheres(s["vertices"], s["indices"])
    assert np.allclose(c[:3], 0.0, atol=1e-7) and abs(c[3] - 0.75) < 1e-6
    # a mesh: equidistance, and orc_circumspheres bit for bit on non-degenerate tets
    sc, _ = _scene(7, n=300, deg=0)
    V = sc["vertices"].reshape(-1, 3).astype(np.float64)
    I = sc["indices"].reshape(-1, 4)
    c = ref.circumspheres(sc["vertices"], sc["indices"]).reshape(-1, 4)
    d2 = ((V[I] - c[:, None, :3].astype(np.float64)) ** 2).sum(-1)
    assert np.allclose(d2, c[:, 3:4], rtol=1e-4)
    assert _same(c.ravel(), orc.circumspheres(sc["vertices"], sc["indices"]))
    # degenerate tets: flat (r^2 = 1e20) and tiny (smallest centroid distance^2), as the native parser
    verts = np.array([0, 0, 0, 1, 0, 0, 0, 1, 0, 1, 1, 0,  # flat
                      5, 5, 5, 5 + 2e-8, 5, 5, 5, 5 + 2e-8, 5, 5, 5, 5 + 2e-8], np.float32)
    idx = np.array([0, 1, 2, 3, 4, 5, 6, 7], np.uint32)
    c = ref.circumspheres(verts, idx).reshape(-1, 4)
    assert c[0, 3] == np.float32(1e20) and np.allclose(c[0, :3], [0.5, 0.5, 0.0])
    assert c[1, 3] < 1e-14
    blob = io.save_ply(None, verts, idx, np.ones(2, np.float32), np.zeros(6, np.float32),
                       np.zeros((2, 3, 1), np.float32), 0)
    assert _same(io.load_ply(blob)[0]["circumdata"], c.ravel())


def test_renderer_from_scene_arrays():
    """A loaded scene carries what RMeshRenderer's constructor takes."""
    s, sh = _scene(8, n=80, deg=2)
    blob = io.save_rmesh(None, s["vertices"], s["indices"], s["densities"], s["color_grads"], sh, 2,
                         k_components=27, compress=True)
    scene, shd, _ = io.load_rmesh(blob)
    M = scene["tet_count"]
    assert scene["circumdata"].size == 4 * M and shd["coeffs"].size == 27 * M
    assert scene["indices"].dtype == np.uint32 and scene["vertices"].dtype == np.float32
