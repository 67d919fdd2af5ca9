"""ctypes wrapper of the CPU oracle (liborc.so) — TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
import this module; the product (rmeshvk_amd, rmesh_wgpu) never does.

What pins each oracle function (checks in tests/test_oracle_kats.py,
tests/test_oracle_render.py, tests/test_oracle_grad.py; fixtures in
tests/golden/reference_kats.json):

================  ==========================================================
orc_project       regular-tet known answers (kernel_tests.rs:1547-1671:
                  projection round trip, visible count 1, tiles > 0);
                  brute-force coverage (every pixel whose ray has a
                  non-grazing chord through a tet lies in a binned tile).
orc_pairs         sum(tiles_touched) == P; sentinel fill of the padding
                  (kernel_tests.rs:428-545); keys decode to (tile,
                  ~bits(float32 |c-cam|^2 - r^2)), the depth agreeing with
                  float64.
orc_sort          numpy ``argsort(kind="stable")`` of the masked key (random
                  bits above sorting_bits, heavy ties); the reference's
                  64-key check (kernel_tests.rs:96-167).
orc_sorting_bits  the table of drs.rs:23-26.
orc_ranges        closed form from sorted tile ids; the sentinel behaviour
                  of tile_ranges_compute.wgsl:21-36 on a planted case.
orc_forward       an independent float64 quadrature renderer
                  (tests/npref.py) on the regular tet, two tets and a
                  280-tet Delaunay mesh; the regular-tet closed form
                  RGB = c * alpha; tile-size invariance; min_t; aux normals
                  and depth of a single tet; the aux entropy channels
                  (aux[4..8]) of two overlapping tets against float64
                  chords in depth order.
orc_loss          KAT of kernel_tests.rs:176-305 (+-1/4096 at 64x64); the L2
                  closed form.
orc_backward      float64 central finite differences (eps 1e-6) of the
                  float64 forward on four scenes: relative L2 < 1e-6.
orc_adam          KAT of kernel_tests.rs:313-420 (m = 0.1 g, v = 0.001 g^2,
                  update ~ -lr sign(g)); constant-gradient closed form; NaN.
orc_sh_colors     basis == scipy real spherical harmonics (Condon-Shortley
                  phase), hence orthonormal; degree-0 closed form.
orc_sh_backward   float64 finite differences, degrees 0-3.
orc_circumspheres the regular-tet known answer; equidistance of the vertices.
orc_interval_forward
                  with near/far 0.01/1000 in the camera (the values the
                  reference linearizes with), the image of single tets and
                  two tets equals the float64 quadrature renderer (NDC z is
                  affine over a planar face, so the linearized chord is the
                  ray chord), depth x ray scale equals the quadrature's
                  alpha-weighted chord depth; distortion moments of one
                  tet against float64 quadrature of the weight
                  sigma e^-sigma(t - t0) (m0, m1, self term) and of two tets
                  (the cross term).  All with constant colour per tet.
                  g != 0 (screen-space interpolation of the colour offsets):
                  in the affine limit (a 0.1 tet at 100 units, float64) the
                  image equals the quadrature of the linear colour field
                  to < 5e-4 while g moves it by 3e-2.
orc_interval_backward (image terms), orc_interval_composite
                  float64 central differences of the compositing over frozen
                  decompositions (slot front/back NDC z, colour offsets, base
                  colour, density of 24 tets): relative L2 < 1e-5.
                  PARITY UNPINNED: the depth (xyzd) and distortion terms of
                  the adjoint (backward_interval_tiled.wgsl:544-650), which
                  the reference does not form as derivatives of its forward;
                  only GPU vs this restatement.
orc_interval_chain, orc_interval_attrs, orc_interval_generate
                  float64 central differences of the attributes of moved
                  vertices under frozen choices and weights, with respect to
                  every vertex, colour gradient, base colour and vertex
                  normal: relative L2 < 1e-6 (3- and 4-point silhouettes).
orc_error         closed forms on one tet (weight = sum of the pixels' optical
                  depths, which agree with the forward's alpha = 1 - e^-od;
                  weighted entry/exit points on the tet's planes; centroid
                  moment / weight = centroid); pe = 1 gives weighted L1 ==
                  weighted L1^2 == weight; pixel count and max weight by
                  brute force; a front tet shields a back one exactly as the
                  -5.5 log-transmittance rule and exp(log T) say.
================  ==========================================================
"""

from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_f32p = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")
_u64p = np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_u32 = ctypes.c_uint32
_u64 = ctypes.c_uint64
_f32 = ctypes.c_float
_vp = ctypes.c_void_p


def build() -> str:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return os.path.join(_HERE, "liborc.so")


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liborc.so")
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        for sfx, rp in (("_f32", _f32p), ("_f64", _f64p)):
            getattr(L, "orc_project" + sfx).argtypes = [rp, _u32p, rp, _u32, _f32p, _f32p, _f32p,
                                                         _u32, _u32, _u32, _u32p, _u8p]
            f = getattr(L, "orc_pairs" + sfx)
            f.argtypes = [rp, _u32p, rp, _u32, _f32p, _f32p, _f32p, _u32, _u32, _u32, _u32p, _u64p, _u32p, _u64]
            f.restype = _u64
            getattr(L, "orc_forward" + sfx).argtypes = [rp, _u32p, rp, rp, rp, _u32p, _u32p, _f32p, _f32p, _f32p,
                                                         _u32, _u32, _u32, _f32, _u32, _u32, _u32, rp, _vp, _vp,
                                                         _vp, _vp, _vp]
            f = getattr(L, "orc_loss" + sfx)
            f.argtypes = [rp, rp, _u32, _u32, _u32, rp]
            f.restype = ctypes.c_double
            getattr(L, "orc_backward" + sfx).argtypes = [rp, _u32p, rp, rp, rp, _u32p, _u32p, _f32p, _f32p, _f32p,
                                                          _u32, _u32, _u32, rp, rp, _u32, _u32, _u32, rp, rp, rp, rp,
                                                          _vp, _vp, _vp]
            getattr(L, "orc_sh_colors" + sfx).argtypes = [rp, _u32p, rp, rp, _u32, _u32, _f32p, rp]
            getattr(L, "orc_sh_backward" + sfx).argtypes = [rp, _u32p, rp, rp, _u32, _u32, _f32p, rp, rp, rp, rp]
            getattr(L, "orc_error" + sfx).argtypes = [rp, _u32p, rp, _u32p, _u32p, _f32p, _f32p, _f32p, _u32, _u32,
                                                       _u32, _f32, rp, rp, rp, _i32p]
            getattr(L, "orc_interval_forward" + sfx).argtypes = [rp, _u32p, rp, rp, rp, _u32p, _u32p, _f32p, _f32p,
                                                                  _f32p, _u32, _u32, _u32, _u32, rp, rp, rp]
            getattr(L, "orc_interval_generate" + sfx).argtypes = [rp, _u32p, rp, rp, rp, _vp, _f32p, _f32p, _f32p,
                                                                   _u32, _u32, _u32, rp]
            getattr(L, "orc_interval_attrs" + sfx).argtypes = [rp, _u32p, rp, rp, _vp, _f32p, _f32p, _f32p, _u32,
                                                                _u32, _u32, rp, rp]
            getattr(L, "orc_interval_composite" + sfx).argtypes = [rp, _u32p, _u32p, _f32p, _f32p, _f32p, _u32, _u32,
                                                                    _u32, _u32, rp, rp, rp]
            getattr(L, "orc_interval_backward" + sfx).argtypes = [rp, _u32p, _u32p, _f32p, _f32p, _f32p, _u32, _u32,
                                                                   _u32, _u32, rp, rp, rp, _vp, _vp, rp]
            getattr(L, "orc_interval_chain" + sfx).argtypes = [rp, _u32p, rp, _f32p, _f32p, _f32p, _u32, _u32, _u32,
                                                                rp, rp, rp, rp, rp, rp, rp]
        L.orc_set_threads.argtypes = [ctypes.c_int]
        L.orc_get_threads.restype = ctypes.c_int
        L.orc_sorting_bits.argtypes = [_u32]
        L.orc_sorting_bits.restype = _u32
        L.orc_sort.argtypes = [_u64p, _u32p, _u64, _u32]
        L.orc_ranges.argtypes = [_u64p, _u64, _u32p]
        L.orc_adam.argtypes = [_f32p, _f32p, _f32p, _f32p, _u64, _u32, _f32, _f32, _f32, _f32]
        L.orc_circumspheres.argtypes = [_f32p, _u32p, _u64, _f32p]
        _LIB = L
    return _LIB


def set_threads(n: int) -> None:
    """Host threads of the tile-parallel loops (0: all cores)."""
    lib().orc_set_threads(int(n))


def get_threads() -> int:
    return int(lib().orc_get_threads())


def _sfx(dtype):
    return "_f64" if np.dtype(dtype) == np.float64 else "_f32"


def _a(x, dtype):
    return np.ascontiguousarray(x, dtype=dtype)


def tile_grid(w, h, ts):
    tx = (w + ts - 1) // ts
    ty = (h + ts - 1) // ts
    return tx, ty, tx * ty


def sorting_bits(num_tiles: int) -> int:
    return int(lib().orc_sorting_bits(num_tiles))


def project(scene, cam, w, h, ts=12, dtype=np.float32):
    """K1: (tiles_touched[M], visible[M]) in tet-id order."""
    cp, vp, ci = cam
    M = len(scene["indices"]) // 4
    tt = np.zeros(M, dtype=np.uint32)
    vis = np.zeros(M, dtype=np.uint8)
    getattr(lib(), "orc_project" + _sfx(dtype))(
        _a(scene["vertices"], dtype), _a(scene["indices"], np.uint32), _a(scene["densities"], dtype), M,
        _a(cp, np.float32), _a(vp, np.float32), _a(ci, np.float32), w, h, ts, tt, vis)
    return tt, vis


def pairs(scene, cam, w, h, tiles_touched, ts=12, dtype=np.float32, cap=None):
    """K6+K7: unsorted (keys u64, values u32, P)."""
    cp, vp, ci = cam
    M = len(scene["indices"]) // 4
    P = int(tiles_touched.sum(dtype=np.uint64))
    cap = P if cap is None else cap
    keys = np.zeros(max(cap, 1), dtype=np.uint64)
    vals = np.zeros(max(cap, 1), dtype=np.uint32)
    n = getattr(lib(), "orc_pairs" + _sfx(dtype))(
        _a(scene["vertices"], dtype), _a(scene["indices"], np.uint32), _a(scene["circumdata"], dtype), M,
        _a(cp, np.float32), _a(vp, np.float32), _a(ci, np.float32), w, h, ts,
        _a(tiles_touched, np.uint32), keys, vals, cap)
    return keys[:min(n, cap)], vals[:min(n, cap)], int(n)


def sort(keys, vals, bits):
    k = np.array(keys, dtype=np.uint64, copy=True)
    v = np.array(vals, dtype=np.uint32, copy=True)
    lib().orc_sort(k, v, len(k), bits)
    return k, v


def ranges(sorted_keys, num_tiles):
    r = np.zeros(2 * num_tiles, dtype=np.uint32)
    lib().orc_ranges(_a(sorted_keys, np.uint64), len(sorted_keys), r)
    return r


def checkpoint_slots(n_pairs, num_tiles):
    """Slots of the checkpoint buffer (each TS*TS values) for n_pairs pairs."""
    return (int(n_pairs) >> 5) + int(num_tiles) + 1


def forward(scene, cam, w, h, sorted_vals, rng_, ts=12, min_t=0.0, colors=None, dtype=np.float32,
            tiles=None, want_aux=False, want_stats=False, tile_step=1, want_checkpoints=False):
    """Tiled forward.  want_checkpoints: also the per-pixel transmittance
    checkpoints of the checkpointed adjoint ("ck" [slots][ts*ts], "fin_logt"
    [H, W], "fin_stop" [H, W])."""
    cp, vp, ci = cam
    _, _, T = tile_grid(w, h, ts)
    t0, t1 = (0, T) if tiles is None else tiles
    rgba = np.zeros(h * w * 4, dtype=dtype)
    aux = np.zeros(h * w * 8, dtype=dtype) if want_aux else None
    stats = np.zeros(h * w * 4, dtype=np.uint32) if want_stats else None
    ck = fin = stop = None
    if want_checkpoints:
        ck = np.zeros(checkpoint_slots(len(sorted_vals), T) * ts * ts, dtype=dtype)
        fin = np.zeros(h * w, dtype=dtype)
        stop = np.zeros(h * w, dtype=np.uint32)
    col = scene["base_colors"] if colors is None else colors
    getattr(lib(), "orc_forward" + _sfx(dtype))(
        _a(scene["vertices"], dtype), _a(scene["indices"], np.uint32), _a(col, dtype),
        _a(scene["densities"], dtype), _a(scene["color_grads"], dtype), _a(sorted_vals, np.uint32),
        _a(rng_, np.uint32), _a(cp, np.float32), _a(vp, np.float32), _a(ci, np.float32),
        w, h, ts, min_t, t0, t1, tile_step, rgba,
        None if aux is None else aux.ctypes.data, None if stats is None else stats.ctypes.data,
        None if ck is None else ck.ctypes.data, None if fin is None else fin.ctypes.data,
        None if stop is None else stop.ctypes.data)
    out = {"rgba": rgba.reshape(h, w, 4)}
    if want_aux:
        out["aux"] = aux.reshape(h, w, 8)
    if want_stats:
        out["stats"] = stats.reshape(h, w, 4)
    if want_checkpoints:
        out.update(ck=ck, fin_logt=fin.reshape(h, w), fin_stop=stop.reshape(h, w))
    return out


def loss_grad(rgba, gt, loss_type, dtype=np.float32):
    h, w = rgba.shape[:2]
    dl = np.zeros(h * w * 4, dtype=dtype)
    loss = getattr(lib(), "orc_loss" + _sfx(dtype))(_a(rgba, dtype).ravel(), _a(gt, dtype).ravel(),
                                                     w, h, loss_type, dl)
    return dl.reshape(h, w, 4), float(loss)


def backward(scene, cam, w, h, sorted_vals, rng_, rgba, dl, ts=12, colors=None, dtype=np.float32, tiles=None,
             tile_step=1, checkpoints=None):
    """Tiled adjoint.  checkpoints=None: the reference's undo from the image
    alpha; otherwise the forward's output dict (with want_checkpoints=True):
    the checkpointed adjoint."""
    cp, vp, ci = cam
    _, _, T = tile_grid(w, h, ts)
    t0, t1 = (0, T) if tiles is None else tiles
    N = len(scene["vertices"]) // 3
    M = len(scene["indices"]) // 4
    dV = np.zeros(N * 3, dtype=dtype)
    dD = np.zeros(M, dtype=dtype)
    dG = np.zeros(M * 3, dtype=dtype)
    dC = np.zeros(M * 3, dtype=dtype)
    col = scene["base_colors"] if colors is None else colors
    keep = []
    ckp = [None, None, None]
    if checkpoints is not None:
        keep = [_a(checkpoints["ck"], dtype), _a(checkpoints["fin_logt"], dtype).ravel(),
                _a(checkpoints["fin_stop"], np.uint32).ravel()]
        ckp = [k.ctypes.data for k in keep]
    getattr(lib(), "orc_backward" + _sfx(dtype))(
        _a(scene["vertices"], dtype), _a(scene["indices"], np.uint32), _a(col, dtype),
        _a(scene["densities"], dtype), _a(scene["color_grads"], dtype), _a(sorted_vals, np.uint32),
        _a(rng_, np.uint32), _a(cp, np.float32), _a(vp, np.float32), _a(ci, np.float32),
        w, h, ts, _a(rgba, dtype).ravel(), _a(dl, dtype).ravel(), t0, t1, tile_step, dV, dD, dG, dC, *ckp)
    return {"d_vertices": dV, "d_densities": dD, "d_color_grads": dG, "d_base_colors": dC}


IVL_NF = 80  # floats per tet decomposition (flat layout of oracle_impl.h)


def _ptr(a):
    return None if a is None else a.ctypes.data


def interval_forward(scene, cam, w, h, sorted_vals, rng_, ts=12, colors=None, dtype=np.float32):
    """Interval-tiled forward over the regular binning's tile lists:
    {"image" [h, w, 4], "xyzd" [h, w, 4], "distortion" [h, w, 5]}."""
    cp, vp, ci = cam
    M = len(scene["indices"]) // 4
    img = np.zeros(h * w * 4, dtype=dtype)
    xyzd = np.zeros(h * w * 4, dtype=dtype)
    dist = np.zeros(h * w * 5, dtype=dtype)
    col = scene["base_colors"] if colors is None else colors
    getattr(lib(), "orc_interval_forward" + _sfx(dtype))(
        _a(scene["vertices"], dtype), _a(scene["indices"], np.uint32), _a(col, dtype),
        _a(scene["densities"], dtype), _a(scene["color_grads"], dtype), _a(sorted_vals, np.uint32),
        _a(rng_, np.uint32), _a(cp, np.float32), _a(vp, np.float32), _a(ci, np.float32), w, h, ts, M, img, xyzd,
        dist)
    return {"image": img.reshape(h, w, 4), "xyzd": xyzd.reshape(h, w, 4), "distortion": dist.reshape(h, w, 5)}


def interval_generate(scene, cam, w, h, colors=None, vertex_normals=None, dtype=np.float32):
    """Per-tet silhouette decompositions, [M, IVL_NF] (flat layout)."""
    cp, vp, ci = cam
    M = len(scene["indices"]) // 4
    out = np.zeros(M * IVL_NF, dtype=dtype)
    col = scene["base_colors"] if colors is None else colors
    vn = None if vertex_normals is None else _a(vertex_normals, dtype)
    getattr(lib(), "orc_interval_generate" + _sfx(dtype))(
        _a(scene["vertices"], dtype), _a(scene["indices"], np.uint32), _a(col, dtype),
        _a(scene["densities"], dtype), _a(scene["color_grads"], dtype), _ptr(vn), _a(cp, np.float32),
        _a(vp, np.float32), _a(ci, np.float32), w, h, M, out)
    return out.reshape(M, IVL_NF)


def interval_attrs(scene, cam, w, h, frozen, colors=None, vertex_normals=None, dtype=np.float32):
    """Attributes of the scene's (moved) vertices under frozen discrete
    choices and weights: [M, IVL_NF]."""
    cp, vp, ci = cam
    M = len(scene["indices"]) // 4
    out = np.zeros(M * IVL_NF, dtype=dtype)
    col = scene["base_colors"] if colors is None else colors
    vn = None if vertex_normals is None else _a(vertex_normals, dtype)
    getattr(lib(), "orc_interval_attrs" + _sfx(dtype))(
        _a(scene["vertices"], dtype), _a(scene["indices"], np.uint32), _a(col, dtype),
        _a(scene["color_grads"], dtype), _ptr(vn), _a(cp, np.float32), _a(vp, np.float32), _a(ci, np.float32),
        w, h, M, _a(frozen, dtype).ravel(), out)
    return out.reshape(M, IVL_NF)


def interval_composite(ivl, cam, w, h, sorted_vals, rng_, ts=12, dtype=np.float32):
    """Compositing over given decompositions [M, IVL_NF]."""
    cp, vp, ci = cam
    M = len(ivl)
    img = np.zeros(h * w * 4, dtype=dtype)
    xyzd = np.zeros(h * w * 4, dtype=dtype)
    dist = np.zeros(h * w * 5, dtype=dtype)
    getattr(lib(), "orc_interval_composite" + _sfx(dtype))(
        _a(ivl, dtype).ravel(), _a(sorted_vals, np.uint32), _a(rng_, np.uint32), _a(cp, np.float32),
        _a(vp, np.float32), _a(ci, np.float32), w, h, ts, M, img, xyzd, dist)
    return {"image": img.reshape(h, w, 4), "xyzd": xyzd.reshape(h, w, 4), "distortion": dist.reshape(h, w, 5)}


def interval_backward(ivl, cam, w, h, sorted_vals, rng_, fwd, dl_image, dl_xyzd=None, dl_distortion=None, ts=12,
                      dtype=np.float32):
    """Per-tile adjoint: d_ivl [M, IVL_NF] (slot attributes, normals, base
    colour, density) from the forward's image and distortion."""
    cp, vp, ci = cam
    M = len(ivl)
    d = np.zeros(M * IVL_NF, dtype=dtype)
    dx = None if dl_xyzd is None else _a(dl_xyzd, dtype).ravel()
    dd = None if dl_distortion is None else _a(dl_distortion, dtype).ravel()
    getattr(lib(), "orc_interval_backward" + _sfx(dtype))(
        _a(ivl, dtype).ravel(), _a(sorted_vals, np.uint32), _a(rng_, np.uint32), _a(cp, np.float32),
        _a(vp, np.float32), _a(ci, np.float32), w, h, ts, M, _a(fwd["image"], dtype).ravel(),
        _a(fwd["distortion"], dtype).ravel(), _a(dl_image, dtype).ravel(), _ptr(dx), _ptr(dd), d)
    return d.reshape(M, IVL_NF)


def interval_chain(scene, cam, w, h, ivl, d_ivl, dtype=np.float32):
    """Per-tet chain back of d_ivl to d_vertices, d_densities,
    d_base_colors, d_color_grads, d_vertex_normals."""
    cp, vp, ci = cam
    M = len(ivl)
    N = len(scene["vertices"]) // 3
    dV = np.zeros(N * 3, dtype=dtype)
    dD = np.zeros(M, dtype=dtype)
    dB = np.zeros(M * 3, dtype=dtype)
    dG = np.zeros(M * 3, dtype=dtype)
    dN = np.zeros(N * 3, dtype=dtype)
    getattr(lib(), "orc_interval_chain" + _sfx(dtype))(
        _a(scene["vertices"], dtype), _a(scene["indices"], np.uint32), _a(scene["color_grads"], dtype),
        _a(cp, np.float32), _a(vp, np.float32), _a(ci, np.float32), w, h, M, _a(ivl, dtype).ravel(),
        _a(d_ivl, dtype).ravel(), dV, dD, dB, dG, dN)
    return {"d_vertices": dV, "d_densities": dD, "d_base_colors": dB, "d_color_grads": dG,
            "d_vertex_normals": dN}


def interval_backward_full(scene, cam, w, h, sorted_vals, rng_, dl_image, dl_xyzd=None, dl_distortion=None, ts=12,
                           colors=None, dtype=np.float32):
    """backward_interval_tiled end to end (decomposition, forward, adjoint, chain back)."""
    ivl = interval_generate(scene, cam, w, h, colors=colors, dtype=dtype)
    fwd = interval_composite(ivl, cam, w, h, sorted_vals, rng_, ts=ts, dtype=dtype)
    d_ivl = interval_backward(ivl, cam, w, h, sorted_vals, rng_, fwd, dl_image, dl_xyzd, dl_distortion, ts=ts,
                              dtype=dtype)
    return interval_chain(scene, cam, w, h, ivl, d_ivl, dtype=dtype)


def error_stats(scene, cam, w, h, sorted_vals, rng_, pixel_err, ssim_err, ts=12, min_t=0.0, dtype=np.float32):
    """Per-tet error statistics of the last forward's tile lists:
    (tet_err [M, 16], tet_count [M, 2] int32)."""
    cp, vp, ci = cam
    M = len(scene["indices"]) // 4
    err = np.zeros(M * 16, dtype=dtype)
    cnt = np.zeros(M * 2, dtype=np.int32)
    getattr(lib(), "orc_error" + _sfx(dtype))(
        _a(scene["vertices"], dtype), _a(scene["indices"], np.uint32), _a(scene["densities"], dtype),
        _a(sorted_vals, np.uint32), _a(rng_, np.uint32), _a(cp, np.float32), _a(vp, np.float32),
        _a(ci, np.float32), w, h, ts, float(min_t), _a(pixel_err, dtype).ravel(), _a(ssim_err, dtype).ravel(),
        err, cnt)
    return err.reshape(M, 16), cnt.reshape(M, 2)


def adam(params, grads, m, v, step, lr, beta1=0.9, beta2=0.999, eps=1e-8):
    p = np.array(params, dtype=np.float32, copy=True)
    mm = np.array(m, dtype=np.float32, copy=True)
    vv = np.array(v, dtype=np.float32, copy=True)
    lib().orc_adam(p, _a(grads, np.float32), mm, vv, len(p), step, lr, beta1, beta2, eps)
    return p, mm, vv


def sh_colors(scene, cam_pos, dtype=np.float32):
    M = len(scene["indices"]) // 4
    out = np.zeros(M * 3, dtype=dtype)
    getattr(lib(), "orc_sh_colors" + _sfx(dtype))(
        _a(scene["vertices"], dtype), _a(scene["indices"], np.uint32), _a(scene["sh"], dtype),
        _a(scene["color_grads"], dtype), M, int(scene["sh_degree"]), _a(cam_pos, np.float32), out)
    return out


def sh_backward(scene, cam_pos, d_colors, dtype=np.float32):
    M = len(scene["indices"]) // 4
    N = len(scene["vertices"]) // 3
    d_sh = np.zeros(len(scene["sh"]), dtype=dtype)
    dG = np.zeros(M * 3, dtype=dtype)
    dV = np.zeros(N * 3, dtype=dtype)
    getattr(lib(), "orc_sh_backward" + _sfx(dtype))(
        _a(scene["vertices"], dtype), _a(scene["indices"], np.uint32), _a(scene["sh"], dtype),
        _a(scene["color_grads"], dtype), M, int(scene["sh_degree"]), _a(cam_pos, np.float32),
        _a(d_colors, dtype), d_sh, dG, dV)
    return {"d_sh": d_sh, "d_color_grads": dG, "d_vertices": dV}


def circumspheres(vertices, indices):
    M = len(indices) // 4
    out = np.zeros(M * 4, dtype=np.float32)
    lib().orc_circumspheres(_a(vertices, np.float32), _a(indices, np.uint32), M, out)
    return out


def render(scene, cam, w, h, ts=12, min_t=0.0, dtype=np.float32, colors=None, want_aux=False,
           want_stats=False):
    """Full forward pipeline K1 -> K7 -> sort -> K11 -> K12; returns every intermediate."""
    if colors is None and "sh" in scene:
        colors = sh_colors(scene, cam[0], dtype)
    tt, vis = project(scene, cam, w, h, ts, dtype)
    keys, vals, P = pairs(scene, cam, w, h, tt, ts, dtype)
    _, _, T = tile_grid(w, h, ts)
    sk, sv = sort(keys, vals, sorting_bits(T))
    rg = ranges(sk, T)
    out = forward(scene, cam, w, h, sv, rg, ts, min_t, colors, dtype, want_aux=want_aux, want_stats=want_stats)
    out.update(tiles_touched=tt, visible=vis, keys=keys, vals=vals, P=P, sorted_keys=sk,
               sorted_vals=sv, ranges=rg, colors=colors)
    return out


def render_backward(scene, cam, w, h, gt, loss_type=1, ts=12, dtype=np.float32):
    """Forward + loss + backward (+ SH adjoint when the scene carries SH)."""
    fw = render(scene, cam, w, h, ts, dtype=dtype)
    colors = fw["colors"]
    dl, loss = loss_grad(fw["rgba"], gt, loss_type, dtype)
    g = backward(scene, cam, w, h, fw["sorted_vals"], fw["ranges"], fw["rgba"], dl, ts, colors, dtype)
    if "sh" in scene:
        sg = sh_backward(scene, cam[0], g["d_base_colors"], dtype)
        g["d_sh"] = sg["d_sh"]
        g["d_color_grads"] = g["d_color_grads"] + sg["d_color_grads"]
        g["d_vertices"] = g["d_vertices"] + sg["d_vertices"]
    return fw, dl, loss, g
