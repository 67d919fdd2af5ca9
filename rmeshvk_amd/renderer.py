"""``RMeshRenderer`` — MI355X drop-in for rmeshvk's PyO3 class.

Mirrors ``rmesh_wgpu._native.RMeshRenderer`` (crates/rmesh-python/src/lib.rs:
138-2705): same constructor arguments, method names, argument meaning, return
types and error classes.  Every call goes through the C ABI of
librmesh_hip.so (include/rmesh_hip.h); nothing here computes on the CPU.

Extensions (keyword-only or new methods, positional signatures unchanged):
``tile_size``, ``device_id``, ``sh_coeffs``/``sh_degree`` (SH colour model),
``max_views`` with ``forward_batched`` / ``train_step_batched`` (many views
per launch, optional hipGraph replay), ``lr_sh``/``lr_base`` in
``train_step``, and read-only inspection of the binning state.
"""

from __future__ import annotations

import ctypes

import numpy as np

from . import _lib

_OUT_OF_SCOPE = (
    "forward_raytrace", "backward_raytrace",
    "locate_tets", "read_aux", "set_vertex_normals",
    "get_vertex_normals", "set_aux_data", "get_aux_data",
)


def _f32(name, a, n=None):
    if not isinstance(a, np.ndarray):
        raise TypeError(f"{name}: expected a numpy.ndarray of float32")
    if a.dtype != np.float32:
        raise TypeError(f"{name}: expected dtype float32, got {a.dtype}")
    if not a.flags.c_contiguous:
        raise ValueError(f"{name}: array is not contiguous")
    if n is not None and a.size != n:
        raise ValueError(f"{name}: expected {n} elements, got {a.size}")
    return a


def _u32(name, a, n=None):
    if not isinstance(a, np.ndarray):
        raise TypeError(f"{name}: expected a numpy.ndarray of uint32")
    if a.dtype != np.uint32:
        raise TypeError(f"{name}: expected dtype uint32, got {a.dtype}")
    if not a.flags.c_contiguous:
        raise ValueError(f"{name}: array is not contiguous")
    if n is not None and a.size != n:
        raise ValueError(f"{name}: expected {n} elements, got {a.size}")
    return a


def _p(a):
    return ctypes.c_void_p(a.ctypes.data) if a is not None else None


class RMeshRenderer:
    """Differentiable radiance-mesh renderer on one MI355X (see module doc)."""

    def __init__(self, vertices, indices, base_colors, densities, color_grads, circumdata, width, height,
                 aux_dim=0, *, tile_size=12, device_id=-1, sh_coeffs=None, sh_degree=0, max_views=1,
                 sh_half=False):
        lib = _lib.load()
        _f32("vertices", vertices)
        _u32("indices", indices)
        if vertices.size % 3 or indices.size % 4:
            raise ValueError("vertices must be [N*3] and indices [M*4]")
        n, m = vertices.size // 3, indices.size // 4
        _f32("base_colors", base_colors, 3 * m)
        _f32("densities", densities, m)
        _f32("color_grads", color_grads, 3 * m)
        _f32("circumdata", circumdata, 4 * m)
        if sh_coeffs is not None:
            _f32("sh_coeffs", sh_coeffs, 3 * m * (sh_degree + 1) ** 2)
        self._keep = (vertices, indices, base_colors, densities, color_grads, circumdata, sh_coeffs)
        d = _lib.SceneDesc(
            vertices=vertices.ctypes.data, n_vertices=n, indices=indices.ctypes.data, n_tets=m,
            base_colors=base_colors.ctypes.data, densities=densities.ctypes.data,
            color_grads=color_grads.ctypes.data, circumdata=circumdata.ctypes.data,
            width=int(width), height=int(height), aux_dim=int(aux_dim), tile_size=int(tile_size),
            device_id=int(device_id), sh_coeffs=None if sh_coeffs is None else sh_coeffs.ctypes.data,
            sh_degree=int(sh_degree), max_views=int(max_views), sh_half=int(bool(sh_half)))
        h = ctypes.c_void_p()
        _lib.check(lib.rmesh_create(ctypes.byref(d), ctypes.byref(h)))
        self._keep = None
        self._h = h
        self._lib = lib
        self._n, self._m = n, m
        self._w, self._hgt = int(width), int(height)
        self._sh_degree = int(sh_degree) if sh_coeffs is not None else None
        self._max_views = int(max_views)
        self.tile_size = int(tile_size)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            self._lib.rmesh_destroy(h)
            self._h = None

    # ---- getters (lib.rs:2689-2704) ----
    @property
    def step(self) -> int:
        return int(self._lib.rmesh_get_step(self._h))

    @property
    def width(self) -> int:
        return int(self._lib.rmesh_get_width(self._h))

    @property
    def height(self) -> int:
        return int(self._lib.rmesh_get_height(self._h))

    @property
    def num_tets(self) -> int:
        return self._m

    @property
    def num_vertices(self) -> int:
        return self._n

    def set_min_t(self, min_t):
        _lib.check(self._lib.rmesh_set_min_t(self._h, float(min_t)))

    def set_stream(self, stream_handle):
        """Run on an external HIP stream (an integer handle, e.g. torch's cuda_stream)."""
        _lib.check(self._lib.rmesh_set_stream(self._h, ctypes.c_void_p(stream_handle or 0)))

    # ---- forward (lib.rs:932-1204) ----
    def _cam(self, cam_pos, vp, c2w_intrinsics):
        return _f32("cam_pos", cam_pos, 3), _f32("vp", vp, 16), _f32("c2w_intrinsics", c2w_intrinsics, 16)

    def forward_tiled(self, cam_pos, vp, c2w_intrinsics, render_aux=False):
        c, v, ci = self._cam(cam_pos, vp, c2w_intrinsics)
        rgba = np.empty((self._hgt, self._w, 4), dtype=np.float32)
        aux = np.empty((self._hgt, self._w, 8), dtype=np.float32) if render_aux else None
        _lib.check(self._lib.rmesh_forward_tiled(self._h, _p(c), _p(v), _p(ci), _p(rgba), _p(aux)))
        return (rgba, aux) if render_aux else rgba

    def forward(self, cam_pos, vp, c2w_intrinsics):
        c, v, ci = self._cam(cam_pos, vp, c2w_intrinsics)
        rgba = np.empty((self._hgt, self._w, 4), dtype=np.float32)
        _lib.check(self._lib.rmesh_forward(self._h, _p(c), _p(v), _p(ci), _p(rgba)))
        return rgba

    def read_instance_count(self) -> int:
        out = ctypes.c_uint32()
        _lib.check(self._lib.rmesh_read_instance_count(self._h, ctypes.byref(out)))
        return int(out.value)

    # ---- forward_interval_tiled (lib.rs:1218-1390) ----
    def forward_interval_tiled(self, cam_pos, vp, c2w_intrinsics):
        """Interval-tiled forward: {'image' [H, W, 4], 'xyzd' [H, W, 4],
        'distortion' [H, W, 5]} (no 'aux': aux_dim is always 0 here)."""
        c, v, ci = self._cam(cam_pos, vp, c2w_intrinsics)
        rgba = np.empty((self._hgt, self._w, 4), dtype=np.float32)
        xyzd = np.empty((self._hgt, self._w, 4), dtype=np.float32)
        dist = np.empty((self._hgt, self._w, 5), dtype=np.float32)
        _lib.check(self._lib.rmesh_forward_interval_tiled(self._h, _p(c), _p(v), _p(ci), _p(rgba), _p(xyzd),
                                                          _p(dist)))
        return {"image": rgba, "xyzd": xyzd, "distortion": dist}

    # ---- backward_interval_tiled (lib.rs:2049-2240) ----
    def backward_interval_tiled(self, dl_d_image, dl_d_xyzd=None, dl_d_distortion=None, dl_d_aux=None):
        """Adjoint of the last forward_interval_tiled: {'d_vertices' [N*3],
        'd_densities' [M], 'd_base_colors' [M*3], 'd_color_grads' [M*3],
        'd_vertex_normals' [N*3]} (no 'd_aux_data': aux_dim is always 0 here)."""
        n = self._hgt * self._w
        dl = _f32("dl_d_image", dl_d_image, 4 * n)
        dx = None if dl_d_xyzd is None else _f32("dl_d_xyzd", dl_d_xyzd, 4 * n)
        dd = None if dl_d_distortion is None else _f32("dl_d_distortion", dl_d_distortion, 5 * n)
        out = {"d_vertices": np.empty(3 * self._n, np.float32), "d_densities": np.empty(self._m, np.float32),
               "d_base_colors": np.empty(3 * self._m, np.float32), "d_color_grads": np.empty(3 * self._m, np.float32),
               "d_vertex_normals": np.empty(3 * self._n, np.float32)}
        _lib.check(self._lib.rmesh_backward_interval_tiled(
            self._h, _p(dl), _p(dx), _p(dd), _p(out["d_vertices"]), _p(out["d_densities"]), _p(out["d_base_colors"]),
            _p(out["d_color_grads"]), _p(out["d_vertex_normals"])))
        return out

    # ---- forward_error (lib.rs:1621-1711) ----
    def forward_error(self, pixel_err, ssim_err):
        """Per-tet error statistics of the last forward_tiled (densification):
        returns (tet_err float32 [M, 16], tet_count int32 [M, 2])."""
        n = self._hgt * self._w
        for name, a in (("pixel_err", pixel_err), ("ssim_err", ssim_err)):
            if isinstance(a, np.ndarray) and a.dtype == np.float32 and a.size != n:
                raise ValueError(f"{name} must have {n} elements, got {a.size}")
        _f32("pixel_err", pixel_err, n)
        _f32("ssim_err", ssim_err, n)
        err = np.empty((self._m, 16), dtype=np.float32)
        cnt = np.empty((self._m, 2), dtype=np.int32)
        _lib.check(self._lib.rmesh_forward_error(self._h, _p(pixel_err), _p(ssim_err), _p(err), _p(cnt)))
        return err, cnt

    # ---- backward (lib.rs:1894-2031) ----
    def backward(self, dl_d_image):
        _f32("dl_d_image", dl_d_image, self._hgt * self._w * 4)
        d_v = np.empty(3 * self._n, dtype=np.float32)
        d_d = np.empty(self._m, dtype=np.float32)
        d_b = np.empty(3 * self._m, dtype=np.float32)
        d_g = np.empty(3 * self._m, dtype=np.float32)
        _lib.check(self._lib.rmesh_backward(self._h, _p(dl_d_image), _p(d_v), _p(d_d), _p(d_b), _p(d_g)))
        out = {"d_vertices": d_v, "d_densities": d_d, "d_base_colors": d_b, "d_color_grads": d_g}
        if self._sh_degree is not None:
            d_sh = np.empty(3 * self._m * (self._sh_degree + 1) ** 2, dtype=np.float32)
            _lib.check(self._lib.rmesh_get_sh_grad(self._h, _p(d_sh)))
            out["d_sh"] = d_sh
        return out

    # ---- device-pointer variants (torch tensors on the GPU) ----
    def update_params_device(self, vertices, base_colors, densities, color_grads):
        """Parameters from device pointers (integers, e.g. tensor.data_ptr()),
        copied on the renderer's stream."""
        _lib.check(self._lib.rmesh_update_params_device(self._h, ctypes.c_void_p(vertices), ctypes.c_void_p(base_colors),
                                                        ctypes.c_void_p(densities), ctypes.c_void_p(color_grads)))

    def forward_tiled_device(self, cam_pos, vp, c2w_intrinsics, rgba_ptr):
        """forward_tiled into a device buffer of H*W*4 floats."""
        c, v, ci = self._cam(cam_pos, vp, c2w_intrinsics)
        _lib.check(self._lib.rmesh_forward_tiled_device(self._h, _p(c), _p(v), _p(ci), ctypes.c_void_p(rgba_ptr)))

    def backward_device(self, dl_ptr, d_vertices, d_densities, d_base_colors, d_color_grads):
        """backward from/to device buffers (dl: H*W*4; outputs as backward's)."""
        _lib.check(self._lib.rmesh_backward_device(self._h, ctypes.c_void_p(dl_ptr), ctypes.c_void_p(d_vertices),
                                                   ctypes.c_void_p(d_densities), ctypes.c_void_p(d_base_colors),
                                                   ctypes.c_void_p(d_color_grads)))

    # ---- backward_hw (lib.rs:2255-2378) ----
    def backward_hw(self, dl_d_image):
        """Adjoint of the last forward() (served by the tiled forward, so the
        tiled adjoint over its state)."""
        _f32("dl_d_image", dl_d_image, self._hgt * self._w * 4)
        d_v = np.empty(3 * self._n, dtype=np.float32)
        d_d = np.empty(self._m, dtype=np.float32)
        d_b = np.empty(3 * self._m, dtype=np.float32)
        d_g = np.empty(3 * self._m, dtype=np.float32)
        _lib.check(self._lib.rmesh_backward_hw(self._h, _p(dl_d_image), _p(d_v), _p(d_d), _p(d_b), _p(d_g)))
        return {"d_vertices": d_v, "d_densities": d_d, "d_base_colors": d_b, "d_color_grads": d_g}

    # ---- training (lib.rs:2405-2636) ----
    def train_step(self, cam_pos, vp, c2w_intrinsics, gt_image, lr_verts, lr_dens, lr_grads, loss_type,
                   *, lr_sh=0.0, lr_base=0.0):
        c, v, ci = self._cam(cam_pos, vp, c2w_intrinsics)
        _f32("gt_image", gt_image, self._hgt * self._w * 3)
        _lib.check(self._lib.rmesh_set_extra_lrs(self._h, float(lr_sh), float(lr_base)))
        loss = ctypes.c_float()
        _lib.check(self._lib.rmesh_train_step(self._h, _p(c), _p(v), _p(ci), _p(gt_image), float(lr_verts),
                                              float(lr_dens), float(lr_grads), int(loss_type), ctypes.byref(loss)))
        return float(loss.value)

    # ---- parameters (lib.rs:2641-2686) ----
    def update_params(self, vertices, base_colors, densities, color_grads):
        _f32("vertices", vertices, 3 * self._n)
        _f32("base_colors", base_colors, 3 * self._m)
        _f32("densities", densities, self._m)
        _f32("color_grads", color_grads, 3 * self._m)
        _lib.check(self._lib.rmesh_update_params(self._h, _p(vertices), _p(base_colors), _p(densities),
                                                 _p(color_grads)))

    def get_params(self):
        v = np.empty(3 * self._n, dtype=np.float32)
        b = np.empty(3 * self._m, dtype=np.float32)
        d = np.empty(self._m, dtype=np.float32)
        g = np.empty(3 * self._m, dtype=np.float32)
        _lib.check(self._lib.rmesh_get_params(self._h, _p(v), _p(b), _p(d), _p(g)))
        out = {"vertices": v, "base_colors": b, "densities": d, "color_grads": g}
        if self._sh_degree is not None:
            out["sh_coeffs"] = self.get_sh()
        return out

    def update_sh(self, sh_coeffs):
        _f32("sh_coeffs", sh_coeffs, 3 * self._m * (self._sh_degree + 1) ** 2 if self._sh_degree is not None else None)
        _lib.check(self._lib.rmesh_update_sh(self._h, _p(sh_coeffs)))

    def get_sh(self):
        if self._sh_degree is None:
            raise ValueError("renderer has no SH coefficients")
        out = np.empty(3 * self._m * (self._sh_degree + 1) ** 2, dtype=np.float32)
        _lib.check(self._lib.rmesh_get_sh(self._h, _p(out)))
        return out

    # ---- batched extension ----
    @staticmethod
    def _stack_cams(cams):
        cp = np.ascontiguousarray(np.stack([c[0] for c in cams]).astype(np.float32))
        vp = np.ascontiguousarray(np.stack([c[1] for c in cams]).astype(np.float32))
        ci = np.ascontiguousarray(np.stack([c[2] for c in cams]).astype(np.float32))
        return cp, vp, ci

    def forward_batched(self, cams, copy_out=True):
        """Render a list of ``(cam_pos, vp, c2w_intrinsics)``; returns [V, H, W, 4] (or None)."""
        cp, vp, ci = self._stack_cams(cams)
        out = np.empty((len(cams), self._hgt, self._w, 4), dtype=np.float32) if copy_out else None
        _lib.check(self._lib.rmesh_forward_batched(self._h, len(cams), _p(cp), _p(vp), _p(ci), _p(out)))
        return out

    def train_step_batched(self, cams, loss_type, lr_verts, lr_dens, lr_grads, lr_sh=0.0, gt_device_ptr=None,
                           lr_base=0.0):
        """One step over all views.  Targets come from ``gt_device_ptr`` (device
        [V][H][W][3]) or, when None, from the renderer's own target buffer
        (``device_buffer(BUF_GT)`` / ``set_targets``).  ``lr_base`` is the
        base-colour group's learning rate (scenes without SH; 0 leaves the
        group out), set on every call."""
        cp, vp, ci = self._stack_cams(cams)
        _lib.check(self._lib.rmesh_set_extra_lrs(self._h, float(lr_sh), float(lr_base)))
        loss = ctypes.c_float()
        _lib.check(self._lib.rmesh_train_step_batched(
            self._h, len(cams), _p(cp), _p(vp), _p(ci), ctypes.c_void_p(gt_device_ptr or 0), int(loss_type),
            float(lr_verts), float(lr_dens), float(lr_grads), float(lr_sh), ctypes.byref(loss)))
        return float(loss.value)

    def set_adjoint_mode(self, mode: str):
        """"undo" (default): the reference's adjoint, log T rebuilt from the
        image alpha; "checkpoint": the forward stores log T every 32 tets per
        pixel and the adjoint restarts from it (the exact adjoint of the
        forward).  A backward needs a forward run in the same mode."""
        m = {"undo": _lib.ADJOINT_UNDO, "checkpoint": _lib.ADJOINT_CHECKPOINT}.get(mode)
        if m is None:
            raise ValueError("adjoint mode must be 'undo' or 'checkpoint'")
        _lib.check(self._lib.rmesh_set_adjoint_mode(self._h, m))

    def get_adjoint_mode(self) -> str:
        return "checkpoint" if self._lib.rmesh_get_adjoint_mode(self._h) == _lib.ADJOINT_CHECKPOINT else "undo"

    def device_bytes(self) -> int:
        """Device memory held by the renderer, in bytes."""
        n = ctypes.c_uint64()
        _lib.check(self._lib.rmesh_device_bytes(self._h, ctypes.byref(n)))
        return int(n.value)

    def set_graph_mode(self, enable: bool):
        _lib.check(self._lib.rmesh_set_graph_mode(self._h, 1 if enable else 0))

    # ---- tuning extension ----
    def set_tuning(self, tile_size: int, fwd_variant: int = 0, bwd_variant: int = 0):
        """Tile size (8/12/16/32) and batch-geometry variants of the compositing kernels."""
        _lib.check(self._lib.rmesh_set_tuning(self._h, int(tile_size), int(fwd_variant), int(bwd_variant)))

    def set_sort_digit_bits(self, bits: int):
        """Radix-sort digit width, 8 (default) or 4 bits."""
        _lib.check(self._lib.rmesh_set_sort_digit_bits(self._h, int(bits)))

    def get_sort_digit_bits(self) -> int:
        return int(self._lib.rmesh_get_sort_digit_bits(self._h))

    def get_tuning(self):
        """(tile_size, fwd_variant, bwd_variant) in use."""
        out = np.zeros(3, dtype=np.int32)
        _lib.check(self._lib.rmesh_get_tuning(self._h, _p(out)))
        return tuple(int(x) for x in out)

    @staticmethod
    def num_variants():
        """(forward variants, adjoint variants)."""
        lib = _lib.load()
        return int(lib.rmesh_num_variants(0)), int(lib.rmesh_num_variants(1))

    def autotune(self, cams, loss_type=0, reps=3, gt_device_ptr=None):
        """Pick tile size and kernel variants by timing them on these views (targets
        as in ``train_step_batched``); returns {"tuning": (ts, fwd, bwd), "ms": ...}."""
        cp, vp, ci = self._stack_cams(cams)
        ms = np.zeros(4, dtype=np.float32)
        _lib.check(self._lib.rmesh_autotune(self._h, len(cams), _p(cp), _p(vp), _p(ci),
                                            ctypes.c_void_p(gt_device_ptr or 0), int(loss_type), int(reps), _p(ms)))
        return {"tuning": self.get_tuning(), "sort_digit_bits": self.get_sort_digit_bits(),
                "ms": dict(zip(("total", "binning_sort", "forward", "backward"), ms.tolist()))}

    def device_buffer(self, which: int) -> int:
        return int(self._lib.rmesh_device_buffer(self._h, which) or 0)

    def synchronize(self):
        _lib.check(self._lib.rmesh_synchronize(self._h))

    def set_targets(self, gt):
        """Upload target images [V, H, W, 3] (float32) into the renderer's target buffer."""
        _f32("gt", gt)
        if gt.size % (self._hgt * self._w * 3):
            raise ValueError("targets must be [V, H, W, 3]")
        _lib.check(self._lib.rmesh_set_targets(self._h, gt.size // (self._hgt * self._w * 3), _p(gt)))

    STAGES = ("binning", "pairs", "sort", "ranges", "forward", "backward", "adam")

    def set_timing(self, enable: bool):
        _lib.check(self._lib.rmesh_set_timing(self._h, 1 if enable else 0))

    def stage_times(self):
        """Durations (ms) of the stages of the last train step (see STAGES)."""
        out = np.empty(7, dtype=np.float32)
        _lib.check(self._lib.rmesh_stage_times(self._h, _p(out)))
        return dict(zip(self.STAGES, out.tolist()))

    def read_buffer(self, which: int, count: int, dtype=np.float32):
        """Host copy of the first `count` elements of an internal buffer (_lib.BUF_*)."""
        out = np.empty(count, dtype=dtype)
        _lib.check(self._lib.rmesh_read_buffer(self._h, int(which), _p(out), out.nbytes))
        return out

    def gradients(self):
        """Gradients left by the last backward / train step (device buffers)."""
        return {"d_vertices": self.read_buffer(_lib.BUF_D_VERTICES, 3 * self._n),
                "d_densities": self.read_buffer(_lib.BUF_D_DENSITIES, self._m),
                "d_color_grads": self.read_buffer(_lib.BUF_D_COLOR_GRADS, 3 * self._m)}

    def sh_gradient(self):
        """SH-coefficient gradient [M*3*nc] left by the last backward or by a
        train step run without an SH learning rate (with one, the fused Adam
        consumes the gradient in registers and this buffer is not written)."""
        if self._sh_degree is None:
            raise ValueError("renderer has no SH coefficients")
        out = np.empty(3 * self._m * (self._sh_degree + 1) ** 2, dtype=np.float32)
        _lib.check(self._lib.rmesh_get_sh_grad(self._h, _p(out)))
        return out

    def pair_count(self) -> int:
        n = ctypes.c_uint64()
        _lib.check(self._lib.rmesh_pair_count(self._h, ctypes.byref(n)))
        return int(n.value)

    # ---- inspection of the binning state ----
    def debug_tiles_touched(self, view=0):
        out = np.empty(self._m, dtype=np.uint32)
        _lib.check(self._lib.rmesh_debug_tiles_touched(self._h, int(view), _p(out)))
        return out

    def debug_pairs(self, cam_pos, vp, c2w_intrinsics, cap=None):
        c, v, ci = self._cam(cam_pos, vp, c2w_intrinsics)
        n = ctypes.c_uint64()
        _lib.check(self._lib.rmesh_debug_pairs(self._h, _p(c), _p(v), _p(ci), None, None, 0, ctypes.byref(n)))
        cap = int(n.value) if cap is None else cap
        keys = np.empty(max(cap, 1), dtype=np.uint64)
        vals = np.empty(max(cap, 1), dtype=np.uint32)
        _lib.check(self._lib.rmesh_debug_pairs(self._h, _p(c), _p(v), _p(ci), _p(keys), _p(vals), cap,
                                               ctypes.byref(n)))
        k = min(cap, int(n.value))
        return keys[:k], vals[:k], int(n.value)

    def debug_sorted(self, views=1):
        n = ctypes.c_uint64()
        _lib.check(self._lib.rmesh_debug_sorted(self._h, None, None, 0, ctypes.byref(n), None, 0))
        P = int(n.value)
        keys = np.empty(max(P, 1), dtype=np.uint64)
        vals = np.empty(max(P, 1), dtype=np.uint32)
        tx = (self._w + self.tile_size - 1) // self.tile_size
        ty = (self._hgt + self.tile_size - 1) // self.tile_size
        ranges = np.zeros(2 * tx * ty * views, dtype=np.uint32)
        _lib.check(self._lib.rmesh_debug_sorted(self._h, _p(keys), _p(vals), P, ctypes.byref(n), _p(ranges),
                                                ranges.size))
        return keys[:P], vals[:P], ranges


def _not_implemented(name):
    def f(self, *a, **k):
        raise NotImplementedError(f"{name} is outside the tiled train-step path served by this renderer")
    f.__name__ = name
    return f


for _n in _OUT_OF_SCOPE:
    setattr(RMeshRenderer, _n, _not_implemented(_n))


def sort_pairs(keys: np.ndarray, vals: np.ndarray, sorting_bits: int, device_id: int = -1):
    """GPU radix sort of (u64 key, u32 value) pairs (the pipeline's sort)."""
    lib = _lib.load()
    k = np.array(keys, dtype=np.uint64, copy=True)
    v = np.array(vals, dtype=np.uint32, copy=True)
    _lib.check(lib.rmesh_sort_pairs(_p(k), _p(v), k.size, int(sorting_bits), int(device_id)))
    return k, v


def sorting_bits(num_tiles: int) -> int:
    return int(_lib.load().rmesh_sorting_bits(int(num_tiles)))
