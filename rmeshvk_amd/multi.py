"""Many scenes of uneven size, resolution and SH degree in one renderer.

Extension with no counterpart in the reference (which holds one scene per
``RMeshRenderer``, crates/rmesh-python/src/lib.rs:265-277): a step runs one
launch per stage over every scene's views (binning, one sort, the compositing
grids over all views' tiles, one Adam update; the SH stages once per SH degree
present).  Scenes keep their own arrays, image size, SH degree and view
capacity; cameras, images and targets are listed view by view, scene by
scene, in the order the scenes were given.

The handle-wide methods of ``RMeshRenderer`` (graph mode, timing, tuning,
adjoint mode, device memory) apply; its single-scene methods raise.
"""

from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from .renderer import RMeshRenderer, _f32, _p, _u32

_BUF = {"vertices": _lib.BUF_VERTICES, "densities": _lib.BUF_DENSITIES, "color_grads": _lib.BUF_COLOR_GRADS,
        "base_colors": _lib.BUF_BASE_COLORS, "sh_coeffs": _lib.BUF_SH, "d_vertices": _lib.BUF_D_VERTICES,
        "d_densities": _lib.BUF_D_DENSITIES, "d_color_grads": _lib.BUF_D_COLOR_GRADS, "d_sh": _lib.BUF_D_SH}


class MultiSceneRenderer(RMeshRenderer):
    """``scenes``: dicts with vertices, indices, base_colors, densities,
    color_grads, circumdata, width, height and, for SH colours, sh (channel-major
    [M][3][(d+1)^2]) and sh_degree; optional max_views (default ``max_views``)."""

    def __init__(self, scenes, *, tile_size=12, device_id=-1, max_views=4):
        lib = _lib.load()
        if not scenes:
            raise ValueError("no scenes")
        descs = (_lib.SceneDesc * len(scenes))()
        keep = []
        self._info = []
        for i, s in enumerate(scenes):
            v, ix = s["vertices"], s["indices"]
            _f32("vertices", v)
            _u32("indices", ix)
            if v.size % 3 or ix.size % 4:
                raise ValueError(f"scene {i}: vertices must be [N*3] and indices [M*4]")
            n, m = v.size // 3, ix.size // 4
            for k, size in (("base_colors", 3 * m), ("densities", m), ("color_grads", 3 * m), ("circumdata", 4 * m)):
                _f32(k, s[k], size)
            sh = s.get("sh")
            deg = int(s.get("sh_degree", 0)) if sh is not None else 0
            if sh is not None:
                _f32("sh", sh, 3 * m * (deg + 1) ** 2)
            keep.append((v, ix, s["base_colors"], s["densities"], s["color_grads"], s["circumdata"], sh))
            mv = int(s.get("max_views", max_views))
            descs[i] = _lib.SceneDesc(
                vertices=v.ctypes.data, n_vertices=n, indices=ix.ctypes.data, n_tets=m,
                base_colors=s["base_colors"].ctypes.data, densities=s["densities"].ctypes.data,
                color_grads=s["color_grads"].ctypes.data, circumdata=s["circumdata"].ctypes.data,
                width=int(s["width"]), height=int(s["height"]), aux_dim=0, tile_size=int(tile_size),
                device_id=int(device_id), sh_coeffs=None if sh is None else sh.ctypes.data, sh_degree=deg,
                max_views=mv, sh_half=0)
            self._info.append({"n": n, "m": m, "w": int(s["width"]), "h": int(s["height"]),
                               "sh_degree": deg if sh is not None else None, "max_views": mv})
        h = ctypes.c_void_p()
        _lib.check(lib.rmesh_create_multi(descs, len(scenes), ctypes.byref(h)))
        del keep
        self._h = h
        self._lib = lib
        self._n = sum(i["n"] for i in self._info)
        self._m = sum(i["m"] for i in self._info)
        self._w, self._hgt = self._info[0]["w"], self._info[0]["h"]
        self._sh_degree = self._info[0]["sh_degree"]
        self._max_views = self._info[0]["max_views"]
        self.tile_size = int(tile_size)

    @property
    def num_scenes(self) -> int:
        return int(self._lib.rmesh_scene_count(self._h))

    def _layout(self, views_per_scene, cams):
        vps = np.ascontiguousarray(np.asarray(views_per_scene, dtype=np.uint32))
        if vps.size != len(self._info):
            raise ValueError("views_per_scene must list every scene")
        if int(vps.sum()) != len(cams):
            raise ValueError("one camera per view")
        cp, vp, ci = self._stack_cams(cams)
        return vps, cp, vp, ci

    def _split(self, flat, views_per_scene, ch):
        out, o = [], 0
        for info, nv in zip(self._info, views_per_scene):
            n = int(nv) * info["h"] * info["w"] * ch
            out.append(flat[o:o + n].reshape(int(nv), info["h"], info["w"], ch))
            o += n
        return out

    def forward(self, views_per_scene, cams):
        """Images of every view: a list (per scene) of [V_s, H_s, W_s, 4]."""
        vps, cp, vp, ci = self._layout(views_per_scene, cams)
        npx = sum(int(nv) * i["h"] * i["w"] for i, nv in zip(self._info, vps))
        out = np.empty(npx * 4, dtype=np.float32)
        _lib.check(self._lib.rmesh_forward_multi(self._h, _p(vps), _p(cp), _p(vp), _p(ci), _p(out)))
        return self._split(out, vps, 4)

    def set_targets(self, views_per_scene, targets):
        """targets: per scene [V_s, H_s, W_s, 3] float32."""
        vps = np.asarray(views_per_scene, dtype=np.uint32)
        flat = np.ascontiguousarray(np.concatenate([np.asarray(t, np.float32).ravel() for t in targets]))
        _lib.check(self._lib.rmesh_set_targets_multi(self._h, _p(np.ascontiguousarray(vps)), _p(flat)))

    def train_step(self, views_per_scene, cams, loss_type, lr_verts, lr_dens, lr_grads, lr_sh=0.0):
        """One step over every scene's views against the targets; returns
        (mean loss, per-view losses)."""
        vps, cp, vp, ci = self._layout(views_per_scene, cams)
        lv = np.empty(len(cams), dtype=np.float32)
        loss = ctypes.c_float()
        _lib.check(self._lib.rmesh_train_step_multi(self._h, _p(vps), _p(cp), _p(vp), _p(ci), None, int(loss_type),
                                                    float(lr_verts), float(lr_dens), float(lr_grads), float(lr_sh),
                                                    _p(lv), ctypes.byref(loss)))
        return float(loss.value), lv

    def autotune(self, views_per_scene, cams, loss_type=0, reps=3):
        vps, cp, vp, ci = self._layout(views_per_scene, cams)
        ms = np.zeros(4, dtype=np.float32)
        _lib.check(self._lib.rmesh_autotune_multi(self._h, _p(vps), _p(cp), _p(vp), _p(ci), None, int(loss_type),
                                                  int(reps), _p(ms)))
        return {"tuning": self.get_tuning(), "sort_digit_bits": self.get_sort_digit_bits(),
                "ms": dict(zip(("total", "binning_sort", "forward", "backward"), ms.tolist()))}

    def read_scene(self, scene: int, name: str):
        """Scene ``scene``'s parameters or gradients: vertices, densities,
        color_grads, base_colors, sh_coeffs, d_vertices, d_densities,
        d_color_grads, d_sh."""
        info = self._info[scene]
        n, m = info["n"], info["m"]
        nc = (info["sh_degree"] + 1) ** 2 if info["sh_degree"] is not None else 0
        size = {"vertices": 3 * n, "d_vertices": 3 * n, "densities": m, "d_densities": m, "color_grads": 3 * m,
                "d_color_grads": 3 * m, "base_colors": 3 * m, "sh_coeffs": 3 * m * nc, "d_sh": 3 * m * nc}[name]
        out = np.empty(size, dtype=np.float32)
        _lib.check(self._lib.rmesh_read_scene(self._h, int(scene), _BUF[name], _p(out), out.nbytes))
        return out

    def set_scene_timing(self, enable: bool):
        _lib.check(self._lib.rmesh_set_scene_timing(self._h, 1 if enable else 0))

    def set_scene_order(self, order: str):
        """"interleaved" (default: every scene's tiles heaviest first, best
        balance) or "by_scene" (scenes one after another, fewest pairs first:
        early scenes finish early).  Results are the same either way."""
        code = {"interleaved": 0, "by_scene": 1}.get(order)
        if code is None:
            raise ValueError("order must be 'interleaved' or 'by_scene'")
        _lib.check(self._lib.rmesh_set_scene_order(self._h, code))

    def get_scene_order(self) -> str:
        return ("interleaved", "by_scene")[int(self._lib.rmesh_get_scene_order(self._h))]

    def scene_latency(self):
        """ms from the last step's start to each scene's last adjoint tile."""
        out = np.empty(len(self._info), dtype=np.float32)
        _lib.check(self._lib.rmesh_scene_latency(self._h, _p(out)))
        return out
