"""ctypes binding of librmesh_hip.so (C ABI declared in include/rmesh_hip.h).

The library is built in-tree by ``rmeshvk_amd.build.build()`` (hipcc, gfx950).
There is no CPU fallback: importing a renderer without the built library
raises, so a missing extension can never pass silently.
"""

from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "librmesh_hip.so")

RMESH_OK = 0
RMESH_ERR_INVALID = 1
RMESH_ERR_DEVICE = 2
RMESH_ERR_UNSUPPORTED = 3
RMESH_ERR_CAPACITY = 4

BUF_RGBA, BUF_GT, BUF_VERTICES, BUF_DENSITIES, BUF_COLOR_GRADS, BUF_BASE_COLORS, BUF_SH = range(7)
BUF_D_VERTICES, BUF_D_DENSITIES, BUF_D_COLOR_GRADS = 7, 8, 9
ADJOINT_UNDO, ADJOINT_CHECKPOINT = 0, 1
BUF_D_SH, BUF_D_BASE_COLORS = 10, 11

# Every symbol include/rmesh_hip.h declares, with (restype, argtypes).
_vp = ctypes.c_void_p
_u32 = ctypes.c_uint32
_u64 = ctypes.c_uint64
_i32 = ctypes.c_int32
_f32 = ctypes.c_float
_st = ctypes.c_int32


class SceneDesc(ctypes.Structure):
    _fields_ = [
        ("vertices", _vp), ("n_vertices", _u64),
        ("indices", _vp), ("n_tets", _u64),
        ("base_colors", _vp), ("densities", _vp), ("color_grads", _vp), ("circumdata", _vp),
        ("width", _u32), ("height", _u32), ("aux_dim", _u32), ("tile_size", _u32),
        ("device_id", _i32), ("sh_coeffs", _vp), ("sh_degree", _u32), ("max_views", _u32), ("sh_half", _u32),
    ]


class SceneInfo(ctypes.Structure):
    _fields_ = [
        ("n_vertices", _u64), ("n_tets", _u64), ("sh_degree", _u32), ("k_components", _u32),
        ("start_pose", _f32 * 7), ("n_tagged", _u32),
    ]


FILE_RMESH, FILE_PLY, FILE_RMESH_RAW = 0, 1, 2

SIGNATURES = {
    "rmesh_scene_parse": (_st, [_vp, _u64, _i32, ctypes.POINTER(_vp), ctypes.POINTER(SceneInfo)]),
    "rmesh_scene_copy": (_st, [_vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "rmesh_scene_free": (None, [_vp]),
    "rmesh_create": (_st, [ctypes.POINTER(SceneDesc), ctypes.POINTER(_vp)]),
    "rmesh_destroy": (None, [_vp]),
    "rmesh_last_error": (ctypes.c_char_p, []),
    "rmesh_set_min_t": (_st, [_vp, _f32]),
    "rmesh_set_stream": (_st, [_vp, _vp]),
    "rmesh_forward_tiled": (_st, [_vp, _vp, _vp, _vp, _vp, _vp]),
    "rmesh_forward": (_st, [_vp, _vp, _vp, _vp, _vp]),
    "rmesh_read_instance_count": (_st, [_vp, ctypes.POINTER(_u32)]),
    "rmesh_forward_error": (_st, [_vp, _vp, _vp, _vp, _vp]),
    "rmesh_forward_interval_tiled": (_st, [_vp] * 7),
    "rmesh_backward_interval_tiled": (_st, [_vp] * 9),
    "rmesh_backward_hw": (_st, [_vp] * 6),
    "rmesh_backward": (_st, [_vp, _vp, _vp, _vp, _vp, _vp]),
    "rmesh_get_sh_grad": (_st, [_vp, _vp]),
    "rmesh_train_step": (_st, [_vp, _vp, _vp, _vp, _vp, _f32, _f32, _f32, _u32, ctypes.POINTER(_f32)]),
    "rmesh_set_extra_lrs": (_st, [_vp, _f32, _f32]),
    "rmesh_update_params": (_st, [_vp, _vp, _vp, _vp, _vp]),
    "rmesh_get_params": (_st, [_vp, _vp, _vp, _vp, _vp]),
    "rmesh_update_sh": (_st, [_vp, _vp]),
    "rmesh_get_sh": (_st, [_vp, _vp]),
    "rmesh_get_step": (_u32, [_vp]),
    "rmesh_get_width": (_u32, [_vp]),
    "rmesh_get_height": (_u32, [_vp]),
    "rmesh_forward_batched": (_st, [_vp, _u32, _vp, _vp, _vp, _vp]),
    "rmesh_train_step_batched": (_st, [_vp, _u32, _vp, _vp, _vp, _vp, _u32, _f32, _f32, _f32, _f32,
                                       ctypes.POINTER(_f32)]),
    "rmesh_set_graph_mode": (_st, [_vp, ctypes.c_int]),
    "rmesh_set_tuning": (_st, [_vp, _u32, _i32, _i32]),
    "rmesh_get_tuning": (_st, [_vp, _vp]),
    "rmesh_num_variants": (_i32, [_i32]),
    "rmesh_autotune": (_st, [_vp, _u32, _vp, _vp, _vp, _vp, _u32, _u32, _vp]),
    "rmesh_device_buffer": (_vp, [_vp, ctypes.c_int]),
    "rmesh_synchronize": (_st, [_vp]),
    "rmesh_set_targets": (_st, [_vp, _u32, _vp]),
    "rmesh_set_timing": (_st, [_vp, ctypes.c_int]),
    "rmesh_stage_times": (_st, [_vp, _vp]),
    "rmesh_pair_count": (_st, [_vp, ctypes.POINTER(_u64)]),
    "rmesh_read_buffer": (_st, [_vp, ctypes.c_int, _vp, _u64]),
    "rmesh_copy_probe": (_st, [_i32, _u64, _u32, ctypes.POINTER(_f32)]),
    "rmesh_atomic_probe": (_st, [_i32, _u32, _u64, ctypes.POINTER(_f32)]),
    "rmesh_debug_tiles_touched": (_st, [_vp, _u32, _vp]),
    "rmesh_debug_pairs": (_st, [_vp, _vp, _vp, _vp, _vp, _vp, _u64, ctypes.POINTER(_u64)]),
    "rmesh_debug_sorted": (_st, [_vp, _vp, _vp, _u64, ctypes.POINTER(_u64), _vp, _u64]),
    "rmesh_sort_pairs": (_st, [_vp, _vp, _u64, _u32, _i32]),
    "rmesh_sort_pairs_noalloc": (_st, [_vp, _vp, _u64, _u32, _i32, _vp]),
    "rmesh_set_adjoint_mode": (_st, [_vp, _i32]),
    "rmesh_update_params_device": (_st, [_vp, _vp, _vp, _vp, _vp]),
    "rmesh_forward_tiled_device": (_st, [_vp, _vp, _vp, _vp, _vp]),
    "rmesh_backward_device": (_st, [_vp, _vp, _vp, _vp, _vp, _vp]),
    "rmesh_create_multi": (_st, [ctypes.POINTER(SceneDesc), _u32, ctypes.POINTER(_vp)]),
    "rmesh_scene_count": (_u32, [_vp]),
    "rmesh_set_scene_order": (_st, [_vp, _u32]),
    "rmesh_get_scene_order": (_u32, [_vp]),
    "rmesh_forward_multi": (_st, [_vp, _vp, _vp, _vp, _vp, _vp]),
    "rmesh_set_targets_multi": (_st, [_vp, _vp, _vp]),
    "rmesh_train_step_multi": (_st, [_vp, _vp, _vp, _vp, _vp, _vp, _u32, _f32, _f32, _f32, _f32, _vp,
                                     ctypes.POINTER(_f32)]),
    "rmesh_autotune_multi": (_st, [_vp, _vp, _vp, _vp, _vp, _vp, _u32, _u32, _vp]),
    "rmesh_read_scene": (_st, [_vp, _u32, ctypes.c_int, _vp, _u64]),
    "rmesh_set_scene_timing": (_st, [_vp, ctypes.c_int]),
    "rmesh_scene_latency": (_st, [_vp, _vp]),
    "rmesh_set_sort_digit_bits": (_st, [_vp, _u32]),
    "rmesh_get_sort_digit_bits": (_u32, [_vp]),
    "rmesh_get_adjoint_mode": (_i32, [_vp]),
    "rmesh_device_bytes": (_st, [_vp, ctypes.POINTER(_u64)]),
    "rmesh_sorting_bits": (_u32, [_u32]),
}

_LIB = None


class RMeshError(RuntimeError):
    pass


def load() -> ctypes.CDLL:
    """Load the HIP library; raise if it has not been built."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"rmeshvk_amd HIP library not found at {LIB_PATH}; build it with "
            "`python -c 'import __graft_entry__ as g; g.build()'` (hipcc, gfx950)")
    # One HIP runtime per process: the library's libamdhip64.so.7 binds to a
    # copy already loaded, so torch (which loads its own build under that
    # soname) goes first when it is installed; otherwise torch, imported later,
    # would start a second runtime that finds no GPU.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _LIB = lib
    return lib


def check(status: int) -> None:
    if status == RMESH_OK:
        return
    msg = (load().rmesh_last_error() or b"").decode("utf-8", "replace")
    if status == RMESH_ERR_INVALID:
        raise ValueError(msg)
    if status == RMESH_ERR_UNSUPPORTED:
        raise NotImplementedError(msg)
    raise RMeshError(msg)
