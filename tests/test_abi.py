"""CPU: the C-ABI boundary — the library loads, exports exactly what
include/rmesh_hip.h declares, its host-only entry points answer, it reports a
missing device as an error, and the product path never reaches the oracle.
No GPU compute is launched here."""

from __future__ import annotations

import ast
import ctypes
import json
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "rmesh_hip.h")


def _declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"//[^\n]*", "", text)
    return set(re.findall(r"\b(rmesh_[a-z0-9_]+)\s*\(", text))


@pytest.fixture(scope="module")
def lib():
    from rmeshvk_amd import _lib

    if not os.path.exists(_lib.LIB_PATH):
        from rmeshvk_amd.build import build

        build(jobs=4, force=False)
    return _lib.load()


def test_header_matches_binding_table():
    from rmeshvk_amd import _lib

    assert _declared() == set(_lib.SIGNATURES), _declared() ^ set(_lib.SIGNATURES)


def test_library_exports_every_declared_symbol(lib):
    for name in sorted(_declared()):
        assert hasattr(lib, name), name


def test_sorting_bits_through_abi(lib):
    with open(os.path.join(ROOT, "tests", "golden", "reference_kats.json")) as f:
        table = json.load(f)["sorting_bits"]["table"]
    for num_tiles, bits in table:
        assert lib.rmesh_sorting_bits(num_tiles) == bits


def _gpu_present():
    return os.path.exists("/dev/kfd")


@pytest.mark.skipif(_gpu_present(), reason="a GPU is present; the no-device path is not reachable")
def test_create_without_device_is_an_error(lib):
    import numpy as np

    from rmeshvk_amd import _lib
    from rmeshvk_amd.scenes import regular_tet_scene

    s = regular_tet_scene()
    keep = {k: np.ascontiguousarray(v) for k, v in s.items()}
    d = _lib.SceneDesc()
    d.vertices, d.n_vertices = keep["vertices"].ctypes.data, 4
    d.indices, d.n_tets = keep["indices"].ctypes.data, 1
    d.base_colors = keep["base_colors"].ctypes.data
    d.densities = keep["densities"].ctypes.data
    d.color_grads = keep["color_grads"].ctypes.data
    d.circumdata = keep["circumdata"].ctypes.data
    d.width = d.height = 16
    d.tile_size = 12
    d.device_id = 0
    h = ctypes.c_void_p()
    st = lib.rmesh_create(ctypes.byref(d), ctypes.byref(h))
    assert st == _lib.RMESH_ERR_DEVICE
    assert not h.value
    assert lib.rmesh_last_error()


def test_invalid_arguments_rejected_on_host(lib):
    from rmeshvk_amd import _lib

    h = ctypes.c_void_p()
    assert lib.rmesh_create(None, ctypes.byref(h)) == _lib.RMESH_ERR_INVALID
    assert lib.rmesh_set_min_t(None, 0.0) == _lib.RMESH_ERR_INVALID


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    from rmeshvk_amd import _lib

    monkeypatch.setattr(_lib, "LIB_PATH", str(tmp_path / "absent.so"))
    monkeypatch.setattr(_lib, "_LIB", None)
    with pytest.raises(ImportError):
        _lib.load()


def test_product_never_imports_the_oracle():
    """rmeshvk_amd/ and rmesh_wgpu/ hold no import of the CPU oracle."""
    for pkg in ("rmeshvk_amd", "rmesh_wgpu"):
        for dirpath, _, files in os.walk(os.path.join(ROOT, pkg)):
            for fn in files:
                if not fn.endswith(".py"):
                    continue
                tree = ast.parse(open(os.path.join(dirpath, fn)).read())
                for node in ast.walk(tree):
                    names = []
                    if isinstance(node, ast.Import):
                        names = [a.name for a in node.names]
                    elif isinstance(node, ast.ImportFrom):
                        names = [node.module or ""]
                    assert not any(n == "oracle" or n.startswith("oracle.") for n in names), (fn, names)
