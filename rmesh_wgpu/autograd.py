"""Drop-in for rmesh_wgpu.autograd (crates/rmesh-python/python/rmesh_wgpu/autograd.py)."""

from rmeshvk_amd.autograd import RMeshForward, RMeshModule  # noqa: F401

__all__ = ["RMeshForward", "RMeshModule"]
