"""Drop-in import path of the reference package (crates/rmesh-python/python/rmesh_wgpu/__init__.py:1-5):
``from rmesh_wgpu import RMeshRenderer`` resolves to the MI355X renderer."""

from rmeshvk_amd.renderer import RMeshRenderer  # noqa: F401

__all__ = ["RMeshRenderer"]
