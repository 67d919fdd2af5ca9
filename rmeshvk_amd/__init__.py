"""rmeshvk_amd — MI355X-native differentiable radiance-mesh train step.

Hand-written HIP kernels for gfx950 behind a C ABI (include/rmesh_hip.h),
with a Python mirror of rmeshvk's ``RMeshRenderer`` and autograd wrapper.
"""

from .multi import MultiSceneRenderer  # noqa: F401
from .renderer import RMeshRenderer, sort_pairs, sorting_bits  # noqa: F401

__all__ = ["RMeshRenderer", "MultiSceneRenderer", "sort_pairs", "sorting_bits"]
