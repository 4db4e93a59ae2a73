"""MI355X (gfx950) energetic space-time Galerkin BEM single-layer hot path.

The compute path is the native HIP library libebem_hip.so (C ABI in
include/ebem.h); this package is its Python binding plus the host-side mirror
of the reference's MATLAB entry points (api.py) and the mesh reader.
"""
from .mesh import SpaceMesh, read_mesh, screen_uniform, icosphere  # noqa: F401

__all__ = ["SpaceMesh", "read_mesh", "screen_uniform", "icosphere"]
