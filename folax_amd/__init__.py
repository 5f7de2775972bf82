"""folax_amd: MI355X-native (gfx950) finite-element residual hot path behind folax's loss API.

The compute path is libfemhot.so (hand-written HIP kernels behind a C ABI, include/femhot.h);
this package holds the PyTorch-ROCm host mirror of the reference's loss / solver objects.
"""
from . import _lib  # noqa: F401
from .mesh import Mesh, create_2D_square_mesh, create_3D_box_mesh, create_3D_tet_box_mesh  # noqa: F401

__all__ = ["Mesh", "create_2D_square_mesh", "create_3D_box_mesh", "create_3D_tet_box_mesh"]
__version__ = "0.1.0"
