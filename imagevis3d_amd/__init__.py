"""imagevis3d_amd -- MI355X (gfx950) renderer backend for Tuvok's raycasting
and brick/LOD hot path.

The product is ``libtvkhip.so`` (C ABI in ``include/tvk_hip.h``): hand-written
HIP kernels for ray-brick traversal + compositing, a device brick pool with a
page table, and the bricking / LOD / histogram kernels, behind a C++
AbstrRenderer-shaped host class.  This package holds its build script, the
ctypes binding, and a Python mirror of Tuvok's Lua renderer interface.
"""
from ._lib import TvkError, load  # noqa: F401
from .renderer import (BlendPrecision, Context, Interpolant, RenderMode, Renderer, brick_slot_bytes,  # noqa: F401
                       deleteClass, generate_volume, generate_volume_gpu, tuvok)

__all__ = ["TvkError", "load", "BlendPrecision", "Context", "Interpolant", "RenderMode", "Renderer",
           "brick_slot_bytes", "deleteClass", "generate_volume", "generate_volume_gpu", "tuvok"]
