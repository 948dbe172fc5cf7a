"""MI355X (gfx950) wavefront path tracer.

Hot path in HIP (libpathtracer_amd.so, C ABI in include/pathtracer_amd.h);
this package holds the ctypes binding, the scene descriptions of the
benchmark workloads and a thin renderer wrapper.
"""
from . import _lib  # noqa: F401  (raises if the HIP library is not built)
from .scenes import SceneSpec  # noqa: F401

__all__ = ["SceneSpec"]
