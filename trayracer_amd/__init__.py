"""trayracer_amd — MI355X (gfx950) wavefront path tracer.

The compute path is libtrt.so (hand-written HIP kernels behind the C ABI of
include/trt.h); this package holds its Python binding, the Renderer façade and the
procedural benchmark scenes.
"""
from . import scenes  # noqa: F401
from ._lib import LIB_PATH, TrtError, load  # noqa: F401
from .renderer import KernelSettings, Renderer  # noqa: F401

__all__ = ["Renderer", "KernelSettings", "TrtError", "scenes", "load", "LIB_PATH"]
