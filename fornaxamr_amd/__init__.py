"""MI355X (gfx950) implementation of FornaxAMR's per-block Godunov RK2 stage.

The product is the C ABI in ``include/fornax_gpu.h`` (``libfornax_gpu.so``);
this package holds its HIP sources (``csrc/``), the build recipe, a ctypes
binding and the host-side RK2 driver mirror (``GpuMesh``).
"""
from .problems import Problem  # noqa: F401

__all__ = ["Problem"]
