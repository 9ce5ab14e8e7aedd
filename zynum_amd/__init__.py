"""zynum_amd: MI355X (gfx950) Level-3 GEMM tier for Zynum's BLAS ABI.

The compute lives in ``libzynum_gpu.so`` (hand-written HIP for gfx950, C ABI in
``include/zynum_gpu.h``).  This package holds its ctypes binding, the typed
view API mirror (``blas``), the grouped-GEMM autotuner and the synthetic input
generator.
"""
from . import _lib  # noqa: F401

__all__ = ["_lib"]
