"""mrmath_amd -- MI355X (gfx950) backend for mrMath's FP64 multiply and LU/solve path.

The product is libmrmath_gpu.so (hand-written HIP kernels behind the C ABI in
include/mrmath_gpu.h).  This package holds its sources (csrc/), the build
script, a ctypes binding and a host-side mirror of the reference's stub-switch
surface for callers and tests.
"""
from . import _lib
from ._lib import MrmError, load

__all__ = ["MrmError", "load", "_lib"]
