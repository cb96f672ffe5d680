"""x3div_amd: MI355X-native (gfx950 HIP) x3div Euler-timestep hot path.

The product is libx3dhip.so (C ABI in include/x3dhip.h); this package is its
host-side mirror of the reference's Fortran interfaces for tests and benches.
"""
from ._lib import LIB_PATH, X3DError, load  # noqa: F401
from .context import OP, OPS, OPS2, X3D, X3DBatch, FIELD_FOURIER8, FIELD_REFERENCE_VEL, FIELD_TG_NOISE  # noqa: F401

__all__ = ["X3D", "X3DBatch", "OP", "OPS", "OPS2", "load", "LIB_PATH", "X3DError",
           "FIELD_REFERENCE_VEL", "FIELD_TG_NOISE", "FIELD_FOURIER8"]
