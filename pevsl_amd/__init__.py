"""pevsl_amd: MI355X-native spectrum-slicing path of pEVSL.

The product is the C-ABI shared library libpevsl_amd.so (include/pevsl.h,
include/pevsl_amd.h) built from pevsl_amd/csrc for gfx950.  This package is
its Python binding (capi) plus the host driver of the Laplacian run (driver).
"""
from . import capi  # noqa: F401
from .capi import load  # noqa: F401

__all__ = ["capi", "load", "driver"]
