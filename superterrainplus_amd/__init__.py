"""MI355X-native SuperTerrain+ heightfield hot path.

The product is the HIP library ``libstp_amd.so`` (C ABI in ``include/stp_amd.h``, C++ API in
``include/stp/SuperTerrainPlus.hpp``). This package holds its sources (``csrc/``), the build
script and the ctypes binding used by Python callers.
"""
from . import capi  # noqa: F401
from .capi import STPError, lib  # noqa: F401

__all__ = ["capi", "lib", "STPError"]
