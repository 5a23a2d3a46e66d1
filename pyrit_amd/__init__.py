"""pyrit_amd - MI355X-native sparse-voxel-DAG build and software traversal.

The product is the HIP library ``libpyrit.so`` (C ABI: include/pyrit.h and
include/pyrit_hip_build.h). This package loads it with ctypes and mirrors the
reference's build / trace interface for Python callers (tests, bench). There is
no fallback: if the library is missing or cannot be loaded, every entry point
raises.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpyrit.so")
TERRAIN_PATH = os.path.join(HERE, "libpyrit_terrain.so")

_lib = None
_terrain = None


class NativeLibraryMissing(RuntimeError):
    pass


def lib() -> ctypes.CDLL:
    """The loaded libpyrit.so (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise NativeLibraryMissing(
                f"{LIB_PATH} is not built; run `python -m pyrit_amd.build` (hipcc, gfx950)")
        _lib = ctypes.CDLL(LIB_PATH)
        from . import abi

        abi.declare(_lib)
    return _lib


def terrain_lib() -> ctypes.CDLL:
    global _terrain
    if _terrain is None:
        if not os.path.exists(TERRAIN_PATH):
            raise NativeLibraryMissing(f"{TERRAIN_PATH} is not built; run `python -m pyrit_amd.build`")
        _terrain = ctypes.CDLL(TERRAIN_PATH)
        from . import abi

        abi.declare_terrain(_terrain)
    return _terrain
