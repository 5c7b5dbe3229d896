"""MI355X-native (gfx950) Domain Transform Solver and HFBS.

The compute lives in libdts_amd.so (hand-written HIP kernels behind the C ABI
in include/dts_capi.h).  This package holds the build recipe, the ctypes
binding, a Python mirror of the reference's solver classes and the synthetic
scene generators.
"""
from . import scenes  # noqa: F401

__all__ = ["scenes", "load_library"]


def load_library():
    from . import _lib
    return _lib.load()
