"""Python binding of the MI355X Crank-Nicolson step (C ABI in include/gm.h).

`CrankNicolsonModes(...)` opens one solver handle on the GPU through
gm/libgm.so. There is no CPU fallback: when the library is missing or no GPU
is visible, constructing a solver raises.
"""
import ctypes as C
import os

from . import capi  # noqa: F401
from .capi import GM_PRECISION_FP64, GM_PRECISION_MIXED, GMError, InvalidArgument, LogicError  # noqa: F401

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libgm.so")
_lib = None


def lib():
    """Load libgm.so (raises OSError when it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise OSError("libgm.so is not built: run boson-cloud-binary_amd/build.py")
        _lib = C.CDLL(LIB_PATH)
        capi.bind(_lib, "gm_")
        _lib.gm_version.restype = C.c_char_p
        _lib.gm_device_count.restype = C.c_int
    return _lib


def device_count():
    return lib().gm_device_count()


class CrankNicolsonModes(capi.Solver):
    """The reference class CrankNicolsonModes (include/CrankNicolsonModes.hpp:84)
    on one MI355X."""

    def __init__(self, *args, **kw):
        super().__init__(lib(), "gm_", *args, **kw)
