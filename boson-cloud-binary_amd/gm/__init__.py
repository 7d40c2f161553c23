"""Python binding of the MI355X Crank-Nicolson step (C ABI in include/gm.h)."""
from . import capi  # noqa: F401
