"""Python handle on the FP64 CPU oracle (test infrastructure only).

Loads oracle/_build/liboracle.so (built by oracle/Makefile) and exposes the
same `Solver` interface as the product's C ABI, with the om_ prefix. Only
tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this.
"""
import ctypes as C
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
sys.path.insert(0, os.path.join(REPO, "boson-cloud-binary_amd"))

from gm import capi  # noqa: E402

LIB_PATH = os.path.join(HERE, "_build", "liboracle.so")
_lib = None

OP_A, OP_B, OP_H, OP_V, OP_DVDT = range(5)


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = C.CDLL(LIB_PATH)
        capi.bind(_lib, "om_")
        _lib.om_wigner3j.restype = C.c_double
        _lib.om_wigner3j.argtypes = [C.c_int] * 6
        _lib.om_sphPlm0.restype = C.c_double
        _lib.om_sphPlm0.argtypes = [C.c_int, C.c_int]
        _lib.om_gauss_legendre01.restype = None
        _lib.om_gauss_legendre01.argtypes = [C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        _lib.om_apply_op.restype = C.c_int
        _lib.om_apply_op.argtypes = [C.c_void_p, C.c_int32, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        _lib.om_sink_vector.restype = C.c_int
        _lib.om_set_multipole_cut.restype = C.c_int
        _lib.om_set_multipole_cut.argtypes = [C.c_void_p, C.POINTER(C.c_int32), C.c_int32]
        _lib.om_sink_vector.argtypes = [C.c_void_p, C.c_int32, C.POINTER(C.c_double)]
        ip = C.POINTER(C.c_int32)
        _lib.om_coupling_terms.restype = C.c_int
        _lib.om_coupling_terms.argtypes = [C.c_void_p, ip, ip, ip, C.POINTER(C.c_double), C.c_int32]
        dp = C.POINTER(C.c_double)
        _lib.om_thomas_probe.restype = C.c_int
        _lib.om_thomas_probe.argtypes = [C.c_int32, C.c_int32, dp, dp, dp, dp, dp]
        _lib.om_rolling_new.restype = C.c_void_p
        _lib.om_rolling_new.argtypes = [C.c_double]
        _lib.om_rolling_update.restype = C.c_int
        _lib.om_rolling_update.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_double, dp]
        _lib.om_rolling_free.restype = None
        _lib.om_rolling_free.argtypes = [C.c_void_p]
        _lib.om_rolling_ready.restype = C.c_int
        _lib.om_rolling_ready.argtypes = [C.c_void_p]
        _lib.om_rolling_value.restype = C.c_double
        _lib.om_rolling_value.argtypes = [C.c_void_p]
        _lib.om_rolling_reset.restype = None
        _lib.om_rolling_reset.argtypes = [C.c_void_p]
        _lib.om_rolling_save.restype = C.c_int
        _lib.om_rolling_save.argtypes = [C.c_void_p, C.c_char_p]
        _lib.om_rolling_load.restype = C.c_int
        _lib.om_rolling_load.argtypes = [C.c_void_p, C.c_char_p]
    return _lib


def wigner3j(j1, j2, j3, m1, m2, m3):
    return lib().om_wigner3j(j1, j2, j3, m1, m2, m3)


def sphPlm0(l, m):
    return lib().om_sphPlm0(l, m)


def gauss_legendre01(n):
    x, w = np.zeros(n), np.zeros(n)
    lib().om_gauss_legendre01(n, x.ctypes.data_as(C.POINTER(C.c_double)),
                              w.ctypes.data_as(C.POINTER(C.c_double)))
    return x, w


class Oracle(capi.Solver):
    def __init__(self, *args, **kw):
        super().__init__(lib(), "om_", *args, **kw)

    def apply_op(self, op, x):
        a = np.ascontiguousarray(np.asarray(x, dtype=np.complex128).ravel())
        out = np.zeros_like(a)
        dp = C.POINTER(C.c_double)
        self._check(self._lib.om_apply_op(self._h, op, a.view(np.float64).ctypes.data_as(dp),
                                          out.view(np.float64).ctypes.data_as(dp)))
        return out

    def sink_vector(self, body):
        out = np.zeros(self.size, dtype=np.complex128)
        self._lib.om_sink_vector(self._h, body, out.view(np.float64).ctypes.data_as(C.POINTER(C.c_double)))
        return out

    def set_multipole_cut(self, lcut, tile=64):
        """Evaluate only the multipoles l* <= lcut[i // tile] at node i (None:
        every term), to check a truncated coupling term for term."""
        if lcut is None:
            self._check(self._lib.om_set_multipole_cut(self._h, None, 0))
            return
        self._lcut = np.ascontiguousarray(np.asarray(lcut, dtype=np.int32))
        assert self._lcut.size * tile >= self.Nr
        self._check(self._lib.om_set_multipole_cut(self._h, self._lcut.ctypes.data_as(C.POINTER(C.c_int32)),
                                                   int(tile)))

    def coupling_terms(self):
        n = self._lib.om_coupling_terms(self._h, None, None, None, None, 0)
        rows, cols, ls = (np.zeros(n, dtype=np.int32) for _ in range(3))
        ang = np.zeros(n)
        ip = C.POINTER(C.c_int32)
        self._lib.om_coupling_terms(self._h, rows.ctypes.data_as(ip), cols.ctypes.data_as(ip),
                                    ls.ctypes.data_as(ip), ang.ctypes.data_as(C.POINTER(C.c_double)), n)
        return rows, cols, ls, ang


def thomas(diag, lower, upper, rhs):
    """Solve K independent tridiagonal systems (arrays of shape (K, N))."""
    K, N = diag.shape
    dp = C.POINTER(C.c_double)
    conv = [np.ascontiguousarray(np.asarray(a, dtype=np.complex128)) for a in (diag, lower, upper, rhs)]
    x = np.zeros((K, N), dtype=np.complex128)
    rc = lib().om_thomas_probe(N, K, *[a.view(np.float64).ctypes.data_as(dp) for a in conv],
                               x.view(np.float64).ctypes.data_as(dp))
    return rc, x


class Rolling:
    """The oracle's RollingPhaseAverage (src/RollingPhase.cpp)."""

    def __init__(self, n_orbits):
        self._h = lib().om_rolling_new(float(n_orbits))
        if not self._h:
            raise ValueError("RollingPhaseAverage: n_orbits must be positive")

    def update(self, t, phase, value):
        out = C.c_double()
        rc = lib().om_rolling_update(self._h, t, phase, value, C.byref(out))
        if rc == -1:
            raise RuntimeError("RollingPhaseAverage: time must increase")
        if rc != 0:
            raise RuntimeError("RollingPhaseAverage: phase must increase")
        return out.value

    def ready(self):
        return bool(lib().om_rolling_ready(self._h))

    def value(self):
        return lib().om_rolling_value(self._h)

    def reset(self):
        lib().om_rolling_reset(self._h)

    def save(self, path):
        if lib().om_rolling_save(self._h, str(path).encode()) != 0:
            raise RuntimeError("RollingPhaseAverage: could not write")

    def load(self, path):
        if lib().om_rolling_load(self._h, str(path).encode()) != 0:
            raise RuntimeError("RollingPhaseAverage: truncated checkpoint")

    def __del__(self):
        try:
            if self._h:
                lib().om_rolling_free(self._h)
        except Exception:
            pass
