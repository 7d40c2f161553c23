"""Python binding of the MI355X Crank-Nicolson step (C ABI in include/gm.h).

`CrankNicolsonModes(...)` opens one solver handle on the GPU through
gm/libgm.so. There is no CPU fallback: when the library is missing or no GPU
is visible, constructing a solver raises.
"""
import ctypes as C
import os

import numpy as np

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
        vp = C.c_void_p
        _lib.gm_set_coupling_tolerance.argtypes = [vp, C.c_double]
        _lib.gm_coupling_work.argtypes = [vp, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        _lib.gm_multipole_cut.argtypes = [vp, C.POINTER(C.c_int32), C.c_int32]
        _lib.gm_profile_enable.argtypes = [vp, C.c_int32]
        _lib.gm_profile_read.argtypes = [vp, C.POINTER(C.c_double), C.POINTER(C.c_int64), C.c_int32]
        _lib.gm_ensemble_step.argtypes = [C.POINTER(vp), C.c_int32, C.c_int32, C.POINTER(capi.Diag)]
        _lib.gm_apply_config.argtypes = [vp, C.c_int32, C.POINTER(C.c_int32), C.POINTER(C.c_double)]
        ip, dp = C.POINTER(C.c_int32), C.POINTER(C.c_double)
        _lib.gm_ilut_bicgstab.argtypes = [C.c_int32, ip, ip, dp, dp, dp, C.c_double, C.c_int32, C.c_double,
                                          C.c_int32, C.c_int32, ip]
        _lib.gm_last_error.restype = C.c_char_p
    return _lib


def ilut_bicgstab(ptr, col, val, b, x0=None, droptol=1e-9, fill_factor=50, tol=1e-10, max_iterations=5000,
                  eigen_mode=True):
    """gm_ilut_bicgstab: solve the complex CSR system (ptr, col, val) for b with
    ILUT-preconditioned BiCGSTAB on the host. Returns (x, iterations); raises
    GMError when the factorization fails or the tolerance is not reached."""
    import numpy as np
    ptr = np.ascontiguousarray(ptr, dtype=np.int32)
    col = np.ascontiguousarray(col, dtype=np.int32)
    val = np.ascontiguousarray(val, dtype=np.complex128)
    b = np.ascontiguousarray(b, dtype=np.complex128)
    n = ptr.size - 1
    x = np.zeros(n, dtype=np.complex128) if x0 is None else np.array(x0, dtype=np.complex128)
    its = C.c_int32()
    ip, dp = C.POINTER(C.c_int32), C.POINTER(C.c_double)
    L = lib()
    rc = L.gm_ilut_bicgstab(n, ptr.ctypes.data_as(ip), col.ctypes.data_as(ip), val.view(np.float64).ctypes.data_as(dp),
                            b.view(np.float64).ctypes.data_as(dp), x.view(np.float64).ctypes.data_as(dp),
                            float(droptol), int(fill_factor), float(tol), int(max_iterations), int(bool(eigen_mode)),
                            C.byref(its))
    if rc != 0:
        err = capi._ERRORS.get(rc, GMError)
        raise err(rc, (L.gm_last_error(None) or b"").decode())
    return x, its.value


def device_count():
    return lib().gm_device_count()


class CrankNicolsonModes(capi.Solver):
    """The reference class CrankNicolsonModes (include/CrankNicolsonModes.hpp:84)
    on one MI355X."""

    def __init__(self, *args, **kw):
        super().__init__(lib(), "gm_", *args, **kw)

    def apply_sink_device(self, d_in, d_out, stream=None):
        """gm_apply_sink_device on device pointers (ints): allocates nothing.
        With stream None the call waits for the result."""
        fn = self._lib.gm_apply_sink_device
        fn.restype = C.c_int
        fn.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        self._check(fn(self._h, C.c_void_p(int(d_in)), C.c_void_p(int(d_out)),
                       None if stream is None else C.c_void_p(int(stream))))

    OPERATORS = ("H", "H0", "A", "B", "V1", "V2", "V", "dVdt", "Id", "CAP")

    def operator(self, name):
        """The public sparse member `name` of the reference class (H, H0, A, B,
        V1, V2, V, dVdt, Id, CAP; include/CrankNicolsonModes.hpp:197-206) at the
        current step, as a scipy CSR matrix in the reference's stacking (row
        j*Nr + i). Small systems only (gm_operator_csr)."""
        import scipy.sparse
        which = self.OPERATORS.index(name)
        fn = self._lib.gm_operator_csr
        fn.restype = C.c_int
        fn.argtypes = [C.c_void_p, C.c_int32, C.POINTER(C.c_int64), C.POINTER(C.c_int64),
                       C.POINTER(C.c_int32), C.POINTER(C.c_double)]
        nnz = C.c_int64(0)
        self._check(fn(self._h, which, C.byref(nnz), None, None, None))
        n = self.size
        ptr = np.zeros(n + 1, dtype=np.int64)
        col = np.zeros(max(nnz.value, 1), dtype=np.int32)
        val = np.zeros(max(nnz.value, 1), dtype=np.complex128)
        self._check(fn(self._h, which, C.byref(nnz), ptr.ctypes.data_as(C.POINTER(C.c_int64)),
                       col.ctypes.data_as(C.POINTER(C.c_int32)), val.view(np.float64).ctypes.data_as(C.POINTER(C.c_double))))
        m = nnz.value
        return scipy.sparse.csr_matrix((val[:m], col[:m], ptr), shape=(n, n))

    def set_coupling_tolerance(self, tol):
        """Multipole truncation of the coupling (0 keeps every term)."""
        self._check(self._lib.gm_set_coupling_tolerance(self._h, float(tol)))

    def coupling_work(self):
        """(Gaunt terms, mode pairs) evaluated by one coupling apply."""
        t, p = C.c_double(), C.c_double()
        self._check(self._lib.gm_coupling_work(self._h, C.byref(t), C.byref(p)))
        return t.value, p.value

    def multipole_cut(self):
        """Largest multipole l* evaluated at each 64-node tile (int32 array)."""
        n = (self.Nr + 63) // 64
        out = np.zeros(n, dtype=np.int32)
        self._check(self._lib.gm_multipole_cut(self._h, out.ctypes.data_as(C.POINTER(C.c_int32)), n))
        return out

    def apply_config(self, precision=GM_PRECISION_FP64):
        """(apply form chosen by the autotuner: tile nodes 16, 32, 64, or 8 for the
        polar-quadrature form; {16: ms, 32: ms, 64: ms, 8: ms}, 0 for a form not timed)."""
        t = C.c_int32()
        ms = (C.c_double * 4)()
        self._check(self._lib.gm_apply_config(self._h, int(precision), C.byref(t), ms))
        return t.value, {16: ms[0], 32: ms[1], 64: ms[2], 8: ms[3]}

    def profile(self, on):
        self._check(self._lib.gm_profile_enable(self._h, int(on)))

    PROFILE_CATEGORIES = ("apply_krylov", "apply_other", "precond", "precond_setup", "update", "factor", "small",
                          "diag")

    def profile_read(self):
        """{category: (device ms, launches)} since profile(True)."""
        nc = len(self.PROFILE_CATEGORIES)
        ms = (C.c_double * nc)()
        n = (C.c_int64 * nc)()
        self._lib.gm_profile_read(self._h, ms, n, nc)
        return {name: (ms[c], n[c]) for c, name in enumerate(self.PROFILE_CATEGORIES)}


def ensemble_step(members, step_id=1, diag=True):
    """One step of every solver in `members` (gm_ensemble_step): the solvers run
    on their own streams and overlap on the device. Returns one diagnostics
    dict per member (None when diag is False). Raises the first member's error."""
    import numpy as np
    n = len(members)
    handles = (C.c_void_p * n)(*[m._h for m in members])
    arrs, diags = [], None
    if diag:
        diags = (capi.Diag * n)()
        for i, m in enumerate(members):
            a = {k: np.zeros(m.n_modes) for k in ("flux", "energy_flux", "norms", "energies")}
            for k, v in a.items():
                setattr(diags[i], k, capi._ptr(v))
            arrs.append(a)
    st = lib().gm_ensemble_step(handles, n, int(step_id), diags)
    if st != 0:
        bad = next((m for m in members if lib().gm_last_error(m._h)), members[0])
        bad._check(st)
    if not diag:
        return [None] * n
    out = []
    for i in range(n):
        d = {k: getattr(diags[i], k) for k in ("energy", "p_bin", "p_bin_inst", "dV_bin_dt", "cloud_mass",
                                                "absorbed")}
        d.update(arrs[i])
        out.append(d)
    return out


class Batch:
    """gm_batch: the members stepped together, every launch of the step one
    ragged grid over all of them, or over each of `lanes` subsets on their
    own streams (each member's result bitwise its own step's). Members stay
    usable alone between batch steps."""

    def __init__(self, members, lanes=1):
        L = lib()
        L.gm_batch_create_lanes.argtypes = [C.c_void_p, C.c_int32, C.c_int32, C.POINTER(C.c_void_p)]
        L.gm_batch_step.argtypes = [C.c_void_p, C.c_int32, C.c_void_p]
        L.gm_batch_destroy.argtypes = [C.c_void_p]
        L.gm_batch_destroy.restype = None
        L.gm_batch_last_error.argtypes = [C.c_void_p]
        L.gm_batch_last_error.restype = C.c_char_p
        self._lib = L
        self.members = list(members)
        n = len(self.members)
        self._handles = (C.c_void_p * n)(*[m._h for m in self.members])
        h = C.c_void_p()
        rc = L.gm_batch_create_lanes(self._handles, n, int(lanes), C.byref(h))
        if rc != 0:
            raise capi.GMError(rc, L.gm_last_error(None).decode())
        self._h = h

    def step(self, step_id=1, diag=True):
        """One step of every member; diagnostics dicts as CrankNicolsonModes.step."""
        import numpy as np
        n = len(self.members)
        arrs, diags = [], None
        if diag:
            diags = (capi.Diag * n)()
            for i, m in enumerate(self.members):
                a = {k: np.zeros(m.n_modes) for k in ("flux", "energy_flux", "norms", "energies")}
                for k, v in a.items():
                    setattr(diags[i], k, capi._ptr(v))
                arrs.append(a)
        st = self._lib.gm_batch_step(self._h, int(step_id), C.cast(diags, C.c_void_p) if diags else None)
        if st != 0:
            raise capi.GMError(st, self._lib.gm_batch_last_error(self._h).decode())
        if not diag:
            return [None] * n
        out = []
        for i in range(n):
            d = {k: getattr(diags[i], k) for k in ("energy", "p_bin", "p_bin_inst", "dV_bin_dt", "cloud_mass",
                                                    "absorbed")}
            d.update(arrs[i])
            out.append(d)
        return out

    def close(self):
        if getattr(self, "_h", None):
            self._lib.gm_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
