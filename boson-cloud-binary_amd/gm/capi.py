"""ctypes view of the C ABI declared in include/gm.h.

`Solver` wraps one handle of a library that exports the gm_* entry points
(or the same entry points under another prefix) and mirrors the public API of
the reference class CrankNicolsonModes (include/CrankNicolsonModes.hpp:60-455
of rdellamonica/boson-cloud-binary): constructor arguments, initialize, the
public flags, step, psi and the orbit scalars.
"""
import ctypes as C

import numpy as np

GM_OK, GM_E_INVALID, GM_E_LOGIC, GM_E_RUNTIME, GM_E_DEVICE = range(5)
GM_PRECISION_FP64, GM_PRECISION_MIXED = 0, 1


class Params(C.Structure):
    _fields_ = [("Nr", C.c_int32), ("L", C.c_double), ("dt", C.c_double),
                ("epsilon", C.c_double), ("l_max", C.c_int32), ("mu", C.c_double),
                ("M_c", C.c_double), ("q", C.c_double), ("a", C.c_double),
                ("R", C.c_double), ("counter", C.c_int32), ("absorption", C.c_int32),
                ("save_every", C.c_int32), ("ramp_periods", C.c_double),
                ("epsilon_smear", C.c_double)]


class Flags(C.Structure):
    _fields_ = [("evolve_orbit", C.c_int32), ("ramp_secondary_potential", C.c_int32),
                ("steps_per_period", C.c_double), ("use_block_preconditioner", C.c_int32),
                ("power_average_orbits", C.c_double), ("precision", C.c_int32),
                ("use_graphs", C.c_int32), ("apply_tile", C.c_int32)]


SCALAR_NAMES = ("t", "theta", "phase", "R", "M1", "M2", "q", "R1", "R2", "Omega", "P",
                "dt", "p_bin", "Omega_0", "last_solver_residual")


class Scalars(C.Structure):
    _fields_ = [(n, C.c_double) for n in SCALAR_NAMES] + [
        ("last_solver_iterations", C.c_int32), ("sink_active", C.c_int32)]


class Diag(C.Structure):
    _fields_ = [("energy", C.c_double), ("p_bin", C.c_double), ("p_bin_inst", C.c_double),
                ("dV_bin_dt", C.c_double), ("cloud_mass", C.c_double),
                ("absorbed", C.c_double),
                ("flux", C.POINTER(C.c_double)), ("energy_flux", C.POINTER(C.c_double)),
                ("norms", C.POINTER(C.c_double)), ("energies", C.POINTER(C.c_double))]


class GMError(RuntimeError):
    """A failed gm_* call; `code` is the C status (GM_E_*)."""

    def __init__(self, code, message):
        super().__init__(message)
        self.code = code


class InvalidArgument(GMError, ValueError):
    pass


class LogicError(GMError):
    pass


_ERRORS = {GM_E_INVALID: InvalidArgument, GM_E_LOGIC: LogicError}

_DP = C.POINTER(C.c_double)


def bind(lib, prefix):
    """Declare argument and return types of the entry points of `lib`."""
    vp = C.c_void_p

    def f(name, res, *args):
        fn = getattr(lib, prefix + name)
        fn.restype = res
        fn.argtypes = list(args)
        return fn

    f("create", C.c_int, C.POINTER(Params), C.POINTER(vp))
    f("destroy", None, vp)
    f("last_error", C.c_char_p, vp)
    f("n_modes", C.c_int, vp)
    f("modes", C.c_int, vp, C.POINTER(C.c_int32))
    f("grid", C.c_int, vp, _DP)
    f("set_flags", C.c_int, vp, C.POINTER(Flags))
    f("get_flags", C.c_int, vp, C.POINTER(Flags))
    f("configure_isolated_black_hole", C.c_int, vp, C.c_double)
    f("set_initial_state_full", C.c_int, vp, _DP, C.c_int64)
    f("initialize", C.c_int, vp)
    f("build_damping_mask", C.c_int, vp, C.c_double, C.c_double, C.c_int32)
    f("build_cap", C.c_int, vp, C.c_double, C.c_double, C.c_int32)
    f("begin_run", C.c_int, vp, C.c_double, C.c_int32)
    f("step", C.c_int, vp, C.c_int32, C.POINTER(Diag))
    f("get_state", C.c_int, vp, _DP)
    f("set_state", C.c_int, vp, _DP)
    f("get_scalars", C.c_int, vp, C.POINTER(Scalars))
    f("set_scalars", C.c_int, vp, C.POINTER(Scalars))
    f("apply_H", C.c_int, vp, _DP, _DP)
    f("apply_sink", C.c_int, vp, _DP, _DP)
    f("schrodinger_rhs", C.c_int, vp, _DP, _DP)
    f("sink_strength", C.c_double, vp, C.c_int32)
    f("save_checkpoint", C.c_int, vp, C.c_char_p, C.c_int32)
    f("load_checkpoint", C.c_int, vp, C.c_char_p, C.c_int32, C.POINTER(C.c_int32))
    f("restore_from_diagnostics", C.c_int, vp, C.c_int32, _DP)
    f("set_restart_thresholds", C.c_int, vp, C.c_double, C.c_double)
    f("restart_count", C.c_int, vp, C.POINTER(C.c_int32))
    return lib


def _ptr(a):
    return a.ctypes.data_as(_DP)


class Solver:
    """One solver handle; keyword arguments are the constructor's
    (include/CrankNicolsonModes.hpp:84)."""

    def __init__(self, lib, prefix, Nr, L, dt, epsilon, l_max, mu, M_c, q, a, R,
                 counter=False, absorption=False, save_every=50, ramp_periods=0.0,
                 epsilon_smear=2.0):
        self._lib, self._p = lib, prefix
        p = Params(int(Nr), float(L), float(dt), float(epsilon), int(l_max), float(mu),
                   float(M_c), float(q), float(a), float(R), int(bool(counter)),
                   int(bool(absorption)), int(save_every), float(ramp_periods),
                   float(epsilon_smear))
        h = C.c_void_p()
        rc = self._fn("create")(C.byref(p), C.byref(h))
        if rc != GM_OK:
            msg = self._fn("last_error")(None).decode()
            raise _ERRORS.get(rc, GMError)(rc, msg)
        self._h = h
        self.params = p
        self.Nr = int(Nr)
        self.n_modes = self._fn("n_modes")(h)
        lm = np.zeros(2 * self.n_modes, dtype=np.int32)
        self._fn("modes")(h, lm.ctypes.data_as(C.POINTER(C.c_int32)))
        self.modes = [(int(lm[2 * j]), int(lm[2 * j + 1])) for j in range(self.n_modes)]
        self.r = np.zeros(self.Nr)
        self._fn("grid")(h, _ptr(self.r))
        self.dr = float(L) / (int(Nr) - 1)
        self.isolated = False

    def _fn(self, name):
        return getattr(self._lib, self._p + name)

    def _check(self, rc):
        if rc != GM_OK:
            msg = self._fn("last_error")(self._h).decode()
            raise _ERRORS.get(rc, GMError)(rc, msg)

    def close(self):
        if getattr(self, "_h", None):
            self._fn("destroy")(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def size(self):
        return self.Nr * self.n_modes

    # flags ---------------------------------------------------------------
    def get_flags(self):
        f = Flags()
        self._check(self._fn("get_flags")(self._h, C.byref(f)))
        return f

    def set_flags(self, **kw):
        f = self.get_flags()
        for k, v in kw.items():
            setattr(f, k, v)
        self._check(self._fn("set_flags")(self._h, C.byref(f)))

    # lifecycle -------------------------------------------------------------
    def configure_isolated_black_hole(self, x):
        self._check(self._fn("configure_isolated_black_hole")(self._h, float(x)))
        self.isolated = True

    def set_initial_state_full(self, psi):
        a = np.ascontiguousarray(np.asarray(psi, dtype=np.complex128).ravel())
        self._check(self._fn("set_initial_state_full")(
            self._h, a.view(np.float64).ctypes.data_as(_DP), a.size))

    def initialize(self):
        self._check(self._fn("initialize")(self._h))

    def build_damping_mask(self, f, W, n):
        self._check(self._fn("build_damping_mask")(self._h, float(f), float(W), int(n)))

    def build_cap(self, f, W, n):
        self._check(self._fn("build_cap")(self._h, float(f), float(W), int(n)))

    def begin_run(self, r_ex, n_steps=2):
        self._check(self._fn("begin_run")(self._h, float(r_ex), int(n_steps)))

    def step(self, step_id=1, diag=True):
        if not diag:
            self._check(self._fn("step")(self._h, int(step_id), None))
            return None
        K = self.n_modes
        arrs = {k: np.zeros(K) for k in ("flux", "energy_flux", "norms", "energies")}
        d = Diag()
        for k, a in arrs.items():
            setattr(d, k, _ptr(a))
        self._check(self._fn("step")(self._h, int(step_id), C.byref(d)))
        out = {k: getattr(d, k) for k in ("energy", "p_bin", "p_bin_inst", "dV_bin_dt",
                                         "cloud_mass", "absorbed")}
        out.update(arrs)
        return out

    # checkpoint -------------------------------------------------------------
    def save_checkpoint(self, path, step_id):
        """save_checkpoint (restart.cpp:139-176) to `path`."""
        self._check(self._fn("save_checkpoint")(self._h, str(path).encode(), int(step_id)))

    def load_checkpoint(self, path, step_id):
        """load_checkpoint (restart.cpp:178-233): True when restored."""
        ok = C.c_int32(0)
        self._check(self._fn("load_checkpoint")(self._h, str(path).encode(), int(step_id), C.byref(ok)))
        return bool(ok.value)

    def restore_from_diagnostics(self, rows):
        """reconstruct_state_from_diagnostics (restart.cpp:238-285): rows is
        (n, 8) = t, R, theta, Omega, M1, q, p_bin, p_bin_inst per step."""
        a = np.ascontiguousarray(np.asarray(rows, dtype=np.float64).reshape(-1, 8))
        self._check(self._fn("restore_from_diagnostics")(self._h, a.shape[0], _ptr(a)))

    # state -----------------------------------------------------------------
    @property
    def psi(self):
        a = np.zeros(self.size, dtype=np.complex128)
        self._check(self._fn("get_state")(self._h, a.view(np.float64).ctypes.data_as(_DP)))
        return a

    @psi.setter
    def psi(self, value):
        a = np.ascontiguousarray(np.asarray(value, dtype=np.complex128).ravel())
        assert a.size == self.size
        self._check(self._fn("set_state")(self._h, a.view(np.float64).ctypes.data_as(_DP)))

    def scalars(self):
        s = Scalars()
        self._check(self._fn("get_scalars")(self._h, C.byref(s)))
        out = {n: getattr(s, n) for n in SCALAR_NAMES}
        out["last_solver_iterations"] = s.last_solver_iterations
        out["sink_active"] = s.sink_active
        return out

    def set_scalars(self, **kw):
        s = Scalars()
        self._check(self._fn("get_scalars")(self._h, C.byref(s)))
        for k, v in kw.items():
            setattr(s, k, v)
        self._check(self._fn("set_scalars")(self._h, C.byref(s)))

    def _op(self, name, x):
        a = np.ascontiguousarray(np.asarray(x, dtype=np.complex128).ravel())
        out = np.zeros_like(a)
        self._check(self._fn(name)(self._h, a.view(np.float64).ctypes.data_as(_DP),
                                   out.view(np.float64).ctypes.data_as(_DP)))
        return out

    def apply_H(self, x):
        return self._op("apply_H", x)

    def apply_sink(self, x):
        return self._op("apply_sink", x)

    def schrodinger_rhs(self, x):
        return self._op("schrodinger_rhs", x)

    def set_restart_thresholds(self, eigen_eps2, sink_factor):
        """BiCGSTAB restart thresholds as fractions of |r0|^2 (a test hook)."""
        self._check(self._fn("set_restart_thresholds")(self._h, float(eigen_eps2), float(sink_factor)))

    def restart_count(self):
        """BiCGSTAB restarts of the last solve."""
        n = C.c_int32()
        self._check(self._fn("restart_count")(self._h, C.byref(n)))
        return n.value

    def sink_strength(self, body):
        return self._fn("sink_strength")(self._h, int(body))

    def __getattr__(self, name):
        if name in SCALAR_NAMES or name in ("last_solver_iterations", "sink_active"):
            return self.scalars()[name]
        raise AttributeError(name)
