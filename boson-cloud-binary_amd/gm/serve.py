"""Serving: a request loop over device-resident solvers (gm_server_* in gm.h).

`Server` wraps the native loop: requests queue in arrival order, are admitted
by HBM footprint and free slots, keep their state on the device, and step as
soon as their own stream is idle, so steps of different sizes from different
requests overlap on the GPU and a request that arrives mid-run joins at once.

`served_stream` drives it with Poisson arrivals of synthetic binaries: mass
ratio log-uniform in [0.05, 1], separation uniform in [20, 60], boson mass
uniform in [0.1, 0.3], N from {4096, 8192, 16384, 32768} with L = N / 4, l_max
from {8, 12, 16, 20, 24}, back-reaction on (gravitational-wave decay and
accretion on both bodies), cloud mass uniform in [0.001, 0.05], a startup
ramp of two orbital periods, 200-2000 steps of dt = 0.25, mixed precision.
Initial state of each request: the hydrogenic 2p (n=2, l=1, m=1) atom of its
own boson mass at the centre of mass plus a seeded 1e-3 complex Gaussian
perturbation in every mode (the perturbation's noise is drawn once per
(N, l_max) shape). Every `probe_every`-th request reports its state after
`probe_step` steps, for a spot check against the FP64 oracle by the caller.
"""
import functools
import ctypes as C
import math
import time

import numpy as np

from . import capi, initial

GM_SERVE_QUEUED, GM_SERVE_RUNNING, GM_SERVE_DONE, GM_SERVE_FAILED = 0, 1, 2, 3


class Request(C.Structure):
    _fields_ = [("params", capi.Params), ("flags", capi.Flags), ("psi0", C.POINTER(C.c_double)),
                ("mask_L_factor", C.c_double), ("mask_W", C.c_double), ("mask_n", C.c_int32),
                ("r_ex", C.c_double), ("n_steps", C.c_int32), ("probe_step", C.c_int32),
                ("probe_psi", C.POINTER(C.c_double))]


class ServeStats(C.Structure):
    _fields_ = [("steps", C.c_int64), ("mode_node_steps", C.c_double), ("busy_seconds", C.c_double),
                ("p50_ms", C.c_double), ("p99_ms", C.c_double), ("max_ms", C.c_double),
                ("queued", C.c_int32), ("active", C.c_int32), ("hbm_in_use", C.c_double),
                ("completed", C.c_int64), ("req_p50_ms", C.c_double), ("req_p99_ms", C.c_double),
                ("req_max_ms", C.c_double)]


def _bind(lib):
    vp = C.c_void_p
    lib.gm_footprint.argtypes = [C.POINTER(capi.Params), C.POINTER(C.c_double)]
    lib.gm_server_create.argtypes = [C.c_double, C.c_int32, C.POINTER(vp)]
    lib.gm_server_destroy.argtypes = [vp]
    lib.gm_server_destroy.restype = None
    lib.gm_server_submit.argtypes = [vp, C.POINTER(Request), C.POINTER(C.c_int64)]
    lib.gm_server_run.argtypes = [vp, C.c_double, C.POINTER(C.c_int32)]
    lib.gm_server_status.argtypes = [vp, C.c_int64, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]
    lib.gm_server_take.argtypes = [vp, C.c_int64, C.POINTER(C.c_double), C.POINTER(capi.Scalars), C.c_char_p,
                                   C.c_int32]
    lib.gm_server_stats.argtypes = [vp, C.POINTER(ServeStats)]


def _params(kw):
    p = capi.Params()
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def default_flags(**kw):
    f = capi.Flags(evolve_orbit=1, ramp_secondary_potential=1, steps_per_period=50.0,
                   use_block_preconditioner=1, power_average_orbits=2.0, precision=0, use_graphs=1,
                   apply_tile=0)
    for k, v in kw.items():
        setattr(f, k, v)
    return f


class Server:
    """gm_server: hbm_budget in bytes, at most max_active solvers at once."""

    def __init__(self, hbm_budget, max_active=16):
        from . import lib
        self._lib = lib()
        _bind(self._lib)
        h = C.c_void_p()
        rc = self._lib.gm_server_create(float(hbm_budget), int(max_active), C.byref(h))
        if rc != 0:
            raise capi.InvalidArgument(rc, "gm_server_create")
        self._h = h
        self._keep = {}
        self._probes = {}

    def close(self):
        if getattr(self, "_h", None):
            self._lib.gm_server_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def footprint(self, **params):
        b = C.c_double()
        if self._lib.gm_footprint(C.byref(_params(params)), C.byref(b)) != 0:
            raise capi.InvalidArgument(1, "gm_footprint")
        return b.value

    def submit(self, params, psi0, n_steps, r_ex, flags=None, mask=None, probe_step=0):
        """Queue a request (params: constructor keywords; mask: (L_factor, W, n));
        probe_step > 0 keeps the state after that step (probe())."""
        a = np.ascontiguousarray(np.asarray(psi0, dtype=np.complex128).ravel())
        r = Request()
        r.params = _params(params)
        r.flags = flags if flags is not None else default_flags()
        r.psi0 = a.view(np.float64).ctypes.data_as(C.POINTER(C.c_double))
        if mask is not None:
            r.mask_L_factor, r.mask_W, r.mask_n = mask
        r.r_ex = float(r_ex)
        r.n_steps = int(n_steps)
        K = (params["l_max"] + 1) * (params["l_max"] + 2) // 2
        probe = None
        if probe_step > 0:
            probe = np.full(K * params["Nr"], np.nan, dtype=np.complex128)
            r.probe_step = int(probe_step)
            r.probe_psi = probe.view(np.float64).ctypes.data_as(C.POINTER(C.c_double))
        t = C.c_int64()
        rc = self._lib.gm_server_submit(self._h, C.byref(r), C.byref(t))
        if rc != 0:
            raise capi.InvalidArgument(rc, "request does not fit the HBM budget or is invalid")
        self._keep[t.value] = (K, params["Nr"])
        if probe is not None:
            self._probes[t.value] = probe
        return t.value

    def probe(self, ticket):
        """The state a request reported after its probe step (NaN until then)."""
        return self._probes[ticket]

    def run(self, seconds):
        done = C.c_int32()
        self._lib.gm_server_run(self._h, float(seconds), C.byref(done))
        return done.value

    def status(self, ticket):
        st, n = C.c_int32(), C.c_int32()
        if self._lib.gm_server_status(self._h, int(ticket), C.byref(st), C.byref(n)) != 0:
            raise KeyError(ticket)
        return st.value, n.value

    def take(self, ticket):
        """(psi, scalars) of a finished request; raises for a failed one."""
        K, N = self._keep.pop(ticket)
        psi = np.zeros(K * N, dtype=np.complex128)
        sc = capi.Scalars()
        err = C.create_string_buffer(512)
        rc = self._lib.gm_server_take(self._h, int(ticket), psi.view(np.float64).ctypes.data_as(C.POINTER(C.c_double)),
                                      C.byref(sc), err, 512)
        if rc != 0:
            raise capi.GMError(rc, err.value.decode())
        return psi, {k: getattr(sc, k) for k, _ in capi.Scalars._fields_}

    def stats(self):
        s = ServeStats()
        self._lib.gm_server_stats(self._h, C.byref(s))
        return {k: getattr(s, k) for k, _ in ServeStats._fields_}


@functools.lru_cache(maxsize=16)
def _noise(Nr, K, seed):
    rng = np.random.default_rng(seed)
    noise = rng.standard_normal((K, Nr)) + 1j * rng.standard_normal((K, Nr))
    noise[:, 0] = noise[:, -1] = 0
    noise.setflags(write=False)
    return noise


def atom_state(Nr, L, l_max, mu, seed=0, rel=1e-3, epsilon=0.0):
    """2p (n=2, l=1, m=1) hydrogenic atom at the origin (Bohr radius 1/mu^2)
    plus a seeded complex Gaussian perturbation of relative amplitude `rel` in
    every mode (the recipe of initial.perturbed_state, with the noise of a
    shape drawn once), normalised to dr |U|^2 = 1."""
    dr = L / (Nr - 1)
    r = epsilon + dr * np.arange(Nr)
    K = (l_max + 1) * (l_max + 2) // 2
    U = np.zeros((K, Nr), dtype=np.complex128)
    U[2] = r * initial.hydrogen_radial(2, 1, r, 1.0 / mu ** 2)  # mode (l, m) = (1, 1)
    U[:, 0] = U[:, -1] = 0
    U /= math.sqrt(np.vdot(U, U).real * dr)
    rms = math.sqrt(np.mean(np.abs(U) ** 2))
    U = U + rel * rms / math.sqrt(2) * _noise(Nr, K, seed)
    U /= math.sqrt(np.vdot(U, U).real * dr)
    return U.ravel()


def draw_request(rng, seed):
    """One served binary (parameters drawn as in the module docstring)."""
    q = math.exp(rng.uniform(math.log(0.05), math.log(1.0)))
    N = int(rng.choice([4096, 8192, 16384, 32768]))
    lm = int(rng.choice([8, 12, 16, 20, 24]))
    params = dict(Nr=N, L=N / 4.0, dt=0.25, epsilon=0.0, l_max=lm, mu=float(rng.uniform(0.1, 0.3)),
                  M_c=float(rng.uniform(0.001, 0.05)), q=q, a=0.0, R=float(rng.uniform(20.0, 60.0)),
                  counter=False, absorption=True, save_every=1000, ramp_periods=2.0, epsilon_smear=2.0)
    steps = int(rng.integers(200, 2001))
    return params, steps, seed


def served_stream(duration=20.0, rate=20.0, seed=0, hbm_budget=160e9, max_active=16, drain=False,
                  sizes=None, steps=None, probe_every=20, probe_step=2, drain_limit=None, precision=1,
                  probe_max_work=None):
    """Poisson(rate) arrivals for `duration` seconds of wall time, served as
    they arrive (and, with `drain`, until every request is done or
    `drain_limit` seconds have passed). Returns the server statistics
    (per-step p50/p99, request p50/p99 from submission to completion,
    aggregate mode-node-steps/s over the serving time, admission counts) and,
    under "probes", every probe_every-th request (the first included) with its
    parameters, initial state, the state it reported after probe_step steps
    and, when it finished, its final state and scalars; with probe_max_work
    only requests with K^2 N at most that count (an oracle of bounded host
    cost) are probe candidates. `sizes`/`steps` override the drawn (N, l_max)
    and step counts (`steps` a number or an inclusive range)."""
    rng = np.random.default_rng(seed)
    srv = Server(hbm_budget, max_active)
    t_arr, arrivals = 0.0, []
    while t_arr < duration:
        t_arr += rng.exponential(1.0 / rate)
        params, n, s = draw_request(rng, len(arrivals))
        if sizes is not None:
            params["Nr"], params["l_max"] = sizes[len(arrivals) % len(sizes)]
            params["L"] = params["Nr"] / 4.0
        if steps is not None:
            n = steps if np.isscalar(steps) else int(rng.integers(steps[0], steps[1] + 1))
        arrivals.append((t_arr, params, n, s))
    # the clients' initial states, built before the stream starts: generating
    # inputs is not part of the served work
    states = [atom_state(p["Nr"], p["L"], p["l_max"], p["mu"]) for _, p, _, _ in arrivals]
    flags = default_flags(precision=precision)
    t0 = time.perf_counter()
    submitted, tickets, n_probe_cand = 0, [], 0
    while True:
        now = time.perf_counter() - t0
        while submitted < len(arrivals) and arrivals[submitted][0] <= now:
            _, params, n, s = arrivals[submitted]
            probe = 0
            K = (params["l_max"] + 1) * (params["l_max"] + 2) // 2
            fits = probe_max_work is None or K * K * params["Nr"] <= probe_max_work
            if probe_every and fits and n >= probe_step and n_probe_cand % probe_every == 0:
                probe = probe_step
            n_probe_cand += fits
            tickets.append(srv.submit(params, states[submitted], n, 0.8 * params["L"], flags=flags,
                                      mask=(0.85, 1.0, 6), probe_step=probe))
            submitted += 1
        nxt = arrivals[submitted][0] - now if submitted < len(arrivals) else 0.05
        srv.run(max(0.002, min(nxt, 0.05)))
        now = time.perf_counter() - t0
        if now >= duration and submitted >= len(arrivals):
            st = srv.stats()
            idle = st["active"] + st["queued"] == 0
            if not drain or idle or (drain_limit is not None and now >= duration + drain_limit):
                break
    st = srv.stats()
    st["submitted"] = submitted
    st["wall_seconds"] = time.perf_counter() - t0
    st["mode_node_steps_per_s"] = st["mode_node_steps"] / max(st["busy_seconds"], 1e-9)
    st["probes"] = []
    for k, t in enumerate(tickets):
        if t in srv._probes:
            status, done = srv.status(t)
            _, params, n, _ = arrivals[k]
            rec = {"ticket": t, "params": params, "n_steps": n, "psi0": states[k], "probe_step": probe_step,
                   "steps_done": done, "psi": srv.probe(t).copy() if done >= probe_step else None,
                   "final": None}
            if status == GM_SERVE_DONE:
                rec["final"], rec["scalars"] = srv.take(t)
            st["probes"].append(rec)
    srv.close()
    return st
