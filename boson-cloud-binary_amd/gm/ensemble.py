"""Batched ragged ensemble: many independent fixed-orbit binaries stepped
together, either as one ragged grid per launch of the step over all members
(gm_batch, graph-captured) or with each member on its own HIP stream
(gm_ensemble_step); the apply tile is autotuned per shape.

Members are drawn from one seed: mass ratio log-uniform in [0.05, 1],
separation uniform in [20, 60], boson mass uniform in [0.1, 0.3], N from
{4096, 8192, 16384, 32768} with L = N / 4, l_max from {8, 12, 16, 20, 24},
dt = 0.25, absorbers on, damping mask (0.85, 1, 6), r_ex = 0.8 L. Initial
state of each member: the hydrogenic 2p atom (n=2, l=1, m=1) of its own boson
mass on the primary, translated to the centre of mass, plus a 1e-3 complex
Gaussian perturbation seeded by the member's index.
"""
import math
import time

import numpy as np

from . import GM_PRECISION_MIXED, Batch, CrankNicolsonModes, ensemble_step
from . import initial


def draw_members(n, seed=0):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        q = math.exp(rng.uniform(math.log(0.05), math.log(1.0)))
        R = float(rng.uniform(20.0, 60.0))
        mu = float(rng.uniform(0.1, 0.3))
        N = int(rng.choice([4096, 8192, 16384, 32768]))
        lm = int(rng.choice([8, 12, 16, 20, 24]))
        out.append(dict(Nr=N, L=N / 4.0, dt=0.25, epsilon=0.0, l_max=lm, mu=mu, M_c=0.01, q=q, a=0.0, R=R,
                        counter=False, absorption=True, save_every=1000, ramp_periods=0.0, epsilon_smear=2.0))
    return out


_STATES = {}


def member_state(p, index):
    """The member's 2p atom on the primary (offset q/(1+q) R from the centre
    of mass) plus the perturbation seeded by its index (kept for the process:
    the quadrature of the translated atom is the slow part of a build)."""
    key = (index,) + tuple(sorted(p.items()))
    if key not in _STATES:
        _STATES[key] = _member_state(p, index)
    return _STATES[key]


def _member_state(p, index):
    dr = p["L"] / (p["Nr"] - 1)
    r = p["epsilon"] + dr * np.arange(p["Nr"])
    base = initial.translated_atom(r, p["l_max"], p["mu"], p["q"] / (1 + p["q"]) * p["R"])
    base /= math.sqrt(np.vdot(base, base).real * dr)
    K = (p["l_max"] + 1) * (p["l_max"] + 2) // 2
    return initial.perturbed_state(base, p["Nr"], K, rel=1e-3, seed=index, dr=dr)


def build(members, precision=GM_PRECISION_MIXED, states=None):
    solvers = []
    for i, p in enumerate(members):
        s = CrankNicolsonModes(**p)
        s.set_flags(evolve_orbit=0, precision=precision)
        s.set_initial_state_full(states[i] if states is not None else member_state(p, i))
        s.initialize()
        s.build_damping_mask(0.85, 1.0, 6)
        s.begin_run(0.8 * p["L"], 100000)
        solvers.append(s)
    return solvers


def run(n=64, steps=10, warmup=3, seed=0, precision=GM_PRECISION_MIXED, mode="batch", lanes=1):
    """Aggregate mode.node-steps/s of `steps` ensemble steps after `warmup`;
    mode "batch" (ragged grids per launch over `lanes` member subsets,
    gm_batch) or "streams" (a stream per member, gm_ensemble_step)."""
    members = draw_members(n, seed)
    t0 = time.perf_counter()
    solvers = build(members, precision)
    t_build = time.perf_counter() - t0
    batch = Batch(solvers, lanes) if mode == "batch" else None

    def one(k):
        if batch is not None:
            batch.step(k, diag=False)
        else:
            ensemble_step(solvers, k, diag=False)

    k = 0
    for _ in range(warmup):
        k += 1
        one(k)
    t0 = time.perf_counter()
    for _ in range(steps):
        k += 1
        one(k)
    el = time.perf_counter() - t0
    work = sum(s.n_modes * s.Nr for s in solvers) * steps
    res = {"mode": mode, "lanes": lanes if mode == "batch" else n, "members": n, "steps": steps, "warmup": warmup, "seconds": el, "build_seconds": t_build,
           "mode_node_steps_per_s": work / el, "ms_per_ensemble_step": el / steps * 1e3,
           "mode_nodes_per_step": work / steps, "iterations_per_step": float(np.mean(
               [s.last_solver_iterations for s in solvers]))}
    if batch is not None:
        batch.close()
    for s in solvers:
        s.close()
    return res
