"""Batched ragged ensemble: many independent fixed-orbit binaries stepped
together by gm_ensemble_step (each member on its own HIP stream, captured
graphs replayed per member, apply tile autotuned per shape).

Members are drawn from one seed: mass ratio log-uniform in [0.05, 1],
separation uniform in [20, 60], boson mass uniform in [0.1, 0.3], N from
{4096, 8192, 16384, 32768} with L = N / 4, l_max from {8, 12, 16, 20, 24},
dt = 0.25, absorbers on, damping mask (0.85, 1, 6), r_ex = 0.8 L. Initial
state: the 2p atom at the origin plus a seeded 1e-3 perturbation, one per
(N, l_max) shape.
"""
import math
import time

import numpy as np

from . import GM_PRECISION_MIXED, CrankNicolsonModes, ensemble_step
from .serve import atom_state


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


def build(members, precision=GM_PRECISION_MIXED):
    solvers, states = [], {}
    for p in members:
        shape = (p["Nr"], p["l_max"])
        if shape not in states:
            states[shape] = atom_state(p["Nr"], p["L"], p["l_max"], p["mu"], len(states))
        s = CrankNicolsonModes(**p)
        s.set_flags(evolve_orbit=0, precision=precision)
        s.set_initial_state_full(states[shape])
        s.initialize()
        s.build_damping_mask(0.85, 1.0, 6)
        s.begin_run(0.8 * p["L"], 100000)
        solvers.append(s)
    return solvers


def run(n=64, steps=10, warmup=3, seed=0, precision=GM_PRECISION_MIXED):
    """Aggregate mode.node-steps/s of `steps` ensemble steps after `warmup`."""
    members = draw_members(n, seed)
    t0 = time.perf_counter()
    solvers = build(members, precision)
    t_build = time.perf_counter() - t0
    k = 0
    for _ in range(warmup):
        k += 1
        ensemble_step(solvers, k, diag=False)
    t0 = time.perf_counter()
    for _ in range(steps):
        k += 1
        ensemble_step(solvers, k, diag=False)
    el = time.perf_counter() - t0
    work = sum(s.n_modes * s.Nr for s in solvers) * steps
    res = {"members": n, "steps": steps, "warmup": warmup, "seconds": el, "build_seconds": t_build,
           "mode_node_steps_per_s": work / el, "ms_per_ensemble_step": el / steps * 1e3,
           "mode_nodes_per_step": work / steps, "iterations_per_step": float(np.mean(
               [s.last_solver_iterations for s in solvers]))}
    for s in solvers:
        s.close()
    return res
