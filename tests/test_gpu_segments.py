"""The segmented preconditioner sweeps against the FP64 oracle (needs an MI355X).

Above one tile (2048 nodes) each mode's row is split into segments of
PSEG = 1664 nodes, solved over windows with PHALO = 192 halo nodes on either
side whose dropped carries decay by the product of |l / pivot| over a halo;
k_factor checks that product per mode and boundary, and a mode that fails
the check is solved whole by one workgroup. Both paths are stepped here
against the oracle's sequential Thomas solve: psi within 1e-10 (FP64) and
1e-8 (mixed precision), BiCGSTAB iterations within one.
"""
import math
import sys

import numpy as np
import pytest

import gm
import oracle

pytestmark = pytest.mark.gpu

PSEG, PHALO = 1664, 192


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def kinetic_ratio(dt, mu, dr):
    """|l / pivot| of the stencil-only CN matrix far along the row: the fixed
    point of pivot = d - l^2 / pivot with d = 1 + i dt / (2 mu dr^2) and
    l = -i dt / (4 mu dr^2) (the potential and the centrifugal term only
    enlarge |d|)."""
    c = dt / (4.0 * mu * dr * dr)
    d, ell = complex(1.0, 2.0 * c), complex(0.0, -c)
    p = d
    for _ in range(10000):
        p = d - ell * ell / p
    return abs(ell / p)


def spread_state(cn, seed):
    """Seeded smooth content over the whole grid (every segment and halo),
    zero at the Dirichlet nodes."""
    rng = np.random.default_rng(seed)
    Nr = cn.Nr
    i = np.arange(Nr)
    psi = np.zeros(cn.size, dtype=complex)
    for j in range(cn.n_modes):
        f = np.zeros(Nr, dtype=complex)
        for k in range(1, 6):
            f += (rng.standard_normal() + 1j * rng.standard_normal()) * np.sin(np.pi * k * i / (Nr - 1) + rng.uniform(0, 0.1))
        f *= np.sin(np.pi * i / (Nr - 1)) / (1.0 + j)
        f[0] = f[-1] = 0.0
        psi[j * Nr:(j + 1) * Nr] = f
    return psi / math.sqrt(np.vdot(psi, psi).real * cn.dr)


def step_against_oracle(args, steps, seed):
    g64, g32 = gm.CrankNicolsonModes(**args), gm.CrankNicolsonModes(**args)
    o = oracle.Oracle(**args)
    psi = spread_state(o, seed)
    g32.set_flags(precision=gm.GM_PRECISION_MIXED)
    for s in (g64, g32, o):
        s.set_flags(evolve_orbit=0)
        s.set_initial_state_full(psi)
        s.initialize()
        s.build_damping_mask(0.85, 1.0, 6)
        s.begin_run(0.8 * args["L"], 6)
    for n in range(1, steps + 1):
        for s in (g64, g32, o):
            s.step(n, diag=False)
        it = o.scalars()["last_solver_iterations"]
        assert abs(g64.scalars()["last_solver_iterations"] - it) <= 1
        assert abs(g32.scalars()["last_solver_iterations"] - it) <= 1
    assert rel(g64.psi, o.psi) < 1e-10
    assert rel(g32.psi, o.psi) < 1e-8
    return g64


def test_segment_windows_with_a_sink_across_a_boundary():
    """Three segments per mode; the secondary's absorber sits on the first
    segment boundary (node 1664), so two segments of every mode hold
    sink-support nodes and both store Woodbury partials."""
    print("segment windows, sink across a boundary", flush=True)
    dr = 0.3
    args = dict(Nr=5001, L=5000 * dr, dt=0.5, epsilon=0.0, l_max=4, mu=0.2, M_c=0.01, q=0.5, a=0.0,
                R=1.5 * PSEG * dr, counter=False, absorption=True, save_every=1000, ramp_periods=0.0,
                epsilon_smear=2.0)
    assert kinetic_ratio(args["dt"], args["mu"], dr) ** PHALO < 2.0 ** -56
    # the secondary at R / (1 + q) lies on node PSEG
    assert abs(args["R"] / (1 + args["q"]) / dr - PSEG) < 1
    g = step_against_oracle(args, 4, seed=11)
    assert g.Nr > 2 * PSEG


def test_whole_row_fallback_for_slow_decay():
    """dt = 60 makes |l / pivot| close to 1: no halo of 192 nodes drops its
    carry below 2^-56, every mode is solved whole by its segment-0
    workgroup."""
    print("whole-row fallback", flush=True)
    dr = 0.3
    args = dict(Nr=4001, L=4000 * dr, dt=60.0, epsilon=0.0, l_max=2, mu=0.2, M_c=0.01, q=0.5, a=0.0, R=40.0,
                counter=False, absorption=True, save_every=1000, ramp_periods=0.0, epsilon_smear=2.0)
    assert kinetic_ratio(args["dt"], args["mu"], dr) ** PHALO > 2.0 ** -40
    step_against_oracle(args, 3, seed=12)


@pytest.mark.parametrize("Nr", [2049, 3330])
def test_short_last_segment(Nr):
    """A last segment shorter than a halo (2049 nodes: segments of 1664 and
    385) and one that ends exactly on a segment size (3330 = 2 x 1664 + 2)."""
    print(f"short last segment, Nr={Nr}", flush=True)
    dr = 0.3
    args = dict(Nr=Nr, L=(Nr - 1) * dr, dt=0.5, epsilon=0.0, l_max=3, mu=0.2, M_c=0.01, q=0.5, a=0.0, R=30.0,
                counter=False, absorption=True, save_every=1000, ramp_periods=0.0, epsilon_smear=2.0)
    step_against_oracle(args, 3, seed=13 + Nr)
