"""Parity of the HIP path with the FP64 oracle (needs an MI355X).

Every case runs the same seeded inputs through gm (libgm.so, gfx950 kernels)
and through the oracle, and compares psi, the diagnostics that step() writes
and the BiCGSTAB iteration counts. Tolerances: the plain FP64 drop-in target
is 1e-10 relative L2 on psi; operators agree to rounding (1e-13).
"""
import math
import sys

import numpy as np
import pytest

import gm
import oracle
from test_oracle import fixture_state, generic_state

pytestmark = pytest.mark.gpu


def pair(**kw):
    """The same solver on the GPU and in the oracle."""
    args = dict(Nr=65, L=64.0, dt=2.0, epsilon=0.0, l_max=2, mu=0.2, M_c=0.1, q=0.5, a=0.0, R=20.0,
                counter=False, absorption=False, save_every=1000, ramp_periods=0.0, epsilon_smear=4.0)
    args.update(kw)
    return gm.CrankNicolsonModes(**args), oracle.Oracle(**args)


def both(objs, name, *a, **kw):
    return [getattr(o, name)(*a, **kw) for o in objs]


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def announce(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("l_max,absorb,cap", [(0, True, False), (1, False, True), (2, False, False), (4, True, False),
                                               (3, True, True), (6, False, False)])
def test_apply_H_matches_oracle(l_max, absorb, cap):
    announce(f"apply_H l_max={l_max} absorb={absorb} cap={cap}")
    objs = pair(l_max=l_max, absorption=absorb, Nr=97, L=96.0)
    both(objs, "set_flags", evolve_orbit=0)
    g, o = objs
    both(objs, "set_initial_state_full", generic_state(o))
    both(objs, "initialize")
    if cap:
        both(objs, "build_cap", 0.8, 0.5, 2)
    both(objs, "begin_run", 40.0, 2)
    rng = np.random.default_rng(l_max)
    x = rng.standard_normal(o.size) + 1j * rng.standard_normal(o.size)
    assert rel(g.apply_H(x), o.apply_H(x)) < 1e-13
    if absorb:
        assert rel(g.apply_sink(x), o.apply_sink(x)) < 1e-12
        assert rel(g.schrodinger_rhs(x), o.schrodinger_rhs(x)) < 1e-13


def correctness_pair(scenario):
    """Scenario `scenario` of tests/parallel_correctness.cpp:109-131 on both."""
    nr, dt, smear = 32, 0.2, 10.0
    absorb = scenario in (1, 2)
    P0 = 2 * math.pi / math.sqrt(1.2 / 30.0 ** 3)
    ramp_periods = 2 * dt / P0 if scenario == 3 else 0.0
    objs = pair(Nr=nr, L=120.0, dt=dt, epsilon=0.1, l_max=3, mu=0.2, M_c=1e-20 if scenario == 4 else 0.01, q=0.2,
                R=30.0, counter=scenario == 2, absorption=absorb, save_every=10, ramp_periods=ramp_periods,
                epsilon_smear=smear)
    g, o = objs
    both(objs, "set_flags", evolve_orbit=int(absorb))
    both(objs, "set_scalars", theta=0.3, Omega=o.Omega_0)
    both(objs, "set_initial_state_full", fixture_state(nr, o.n_modes))
    both(objs, "initialize")
    if scenario == 2:
        both(objs, "build_cap", 0.8, 0.01, 3)
        both(objs, "build_damping_mask", 0.8, 0.01, 3)
    both(objs, "begin_run", 60.0, 2)
    return objs


@pytest.mark.parametrize("scenario", range(5))
def test_one_step_matches_oracle(scenario):
    announce(f"one step scenario {scenario}")
    g, o = correctness_pair(scenario)
    dg, do = g.step(1), o.step(1)
    assert rel(g.psi, o.psi) < 1e-10
    sg, so = g.scalars(), o.scalars()
    assert abs(sg["last_solver_iterations"] - so["last_solver_iterations"]) <= 1
    assert sg["last_solver_residual"] <= 1e-9
    for k in ("R", "M1", "M2", "q", "R1", "R2", "Omega", "P", "dt", "theta", "phase"):
        assert sg[k] == pytest.approx(so[k], rel=1e-12, abs=1e-300), k
    for k in ("energy", "p_bin_inst", "dV_bin_dt", "cloud_mass"):
        assert dg[k] == pytest.approx(do[k], rel=1e-9, abs=1e-14 * (1 + abs(do[k]))), k
    for k in ("flux", "energy_flux", "norms", "energies"):
        assert rel(dg[k], do[k]) < 1e-9 or np.linalg.norm(do[k]) < 1e-300, k


@pytest.mark.parametrize("scenario", range(5))
def test_ilut_key_matches_radial_path(scenario):
    """tests/parallel_correctness.cpp:189-193: with use_block_preconditioner
    off (the "ilut" key) the step equals the radial path's to 1e-8 and passes
    the 1e-9 acceptance check. Here the GPU's ILUT step (operator read off the
    device, factor and BiCGSTAB on the host, acceptance, mask and diagnostics
    on the device) against the oracle's radial step, over five steps."""
    announce(f"ilut key scenario {scenario}")
    g, o = correctness_pair(scenario)
    g.set_flags(use_block_preconditioner=0)
    for n in range(1, 6):
        dg, do = g.step(n), o.step(n)
        assert rel(g.psi, o.psi) < 1e-8, n
        assert g.scalars()["last_solver_residual"] <= 1e-9
        assert dg["energy"] == pytest.approx(do["energy"], rel=1e-7, abs=1e-14), n
    assert 1 <= g.scalars()["last_solver_iterations"] <= 10


def test_ilut_key_then_radial_path():
    """Switching the key between steps: an ILUT step followed by radial steps
    tracks the oracle's radial run."""
    announce("ilut then radial")
    g, o = correctness_pair(1)
    g.set_flags(use_block_preconditioner=0)
    g.step(1), o.step(1)
    g.set_flags(use_block_preconditioner=1)
    for n in range(2, 6):
        g.step(n), o.step(n)
    assert rel(g.psi, o.psi) < 1e-8


def run_both(objs, steps, r_ex):
    g, o = objs
    both(objs, "begin_run", r_ex, steps + 1)
    hist = []
    for n in range(1, steps + 1):
        dg, do = g.step(n), o.step(n)
        hist.append((dg, do, g.scalars(), o.scalars()))
    return hist


def check_history(objs, hist, psi_tol=1e-10, diag_tol=1e-8):
    g, o = objs
    assert rel(g.psi, o.psi) < psi_tol
    for dg, do, sg, so in hist:
        assert abs(sg["last_solver_iterations"] - so["last_solver_iterations"]) <= 1
        for k in ("energy", "p_bin_inst", "p_bin", "dV_bin_dt", "cloud_mass", "absorbed"):
            assert abs(dg[k] - do[k]) <= diag_tol * (abs(do[k]) + 1e-12 * abs(do["energy"]) + 1e-300), k
        for k in ("flux", "energy_flux", "norms", "energies"):
            assert np.linalg.norm(dg[k] - do[k]) <= diag_tol * (np.linalg.norm(do[k]) + 1e-300), k
        for k in ("R", "M1", "M2", "Omega", "dt"):
            assert sg[k] == pytest.approx(so[k], rel=1e-11), k


def test_fixed_orbit_with_ramp_mask_and_sinks():
    """A small config-1 shape: 2p-like state, ramp, mask and absorbers."""
    announce("fixed orbit ramp/mask/sinks")
    objs = pair(Nr=257, L=128.0, dt=2.0, l_max=4, mu=0.2, M_c=0.01, q=0.1, R=30.0, absorption=True,
                ramp_periods=0.05, epsilon_smear=2.0)
    g, o = objs
    both(objs, "set_flags", evolve_orbit=0)
    both(objs, "set_initial_state_full", generic_state(o))
    both(objs, "initialize")
    both(objs, "build_damping_mask", 0.85, 1.0, 6)
    hist = run_both(objs, 30, 100.0)
    assert o.scalars()["sink_active"] == 1
    check_history(objs, hist)


def test_counter_rotating_with_cap():
    announce("counter rotating with CAP")
    objs = pair(Nr=129, L=128.0, l_max=3, counter=True, absorption=True, ramp_periods=0.3)
    both(objs, "set_flags", evolve_orbit=0)
    g, o = objs
    both(objs, "set_initial_state_full", generic_state(o))
    both(objs, "initialize")
    both(objs, "build_cap", 0.8, 0.5, 2)
    hist = run_both(objs, 20, 40.0)
    check_history(objs, hist)


def test_evolving_orbit_accretion():
    """tests/physics.cpp:354-382 setup: back-reaction and accretion on both bodies."""
    announce("evolving orbit accretion")
    objs = pair(q=1.0, R=20.0, l_max=2, M_c=0.05, absorption=True)
    g, o = objs
    both(objs, "set_flags", evolve_orbit=1, steps_per_period=16.0)
    both(objs, "set_scalars", dt=o.P / 16)
    both(objs, "set_initial_state_full", generic_state(o))
    both(objs, "initialize")
    hist = run_both(objs, 40, 40.0)
    check_history(objs, hist, diag_tol=1e-8)
    assert g.M1 > 1 and g.M2 > 1


@pytest.mark.parametrize("spp", [256.0, 24.0])
def test_evolving_orbit_mixed_precision_tracks_the_oracle(spp):
    """The evolving orbit in FP64 and in mixed precision: every step rebuilds
    the profiles, the multipole cut, the FP32 and FP64 pivots and the Woodbury
    columns, and the orbit's back-reaction reads the step's diagnostics. Over
    30 steps at l_max = 4 with accretion on both bodies and the mask: FP64
    within 1e-10 of the oracle with iterations +-1; mixed precision within
    1e-8 on psi and 1e-9 on the orbit (R, M1, M2, Omega, theta), every step
    accepted by the FP64 residual check. At 24 steps per period (dt ~ 23) the
    FP32 coupling error lies above that check's 1e-9, so every mixed step is
    solved again with the FP64 operator, and matches the FP64 run."""
    announce(f"evolving orbit, FP64 and mixed precision, {spp} steps per period")
    args = dict(Nr=257, L=128.0, dt=1.0, epsilon=0.0, l_max=4, mu=0.2, M_c=0.05, q=0.5, a=0.0, R=20.0,
                counter=False, absorption=True, save_every=1000, ramp_periods=0.0, epsilon_smear=2.0)
    g64, g32, o = gm.CrankNicolsonModes(**args), gm.CrankNicolsonModes(**args), oracle.Oracle(**args)
    psi = generic_state(o)
    for s in (g64, g32, o):
        s.set_flags(evolve_orbit=1, steps_per_period=spp)
        s.set_scalars(dt=o.P / spp)
        s.set_initial_state_full(psi)
        s.initialize()
        s.build_damping_mask(0.85, 1.0, 6)
        s.begin_run(100.0, 31)
    g32.set_flags(evolve_orbit=1, steps_per_period=spp, precision=gm.GM_PRECISION_MIXED)
    m0 = (o.M1, o.M2)
    for n in range(1, 31):
        for s in (g64, g32, o):
            s.step(n, diag=True)
        it = o.scalars()["last_solver_iterations"]
        assert abs(g64.scalars()["last_solver_iterations"] - it) <= 1, n
        assert g32.last_solver_residual <= 1e-9, n
        if spp < 100:
            assert g32.scalars()["last_solver_iterations"] == g64.scalars()["last_solver_iterations"], n
    so = o.scalars()
    for k in ("R", "M1", "M2", "Omega", "theta"):
        assert g64.scalars()[k] == pytest.approx(so[k], rel=1e-11), k
        assert g32.scalars()[k] == pytest.approx(so[k], rel=1e-9), k
    assert rel(g64.psi, o.psi) < 1e-10
    assert rel(g32.psi, o.psi) < 1e-8
    if spp < 100:
        assert rel(g32.psi, g64.psi) < 1e-13
    assert g64.M1 > m0[0] and g64.M2 > m0[1]  # both bodies accreted


def test_isolated_black_hole():
    announce("isolated BH")
    objs = pair(Nr=257, L=40.0, dt=0.02, l_max=6, mu=0.5, M_c=0.5, q=0.0, R=1.0, absorption=True,
                epsilon_smear=2.0)
    g, o = objs
    both(objs, "configure_isolated_black_hole", 2.0)
    rng = np.random.default_rng(3)
    psi = np.zeros(o.size, dtype=complex)
    for j in range(o.n_modes):
        psi[j * o.Nr + 1:(j + 1) * o.Nr - 1] = o.r[1:-1] * np.exp(-o.r[1:-1] / 4.0) * (rng.standard_normal() + 0.5j)
    psi /= math.sqrt(np.vdot(psi, psi).real * o.dr)
    both(objs, "set_initial_state_full", psi)
    both(objs, "initialize")
    hist = run_both(objs, 8, 30.0)
    check_history(objs, hist)


@pytest.mark.parametrize("l_max", [0, 1])
def test_low_lmax_steps_match_oracle(l_max):
    """One mode (only the l* = 0 self-coupling) and three modes: an evolving
    orbit with active absorbers, 15 steps against the oracle."""
    announce(f"low l_max {l_max}")
    objs = pair(Nr=129, L=128.0, dt=2.0, l_max=l_max, mu=0.2, M_c=0.05, q=0.5, R=20.0, absorption=True,
                ramp_periods=0.05, epsilon_smear=4.0)
    g, o = objs
    both(objs, "set_flags", evolve_orbit=1, steps_per_period=32.0)
    both(objs, "set_initial_state_full", generic_state(o))
    both(objs, "initialize")
    both(objs, "build_damping_mask", 0.85, 1.0, 6)
    hist = run_both(objs, 15, 100.0)
    assert o.scalars()["sink_active"] == 1
    check_history(objs, hist)


def test_exact_preconditioner_one_iteration_on_the_gpu():
    """One mode and no absorbers: the radial preconditioner is A's exact
    inverse, so both paths stop after exactly one BiCGSTAB iteration."""
    announce("exact preconditioner")
    objs = pair(l_max=0, ramp_periods=0.2)
    g, o = objs
    both(objs, "set_flags", evolve_orbit=0)
    both(objs, "set_initial_state_full", generic_state(o))
    both(objs, "initialize")
    both(objs, "begin_run", 40.0, 11)
    for n in range(1, 11):
        g.step(n, diag=False), o.step(n, diag=False)
        assert g.last_solver_iterations == 1 and o.last_solver_iterations == 1, n
    assert rel(g.psi, o.psi) < 1e-12


def test_gpu_unitarity_without_absorption():
    """tests/physics.cpp:106-124 on the GPU path."""
    announce("unitarity")
    g, o = pair(l_max=4, ramp_periods=0.2)
    g.set_flags(evolve_orbit=0)
    g.set_initial_state_full(generic_state(o))
    g.initialize()
    g.begin_run(40, 31)
    n0 = np.vdot(g.psi, g.psi).real
    for step in range(1, 31):
        g.step(step, diag=False)
        assert abs(np.vdot(g.psi, g.psi).real - n0) <= 1e-9 * max(1.0, n0)


@pytest.mark.parametrize("l_max,Nr", [(12, 193), (20, 129), (30, 97)])
def test_apply_H_high_lmax_matches_oracle(l_max, Nr):
    """The coupling stream at many row chunks (m-major chunks, per-column slot
    groups, the multipole cut and its zero LDS rows) against the oracle's
    sum over every Gaunt term, for H and for the FP32 Krylov operator."""
    announce(f"apply_H l_max={l_max} Nr={Nr}")
    objs = pair(l_max=l_max, absorption=True, Nr=Nr, L=float(8 * (Nr - 1)), R=40.0)
    both(objs, "set_flags", evolve_orbit=0)
    g, o = objs
    both(objs, "set_initial_state_full", generic_state(o))
    both(objs, "initialize")
    both(objs, "begin_run", 0.6 * 8 * (Nr - 1), 2)
    rng = np.random.default_rng(l_max)
    x = rng.standard_normal(o.size) + 1j * rng.standard_normal(o.size)
    assert rel(g.apply_H(x), o.apply_H(x)) < 1e-13
    g.set_coupling_tolerance(0.0)
    assert rel(g.apply_H(x), o.apply_H(x)) < 1e-13


CUT_ARGS = dict(Nr=1025, L=2048.0, dt=0.5, epsilon=0.0, mu=0.2, M_c=0.01, q=0.5, a=0.0, R=40.0,
                counter=False, absorption=True, save_every=1000, ramp_periods=0.0, epsilon_smear=2.0)


def cut_pair(l_max, Nr=1025, tol=1e-12):
    """GPU and oracle under a coarse multipole truncation: at tol = 1e-12 of a
    tile's leading profile the 64-node tiles keep from 2 l_max multipoles near
    the bodies down to a handful at the far end, so every chunk's pair list
    runs as whole pairs, as pairs cut after their first group, or both,
    depending on the tile. Returns the GPU's per-tile cut; the GPU's
    untruncated apply of x is taken first when x is given."""
    args = dict(CUT_ARGS, l_max=l_max, Nr=Nr, L=2.0 * (Nr - 1))
    g, o = gm.CrankNicolsonModes(**args), oracle.Oracle(**args)
    psi = generic_state(o)
    for s in (g, o):
        s.set_flags(evolve_orbit=0)
        s.set_initial_state_full(psi)
        s.initialize()
        s.build_damping_mask(0.85, 1.0, 6)
        s.begin_run(0.6 * args["L"], 6)
    g.set_coupling_tolerance(tol)
    lc = g.multipole_cut()
    assert lc.max() == 2 * l_max and lc.min() <= 2 * l_max - 16 and len(set(lc.tolist())) >= 4, lc
    return g, o, lc


@pytest.mark.parametrize("l_max,Nr,tol", [(12, 1025, 1e-12), (20, 513, 1e-12), (12, 1025, 9e-7)])
def test_coarse_multipole_cut_matches_oracle_term_for_term(l_max, Nr, tol):
    """The coupling streams under a cut that falls inside many slot-group
    pairs. The apply equals the oracle's sum over the multipoles the GPU's
    per-tile cut keeps, to rounding (a kept term run twice or not at all would
    move the weak coupling part far above that). And the change the cut makes
    (cut minus untruncated apply, in which the stencil cancels) is the same on
    both sides: the GPU drops exactly the oracle's terms."""
    announce(f"coarse cut apply l_max={l_max} Nr={Nr} tol={tol}")
    g, o, lc = cut_pair(l_max, Nr, tol)
    rng = np.random.default_rng(7 + l_max)
    x = rng.standard_normal(o.size) + 1j * rng.standard_normal(o.size)
    g_cut = g.apply_H(x)
    g.set_coupling_tolerance(0.0)
    g_full = g.apply_H(x)
    o_full = o.apply_H(x)
    o.set_multipole_cut(lc, 64)
    o_cut = o.apply_H(x)
    assert rel(g_cut, o_cut) < 1e-13
    assert rel(g_full, o_full) < 1e-13
    d_o = o_cut - o_full
    if np.linalg.norm(d_o) > 1e-11 * np.linalg.norm(o_full):
        # resolvable above the rounding of the two applies
        assert rel(g_cut - g_full, d_o) < 1e-3
    else:
        assert tol < 1e-9, "the coarsest cut must change the apply measurably"


def test_coarse_multipole_cut_steps_match_oracle():
    """Three steps under the coarse cut, FP64 within 1e-10 of the oracle with
    the same cut and mixed precision within 1e-8 (the FP32 Krylov applies run
    the same stream tables), iterations +-1."""
    announce("coarse cut steps")
    g, o, lc = cut_pair(12)
    g32 = gm.CrankNicolsonModes(**dict(CUT_ARGS, l_max=12))
    g32.set_flags(evolve_orbit=0, precision=gm.GM_PRECISION_MIXED)
    g32.set_initial_state_full(generic_state(o))
    g32.initialize()
    g32.build_damping_mask(0.85, 1.0, 6)
    g32.begin_run(0.6 * CUT_ARGS["L"], 6)
    g32.set_coupling_tolerance(1e-12)
    assert np.array_equal(g32.multipole_cut(), lc)
    o.set_multipole_cut(lc, 64)
    for n in range(1, 4):
        for s in (g, g32, o):
            s.step(n, diag=False)
        it = o.scalars()["last_solver_iterations"]
        assert abs(g.scalars()["last_solver_iterations"] - it) <= 1
        assert abs(g32.scalars()["last_solver_iterations"] - it) <= 1
    assert rel(g.psi, o.psi) < 1e-10
    assert rel(g32.psi, o.psi) < 1e-8


def test_high_lmax_steps_match_oracle():
    """Five steps at l_max = 16 with both absorbers and the mask: FP64 within
    1e-10 of the oracle, mixed precision within 1e-8, iterations +-1."""
    announce("l_max=16 steps")
    args = dict(Nr=257, L=768.0, dt=0.5, epsilon=0.0, l_max=16, mu=0.2, M_c=0.01, q=0.5, a=0.0, R=40.0,
                counter=False, absorption=True, save_every=1000, ramp_periods=0.0, epsilon_smear=2.0)
    g64, g32 = gm.CrankNicolsonModes(**args), gm.CrankNicolsonModes(**args)
    o = oracle.Oracle(**args)
    psi = generic_state(o)
    g32.set_flags(precision=gm.GM_PRECISION_MIXED)
    for s in (g64, g32, o):
        s.set_flags(evolve_orbit=0)
        s.set_initial_state_full(psi)
        s.initialize()
        s.build_damping_mask(0.85, 1.0, 6)
        s.begin_run(600.0, 6)
    for n in range(1, 6):
        for s in (g64, g32, o):
            s.step(n, diag=False)
        it = o.scalars()["last_solver_iterations"]
        assert abs(g64.scalars()["last_solver_iterations"] - it) <= 1
        assert abs(g32.scalars()["last_solver_iterations"] - it) <= 1
    assert rel(g64.psi, o.psi) < 1e-10
    assert rel(g32.psi, o.psi) < 1e-8


def test_multipole_truncation_is_below_rounding():
    """The default 2^-64 multipole cut changes psi only at rounding level, and
    both agree with the untruncated oracle; the cut must actually skip work."""
    announce("multipole truncation")
    args = dict(Nr=2048, L=1500.0, dt=0.5, epsilon=0.0, l_max=8, mu=0.15, M_c=0.01, q=0.5, a=0.0, R=50.0,
                counter=False, absorption=True, save_every=1000, ramp_periods=0.0, epsilon_smear=2.0)
    cut, full = gm.CrankNicolsonModes(**args), gm.CrankNicolsonModes(**args)
    o = oracle.Oracle(**args)
    psi = generic_state(o)
    for s in (cut, full, o):
        s.set_flags(evolve_orbit=0)
        s.set_initial_state_full(psi)
        s.initialize()
        s.begin_run(1200.0, 6)
    full.set_coupling_tolerance(0.0)
    tc, _ = cut.coupling_work()
    tf, _ = full.coupling_work()
    assert tc < 0.95 * tf, (tc, tf)
    for n in range(1, 6):
        for s in (cut, full, o):
            s.step(n, diag=False)
    assert rel(cut.psi, full.psi) < 1e-14
    assert rel(cut.psi, o.psi) < 1e-10
    assert rel(full.psi, o.psi) < 1e-10


def test_mixed_precision_tracks_the_oracle():
    """FP32 coupling in the Krylov applies (rhs, residual check and every
    diagnostic stay FP64): after 40 steps psi is within 1e-8 of the FP64
    oracle, the FP64 path within 1e-10, and iteration counts agree to +-1."""
    announce("mixed precision 40 steps")
    args = dict(Nr=513, L=256.0, dt=0.5, epsilon=0.0, l_max=6, mu=0.2, M_c=0.01, q=0.1, a=0.0, R=30.0,
                counter=False, absorption=True, save_every=1000, ramp_periods=0.0, epsilon_smear=2.0)
    f64, mix = gm.CrankNicolsonModes(**args), gm.CrankNicolsonModes(**args)
    o = oracle.Oracle(**args)
    psi = generic_state(o)
    for s, prec in ((f64, gm.GM_PRECISION_FP64), (mix, gm.GM_PRECISION_MIXED), (o, None)):
        if prec is None:
            s.set_flags(evolve_orbit=0)
        else:
            s.set_flags(evolve_orbit=0, precision=prec)
        s.set_initial_state_full(psi)
        s.initialize()
        s.build_damping_mask(0.85, 1.0, 6)
        s.begin_run(200.0, 41)
    for n in range(1, 41):
        df, dm, do = f64.step(n), mix.step(n), o.step(n)
        assert abs(mix.last_solver_iterations - f64.last_solver_iterations) <= 1
    assert rel(f64.psi, o.psi) < 1e-10
    assert rel(mix.psi, o.psi) < 1e-8
    for k in ("energy", "cloud_mass", "p_bin_inst"):
        assert dm[k] == pytest.approx(do[k], rel=1e-7, abs=1e-14), k


def test_config1_thousand_steps_against_the_oracle():
    """The minimal example's setup (examples/minimal/config.toml: q = 0.1,
    R = 30, mu = 0.2, N = 2000, L = 1000, l_max = 6, dt = 0.5, a two-period
    ramp, absorption on, fixed orbit) from the translated 2p atom, 1000 steps
    (t = 500, inside the ramp, so the sinks are not yet active) in three
    solvers in lockstep: plain FP64 within 1e-10 and mixed precision
    within 1e-8 of the FP64 oracle, BiCGSTAB iterations within +-1 on every
    step, and the step-1000 diagnostics."""
    from gm import initial
    args = dict(Nr=2000, L=1000.0, dt=0.5, epsilon=0.0, l_max=6, mu=0.2, M_c=0.01, q=0.1, a=0.0, R=30.0,
                counter=False, absorption=True, save_every=1000, ramp_periods=2.0, epsilon_smear=2.0)
    dr = args["L"] / (args["Nr"] - 1)
    r = dr * np.arange(args["Nr"])
    psi = initial.translated_atom(r, args["l_max"], args["mu"], args["q"] / (1 + args["q"]) * args["R"])
    psi /= math.sqrt(np.vdot(psi, psi).real * dr)
    f64, mix = gm.CrankNicolsonModes(**args), gm.CrankNicolsonModes(**args)
    o = oracle.Oracle(**args)
    for s, prec in ((f64, gm.GM_PRECISION_FP64), (mix, gm.GM_PRECISION_MIXED), (o, None)):
        if prec is None:
            s.set_flags(evolve_orbit=0)
        else:
            s.set_flags(evolve_orbit=0, precision=prec)
        s.set_initial_state_full(psi)
        s.initialize()
        s.build_damping_mask(0.85, 1.0, 6)
        s.begin_run(800.0, 1001)
    worst = 0
    for n in range(1, 1001):
        last = n == 1000
        df, dm, do = f64.step(n, diag=last), mix.step(n, diag=last), o.step(n, diag=last)
        it = o.last_solver_iterations
        worst = max(worst, abs(f64.last_solver_iterations - it), abs(mix.last_solver_iterations - it))
        assert worst <= 1, n
        if n % 200 == 0:
            announce(f"config 1: step {n}, fp64 {rel(f64.psi, o.psi):.2e}, mixed {rel(mix.psi, o.psi):.2e}")
    e64, emx = rel(f64.psi, o.psi), rel(mix.psi, o.psi)
    announce(f"config 1 after 1000 steps: psi fp64 {e64:.3e}, mixed {emx:.3e}")
    assert e64 < 1e-10
    assert emx < 1e-8
    for k in ("cloud_mass", "energy", "p_bin_inst", "absorbed"):
        den = max(abs(do[k]), 1e-300)
        announce(f"  {k}: {do[k]:.6e}; fp64 {abs(df[k] - do[k]) / den:.2e}, mixed {abs(dm[k] - do[k]) / den:.2e}")
        assert df[k] == pytest.approx(do[k], rel=1e-9, abs=1e-15), k
        assert dm[k] == pytest.approx(do[k], rel=1e-7, abs=1e-14), k
    # the flux at r_ex = 800 (t = 500, before the cloud arrives) is a small
    # difference of products of size |U(r_ex)|^2 / dr: its error is measured
    # against that scale; the per-mode norms are well conditioned
    ix = int(math.floor(800.0 / dr))
    U = o.psi.reshape(o.n_modes, -1)
    prod = np.max(np.abs(U[:, ix]) * (np.abs(U[:, ix]) + np.abs(U[:, ix + 1]))) / dr
    for k, scale in (("flux", prod), ("norms", np.max(np.abs(np.asarray(do["norms"]))))):
        ref = np.asarray(do[k])
        scale = max(scale, np.max(np.abs(ref)), 1e-300)
        e_f, e_m = np.max(np.abs(np.asarray(df[k]) - ref)) / scale, np.max(np.abs(np.asarray(dm[k]) - ref)) / scale
        announce(f"  {k}: max |ref| {np.max(np.abs(ref)):.3e}, scale {scale:.3e}; fp64 {e_f:.2e}, mixed {e_m:.2e}")
        assert e_f < 1e-9 and e_m < 1e-7, k


def test_config1_sinks_switch_on_against_the_oracle():
    """The minimal example's setup across the end of its two-period ramp
    (t = 1968.8), where the s-wave sinks switch on: the three solvers start
    from the translated 2p atom at t = 1700 and step 800 times to t = 2100 in
    lockstep. Iterations within one on every step; after the last step psi
    within 1e-10 (FP64) and 1e-8 (mixed) of the oracle, and the absorbed mass,
    now non-zero, within 1e-9 and 1e-7."""
    from gm import initial
    args = dict(Nr=2000, L=1000.0, dt=0.5, epsilon=0.0, l_max=6, mu=0.2, M_c=0.01, q=0.1, a=0.0, R=30.0,
                counter=False, absorption=True, save_every=1000, ramp_periods=2.0, epsilon_smear=2.0)
    dr = args["L"] / (args["Nr"] - 1)
    r = dr * np.arange(args["Nr"])
    psi = initial.translated_atom(r, args["l_max"], args["mu"], args["q"] / (1 + args["q"]) * args["R"])
    psi /= math.sqrt(np.vdot(psi, psi).real * dr)
    f64, mix = gm.CrankNicolsonModes(**args), gm.CrankNicolsonModes(**args)
    o = oracle.Oracle(**args)
    t0 = 1700.0
    for s, prec in ((f64, gm.GM_PRECISION_FP64), (mix, gm.GM_PRECISION_MIXED), (o, None)):
        if prec is None:
            s.set_flags(evolve_orbit=0)
        else:
            s.set_flags(evolve_orbit=0, precision=prec)
        s.set_initial_state_full(psi)
        s.initialize()
        s.build_damping_mask(0.85, 1.0, 6)
        s.begin_run(800.0, 801)
        s.set_scalars(t=t0, theta=0.4, phase=0.4)
    worst, active_from = 0, None
    for n in range(1, 801):
        last = n == 800
        df, dm, do = f64.step(n, diag=last), mix.step(n, diag=last), o.step(n, diag=last)
        it = o.last_solver_iterations
        worst = max(worst, abs(f64.last_solver_iterations - it), abs(mix.last_solver_iterations - it))
        assert worst <= 1, n
        if active_from is None and o.scalars()["sink_active"]:
            active_from = n
            assert f64.scalars()["sink_active"] and mix.scalars()["sink_active"]
        if n % 200 == 0:
            announce(f"config 1 ramp end: step {n}, fp64 {rel(f64.psi, o.psi):.2e}, mixed {rel(mix.psi, o.psi):.2e}")
    assert active_from is not None and 500 < active_from < 600
    e64, emx = rel(f64.psi, o.psi), rel(mix.psi, o.psi)
    announce(f"config 1 ramp end: sinks on from step {active_from}; psi fp64 {e64:.3e}, mixed {emx:.3e}, "
             f"absorbed {do['absorbed']:.3e}")
    assert e64 < 1e-10
    assert emx < 1e-8
    assert do["absorbed"] > 0
    for k in ("cloud_mass", "energy", "p_bin_inst", "absorbed"):
        assert df[k] == pytest.approx(do[k], rel=1e-9, abs=1e-15), k
        assert dm[k] == pytest.approx(do[k], rel=1e-7, abs=1e-14), k


@pytest.mark.parametrize("absorb", [False, True])
@pytest.mark.parametrize("precision", [gm.GM_PRECISION_FP64, gm.GM_PRECISION_MIXED])
def test_forced_bicgstab_restart_matches_oracle(absorb, precision):
    """The restart branch of BiCGSTAB on the device (matrices.cpp:67-74 with
    the sinks; Eigen's, with its counter reset on the first restart, without
    them), forced on both paths by raising the restart thresholds to 0.1 of
    |r0|^2: five steps, psi within 1e-10 (FP64) or 1e-8 (mixed) of the
    oracle, iterations within one and the same restart count."""
    announce(f"forced restart: absorb {absorb}, precision {precision}")
    args = dict(Nr=129, L=128.0, dt=4.0, epsilon=0.0, l_max=4, mu=0.2, M_c=0.1, q=0.5, a=0.0, R=20.0,
                counter=False, absorption=absorb, save_every=1000, ramp_periods=0.0, epsilon_smear=4.0)
    g, o = gm.CrankNicolsonModes(**args), oracle.Oracle(**args)
    psi = None
    for s in (g, o):
        s.set_flags(evolve_orbit=0)
        if s is g:
            s.set_flags(precision=precision)
        if psi is None:
            from test_oracle import generic_state
            psi = generic_state(s)
        s.set_initial_state_full(psi)
        s.initialize()
        s.set_restart_thresholds(0.1, 0.1)
        s.begin_run(40.0, 6)
    tol = 1e-10 if precision == gm.GM_PRECISION_FP64 else 1e-8
    total = 0
    for n in range(1, 6):
        dg, do = g.step(n), o.step(n)
        assert g.scalars()["sink_active"] == int(absorb)
        assert abs(g.last_solver_iterations - o.last_solver_iterations) <= 1, n
        assert g.restart_count() == o.restart_count(), n
        total += o.restart_count()
        assert rel(g.psi, o.psi) < tol, n
        assert dg["energy"] == pytest.approx(do["energy"], rel=1e-9 if tol < 1e-9 else 1e-7)
    assert total > 0
    announce(f"  restarts {total}, psi {rel(g.psi, o.psi):.2e}")


def test_checkpoint_resume_is_bitwise(tmp_path):
    """save_checkpoint at step 10, then a fresh solver resumed from it runs
    steps 11-20 exactly as the uninterrupted run (restart.cpp:139-233); the
    file parses with the independent Python reader."""
    from gm import checkpoint
    announce("checkpoint resume")
    objs = pair(Nr=257, L=128.0, dt=2.0, l_max=4, mu=0.2, M_c=0.01, q=0.3, R=30.0, absorption=True,
                ramp_periods=0.05, epsilon_smear=2.0)
    g, o = objs
    psi0 = generic_state(o)

    def setup(s):
        s.set_flags(evolve_orbit=1, steps_per_period=30.0)
        s.set_initial_state_full(psi0)
        s.initialize()
        s.build_damping_mask(0.85, 1.0, 6)
        s.begin_run(100.0, 21)

    setup(g)
    for n in range(1, 11):
        g.step(n, diag=False)
    path = str(tmp_path / "checkpoint.bin")
    g.save_checkpoint(path, 10)
    ck = checkpoint.read_checkpoint(path)
    assert ck["step_id"] == 10 and ck["Nr"] == g.Nr and ck["n_modes"] == g.n_modes
    assert np.array_equal(ck["psi"], g.psi)
    sc = g.scalars()
    for k in checkpoint.SCALARS:
        assert ck["scalars"][k] == sc[k], k
    for n in range(11, 21):
        g.step(n, diag=False)
    h = gm.CrankNicolsonModes(Nr=257, L=128.0, dt=2.0, epsilon=0.0, l_max=4, mu=0.2, M_c=0.01, q=0.3, a=0.0,
                              R=30.0, counter=False, absorption=True, save_every=1000, ramp_periods=0.05,
                              epsilon_smear=2.0)
    setup(h)
    assert not h.load_checkpoint(path, 9)  # step mismatch: load_checkpoint returns false
    assert h.load_checkpoint(path, 10)
    for n in range(11, 21):
        h.step(n, diag=False)
    assert np.array_equal(h.psi, g.psi)
    assert h.scalars()["R"] == g.scalars()["R"] and h.scalars()["p_bin"] == g.scalars()["p_bin"]
