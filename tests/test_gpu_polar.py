"""The operator apply in polar-quadrature form (apply_tile = 8) against the
oracle's sum over every Gaunt term: H at every l_max class, under the
multipole cut term for term, Krylov and residual applies through whole steps
(FP64 and mixed precision), batched launches and the headline shape."""
import numpy as np
import pytest

import gm
import oracle
from test_gpu_parity import CUT_ARGS, announce, both, cut_pair, pair, rel
from test_oracle import generic_state

pytestmark = pytest.mark.gpu

POLAR = 8


@pytest.mark.parametrize("l_max,Nr", [(0, 65), (1, 65), (2, 129), (4, 129), (5, 97), (8, 129), (12, 193),
                                      (17, 97), (20, 129)])
def test_polar_apply_H_matches_oracle(l_max, Nr):
    """H through the polar form at l_max in every class (4 .. 20, runtime l_max
    below the class's as well), with absorbers, cut and untruncated."""
    announce(f"polar apply_H l_max={l_max} Nr={Nr}")
    objs = pair(l_max=l_max, absorption=True, Nr=Nr, L=float(8 * (Nr - 1)), R=40.0)
    both(objs, "set_flags", evolve_orbit=0)
    g, o = objs
    g.set_flags(apply_tile=POLAR)
    both(objs, "set_initial_state_full", generic_state(o))
    both(objs, "initialize")
    both(objs, "begin_run", 0.6 * 8 * (Nr - 1), 2)
    rng = np.random.default_rng(100 + l_max)
    x = rng.standard_normal(o.size) + 1j * rng.standard_normal(o.size)
    assert rel(g.apply_H(x), o.apply_H(x)) < 1e-13
    g.set_coupling_tolerance(0.0)
    assert rel(g.apply_H(x), o.apply_H(x)) < 1e-13


def test_polar_form_needs_its_lmax_class():
    """Above l_max 20 the polar form has no class: the flag is refused and the
    autotuner times the Gaunt-form tiles only."""
    g, _ = pair(l_max=21, Nr=65)
    with pytest.raises(gm.InvalidArgument):
        g.set_flags(apply_tile=POLAR)


@pytest.mark.parametrize("l_max,Nr,tol", [(12, 1025, 1e-12), (20, 513, 1e-12), (12, 1025, 9e-7)])
def test_polar_cut_matches_oracle_term_for_term(l_max, Nr, tol):
    """Under a coarse multipole cut each tile takes the smallest exact rule for
    its cut (12 to 24 points at l_max 12 and 20): the apply equals the oracle's
    sum over the kept multipoles, and the change the cut makes equals the
    oracle's."""
    announce(f"polar coarse cut l_max={l_max} Nr={Nr} tol={tol}")
    g, o, lc = cut_pair(l_max, Nr, tol)
    g.set_flags(apply_tile=POLAR)
    rng = np.random.default_rng(17 + l_max)
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
        assert rel(g_cut - g_full, d_o) < 1e-3


def test_polar_steps_match_oracle():
    """Five steps at l_max 16 with both absorbers and the mask in the polar
    form: FP64 within 1e-10 of the oracle, mixed precision within 1e-8,
    iterations +-1 (every apply mode of the step: rhs, Krylov, residual)."""
    announce("polar steps l_max=16")
    args = dict(Nr=257, L=768.0, dt=0.5, epsilon=0.0, l_max=16, mu=0.2, M_c=0.01, q=0.5, a=0.0, R=40.0,
                counter=False, absorption=True, save_every=1000, ramp_periods=0.0, epsilon_smear=2.0)
    g64, g32 = gm.CrankNicolsonModes(**args), gm.CrankNicolsonModes(**args)
    o = oracle.Oracle(**args)
    psi = generic_state(o)
    g64.set_flags(apply_tile=POLAR)
    g32.set_flags(precision=gm.GM_PRECISION_MIXED, apply_tile=POLAR)
    for s in (g64, g32, o):
        s.set_flags(evolve_orbit=0)
        s.set_initial_state_full(psi)
        s.initialize()
        s.build_damping_mask(0.85, 1.0, 6)
        s.begin_run(600.0, 6)
    for n in range(1, 6):
        dg = g64.step(n)
        g32.step(n, diag=False)
        do = o.step(n)
        it = o.scalars()["last_solver_iterations"]
        assert abs(g64.scalars()["last_solver_iterations"] - it) <= 1
        assert abs(g32.scalars()["last_solver_iterations"] - it) <= 1
        assert abs(dg["energy"] - do["energy"]) <= 1e-9 * abs(do["energy"])
    assert rel(g64.psi, o.psi) < 1e-10
    assert rel(g32.psi, o.psi) < 1e-8
    assert g64.apply_config(gm.GM_PRECISION_FP64)[0] == POLAR
    assert g32.apply_config(gm.GM_PRECISION_MIXED)[0] == POLAR


def test_polar_coarse_cut_steps_match_oracle():
    """Three steps under the coarse cut in the polar form, FP64 within 1e-10
    of the oracle with the same cut and mixed within 1e-8."""
    announce("polar coarse cut steps")
    g, o, lc = cut_pair(12)
    g.set_flags(apply_tile=POLAR)
    g32 = gm.CrankNicolsonModes(**dict(CUT_ARGS, l_max=12))
    g32.set_flags(evolve_orbit=0, precision=gm.GM_PRECISION_MIXED, apply_tile=POLAR)
    g32.set_initial_state_full(generic_state(o))
    g32.initialize()
    g32.build_damping_mask(0.85, 1.0, 6)
    g32.begin_run(0.6 * CUT_ARGS["L"], 6)
    g32.set_coupling_tolerance(1e-12)
    o.set_multipole_cut(lc, 64)
    for n in range(1, 4):
        for s in (g, g32, o):
            s.step(n, diag=False)
        it = o.scalars()["last_solver_iterations"]
        assert abs(g.scalars()["last_solver_iterations"] - it) <= 1
        assert abs(g32.scalars()["last_solver_iterations"] - it) <= 1
    assert rel(g.psi, o.psi) < 1e-10
    assert rel(g32.psi, o.psi) < 1e-8


def test_polar_batch_equals_members_alone():
    """A batched ragged step over members of three l_max classes in the polar
    form (one launch per class) equals each member stepped alone, bitwise."""
    announce("polar batch")
    specs = [dict(Nr=257, L=128.0, dt=1.0, l_max=4, mu=0.2, M_c=0.01, q=0.1, R=30.0),
             dict(Nr=513, L=256.0, dt=0.5, l_max=6, mu=0.25, M_c=0.02, q=0.5, R=40.0),
             dict(Nr=193, L=192.0, dt=0.5, l_max=11, mu=0.15, M_c=0.005, q=0.8, R=20.0)]

    def make(spec, precision):
        a = dict(epsilon=0.0, a=0.0, counter=False, absorption=True, save_every=1000, ramp_periods=0.0,
                 epsilon_smear=2.0)
        a.update(spec)
        s = gm.CrankNicolsonModes(**a)
        s.set_flags(evolve_orbit=0, precision=precision, apply_tile=POLAR)
        s.set_initial_state_full(generic_state(s))
        s.initialize()
        s.begin_run(0.6 * a["L"], 4)
        return s

    for precision in (gm.GM_PRECISION_FP64, gm.GM_PRECISION_MIXED):
        alone = [make(sp, precision) for sp in specs]
        together = [make(sp, precision) for sp in specs]
        for n in range(1, 4):
            for s in alone:
                s.step(n, diag=False)
        b = gm.Batch(together)
        for n in range(1, 4):
            b.step(n)
        for s, t in zip(alone, together):
            assert np.array_equal(s.psi, t.psi)
        b.close()


def test_polar_headline_shape_matches_gaunt_form():
    """At l_max 20, N 16384, L 4000 (the headline shape, default multipole cut)
    the polar-form H equals the Gaunt-form H to rounding, and one step of each
    form agrees within 1e-12."""
    announce("polar headline shape")
    args = dict(Nr=16384, L=4000.0, dt=0.25, epsilon=0.0, l_max=20, mu=0.15, M_c=0.01, q=0.5, a=0.0, R=50.0,
                counter=False, absorption=True, save_every=1000, ramp_periods=0.0, epsilon_smear=2.0)
    gp, gg = gm.CrankNicolsonModes(**args), gm.CrankNicolsonModes(**args)
    gp.set_flags(evolve_orbit=0, apply_tile=POLAR)
    gg.set_flags(evolve_orbit=0, apply_tile=32)
    psi = generic_state(gp)
    for s in (gp, gg):
        s.set_initial_state_full(psi)
        s.initialize()
        s.build_damping_mask(0.85, 1.0, 6)
        s.begin_run(3200.0, 4)
    rng = np.random.default_rng(5)
    x = rng.standard_normal(gp.size) + 1j * rng.standard_normal(gp.size)
    assert rel(gp.apply_H(x), gg.apply_H(x)) < 1e-13
    gp.step(1, diag=False)
    gg.step(1, diag=False)
    assert rel(gp.psi, gg.psi) < 1e-12
