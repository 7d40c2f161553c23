"""Ensemble step (gm_ensemble_step): independent solvers of different shapes
stepped together on their own streams give exactly the states they give when
stepped alone (the kernels are deterministic and members share nothing)."""
import numpy as np
import pytest

import gm
import oracle
from test_oracle import generic_state

pytestmark = pytest.mark.gpu

MEMBERS = [
    dict(Nr=257, L=128.0, dt=1.0, l_max=4, mu=0.2, M_c=0.01, q=0.1, R=30.0, precision=gm.GM_PRECISION_FP64),
    dict(Nr=513, L=256.0, dt=0.5, l_max=6, mu=0.25, M_c=0.02, q=0.5, R=40.0, precision=gm.GM_PRECISION_MIXED),
    dict(Nr=129, L=96.0, dt=0.5, l_max=2, mu=0.15, M_c=0.005, q=0.8, R=20.0, precision=gm.GM_PRECISION_FP64),
]


def make(spec):
    spec = dict(spec)
    precision = spec.pop("precision")
    args = dict(epsilon=0.0, a=0.0, counter=False, absorption=True, save_every=1000, ramp_periods=0.0,
                epsilon_smear=2.0)
    args.update(spec)
    s = gm.CrankNicolsonModes(**args)
    o = oracle.Oracle(**args)
    s.set_flags(evolve_orbit=0, precision=precision)
    s.set_initial_state_full(generic_state(o))
    s.initialize()
    s.build_damping_mask(0.85, 1.0, 6)
    s.begin_run(0.7 * args["L"], 100)
    return s


def test_ensemble_equals_members_alone():
    print("ensemble of 3 ragged members, 12 steps", flush=True)
    together = [make(m) for m in MEMBERS]
    alone = [make(m) for m in MEMBERS]
    for n in range(1, 13):
        dt = gm.ensemble_step(together, n)
        da = [s.step(n) for s in alone]
        for t, a in zip(dt, da):
            for k in ("energy", "p_bin_inst", "cloud_mass", "absorbed"):
                assert t[k] == a[k], k
    for t, a in zip(together, alone):
        assert np.array_equal(t.psi, a.psi)
        assert t.last_solver_iterations == a.last_solver_iterations


def test_ensemble_retries_a_failed_mixed_member_in_fp64():
    """A mixed-precision member whose FP32 coupling error lies above the FP64
    residual check (dt = 23 at R = 20): inside the ensemble each of its steps
    is solved again with the FP64 operator, so it steps exactly as the same
    member in FP64 alone, while its neighbour is unaffected."""
    print("ensemble with a re-solved mixed member", flush=True)
    big = dict(Nr=257, L=128.0, dt=23.0, l_max=4, mu=0.2, M_c=0.05, q=0.5, R=20.0)
    together = [make(dict(big, precision=gm.GM_PRECISION_MIXED)), make(MEMBERS[2])]
    alone = [make(dict(big, precision=gm.GM_PRECISION_FP64)), make(MEMBERS[2])]
    for n in range(1, 6):
        gm.ensemble_step(together, n)
        for s in alone:
            s.step(n)
        assert together[0].last_solver_residual <= 1e-9
        assert together[0].last_solver_iterations == alone[0].last_solver_iterations, n
    assert np.linalg.norm(together[0].psi - alone[0].psi) <= 1e-13 * np.linalg.norm(alone[0].psi)
    assert np.array_equal(together[1].psi, alone[1].psi)


def test_ensemble_rejects_duplicates():
    s = make(MEMBERS[2])
    with pytest.raises(gm.InvalidArgument):
        gm.ensemble_step([s, s], 1)


@pytest.mark.parametrize("precision", [gm.GM_PRECISION_FP64, gm.GM_PRECISION_MIXED])
def test_apply_tiles_agree(precision):
    """Every operator-apply tile (16, 32, 64 nodes) and the autotuned choice
    give the same steps: per-node arithmetic is identical, only the order of
    the fused reductions differs."""
    print(f"apply tiles, precision {precision}", flush=True)
    spec = dict(MEMBERS[1], precision=precision)
    runs = {}
    for tile in (0, 16, 32, 64):
        s = make(spec)
        s.set_flags(apply_tile=tile)
        for n in range(1, 11):
            s.step(n, diag=False)
        runs[tile] = s.psi
        chosen, ms = s.apply_config(precision)
        assert chosen == (tile or chosen) and chosen in (16, 32, 64)
        if tile == 0:
            assert all(v > 0 for v in ms.values())
    ref = runs[32]
    for tile, psi in runs.items():
        assert np.linalg.norm(psi - ref) / np.linalg.norm(ref) < 1e-11, tile


def test_apply_tile_flag_is_validated():
    s = make(MEMBERS[2])
    with pytest.raises(gm.InvalidArgument):
        s.set_flags(apply_tile=48)
