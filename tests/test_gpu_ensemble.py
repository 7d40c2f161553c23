"""Ensembles: gm_ensemble_step (each member on its own stream) and gm_batch
(every launch of the step one ragged grid over all members, graph-captured):
independent solvers of different shapes stepped together give exactly the
states they give when stepped alone (the kernels are deterministic, members
share nothing, and a member's blocks in a batched grid are those of its own
launch). The batched ensemble of 64 binaries drawn from seed 0 is checked
member by member, and its smallest member against the FP64 oracle."""
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
    """Every operator-apply form (Gaunt-form tiles of 16, 32, 64 nodes, the
    polar-quadrature form) and the autotuned choice give the same steps: the
    Gaunt tiles differ only in the order of the fused reductions, the polar
    form evaluates the same coupling through an exact quadrature (rounding
    apart)."""
    print(f"apply tiles, precision {precision}", flush=True)
    spec = dict(MEMBERS[1], precision=precision)
    runs = {}
    for tile in (0, 8, 16, 32, 64):
        s = make(spec)
        s.set_flags(apply_tile=tile)
        for n in range(1, 11):
            s.step(n, diag=False)
        runs[tile] = s.psi
        chosen, ms = s.apply_config(precision)
        assert chosen == (tile or chosen) and chosen in (8, 16, 32, 64)
        if tile == 0:
            assert all(v > 0 for v in ms.values())
    ref = runs[32]
    tol = 1e-11 if precision == gm.GM_PRECISION_FP64 else 1e-9
    for tile, psi in runs.items():
        assert np.linalg.norm(psi - ref) / np.linalg.norm(ref) < (1e-11 if tile in (16, 32, 64) else tol), tile


def test_apply_tile_flag_is_validated():
    s = make(MEMBERS[2])
    with pytest.raises(gm.InvalidArgument):
        s.set_flags(apply_tile=48)


@pytest.mark.parametrize("lanes", [1, 2])
def test_batch_equals_members_alone(lanes):
    print(f"batched ensemble of 3 ragged members in {lanes} lanes, 12 steps", flush=True)
    together = [make(m) for m in MEMBERS]
    alone = [make(m) for m in MEMBERS]
    batch = gm.Batch(together, lanes)
    for n in range(1, 13):
        dt = batch.step(n)
        da = [s.step(n) for s in alone]
        for t, a in zip(dt, da):
            for k in ("energy", "p_bin_inst", "cloud_mass", "absorbed", "dV_bin_dt"):
                assert t[k] == a[k], (n, k)
            for k in ("flux", "norms", "energies", "energy_flux"):
                assert np.array_equal(t[k], a[k]), (n, k)
    for t, a in zip(together, alone):
        assert np.array_equal(t.psi, a.psi)
        assert t.last_solver_iterations == a.last_solver_iterations
    # members step alone between batch steps, and the batch goes on
    together[0].step(13)
    alone[0].step(13)
    batch.step(14)
    for a in alone:
        a.step(14)
    for t, a in zip(together, alone):
        assert np.array_equal(t.psi, a.psi)
    batch.close()


def test_batches_beyond_the_constant_argument_pool():
    """The batched applies read their members' arguments from a constant
    pool of 2048 entries per GPU; a batch's plan takes 4 ragged applies x 64
    members. Nine batches of 64 alive at once need more: the last ones read
    global argument tables instead, and every member still steps as alone.
    Closing the batches frees their entries for a new batch."""
    print("nine batches of 64 small members", flush=True)
    args = dict(Nr=65, L=64.0, dt=1.0, epsilon=0.0, l_max=1, mu=0.2, M_c=0.01, q=0.5, a=0.0, R=20.0,
                counter=False, absorption=True, save_every=1000, ramp_periods=0.0, epsilon_smear=2.0)
    psi = generic_state(oracle.Oracle(**args))

    def small():
        s = gm.CrankNicolsonModes(**args)
        s.set_flags(evolve_orbit=0, apply_tile=32)
        s.set_initial_state_full(psi)
        s.initialize()
        s.build_damping_mask(0.85, 1.0, 6)
        s.begin_run(40.0, 100)
        return s

    ref = small()
    for n in (1, 2):
        ref.step(n, diag=False)
    groups = [[small() for _ in range(64)] for _ in range(9)]
    batches = [gm.Batch(g) for g in groups]
    for n in (1, 2):
        for b in batches:
            b.step(n, diag=False)
    for g in groups:
        for s in g:
            assert np.array_equal(s.psi, ref.psi)
    for b in batches:
        b.close()
    again = [small() for _ in range(64)]
    b = gm.Batch(again)
    for n in (1, 2):
        b.step(n, diag=False)
    assert all(np.array_equal(s.psi, ref.psi) for s in again)
    b.close()
    for g in groups + [again, [ref]]:
        for s in g:
            s.close()


@pytest.mark.timeout(900)
def test_config4_batched_ensemble_64_members():
    """The 64 binaries of the batched ensemble from seed 0 (q log-uniform in
    [0.05, 1], R in [20, 60], mu in [0.1, 0.3], N from 4096 to 32768 with
    L = N/4, l_max from 8 to 24, the 2p atom on the primary plus a 1e-3
    perturbation, mixed precision): three batched steps equal each member's
    own three steps bitwise, and the smallest member's FP64 twin, batched the
    same way, matches the oracle to 1e-10."""
    from gm import ensemble
    print("config 4: 64 batched members", flush=True)
    members = ensemble.draw_members(64, seed=0)
    states = [ensemble.member_state(p, i) for i, p in enumerate(members)]
    batched = ensemble.build(members, states=states)
    print("  built", flush=True)
    small = min(range(64), key=lambda i: members[i]["Nr"] * (members[i]["l_max"] + 1) ** 2)
    p = members[small]
    psi0 = states[small]
    twin = gm.CrankNicolsonModes(**p)
    o = oracle.Oracle(**p)
    for s in (twin, o):
        s.set_flags(evolve_orbit=0)
        s.set_initial_state_full(psi0)
        s.initialize()
        s.build_damping_mask(0.85, 1.0, 6)
        s.begin_run(0.8 * p["L"], 100000)
    batch = gm.Batch(batched + [twin])
    for n in range(1, 4):
        batch.step(n, diag=False)
        o.step(n, diag=False)
    e = np.linalg.norm(twin.psi - o.psi) / np.linalg.norm(o.psi)
    print(f"  FP64 twin of member {small} (N {p['Nr']}, l_max {p['l_max']}) vs oracle {e:.2e}", flush=True)
    assert e < 1e-10
    for i, s in enumerate(batched):
        solo = ensemble.build([members[i]], states=[states[i]])[0]
        for n in range(1, 4):
            solo.step(n, diag=False)
        assert np.array_equal(s.psi, solo.psi), i
        assert s.last_solver_iterations == solo.last_solver_iterations, i
        solo.close()
    batch.close()
