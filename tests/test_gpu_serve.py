"""Serving loop (gm_server_*): requests of different sizes that arrive while
others run give, bit for bit, the result of the same request run alone; a
Poisson stream with the served workload's physics (back-reaction on, cloud
masses drawn, a two-period startup ramp, mixed precision) at small shapes,
with 5 % of its requests spot-checked against the FP64 oracle."""
import math

import numpy as np
import pytest

import gm
import oracle
from gm import serve

pytestmark = pytest.mark.gpu


def spec(Nr, l_max, q, R, mu, seed, steps, precision=1, evolve=1):
    params = dict(Nr=Nr, L=Nr / 4.0, dt=0.25, epsilon=0.0, l_max=l_max, mu=mu, M_c=0.02, q=q, a=0.0, R=R,
                  counter=False, absorption=True, save_every=1000, ramp_periods=0.5, epsilon_smear=2.0)
    psi = serve.atom_state(Nr, params["L"], l_max, mu, seed)
    flags = serve.default_flags(precision=precision, evolve_orbit=evolve, steps_per_period=40.0)
    return params, psi, steps, flags


def alone(params, psi, steps, flags):
    s = gm.CrankNicolsonModes(**params)
    s.set_flags(**{k: getattr(flags, k) for k, _ in gm.capi.Flags._fields_})
    s.set_initial_state_full(psi)
    s.initialize()
    s.build_damping_mask(0.85, 1.0, 6)
    s.begin_run(0.8 * params["L"], steps + 1)
    for n in range(1, steps + 1):
        s.step(n, diag=False)
    return s.psi, s.scalars()


def test_served_results_equal_alone():
    print("serving: 4 ragged requests, two waves", flush=True)
    reqs = [spec(1024, 4, 0.3, 30.0, 0.2, 1, 25), spec(2048, 6, 0.8, 40.0, 0.25, 2, 18, precision=0),
            spec(512, 8, 0.1, 25.0, 0.15, 3, 30), spec(1536, 2, 0.5, 50.0, 0.3, 4, 22, evolve=0)]
    srv = serve.Server(hbm_budget=8e9, max_active=3)
    tickets = [srv.submit(p, psi, n, 0.8 * p["L"], flags=f, mask=(0.85, 1.0, 6)) for p, psi, n, f in reqs[:2]]
    srv.run(0.05)
    tickets += [srv.submit(p, psi, n, 0.8 * p["L"], flags=f, mask=(0.85, 1.0, 6)) for p, psi, n, f in reqs[2:]]
    for _ in range(600):
        srv.run(1.0)
        if all(srv.status(t)[0] >= serve.GM_SERVE_DONE for t in tickets):
            break
    st = srv.stats()
    assert st["steps"] == sum(r[2] for r in reqs)
    assert st["p99_ms"] >= st["p50_ms"] > 0
    for t, (p, psi, n, f) in zip(tickets, reqs):
        got, sc = srv.take(t)
        ref, sref = alone(p, psi, n, f)
        assert np.array_equal(got, ref)
        for k in ("R", "M1", "M2", "Omega", "dt", "theta", "t"):
            assert sc[k] == sref[k], k
    srv.close()


def test_admission_respects_the_budget():
    p, psi, n, f = spec(1024, 4, 0.3, 30.0, 0.2, 1, 5)
    srv = serve.Server(hbm_budget=1e9, max_active=8)
    fp = srv.footprint(**p)
    srv.close()
    srv = serve.Server(hbm_budget=1.5 * fp, max_active=8)
    a = srv.submit(p, psi, n, 0.8 * p["L"], flags=f)
    b = srv.submit(p, psi, n, 0.8 * p["L"], flags=f)
    srv.run(0.0)
    assert srv.stats()["active"] == 1 and srv.stats()["queued"] == 1
    while srv.status(b)[0] < serve.GM_SERVE_DONE:
        srv.run(1.0)
    assert srv.status(a)[0] == serve.GM_SERVE_DONE
    with pytest.raises(gm.InvalidArgument):
        srv.submit(dict(p, Nr=4096, L=1024.0), np.zeros(15 * 4096, dtype=complex), 1, 800.0, flags=f)
    srv.close()


def test_served_mixed_request_retried_in_fp64():
    """A served mixed-precision request whose FP32 coupling error lies above
    the FP64 residual check (dt = 23 at R = 20, as in the ensemble's retry
    test): each of its steps is enqueued again with the FP64 operator while
    the server keeps advancing its neighbour, so it ends as the same request
    run alone in FP64, and the neighbour as its own solo run, bitwise."""
    print("serving: a mixed request re-solved in FP64 beside another", flush=True)
    big = dict(Nr=257, L=128.0, dt=23.0, epsilon=0.0, l_max=4, mu=0.2, M_c=0.05, q=0.5, a=0.0, R=20.0,
               counter=False, absorption=True, save_every=1000, ramp_periods=0.0, epsilon_smear=2.0)
    from test_oracle import generic_state
    psi = generic_state(oracle.Oracle(**big))
    n = 5
    mixed = serve.default_flags(precision=1, evolve_orbit=0)
    fp64 = serve.default_flags(precision=0, evolve_orbit=0)
    other = spec(1024, 4, 0.3, 30.0, 0.2, 1, 8)
    srv = serve.Server(hbm_budget=8e9, max_active=4)
    t_big = srv.submit(big, psi, n, 100.0, flags=mixed, mask=(0.85, 1.0, 6))
    t_other = srv.submit(other[0], other[1], other[2], 0.8 * other[0]["L"], flags=other[3], mask=(0.85, 1.0, 6))
    for _ in range(600):
        srv.run(1.0)
        if all(srv.status(t)[0] >= serve.GM_SERVE_DONE for t in (t_big, t_other)):
            break
    got, _ = srv.take(t_big)
    got_other, _ = srv.take(t_other)
    srv.close()
    solo = {}
    for name, f in (("mixed", mixed), ("fp64", fp64)):
        s = gm.CrankNicolsonModes(**big)
        s.set_flags(**{k: getattr(f, k) for k, _ in gm.capi.Flags._fields_})
        s.set_initial_state_full(psi)
        s.initialize()
        s.build_damping_mask(0.85, 1.0, 6)
        s.begin_run(100.0, n + 1)
        for k in range(1, n + 1):
            s.step(k, diag=False)
            assert s.last_solver_residual <= 1e-9
        assert s.get_flags().precision == f.precision
        solo[name] = s.psi
        s.close()
    # served = alone; and alone every mixed step was re-solved in FP64
    assert np.array_equal(got, solo["mixed"])
    assert np.linalg.norm(solo["mixed"] - solo["fp64"]) <= 1e-13 * np.linalg.norm(solo["fp64"])
    ref, _ = alone(*other)
    assert np.array_equal(got_other, ref)


def oracle_steps(params, psi0, n_steps, steps):
    """The FP64 oracle on a served request's inputs, `steps` steps."""
    o = oracle.Oracle(**params)
    f = serve.default_flags()
    o.set_flags(**{k: getattr(f, k) for k, _ in gm.capi.Flags._fields_})
    o.set_initial_state_full(psi0)
    o.initialize()
    o.build_damping_mask(0.85, 1.0, 6)
    o.begin_run(0.8 * params["L"], n_steps + 1)
    for n in range(1, steps + 1):
        o.step(n, diag=False)
    return o.psi, o.scalars()


def test_served_stream_with_oracle_spot_checks():
    """Poisson arrivals at 10/s for 2 s of 20-40-step requests (N 1024-2048,
    l_max 2-8), served in mixed precision while they arrive; every request
    finishes, per-step and request latencies are reported, and every 20th
    request (5 %) is checked against the FP64 oracle: its state after two
    steps, taken inside the mix, and its final state and orbit, within the
    mixed-precision tolerance 1e-8."""
    print("serving: Poisson stream with oracle spot checks", flush=True)
    st = serve.served_stream(duration=2.0, rate=10.0, seed=3, hbm_budget=8e9, max_active=6, drain=True,
                             drain_limit=240, sizes=[(1024, 4), (2048, 6), (1536, 8), (1024, 2)], steps=(20, 40),
                             probe_every=20, probe_step=2)
    print({k: v for k, v in st.items() if k != "probes"}, flush=True)
    assert st["submitted"] >= 10
    assert st["completed"] == st["submitted"] and st["queued"] == 0 and st["active"] == 0
    assert st["p99_ms"] >= st["p50_ms"] > 0
    assert st["req_p99_ms"] >= st["req_p50_ms"] > 0
    assert len(st["probes"]) == math.ceil(st["submitted"] / 20)
    for rec in st["probes"]:
        p, n = rec["params"], rec["n_steps"]
        ref2, _ = oracle_steps(p, rec["psi0"], n, 2)
        e2 = np.linalg.norm(rec["psi"] - ref2) / np.linalg.norm(ref2)
        ref, sref = oracle_steps(p, rec["psi0"], n, n)
        e = np.linalg.norm(rec["final"] - ref) / np.linalg.norm(ref)
        print(f"  spot check ticket {rec['ticket']}: after 2 steps {e2:.2e}, after {n} steps {e:.2e}", flush=True)
        assert e2 < 1e-8 and e < 1e-8
        for k in ("R", "M1", "M2", "Omega", "theta", "t"):
            assert rec["scalars"][k] == pytest.approx(sref[k], rel=1e-9, abs=1e-12), k
