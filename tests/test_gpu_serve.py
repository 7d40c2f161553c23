"""Serving loop (gm_server_*): requests of different sizes that arrive while
others run give, bit for bit, the result of the same request run alone."""
import numpy as np
import pytest

import gm
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
