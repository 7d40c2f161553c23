"""Pins of the FP64 CPU oracle (CPU only).

Each check anchors a part of the oracle on something outside it: exact sympy
values, the dense restatement of the reference's tests/parallel_correctness.cpp
(tests/reference_dense.py), the synthetic tridiagonal known answer and the
physics invariants of tests/physics.cpp of rdellamonica/boson-cloud-binary.
"""
import math

import numpy as np
import pytest

import oracle
import reference_dense as rd

# ---------------------------------------------------------------- special functions


@pytest.mark.parametrize("args", [(2, 3, 1, 0, 0, 0), (20, 30, 40, 3, -5, 2), (60, 30, 30, -7, 3, 4),
                                  (45, 29, 30, 13, -11, -2), (4, 2, 2, -2, 1, 1), (7, 3, 4, 0, 0, 0)])
def test_wigner3j_matches_sympy(args):
    assert oracle.wigner3j(*args) == pytest.approx(rd.w3j(*args), rel=1e-14, abs=1e-300)


@pytest.mark.parametrize("j1,j2,m1,m2", [(25, 30, 3, -7), (30, 30, 0, 0), (40, 35, -12, 20), (60, 60, 5, 5)])
def test_wigner3j_orthogonality_at_high_degree(j1, j2, m1, m2):
    """sum_j3 (2 j3 + 1) (j1 j2 j3; m1 m2 -m1-m2)^2 = 1 (orthogonality of the
    3j symbols) up to the degrees of l_max = 30..60, where the Racah sums and
    prefactors leave the double range; plus the closed form
    (j j 0; m -m 0) = (-1)^(j-m) / sqrt(2 j + 1)."""
    m3 = -m1 - m2
    s = sum((2 * j3 + 1) * oracle.wigner3j(j1, j2, j3, m1, m2, m3) ** 2 for j3 in range(abs(j1 - j2), j1 + j2 + 1))
    assert s == pytest.approx(1.0, rel=1e-13)
    assert oracle.wigner3j(j1, j1, 0, m1, -m1, 0) == pytest.approx((-1) ** (j1 - m1) / math.sqrt(2 * j1 + 1),
                                                                   rel=1e-14)


def test_wigner3j_selection_rules():
    assert oracle.wigner3j(1, 1, 1, 0, 0, 0) == 0.0      # odd sum with m = 0
    assert oracle.wigner3j(2, 2, 5, 0, 0, 0) == 0.0      # triangle violated
    assert oracle.wigner3j(2, 2, 2, 1, 1, 1) == 0.0      # m sum not zero


@pytest.mark.parametrize("lm", [(0, 0), (1, 1), (2, 0), (3, 1), (20, 4), (40, 10), (59, 1), (7, 3)])
def test_sphPlm0_matches_sympy(lm):
    assert oracle.sphPlm0(*lm) == pytest.approx(rd.sph_plm0(*lm), rel=1e-14, abs=1e-300)


def test_gauss_legendre_is_exact_for_polynomials():
    x, w = oracle.gauss_legendre01(128)
    assert np.all(np.diff(x) > 0) and 0 < x[0] and x[-1] < 1
    for k in (0, 1, 10, 101, 255):
        assert np.sum(w * x ** k) == pytest.approx(1.0 / (k + 1), rel=1e-14)
    xr, wr = np.polynomial.legendre.leggauss(128)
    np.testing.assert_allclose(x, 0.5 * (xr + 1), rtol=0, atol=2e-16)
    np.testing.assert_allclose(w, 0.5 * wr, rtol=0, atol=1e-14)  # numpy weights carry ~1e-14 abs error


# ---------------------------------------------------------------- Thomas


def test_thomas_known_answer():
    """tests/parallel_correctness.cpp:16-47: block sizes 1 and 9, 12 blocks."""
    for block in (1, 9):
        K = 12
        idx = np.arange(K * block).reshape(K, block)
        diag = 3.0 + idx * 0.01 + 0.2j
        lower = np.where(idx % block > 0, -0.2 + 0.1j, 0)
        upper = np.where(idx % block + 1 < block, -0.4 + 0.1j, 0)
        rhs = np.sin(idx) + 1j * np.cos(idx)
        rc, x = oracle.thomas(diag, lower, upper, rhs)
        assert rc == 0
        for j in range(K):
            A = np.diag(diag[j]) + np.diag(lower[j, 1:], -1) + np.diag(upper[j, :-1], 1)
            assert np.linalg.norm(A @ x[j] - rhs[j]) / np.linalg.norm(rhs[j]) < 1e-13
    rc, _ = oracle.thomas(np.zeros((1, 1)), np.zeros((1, 1)), np.zeros((1, 1)), np.ones((1, 1)))
    assert rc != 0, "singular pivot not detected"


# ---------------------------------------------------------------- one CN step vs dense LU


def fixture_state(nr, K):
    n = nr * K
    i = np.arange(n)
    psi = np.where((i % nr == 0) | (i % nr == nr - 1), 0, (0.1 + 0.001 * i) * np.exp(1j * 0.1 * i))
    return psi / np.linalg.norm(psi)


def correctness_case(scenario):
    """Build scenario `scenario` of tests/parallel_correctness.cpp:109-180 on the
    oracle, returning it (initialised) and the dense expected step."""
    nr, dt, smear = 32, 0.2, 10.0
    absorb = scenario in (1, 2)
    P0 = 2 * math.pi / math.sqrt(1.2 / 30.0 ** 3)
    ramp_periods = 2 * dt / P0 if scenario == 3 else 0.0
    cn = oracle.Oracle(nr, 120.0, dt, 0.1, 3, 0.2, 1e-20 if scenario == 4 else 0.01, 0.2, 0.0, 30.0,
                       counter=scenario == 2, absorption=absorb, save_every=10,
                       ramp_periods=ramp_periods, epsilon_smear=smear)
    cn.set_flags(evolve_orbit=int(absorb))
    cn.set_scalars(theta=0.3, Omega=cn.Omega_0)
    cn.set_initial_state_full(fixture_state(nr, cn.n_modes))
    cn.initialize()
    if scenario == 2:
        cn.build_cap(0.8, 0.01, 3)
        cn.build_damping_mask(0.8, 0.01, 3)
    s = cn.scalars()
    theta = s["theta"] + (-1.0 if scenario == 2 else 1.0) * s["Omega"] * dt
    Omega = s["Omega"]
    if dt < ramp_periods * s["P"]:
        Omega = cn.Omega_0 * rd.smoothstep(dt / (ramp_periods * s["P"]))
    x = dt / (ramp_periods * P0) if ramp_periods > 0 else 1.0
    companion_ramp = rd.smoothstep(x) if x < 1 else 1.0
    h0, v1, v2, dvdt = rd.operators(nr, cn.dr, cn.r, 3, 0.2, s["M1"], s["M2"], s["q"], s["R1"], s["R2"],
                                   theta, Omega, scenario == 2, companion_ramp)
    h = h0 + v1 + companion_ramp * v2
    if scenario == 2:
        f, W, nn = 0.8, 0.01, 3
        idx = int(math.floor(cn.params.L * f / cn.dr))
        cap = np.zeros(nr, dtype=complex)
        for i in range(idx, nr):
            cap[i] = -1j * W * ((cn.r[i] - f * 120.0) / (cn.r[i] - (1 - f) * 120.0)) ** nn
        h += np.diag(np.tile(cap, cn.n_modes))
    gamma = np.zeros_like(h)
    if absorb:
        abs_ramp = rd.smoothstep(dt / s["P"]) if dt < s["P"] else 1.0
        gamma = rd.sink(nr, cn.dr, cn.r, 3, smear, s["M1"], s["M2"], s["R1"], s["R2"], theta, abs_ramp)
    h_full = h - 0.5j * gamma
    eye = np.eye(h.shape[0])
    psi0 = cn.psi
    expected = np.linalg.solve(eye + 1j * dt / 2 * h_full, (eye - 1j * dt / 2 * h_full) @ psi0)
    if scenario == 2:
        start = int(math.floor(120.0 * 0.8 / cn.dr))
        for m in range(cn.n_modes):
            for i in range(start, nr):
                expected[m * nr + i] *= math.exp(-0.01 * ((cn.r[i] - 96.0) / 24.0) ** 3)
    return cn, dict(h=h, h0=h0, v1=v1, v2=v2, dvdt=dvdt, gamma=gamma, expected=expected, psi0=psi0,
                    companion_ramp=companion_ramp)


@pytest.mark.parametrize("scenario", range(5))
def test_one_step_matches_dense_solve(scenario):
    cn, ref = correctness_case(scenario)
    cn.begin_run(60.0, 2)
    # H and the sink at the step's end time, before the step moves the orbit.
    cn.step(1)
    psi = cn.psi
    exp = ref["expected"]
    # BiCGSTAB stops at relative residual 1e-10; the reference test allows 1e-8.
    assert np.linalg.norm(psi - exp) / np.linalg.norm(exp) < 1e-9
    assert cn.last_solver_residual <= 1e-9


@pytest.mark.parametrize("scenario", [0, 1])
def test_restarted_solve_matches_dense_solve(scenario):
    """The restart branch of BiCGSTAB (matrices.cpp:67-74 with the sinks,
    Eigen's without), forced by raising the thresholds to 0.1 |r0|^2: the
    restarted solve still meets the dense LU solution of the step."""
    cn, ref = correctness_case(scenario)
    cn.set_restart_thresholds(0.1, 0.1)
    cn.begin_run(60.0, 2)
    cn.step(1)
    assert cn.restart_count() > 0
    exp = ref["expected"]
    assert np.linalg.norm(cn.psi - exp) / np.linalg.norm(exp) < 1e-9
    assert cn.last_solver_residual <= 1e-9
    with pytest.raises(ValueError):
        cn.set_restart_thresholds(0.0, 1e-32)


@pytest.mark.parametrize("scenario", [0, 1, 2, 3])
def test_operators_match_dense(scenario):
    cn, ref = correctness_case(scenario)
    cn.begin_run(60.0, 2)
    # advance the orbit exactly as step() does, then compare H (and dV/dt) on probes
    cn.step(1)
    rng = np.random.default_rng(scenario)
    x = rng.standard_normal(cn.size) + 1j * rng.standard_normal(cn.size)
    # After step(1) the oracle holds H(t_1); the dense h was assembled at t_1 too.
    Hx = cn.apply_H(x)
    assert np.linalg.norm(Hx - ref["h"] @ x) <= 1e-12 * np.linalg.norm(ref["h"] @ x)
    dv = cn.apply_op(oracle.OP_DVDT, x)
    ex = ref["dvdt"] @ x
    assert np.linalg.norm(dv - ex) <= 1e-12 * (1 + np.linalg.norm(ex))
    if scenario in (1, 2):
        sx = cn.apply_sink(x)
        ex = ref["gamma"] @ x
        assert np.linalg.norm(ex) > 0
        assert np.linalg.norm(sx - ex) <= 1e-10 * np.linalg.norm(ex)


# ---------------------------------------------------------------- physics invariants


def generic_state(cn):
    """tests/physics.cpp:42-56."""
    psi = np.zeros(cn.size, dtype=complex)
    Nr = cn.Nr
    for j, (l, m) in enumerate(cn.modes):
        width, centre = 6.0 + 1.5 * l, 18.0 + 2.0 * m
        weight = (1.0 / (1.0 + l)) * np.exp(1j * (0.7 * m + 0.3 * l))
        i = np.arange(1, Nr - 1)
        x = (cn.r[i] - centre) / width
        psi[j * Nr + i] = weight * np.exp(-x * x) * np.sin(np.pi * i / (Nr - 1)) * np.exp(1j * 0.05 * cn.r[i])
    return psi / math.sqrt(np.vdot(psi, psi).real * cn.dr)


def make_physics(**kw):
    s = dict(Nr=65, L=64.0, dt=2.0, epsilon=0.0, l_max=2, mu=0.2, M_c=0.1, q=0.5, a=0.0, R=20.0,
             counter=False, absorption=False, save_every=1000, ramp_periods=0.0, epsilon_smear=4.0)
    s.update(kw)
    return oracle.Oracle(**s)


def test_unitarity_without_absorption():
    """tests/physics.cpp:106-124."""
    cn = make_physics(l_max=4, ramp_periods=0.2)
    cn.set_flags(evolve_orbit=0)
    cn.set_initial_state_full(generic_state(cn))
    cn.initialize()
    cn.begin_run(40, 31)
    n0 = np.vdot(cn.psi, cn.psi).real
    for step in range(1, 31):
        cn.step(step, diag=False)
        assert abs(np.vdot(cn.psi, cn.psi).real - n0) <= 1e-9 * max(1.0, n0)
    assert cn.last_solver_residual <= 1e-9


def test_exact_preconditioner_converges_in_one_iteration():
    """A pin of the BiCGSTAB loop count. With one mode (l_max = 0) and no
    absorbers, A has no inter-mode coupling: the mode's self-coupling is on
    the diagonal the Thomas pivots include, so the radial preconditioner is A's
    exact inverse. Eigen's loop (matrices.cpp:188-196) then stops after exactly
    one iteration, with the residual at rounding."""
    cn = make_physics(l_max=0, ramp_periods=0.2)
    cn.set_flags(evolve_orbit=0)
    cn.set_initial_state_full(generic_state(cn))
    cn.initialize()
    cn.begin_run(40, 11)
    for step in range(1, 11):
        cn.step(step, diag=False)
        assert cn.last_solver_iterations == 1, step
        assert cn.last_solver_residual <= 1e-13


def test_multipole_cut_probe():
    """The oracle's truncation probe (a checker for truncating applies), pinned
    by the Gaunt selection rules: with only l* = 0 kept, the coupling is the
    monopole, which joins each mode to itself alone, so a single-mode input
    stays in its mode; a cut at 2 l_max keeps every term; a cut in between
    changes the apply; no cut restores it."""
    cn = make_physics(l_max=3, Nr=129, L=128.0)
    cn.set_flags(evolve_orbit=0)
    cn.set_initial_state_full(generic_state(cn))
    cn.initialize()
    cn.begin_run(40, 2)
    rng = np.random.default_rng(5)
    x = rng.standard_normal(cn.size) + 1j * rng.standard_normal(cn.size)
    full = cn.apply_H(x)
    tiles = (cn.Nr + 63) // 64
    cn.set_multipole_cut(np.full(tiles, 6))
    assert np.array_equal(cn.apply_H(x), full)
    cn.set_multipole_cut(np.full(tiles, 0))
    K, N = cn.n_modes, cn.Nr
    for j in range(K):
        e = np.zeros(cn.size, dtype=complex)
        e[j * N:(j + 1) * N] = x[j * N:(j + 1) * N]
        y = cn.apply_H(e).reshape(K, N)
        assert np.abs(np.delete(y, j, axis=0)).max() == 0.0, j
        assert np.abs(y[j]).max() > 0.0
    cn.set_multipole_cut([2, 6, 6])
    assert np.linalg.norm(cn.apply_H(x) - full) > 1e-6 * np.linalg.norm(full)
    cn.set_multipole_cut(None)
    assert np.array_equal(cn.apply_H(x), full)


def test_hermitian_and_sink_psd():
    """tests/physics.cpp:77-104 on operator probes."""
    cn = make_physics(l_max=4, absorption=True)
    cn.set_initial_state_full(generic_state(cn))
    cn.initialize()
    cn.begin_run(40, 4)
    for step in range(1, 4):
        cn.step(step, diag=False)
    n = cn.size
    interior = np.ones(n)
    for j in range(cn.n_modes):
        interior[j * cn.Nr] = interior[j * cn.Nr + cn.Nr - 1] = 0
    H = np.array([cn.apply_H(e) for e in np.eye(n)]).T * interior[:, None] * interior[None, :]
    assert np.linalg.norm(H - H.conj().T) < 1e-12 * np.linalg.norm(H)
    a = generic_state(cn)
    b = np.roll(a, 7)
    ab, ba = np.vdot(a, cn.apply_sink(b)), np.vdot(b, cn.apply_sink(a))
    assert abs(ab - np.conj(ba)) < 1e-12 * (1 + abs(ab))
    assert np.vdot(a, cn.apply_sink(a)).real >= 0


def test_hydrogen_energy_and_phase():
    """tests/physics.cpp:126-170 for n = 1 (isolated BH at the origin)."""
    cn = make_physics(Nr=513, L=80.0, mu=0.5, l_max=2, dt=0.5, M_c=0.5)
    cn.configure_isolated_black_hole(0.0)
    a0 = 1 / 0.25
    psi = np.zeros(cn.size, dtype=complex)
    j = cn.modes.index((0, 0))
    psi[j * cn.Nr:(j + 1) * cn.Nr] = cn.r * np.exp(-cn.r / a0)
    psi /= math.sqrt(np.vdot(psi, psi).real * cn.dr)
    cn.set_initial_state_full(psi)
    cn.initialize()
    init = cn.psi
    norm = np.vdot(init, init).real
    energy = np.vdot(init, cn.apply_H(init)).real / norm
    exact = -0.5 ** 3 / 2
    assert abs(energy - exact) < 2e-3 * abs(exact)
    cn.begin_run(60, 41)
    for step in range(1, 41):
        cn.step(step, diag=False)
    overlap = np.vdot(init, cn.psi) / norm
    assert abs(overlap) > 1 - 1e-4
    expected_phase = -2.0 * 40 * math.atan(energy * 0.5 / 2.0)
    assert abs(math.remainder(np.angle(overlap) - expected_phase, 2 * math.pi)) < 1e-3


def test_fixed_orbit_absorption_balance():
    """tests/fixed_orbit_absorption.cpp:37-105, resolved equal-mass case."""
    cn = oracle.Oracle(65, 64.0, 1.0, 0.0, 0, 0.2, 0.1, 1.0, 0.0, 20.0, counter=True, absorption=True,
                       save_every=100, ramp_periods=0.0, epsilon_smear=4.0)
    cn.set_flags(evolve_orbit=0)
    P = cn.P
    cn.set_scalars(dt=P / 16)
    init = np.zeros(cn.Nr)
    init[1:-1] = np.sin(np.pi * np.arange(1, cn.Nr - 1) / (cn.Nr - 1))
    cn.set_initial_state_full(init)
    cn.initialize()
    cn.begin_run(40.0, 41)
    n0 = np.vdot(cn.psi, cn.psi).real
    total = 0.0
    for step in range(1, 41):
        prev = cn.psi
        d = cn.step(step)
        loss = np.vdot(prev, prev).real - np.vdot(cn.psi, cn.psi).real
        mid = 0.5 * (prev + cn.psi)
        predicted = cn.dt * np.vdot(mid, cn.apply_sink(mid)).real
        assert abs(loss - predicted) < 1e-9 * n0
        assert d["absorbed"] == pytest.approx(cn.params.mu * cn.dr * predicted, rel=1e-9, abs=1e-300)
        total += loss
    assert total > 1e-4 * n0


def test_rolling_average_constant_and_window():
    """tests/rolling_average.cpp behaviours: a constant signal averages to itself;
    the window cut keeps the average of a step signal exact."""
    avg = oracle.Rolling(1.0)
    for k in range(50):
        v = avg.update(0.1 * k, 0.2 * k, 3.5)
        assert v == pytest.approx(3.5, rel=1e-15)
    avg = oracle.Rolling(1.0)
    out = None
    for k in range(200):
        out = avg.update(float(k), 0.1 * k, 0.0 if k < 100 else 1.0)
    assert out == pytest.approx(1.0, rel=1e-12)
    with pytest.raises(RuntimeError):
        avg.update(199.0, 100.0, 1.0)
