"""Pins of the oracle's diagnostics, orbit integrator and rolling average (CPU).

Each check restates one of the reference's own tests of
rdellamonica/boson-cloud-binary against the FP64 oracle, or a balance law the
mathematics fixes, so that the oracle's flux, energy flux, norms, energies,
binary power p_bin_inst, dV_bin_dt, absorbed mass and orbit scalars
(R, M1, M2, q, R1, R2, Omega, P, dt) each answer to something outside the
oracle:

- tests/physics.cpp: Jacobi constant, co/counter mirror symmetry, probability
  flux continuity, gravitational-wave inspiral against RK4, accretion mass
  balance, damping mask, complex absorbing potential;
- tests/orbital_ramp.cpp: the startup ramp, back-reaction switch-on and the
  companion-potential weight (operators checked against the dense sympy
  restatement in reference_dense.py);
- tests/flux_time_alignment.cpp: the end-of-step pairing of the energy flux
  and dV_bin_dt with the post-mask state and its derivative;
- tests/rolling_average.cpp: every case of RollingPhaseAverage;
- energy and potential balances: for CN with H(t_{n+1}) on both sides the
  energy changes by <psi_n|H_{n+1} - H_n|psi_n>, which the binary power
  (the dV/dt expectation) predicts; the potential energy changes by
  <dV/dt> + 2 Re<dpsi/dt|V psi>, which fixes dV_bin_dt.
"""
import math

import numpy as np
import pytest

import oracle
import reference_dense as rd
from test_oracle import generic_state, make_physics


def mode_index(cn, l, m):
    return cn.modes.index((l, m))


def mirror(cn, psi):
    """tests/physics.cpp:58-69: U'_{l,m} = (-1)^m U_{l,-m}."""
    out = np.empty_like(psi)
    N = cn.Nr
    for j, (l, m) in enumerate(cn.modes):
        p = mode_index(cn, l, -m)
        out[j * N:(j + 1) * N] = (1.0 if m % 2 == 0 else -1.0) * psi[p * N:(p + 1) * N]
    return out


def cloud_mass(cn, psi):
    return cn.params.mu * cn.dr * np.vdot(psi, psi).real


# ---------------------------------------------------------------- rolling average (tests/rolling_average.cpp)


def test_rolling_first_sample_and_readiness():
    avg = oracle.Rolling(1.0)
    assert not avg.ready()
    assert avg.update(0.0, 0.0, 3.5) == 3.5
    for k in range(1, 100):
        avg.update(float(k), 2 * math.pi * k / 100, 3.5)
    assert not avg.ready()
    avg.update(100.0, 2 * math.pi, 3.5)
    assert avg.ready()
    assert avg.value() == pytest.approx(3.5, abs=1e-14)


def test_rolling_linear_signal_is_exact():
    """The trapezoidal average of a linear signal is the window's midpoint
    value, also when the window edge falls between samples."""
    rate, n_orbits, omega, dt = 0.37, 1.3, 0.05, 0.71
    avg = oracle.Rolling(n_orbits)
    window = 2 * math.pi * n_orbits / omega
    checked = 0
    for k in range(2000):
        t = k * dt
        res = avg.update(t, omega * t, 2.0 + rate * t)
        if t >= window:
            assert res == pytest.approx(2.0 + rate * (t - 0.5 * window), abs=1e-10)
            checked += 1
    assert checked > 1000


def test_rolling_oscillation_averages_out():
    avg = oracle.Rolling(2.0)
    res = None
    for k in range(4001):
        ph = 2 * math.pi * k / 400
        res = avg.update(0.1 * k, ph, 1.0 + math.sin(ph) + 0.5 * math.cos(2 * ph))
    assert res == pytest.approx(1.0, abs=1e-4)


def test_rolling_chirping_orbit():
    avg = oracle.Rolling(1.0)
    phase, res = 0.0, None
    for k in range(3001):
        t, omega = 0.5 * k, 0.01 * (1 + 0.001 * k)
        res = avg.update(t, phase, -2.25)
        phase += omega * 0.5
    assert res == pytest.approx(-2.25, abs=1e-13)


def test_rolling_rejects_non_monotonic_input():
    avg = oracle.Rolling(1.0)
    avg.update(1.0, 1.0, 0.0)
    with pytest.raises(RuntimeError, match="time"):
        avg.update(1.0, 2.0, 0.0)
    with pytest.raises(RuntimeError, match="phase"):
        avg.update(2.0, 1.0, 0.0)
    with pytest.raises(ValueError):
        oracle.Rolling(0.0)


def test_rolling_reset_and_checkpoint(tmp_path):
    avg = oracle.Rolling(0.7)
    for k in range(300):
        avg.update(float(k), 0.03 * k, math.cos(0.1 * k))
    path = tmp_path / "avg.bin"
    avg.save(path)
    restored = oracle.Rolling(5.0)  # the window is part of the saved state
    restored.load(path)
    for k in range(300, 600):
        a = avg.update(float(k), 0.03 * k, math.cos(0.1 * k))
        b = restored.update(float(k), 0.03 * k, math.cos(0.1 * k))
        assert a == b
    assert avg.ready() == restored.ready()
    trunc = tmp_path / "trunc.bin"
    trunc.write_bytes(path.read_bytes()[:20])
    with pytest.raises(RuntimeError, match="truncated"):
        restored.load(trunc)
    avg.reset()
    assert not avg.ready()
    assert avg.update(1000.0, 50.0, 7.0) == 7.0


# ---------------------------------------------------------------- tests/physics.cpp


def jacobi_drift(counter, per_period):
    """tests/physics.cpp:174-208: largest relative drift of <H> - Omega <L_z>
    over two orbits of a fixed circular orbit."""
    cn = make_physics(l_max=4, counter=counter)
    cn.set_flags(evolve_orbit=0)
    cn.set_scalars(dt=cn.P / per_period)
    cn.set_initial_state_full(generic_state(cn))
    cn.initialize()
    cn.begin_run(40, 2 * per_period + 1)
    N = cn.Nr
    ms = np.repeat([m for (_, m) in cn.modes], N)

    def expectation(jacobi):
        psi = cn.psi
        lz = float(np.sum(ms * np.abs(psi) ** 2))
        energy = np.vdot(psi, cn.apply_H(psi)).real
        sign = -1.0 if counter else 1.0
        return (energy - (sign * cn.Omega * lz if jacobi else 0.0)) * cn.dr

    cn.step(1, diag=False)
    j0, e0 = expectation(True), expectation(False)
    drift = energy_change = 0.0
    for step in range(2, 2 * per_period + 1):
        cn.step(step, diag=False)
        drift = max(drift, abs(expectation(True) - j0) / abs(j0))
        energy_change = max(energy_change, abs(expectation(False) - e0) / abs(e0))
    assert energy_change > 20 * drift, "rotating potential did no work on the cloud"
    return drift


@pytest.mark.parametrize("counter", [False, True])
def test_jacobi_constant_on_fixed_orbit(counter):
    """tests/physics.cpp:210-219."""
    coarse, fine = jacobi_drift(counter, 200), jacobi_drift(counter, 800)
    assert coarse < 1e-3
    assert fine < coarse / 3


def test_co_counter_mirror_symmetry():
    """tests/physics.cpp:221-251: y -> -y maps the co-rotating binary onto the
    counter-rotating one, with absorbers and a startup ramp."""
    runs = []
    for counter in (False, True):
        cn = make_physics(l_max=3, absorption=True, ramp_periods=0.3, counter=counter)
        cn.set_flags(evolve_orbit=0)
        runs.append(cn)
    co, ctr = runs
    init = generic_state(co)
    co.set_initial_state_full(init)
    ctr.set_initial_state_full(mirror(ctr, init))
    for cn in runs:
        cn.initialize()
        cn.begin_run(40, 41)
    for step in range(1, 41):
        co.step(step, diag=False)
        ctr.step(step, diag=False)
    assert abs(co.theta + ctr.theta) < 1e-12
    a, b = co.psi, ctr.psi
    assert np.linalg.norm(mirror(co, a) - b) / np.linalg.norm(a) < 1e-7
    assert np.linalg.norm(a - b) > 1e-3 * np.linalg.norm(a), "mirror test is vacuous"


def test_probability_flux_continuity():
    """tests/physics.cpp:253-299 on the oracle's own diagnostics. For CN the
    mass inside node k (half weight on k) changes by -dt times the centred flux
    of the midpoint state. The flux is a quadratic form F, and
    F(mid) = (F(a) + F(b))/2 - F(b - a)/4, so the flux values the oracle
    reports after consecutive steps, with the small F(b - a) term evaluated
    here, must give the change of mass; the per-mode norms the oracle reports
    (nodes below k) must give that mass too."""
    cn = make_physics(l_max=4)
    cn.set_flags(evolve_orbit=0)
    cn.set_scalars(dt=cn.P / 64)
    cn.set_initial_state_full(generic_state(cn))
    cn.initialize()
    r_ex = 30.0
    k = int(math.floor(r_ex / cn.dr))
    cn.begin_run(r_ex, 41)
    mu, dr, N, K = cn.params.mu, cn.dr, cn.Nr, cn.n_modes

    def test_flux(psi):
        u = psi.reshape(K, N)
        dU = (u[:, k + 1] - u[:, k - 1]) / (2 * dr) - u[:, k] / cn.r[k]
        return float(np.sum((np.conj(u[:, k]) * dU).imag))

    def edge(psi):
        return 0.5 * float(np.sum(np.abs(psi.reshape(K, N)[:, k]) ** 2))

    prev_psi, prev_flux, prev_norms = cn.psi, None, None
    largest = 0.0
    for step in range(1, 41):
        d = cn.step(step)
        psi = cn.psi
        total_flux = float(np.sum(d["flux"]))
        assert total_flux == pytest.approx(test_flux(psi), rel=1e-12, abs=1e-300)
        inner = mu * (float(np.sum(d["norms"])) + dr * edge(psi))
        assert inner == pytest.approx(mu * dr * (np.sum(np.abs(psi.reshape(K, N)[:, :k]) ** 2) + edge(psi)),
                                      rel=1e-12)
        if prev_flux is not None:
            change = inner - prev_inner
            predicted = -cn.dt * (0.5 * (prev_flux + total_flux) - 0.25 * test_flux(psi - prev_psi))
            largest = max(largest, abs(predicted))
            assert abs(change - predicted) < 1e-8 * cloud_mass(cn, prev_psi)
        prev_psi, prev_flux, prev_inner = psi, total_flux, inner
    assert largest > 1e-6 * cn.params.M_c, "no flux crossed r_ex"


def test_gravitational_wave_inspiral():
    """tests/physics.cpp:301-352: with a negligible cloud the separation follows
    dR/dt = -(64/5) M1 M2 M / R^3, switched on over one initial period; checked
    against an independent RK4 integration, Omega Keplerian, R1 = R/2 and
    dt = P / steps_per_period after the ramp; R^4 decreases linearly."""
    cn = make_physics(q=1.0, R=16.0, M_c=1e-12, l_max=0)
    cn.set_flags(evolve_orbit=1, steps_per_period=200)
    P0 = cn.P
    cn.set_scalars(dt=P0 / 200)
    cn.set_initial_state_full(generic_state(cn))
    cn.initialize()
    steps = 600
    cn.begin_run(40, steps + 1)

    def ramp(t):
        x = min(max(t / P0, 0.0), 1.0)
        return 3 * x * x - 2 * x * x * x

    def f(t, R):
        return -ramp(t) * (64.0 / 5.0) * 2.0 / (R * R * R)

    R_ref, t_ref = 16.0, 0.0
    R4_start = t_start = 0.0
    for step in range(1, steps + 1):
        cn.step(step, diag=False)
        s = cn.scalars()
        t_end = s["t"]
        sub = 20
        h = (t_end - t_ref) / sub
        for _ in range(sub):
            k1 = f(t_ref, R_ref)
            k2 = f(t_ref + h / 2, R_ref + h * k1 / 2)
            k3 = f(t_ref + h / 2, R_ref + h * k2 / 2)
            k4 = f(t_ref + h, R_ref + h * k3)
            R_ref += h * (k1 + 2 * k2 + 2 * k3 + k4) / 6
            t_ref += h
        t_ref = t_end
        assert abs(s["R"] - R_ref) <= 2e-3 * max(1.0, abs(R_ref)), step
        assert s["Omega"] == pytest.approx(math.sqrt(2.0 / s["R"] ** 3), rel=1e-12)
        assert s["R1"] == pytest.approx(s["R"] / 2, rel=1e-12)
        assert s["dt"] == pytest.approx(s["P"] / 200, rel=1e-12)
        assert s["P"] == pytest.approx(2 * math.pi / s["Omega"], rel=1e-14)
        if t_end > 1.5 * P0 and t_start == 0:
            t_start, R4_start = t_end, s["R"] ** 4
    s = cn.scalars()
    assert s["R"] < 16.0 - 0.5
    expected = R4_start - (256.0 / 5.0) * 2.0 * (s["t"] - t_start)
    assert s["R"] ** 4 == pytest.approx(expected, rel=5e-3)


def test_accretion_balances_absorbed_mass():
    """tests/physics.cpp:354-382: mass the sinks take from the field is added
    to the black holes (1e-10), q tracks the masses, both holes accrete."""
    cn = make_physics(q=1.0, R=20.0, l_max=2, M_c=0.05, absorption=True)
    cn.set_flags(evolve_orbit=1, steps_per_period=16)
    cn.set_scalars(dt=cn.P / 16)
    cn.set_initial_state_full(generic_state(cn))
    cn.initialize()
    cn.begin_run(40, 41)
    s = cn.scalars()
    cloud0 = cloud_mass(cn, cn.psi)
    total0 = cloud0 + s["M1"] + s["M2"]
    for step in range(1, 41):
        d = cn.step(step)
        s = cn.scalars()
        psi = cn.psi
        assert cloud_mass(cn, psi) + s["M1"] + s["M2"] == pytest.approx(total0, abs=1e-10 * max(1.0, total0))
        assert s["q"] == pytest.approx(s["M2"] / s["M1"], rel=1e-14)
        assert d["cloud_mass"] == pytest.approx(cloud_mass(cn, psi), rel=1e-12)
    assert cloud0 - cloud_mass(cn, cn.psi) > 1e-3 * cloud0
    assert s["M1"] > 1 and s["M2"] > 1


def test_damping_mask_acts_outside_only():
    """tests/physics.cpp:384-412."""
    plain, masked = make_physics(), make_physics()
    for cn in (plain, masked):
        cn.set_flags(evolve_orbit=0)
        cn.set_initial_state_full(generic_state(cn))
        cn.initialize()
    masked.build_damping_mask(0.75, 3.0, 2)
    for cn in (plain, masked):
        cn.begin_run(30, 2)
        cn.step(1, diag=False)
    edge = int(math.floor(0.75 * 64.0 / plain.dr))
    a, b = plain.psi.reshape(plain.n_modes, -1), masked.psi.reshape(plain.n_modes, -1)
    i = np.arange(plain.Nr)
    factor = np.where(i < edge, 1.0, np.exp(-3.0 * ((plain.r - 0.75 * 64) / (0.25 * 64)) ** 2))
    assert np.all(np.abs(b - factor * a) < 1e-14 * (1 + np.abs(a)))
    assert np.vdot(b, b).real < np.vdot(a, a).real


def test_complex_absorbing_potential():
    """tests/physics.cpp:414-453: the CAP is diagonal, acts from its start
    radius, is negative imaginary, and |psi'|^2 - |psi|^2 = 2 dt Im<mid|CAP|mid>."""
    plain, cap = make_physics(Nr=129, L=128.0), make_physics(Nr=129, L=128.0)
    for cn in (plain, cap):
        cn.set_flags(evolve_orbit=0)
        cn.set_initial_state_full(generic_state(cn))
        cn.initialize()
    cap.build_cap(0.8, 0.5, 2)
    for cn in (plain, cap):
        cn.begin_run(40, 21)
    N, K = plain.Nr, plain.n_modes
    n = N * K
    # H - H_plain probed with ones on the nodes i = c (mod 3) of one mode j:
    # row (j, i) of the product picks up the diagonal entry alone, so the
    # difference there is that entry of CAP exactly; every other row must not
    # change (the CAP has no off-diagonal entries)
    diff = np.zeros(n, dtype=complex)
    node = np.arange(n) % N
    for j in range(K):
        for c in range(3):
            e = np.zeros(n, dtype=complex)
            sel = (np.arange(n) // N == j) & (node % 3 == c)
            e[sel] = 1.0
            dd = cap.apply_H(e) - plain.apply_H(e)
            assert np.all(dd[~sel] == 0)
            diff[sel] = dd[sel]
    edge = int(math.floor(0.8 * 128.0 / plain.dr))
    nz = np.abs(diff) > 0
    assert np.any(nz)
    assert np.all(node[nz] >= edge)
    assert np.all(diff[nz].real == 0) and np.all(diff[nz].imag <= 0) and np.all(np.isfinite(diff))
    for step in range(1, 21):
        prev = cap.psi
        cap.step(step, diag=False)
        cur = cap.psi
        mid = 0.5 * (prev + cur)
        predicted = 2 * cap.dt * np.vdot(mid, diff * mid).imag
        change = np.vdot(cur, cur).real - np.vdot(prev, prev).real
        assert change <= 1e-12 * np.vdot(prev, prev).real
        assert abs(change - predicted) < 1e-9 * np.vdot(prev, prev).real


# ---------------------------------------------------------------- tests/orbital_ramp.cpp


def smoothstep(x):
    return 0.0 if x <= 0 else 1.0 if x >= 1 else 3 * x * x - 2 * x * x * x


def close11(actual, expected):
    assert abs(actual - expected) <= 1e-11 * max(1.0, abs(expected)), (actual, expected)


def ramp_solver(Nr, l_max, absorb, counter=False, dt=200.0, ramp=5.0):
    cn = oracle.Oracle(Nr, 200.0, dt, 0.1, l_max, 0.2, 0.0, 0.25, 0.0, 50.0, counter=counter, absorption=absorb,
                       save_every=100, ramp_periods=ramp, epsilon_smear=15.0)
    return cn


def prepare(cn):
    cn.set_initial_state_full(np.zeros(cn.size, dtype=complex))
    cn.initialize()
    cn.begin_run(100.0, 32)


def set_radius(cn, radius):
    s = cn.scalars()
    M1, M2 = s["M1"], s["M2"]
    Om = math.sqrt((M1 + M2) / radius ** 3)
    cn.set_scalars(R=radius, R1=M2 / (M1 + M2) * radius, R2=M1 / (M1 + M2) * radius, Omega=Om, P=2 * math.pi / Om)


def expected_radius(cn, duration, ramp=1.0):
    s = cn.scalars()
    return s["R"] - (64.0 / 5.0) * s["M1"] * s["M2"] * (s["M1"] + s["M2"]) / s["R"] ** 3 * duration * ramp


def companion_weight_error(cn, weight, pre=None):
    """|V x - (V1 + weight V2) x| / |V x| with V1, V2 from the dense sympy
    restatement. A step builds V at its end time from the orbit it starts
    with (the orbit update follows the solve), so the positions and masses are
    those of `pre`, the scalars before the last step, and the angle is the
    current one."""
    s = dict(cn.scalars())
    if pre is not None:
        for k in ("R1", "R2", "M1", "M2", "q"):
            s[k] = pre[k]
    _, v1, v2, _ = rd.operators(cn.Nr, cn.dr, cn.r, cn.params.l_max, cn.params.mu, s["M1"], s["M2"], s["q"],
                                s["R1"], s["R2"], s["theta"], s["Omega"], bool(cn.params.counter), 1.0)
    rng = np.random.default_rng(11)
    x = rng.standard_normal(cn.size) + 1j * rng.standard_normal(cn.size)
    vx = cn.apply_op(oracle.OP_V, x)
    return np.linalg.norm(vx - (v1 + weight * v2) @ x) / max(1.0, np.linalg.norm(vx)), np.linalg.norm(v2 @ x)


@pytest.mark.parametrize("absorb", [False, True])
def test_orbital_ramp_expanding_orbit(absorb):
    """orbital_ramp.cpp:58-107: after the startup ramp, back-reaction and
    absorption switch on over one initial period and stay on when the orbit
    expands (a longer period must not re-enter the ramp)."""
    cn = ramp_solver(64, 0, absorb)
    P0 = cn.P
    ramp_end = 5 * P0
    prepare(cn)
    cn.set_scalars(dt=ramp_end + 0.1 * P0)
    cn.step(1)
    t1 = cn.t
    set_radius(cn, 80.0)
    cn.set_scalars(dt=0.1 * P0)
    assert 5 * cn.P > t1 + cn.dt
    expected = expected_radius(cn, cn.dt, smoothstep((t1 + cn.dt - ramp_end) / P0))
    pre = cn.scalars()
    cn.step(2)
    close11(cn.R, expected)
    err, _ = companion_weight_error(cn, 1.0, pre)
    assert err < 1e-13
    partial = cn.sink_strength(1)
    if absorb:
        assert partial > 0
    cn.set_scalars(dt=P0)
    cn.step(3)
    set_radius(cn, 80.0)
    for step in range(4, 9):
        cn.set_scalars(dt=0.05 * P0)
        nxt = expected_radius(cn, cn.dt)
        pre = cn.scalars()
        cn.step(step)
        close11(cn.R, nxt)
        err, _ = companion_weight_error(cn, 1.0, pre)
        assert err < 1e-13
        if absorb and step == 4:
            close11(partial, cn.sink_strength(1) * smoothstep(0.2))


def test_orbital_ramp_crossing_step():
    """orbital_ramp.cpp:109-126: only the part of a step after the ramp's end
    moves the orbit, weighted by the back-reaction ramp; the step keeps its
    time stamp."""
    cn = ramp_solver(32, 0, False)
    P0 = cn.P
    prepare(cn)
    cn.set_scalars(dt=4.9 * P0)
    cn.step(1)
    close11(cn.R, 50.0)
    cn.set_scalars(dt=0.2 * P0)
    expected = expected_radius(cn, 0.1 * P0, smoothstep((5.1 * P0 - 5 * P0) / P0))
    cn.step(2)
    close11(cn.R, expected)
    close11(cn.t, 5.1 * P0)


@pytest.mark.parametrize("ramp_potential", [True, False])
def test_orbital_ramp_frequency_and_companion_weight(ramp_potential):
    """orbital_ramp.cpp:128-154: Omega follows the smoothstep ramp, the
    companion potential carries the ramp weight only when
    ramp_secondary_potential is set, and the orbit does not move during it."""
    cn = ramp_solver(32, 2, False)
    cn.set_flags(ramp_secondary_potential=int(ramp_potential))
    P0 = cn.P
    prepare(cn)
    err, v2n = companion_weight_error(cn, 0.0 if ramp_potential else 1.0)
    assert err < 1e-13 and v2n > 0
    cn.set_scalars(dt=2.0 * P0)
    cn.step(1)
    s_ = smoothstep(0.4)
    close11(cn.Omega, cn.Omega_0 * s_)
    err, _ = companion_weight_error(cn, s_ if ramp_potential else 1.0)
    assert err < 1e-13
    close11(cn.R, 50.0)
    cn.set_scalars(dt=4.0 * P0)
    pre = cn.scalars()
    cn.step(2)
    err, _ = companion_weight_error(cn, 1.0, pre)
    assert err < 1e-13


def test_orbital_ramp_zero_ramp_fixed_orbit():
    """orbital_ramp.cpp:156-168."""
    cn = ramp_solver(32, 0, False, counter=True, dt=20.0, ramp=0.0)
    cn.set_flags(evolve_orbit=0)
    prepare(cn)
    close11(cn.Omega, cn.Omega_0)
    cn.step(1)
    cn.step(2)
    close11(cn.R, 50.0)
    assert cn.phase > 0 and cn.theta < 0


# ---------------------------------------------------------------- tests/flux_time_alignment.cpp


@pytest.mark.parametrize("mask", [False, True])
def test_flux_endpoint_alignment(mask):
    """tests/flux_time_alignment.cpp:29-122: the energy flux and dV_bin_dt a
    step reports pair the post-mask state with its own derivative (the full
    generator, sinks included), at the step's time, with a timestep that
    changes between steps; the fixture is able to detect the old-time and
    pre-mask derivative mistakes."""
    cn = oracle.Oracle(129, 64.0, 2.0, 0.1, 3, 0.2, 0.1, 0.2, 0.0, 12.0, counter=False, absorption=True,
                       save_every=1, ramp_periods=0.02, epsilon_smear=4.0)
    cn.set_flags(evolve_orbit=0)
    N, K = cn.Nr, cn.n_modes
    i = np.arange(N)
    init = np.concatenate([np.sin(np.pi * i / (N - 1)) / (1 + m) * np.exp(1j * (0.15 * i + 0.3 * m))
                           for m in range(K)])
    init /= np.linalg.norm(init)
    cn.set_initial_state_full(init)
    cn.initialize()
    cn.build_cap(0.75, 0.02, 3)
    if mask:
        cn.build_damping_mask(0.3, 3, 2)
    cn.begin_run(42.5, 5)
    ix = int(math.floor(42.5 / cn.dr))
    # the initial derivative is -i H psi with the CAP enabled after initialize()
    psi = cn.psi
    assert np.linalg.norm(cn.schrodinger_rhs(psi) + 1j * cn.apply_H(psi)) <= 1e-14 * np.linalg.norm(cn.apply_H(psi))
    s = cn.scalars()
    h0, v1, _, _ = rd.operators(N, cn.dr, cn.r, 3, 0.2, s["M1"], s["M2"], s["q"], s["R1"], s["R2"], s["theta"],
                                s["Omega"], False, 0.0)
    capd = np.zeros(N, dtype=complex)
    f = 0.75
    for n in range(int(math.floor(64.0 * f / cn.dr)), N):
        capd[n] = -1j * 0.02 * ((cn.r[n] - f * 64.0) / (cn.r[n] - (1 - f) * 64.0)) ** 3
    hx = (h0 + v1) @ psi + np.tile(capd, K) * psi  # the companion weight is 0 at t = 0 of the ramp
    assert np.linalg.norm(cn.apply_H(psi) - hx) <= 1e-12 * np.linalg.norm(hx)

    def flux_at(state, deriv, m):
        k = m * N + ix
        dU = (state[k + 1] - state[k - 1]) / (2 * cn.dr) - state[k] / cn.r[ix]
        return (np.conj(deriv[k]) * dU).real / 0.2

    def close2(actual, expected):
        assert abs(actual - expected) <= 2e-11 * max(1e-10, abs(expected)), (actual, expected)

    detects_prev = detects_premask = False
    for step in range(1, 5):
        cn.set_scalars(dt=3.1 if step % 2 == 0 else 2.0)
        t_expected = cn.t + cn.dt
        prev_deriv = cn.schrodinger_rhs(cn.psi)
        d = cn.step(step)
        psi = cn.psi
        deriv = cn.schrodinger_rhs(psi)
        vpsi = cn.apply_op(oracle.OP_V, psi)
        sel = (np.arange(N * K) % N) <= ix
        exp_dv = np.sum(np.conj(deriv[sel]) * vpsi[sel])
        stale = np.sum(np.conj(prev_deriv[sel]) * vpsi[sel])
        detects_prev |= abs(exp_dv.real - stale.real) > 1e-8
        close2(cn.t, t_expected)
        close2(d["dV_bin_dt"], exp_dv.real * cn.dr)
        for m in range(K):
            close2(d["energy_flux"][m], flux_at(psi, deriv, m))
        if mask:
            before = psi.copy().reshape(K, N)
            start = int(math.floor(0.3 * 64.0 / cn.dr))
            before[:, start:] /= np.exp(-3 * ((cn.r[start:] - 0.3 * 64.0) / (0.7 * 64.0)) ** 2)
            bad = cn.schrodinger_rhs(before.ravel())
            detects_premask |= abs(flux_at(psi, bad, 0) - d["energy_flux"][0]) > 1e-8
    assert detects_prev
    assert detects_premask or not mask


# ---------------------------------------------------------------- energy and potential balances


def test_energy_and_potential_balance_fix_power_diagnostics():
    """Balances that fix the binary power p_bin_inst, the energy and dV_bin_dt.

    A step solves (1 + i dt/2 H') psi' = (1 - i dt/2 H') psi with H' = H(t_{n+1})
    on both sides, so psi' - psi = -i dt H' mid with mid = (psi + psi')/2, and
    for any Hermitian O, <psi'|O|psi'> - <psi|O|psi> = 2 dt Re<-i H' mid|O mid>.
    Without sinks, mask or CAP:

    - O = H': <psi'|H'|psi'> = <psi|H'|psi> exactly, so the energy changes by
      <psi|H' - H|psi> = <psi|V' - V|psi>, which the trapezoid rule in the
      orbital angle gives as dt (P + Q)/2 to O(dt^3), with P = p_bin_inst of
      the last step (dV/dt at the old angle) and Q the same expectation at the
      new angle;
    - O = V': the potential energy changes by 2 dt D'(mid) + dt (P + Q)/2,
      where D'(x) = Re<-i H' x|V' x> is a quadratic form whose value at psi'
      is dV_bin_dt of the step, so D'(mid) = (D'(psi') + D'(psi))/2 -
      D'(psi' - psi)/4.

    The extraction radius is the last interior node, so the restricted sums
    cover the whole field. A wrong sign or factor of Omega or m*, the wrong
    part (real/imaginary), or a derivative of the wrong state or time in
    p_bin_inst or dV_bin_dt breaks these by O(1); the energy relation is exact
    to rounding."""
    cn = make_physics(l_max=4, Nr=129, L=128.0)
    cn.set_flags(evolve_orbit=0)
    cn.set_scalars(dt=cn.P / 400)
    cn.set_initial_state_full(generic_state(cn))
    cn.initialize()
    r_ex = (cn.Nr - 1.5) * cn.dr
    assert int(math.floor(r_ex / cn.dr)) == cn.Nr - 2
    cn.begin_run(r_ex, 41)
    dr = cn.dr

    def expect(op, x):
        return np.vdot(x, cn.apply_op(op, x)).real * dr

    def dform(x):  # D'(x) with the solver's current H' and V'
        return np.vdot(cn.schrodinger_rhs(x), cn.apply_op(oracle.OP_V, x)).real * dr

    prev = None
    for step in range(1, 41):
        old = cn.psi
        d = cn.step(step)
        psi, dt = cn.psi, cn.dt
        v_new = expect(oracle.OP_V, psi)
        if prev is not None:
            e_old, p_old, v_old = prev
            # exact: the new energy is the old state's under the new H
            assert d["energy"] == pytest.approx(np.vdot(old, cn.apply_H(old)).real * dr, rel=1e-12)
            q = expect(oracle.OP_DVDT, old)
            de = d["energy"] - e_old
            pred_e = 0.5 * dt * (p_old + q)
            assert abs(de - pred_e) < 3e-4 * abs(de), (step, de, pred_e)
            dmid = 0.5 * (d["dV_bin_dt"] + dform(old)) - 0.25 * dform(psi - old)
            dv = v_new - v_old
            pred_v = 2 * dt * dmid + pred_e
            assert abs(dv - pred_v) < 3e-4 * abs(dv), (step, dv, pred_v)
            assert abs(2 * dt * d["dV_bin_dt"]) > 0.1 * abs(dv)  # the dV_bin_dt term matters
        prev = (d["energy"], d["p_bin_inst"], v_new)
