"""The reference's public sparse operators (H, H0, A, B, V1, V2, V, dVdt, Id,
CAP; include/CrankNicolsonModes.hpp:197-206) materialised from the device
tables, against the FP64 oracle's matrix-free operators and the relations the
reference builds them by (hamiltonian.cpp:68-86,205-243, matrices.cpp:27-40),
plus the Hermiticity checks of tests/physics.cpp:80-95.

Setup: l_max 4 (15 modes), Nr 121, absorbers on, a CAP, a startup ramp of
one period with the companion's potential ramped, so that V = V1 at t = 0 and
V = V1 + ramp V2 with 0 < ramp < 1 after a few steps. Operators agree with the
oracle to 1e-12 relative on random vectors.
"""
import math
import os
import subprocess
import sys

import numpy as np
import pytest

import gm
import oracle
from test_oracle import generic_state

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))

OP_A, OP_B, OP_H, OP_V, OP_DVDT = range(5)  # the oracle's matrix-free operators

ARGS = dict(Nr=121, L=60.0, dt=0.25, epsilon=0.0, l_max=4, mu=0.2, M_c=0.01, q=0.5, a=0.0, R=12.0,
            counter=False, absorption=True, save_every=1000, ramp_periods=1.0, epsilon_smear=2.0)


def announce(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def setup(steps):
    objs = gm.CrankNicolsonModes(**ARGS), oracle.Oracle(**ARGS)
    for o in objs:
        o.set_flags(evolve_orbit=0, ramp_secondary_potential=1)
    psi = generic_state(objs[1])
    for o in objs:
        o.set_initial_state_full(psi)
        o.initialize()
        o.build_cap(0.8, 0.5, 2)
        o.begin_run(40.0, steps + 1)
    for n in range(1, steps + 1):
        for o in objs:
            o.step(n, diag=False)
    return objs


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def vectors(n, k=3, seed=0):
    rng = np.random.default_rng(seed)
    return [rng.standard_normal(n) + 1j * rng.standard_normal(n) for _ in range(k)]


def ramp(t):
    P0 = 2 * math.pi / math.sqrt((1 + ARGS["q"]) / ARGS["R"] ** 3)
    xx = t / (ARGS["ramp_periods"] * P0)
    return 3 * xx * xx - 2 * xx ** 3


def test_operators_at_start_match_the_oracle():
    announce("operators at t = 0 (companion potential ramped to zero)")
    g, o = setup(0)
    ops = {name: g.operator(name) for name in g.OPERATORS}
    n = o.size
    for x in vectors(n):
        assert rel(ops["V"] @ x, o.apply_op(OP_V, x)) < 1e-12
        # the companion's weight is zero: V is the primary's potential alone
        assert rel(ops["V1"] @ x, o.apply_op(OP_V, x)) < 1e-12
        assert rel(ops["H"] @ x, o.apply_op(OP_H, x)) < 1e-12
    assert abs(ops["Id"] - gm_identity(n)).max() == 0
    g.close()


def gm_identity(n):
    import scipy.sparse
    return scipy.sparse.identity(n, dtype=complex, format="csr")


def test_operators_after_steps_match_the_oracle():
    announce("operators after 6 steps of the startup ramp")
    g, o = setup(6)
    ops = {name: g.operator(name) for name in g.OPERATORS}
    n = o.size
    t = g.scalars()["t"]
    w = ramp(t)
    assert 0.0 < w < 1.0
    for x in vectors(n, seed=1):
        for name, op in (("H", OP_H), ("V", OP_V), ("dVdt", OP_DVDT), ("A", OP_A), ("B", OP_B)):
            assert rel(ops[name] @ x, o.apply_op(op, x)) < 1e-12, name
        # H = H0 + V + CAP (hamiltonian.cpp:224-243)
        h_ref = o.apply_op(OP_H, x)
        assert rel((ops["H0"] + ops["V"] + ops["CAP"]) @ x, h_ref) < 1e-12
        # V = V1 + ramp V2 (hamiltonian.cpp:214-217)
        assert rel((ops["V1"] + w * ops["V2"]) @ x, ops["V"] @ x) < 1e-12
    # the CAP is diagonal, zero inside its radius and negative imaginary beyond
    cap = ops["CAP"]
    assert abs(cap - cap.multiply(gm_identity(n))).max() == 0
    d = cap.diagonal()
    assert np.all(d.real == 0) and np.all(d.imag <= 0) and np.any(d.imag < 0)
    # A = I + i dt/2 H and B = I - i dt/2 H (matrices.cpp:27-40)
    I = gm_identity(n)
    hd = 0.5 * ARGS["dt"]
    assert abs(ops["A"] - (I + 1j * hd * ops["H"])).max() < 1e-14 * abs(ops["H"]).max() + 1e-15
    assert abs(ops["B"] - (I - 1j * hd * ops["H"])).max() < 1e-14 * abs(ops["H"]).max() + 1e-15
    g.close()


def test_hamiltonian_and_dvdt_are_hermitian():
    """tests/physics.cpp:80-95: the conservative Hamiltonian without the CAP is
    Hermitian on the interior nodes (Dirichlet rows only pin the boundary), and
    so is dV/dt."""
    announce("Hermiticity of H0 + V and dV/dt")
    g, o = setup(3)
    H = (g.operator("H0") + g.operator("V")).tocsr()
    Nr, K = g.Nr, g.n_modes
    interior = np.ones(Nr * K)
    for j in range(K):
        interior[j * Nr] = interior[j * Nr + Nr - 1] = 0
    import scipy.sparse
    P = scipy.sparse.diags(interior)
    h = P @ H @ P
    assert abs(h - h.conj().T).max() < 1e-12 * abs(h).max()
    dv = g.operator("dVdt")
    assert dv.nnz > 0
    assert abs(dv - dv.conj().T).max() < 1e-12 * (1 + abs(dv).max())
    g.close()


def test_operator_rejects_large_systems_and_unknown_names():
    announce("operator size limit")
    s = gm.CrankNicolsonModes(Nr=8192, L=2048.0, dt=0.25, epsilon=0.0, l_max=20, mu=0.15, M_c=0.01, q=0.5,
                              a=0.0, R=50.0)
    s.set_initial_state_full(np.zeros(s.size, dtype=complex) + 1e-3)
    s.initialize()
    with pytest.raises(gm.InvalidArgument):
        s.operator("H")
    with pytest.raises(ValueError):
        s.operator("W")
    s.close()


def test_reference_operator_checks_through_the_dropin_class(tmp_path):
    """tests/cpp/physics_ops.cpp: the reference's operator checks
    (tests/physics.cpp:80-102) written against include/gm_cnm.hpp, whose
    refresh_operators() fills the public members H, H0, A, B, V1, V2, V,
    dVdt, Id and CAP."""
    announce("drop-in class: operator checks of the reference's physics test")
    exe = os.path.join(HERE, "cpp", "_bin", "physics_ops")
    res = subprocess.run([exe], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    print(res.stdout, flush=True)
    assert res.returncode == 0, res.stdout + res.stderr
    assert "all checks passed" in res.stdout
