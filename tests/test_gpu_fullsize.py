"""Size-independent properties at the full l_max 20, N 16384 shape (needs an
MI355X). The oracle takes 5-20 s per step here; it is compared with the GPU
over the first 10 steps in test_gpu_config2.py. These checks run more steps
and other states, and rest on what the mathematics fixes:

- H is Hermitian on vectors that vanish at the two Dirichlet nodes (the
  three-point stencil, the centrifugal term and the Gaunt coupling are all
  symmetric under the radial sum);
- the apply is linear over complex coefficients;
- without absorbers, mask or CAP the CN step is unitary: the norm is kept to
  the BiCGSTAB tolerance, in FP64 and in mixed precision;
- with the sinks and the damping mask, every step is a contraction.

The Hermiticity, linearity and unitarity checks also run at l_max 30,
N 32768.
"""
import math
import sys

import numpy as np
import pytest

import gm

pytestmark = pytest.mark.gpu

SHAPE = dict(Nr=16384, L=4000.0, dt=0.25, epsilon=0.0, l_max=20, mu=0.15, M_c=0.01, q=0.5, a=0.0, R=50.0,
             counter=False, absorption=False, save_every=1000, ramp_periods=0.0, epsilon_smear=2.0)
# the second size of the mixed-precision study
LARGE = dict(SHAPE, Nr=32768, L=8000.0, l_max=30)
SHAPES = [pytest.param(SHAPE, id="l20_n16384"), pytest.param(LARGE, id="l30_n32768")]


def announce(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def smooth_state(s, seed):
    """Seeded random modes, smooth in r and zero at both boundary nodes."""
    rng = np.random.default_rng(seed)
    K, N = s.n_modes, s.Nr
    r = s.r
    env = np.exp(-((r - 300.0) / 400.0) ** 2)
    psi = np.zeros((K, N), dtype=np.complex128)
    for j in range(K):
        c = rng.standard_normal(3) + 1j * rng.standard_normal(3)
        psi[j] = env * (c[0] + c[1] * np.cos(r / 37.0) + c[2] * np.sin(r / 53.0))
    psi[:, 0] = psi[:, -1] = 0
    return (psi / math.sqrt(np.vdot(psi, psi).real * s.dr)).ravel()


@pytest.mark.parametrize("shape", SHAPES)
def test_full_size_operator_is_hermitian_and_linear(shape):
    announce(f"full size: Hermiticity and linearity of H, l_max {shape['l_max']}")
    s = gm.CrankNicolsonModes(**shape)
    s.set_flags(evolve_orbit=0)
    x, y = smooth_state(s, 1), smooth_state(s, 2)
    s.set_initial_state_full(x)
    s.initialize()
    s.begin_run(0.8 * shape["L"], 10)
    hx, hy = s.apply_H(x), s.apply_H(y)
    lhs, rhs = np.vdot(x, hy), np.conj(np.vdot(y, hx))
    assert abs(lhs - rhs) <= 1e-12 * np.linalg.norm(x) * np.linalg.norm(hy)
    a, b = 2.0 - 0.5j, -3.0j
    hz = s.apply_H(a * x + b * y)
    assert np.linalg.norm(hz - (a * hx + b * hy)) <= 1e-13 * np.linalg.norm(a * hx + b * hy)
    s.close()


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("precision", [gm.GM_PRECISION_FP64, gm.GM_PRECISION_MIXED])
def test_full_size_steps_are_unitary(precision, shape):
    announce(f"full size: unitarity, precision {precision}, l_max {shape['l_max']}")
    s = gm.CrankNicolsonModes(**shape)
    s.set_flags(evolve_orbit=0, precision=precision)
    s.set_initial_state_full(smooth_state(s, 3))
    s.initialize()
    s.begin_run(0.8 * shape["L"], 10)
    n0 = np.vdot(s.psi, s.psi).real
    for n in range(1, 6):
        s.step(n, diag=False)
        assert s.last_solver_residual <= 1e-9
        assert abs(np.vdot(s.psi, s.psi).real - n0) <= 1e-9 * n0, n
    s.close()


def test_full_size_absorption_only_removes_norm():
    """With the black-hole sinks and the damping mask on, each CN step is a
    contraction: the norm never grows from one step to the next
    (up to the solver tolerance) and falls over the run. The cloud mass the
    step reports is mu dr |psi|^2 of the state it leaves."""
    announce("full size: absorbing steps")
    s = gm.CrankNicolsonModes(**dict(SHAPE, absorption=True))
    s.set_flags(evolve_orbit=0, precision=gm.GM_PRECISION_MIXED)
    s.set_initial_state_full(smooth_state(s, 4))
    s.initialize()
    s.build_damping_mask(0.85, 1.0, 6)
    s.begin_run(3200.0, 10)
    n0 = prev = np.vdot(s.psi, s.psi).real
    for n in range(1, 6):
        d = s.step(n, diag=True)
        assert s.scalars()["sink_active"] == 1
        now = np.vdot(s.psi, s.psi).real
        assert now <= prev * (1 + 1e-9), n
        assert abs(d["cloud_mass"] - SHAPE["mu"] * s.dr * now) <= 1e-12 * d["cloud_mass"], n
        prev = now
    assert prev < n0
    s.close()
