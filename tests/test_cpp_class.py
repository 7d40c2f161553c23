"""The C++17 class of include/gm.hpp, driven by tests/cpp/cnm_driver.cpp.

CPU: the reference's exception types for invalid constructor arguments.
GPU: one step through the C++ class equals the oracle's step (1e-10).
"""
import json
import os
import subprocess

import numpy as np
import pytest

import oracle

HERE = os.path.dirname(os.path.abspath(__file__))
DRIVER = os.path.join(HERE, "cpp", "_bin", "cnm_driver")


def driver():
    if not os.path.exists(DRIVER):
        pytest.fail("tests/cpp/_bin/cnm_driver is not built: run boson-cloud-binary_amd/build.py")
    return DRIVER


def test_constructor_errors_map_to_std_exceptions():
    res = subprocess.run([driver(), "errors"], capture_output=True, text=True, timeout=60)
    assert res.returncode == 0, res.stdout + res.stderr
    assert res.stdout.count("invalid_argument:") == 2


@pytest.mark.gpu
def test_cpp_class_step_matches_oracle(tmp_path):
    args = dict(Nr=129, L=128.0, dt=1.0, epsilon=0.0, l_max=4, mu=0.2, M_c=0.05, q=0.5, a=0.0, R=20.0,
                counter=False, absorption=True, save_every=100, ramp_periods=0.0, epsilon_smear=4.0)
    o = oracle.Oracle(**args)
    K, N = o.n_modes, o.Nr
    rng = np.random.default_rng(7)
    r = o.r
    psi = np.zeros((K, N), dtype=complex)
    for j in range(K):
        psi[j, 1:-1] = r[1:-1] * np.exp(-r[1:-1] / 12.0) * (rng.standard_normal() + 1j * rng.standard_normal())
    psi /= np.sqrt(np.vdot(psi, psi).real * o.dr)
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    psi.astype(np.complex128).tofile(fin)
    print("cpp class step, l_max=4, N=129", flush=True)
    res = subprocess.run([driver(), "step", str(fin), str(fout)], capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stdout + res.stderr
    info = json.loads(res.stdout.strip().splitlines()[-1])
    o.set_flags(evolve_orbit=0)
    o.set_initial_state_full(psi.ravel())
    o.initialize()
    o.begin_run(100.0, 2)
    do = o.step(1)
    g = np.fromfile(fout, dtype=np.complex128)
    assert np.linalg.norm(g - o.psi) / np.linalg.norm(o.psi) < 1e-10
    assert abs(info["iterations"] - o.last_solver_iterations) <= 1
    assert info["energy"] == pytest.approx(do["energy"], rel=1e-9)
