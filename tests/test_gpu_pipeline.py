"""The reference's main path end to end (src/main.cpp:55-106 with
initial_states/get_initial_state_COM_fast.py in front): config.toml, the
MODESBIN generator's command line, the MODESBIN reader for the solver's
modes, then a run folder written by Run.run, on the HIP path and on the
oracle from the same file. Every diagnostic file agrees to 1e-9 of its
scale, and the snapshots agree to their printed digits."""
import os

import numpy as np
import pytest

import gm
import oracle
from gm import initial
from gm import run as gmrun

pytestmark = pytest.mark.gpu

PARAMS = dict(Nr=257, L=256.0, dt=1.0, epsilon=0.0, l_max=4, mu=0.2, M_c=0.02, q=0.5, a=0.0, R=20.0,
              counter=False, absorption=True, save_every=10, ramp_periods=0.2, epsilon_smear=2.0)
STEPS, R_EX = 30, 200.0

CONFIG = f"""[grid]
N = {PARAMS['Nr']}
L = {PARAMS['L']}
[binary]
q = {PARAMS['q']}
R = {PARAMS['R']}
[atom]
mu = {PARAMS['mu']}
n = 2
l = 1
m = 1
[simulation]
epsilon = {PARAMS['epsilon']}
l_max = {PARAMS['l_max']}
"""


def rows(path):
    with open(path) as f:
        return np.array([[float(x) for x in ln.split()] for ln in f if ln.strip()])


def one_run(solver, modesbin, root):
    solver.set_flags(evolve_orbit=0)
    solver.set_initial_state_full(initial.load_modesbin(modesbin, initial.modes_for(PARAMS["l_max"])))
    solver.initialize()
    solver.build_damping_mask(0.85, 1.0, 6)
    return gmrun.Run(solver).run(STEPS, R_EX, root=root)


def test_config_to_run_folder(tmp_path):
    cfg = tmp_path / "config.toml"
    cfg.write_text(CONFIG)
    modesbin = str(tmp_path / "initial_modes.bin")
    initial.main(["--config", str(cfg), "--output", modesbin])
    fg = one_run(gm.CrankNicolsonModes(**PARAMS), modesbin, str(tmp_path / "gpu"))
    fo = one_run(oracle.Oracle(**PARAMS), modesbin, str(tmp_path / "oracle"))
    names = sorted(n for n in os.listdir(fo) if n.endswith(".txt") and n not in ("r.txt", "modes.txt"))
    assert len(names) >= 14
    for name in names:
        a, b = rows(os.path.join(fg, name)), rows(os.path.join(fo, name))
        assert a.shape == b.shape, name
        scale = max(np.max(np.abs(b)), 1e-300)
        assert np.max(np.abs(a - b)) <= 1e-9 * scale, name
    for sub in ("psi", "psi_dt"):
        snaps = sorted(os.listdir(os.path.join(fo, sub)))
        assert snaps and snaps == sorted(os.listdir(os.path.join(fg, sub)))
        for snap in snaps:
            a, b = rows(os.path.join(fg, sub, snap)), rows(os.path.join(fo, sub, snap))
            # six significant digits ("%g"): compare at the printed precision
            assert np.allclose(a, b, rtol=2e-5, atol=1e-6 * np.max(np.abs(b))), (sub, snap)
