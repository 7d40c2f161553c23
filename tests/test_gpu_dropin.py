"""The reference's GravitationalMolecule driver (src/main.cpp:20-106), restated
statement for statement in tests/cpp/gm_main.cpp over the drop-in headers
(include/gm_config.hpp, include/gm_cnm.hpp, include/gm_initial_data.hpp),
against the oracle driven through the same steps by gm.run.Run (needs an
MI355X).

A config.toml with the startup ramp, absorbers and a damping mask goes
through the MODESBIN generator, then `gm_main config.toml initial_modes.bin`
writes its run folder; the oracle's folder from the same inputs must agree:
r.txt, modes.txt and the copied config.toml byte for byte, every diagnostic
file to 1e-9 of its scale, the snapshots to their printed digits. Then both
folders are cut back to an earlier snapshot (the last snapshot and the
checkpoint removed), their config.toml extended to more steps, and resumed:
`gm_main <folder>` on the GPU, gm.run.Run on the oracle; the resumed folders
agree the same way.
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

import oracle
from gm import initial
from gm import run as gmrun

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
MAIN = os.path.join(HERE, "cpp", "_bin", "gm_main")


def config(num_steps):
    return f"""# drop-in driver fixture
[grid]
N = 257
L = 256.0
[binary]
q = 0.5
R = 20.0
counter = false
[atom]
mu = 0.2
n = 2
l = 1
m = 1
M_c = 0.02
[simulation]
dt = 1.0
num_steps = {num_steps}
save_every = 10
l_max = 4
r_ex = 200.0
ramp_periods = 0.2
epsilon = 0.0
evolve_orbit = false
absorption = true
epsilon_smear = 2.0
[mask]
L_factor = 0.85
W = 1
n = 6
"""


def rows(path):
    with open(path) as f:
        return np.array([[float(x) for x in ln.split()] for ln in f if ln.strip()])


def run_main(cwd, *args):
    if not os.path.exists(MAIN):
        pytest.fail("tests/cpp/_bin/gm_main is not built: run boson-cloud-binary_amd/build.py")
    res = subprocess.run([MAIN, *args], cwd=cwd, capture_output=True, text=True, timeout=240)
    print(res.stdout[-2000:], res.stderr[-2000:], flush=True)
    assert res.returncode == 0, res.stderr
    folder = [ln.split(None, 1)[1] for ln in res.stdout.splitlines() if ln.startswith("folder ")][-1]
    return os.path.join(cwd, folder)


def new_oracle():
    return oracle.Oracle(Nr=257, L=256.0, dt=1.0, epsilon=0.0, l_max=4, mu=0.2, M_c=0.02, q=0.5, a=0.0, R=20.0,
                         counter=False, absorption=True, save_every=10, ramp_periods=0.2, epsilon_smear=2.0)


def oracle_run(cfg_path, modesbin, root, num_steps):
    """main.cpp's sequence on the oracle, with gm.run.Run writing the folder."""
    o = new_oracle()
    o.set_initial_state_full(initial.load_modesbin(modesbin, o.modes))
    o.set_flags(evolve_orbit=0, ramp_secondary_potential=1, steps_per_period=50.0, power_average_orbits=2.0,
                use_block_preconditioner=1)
    o.initialize()
    o.build_damping_mask(0.85, 1.0, 6)
    os.makedirs(root, exist_ok=True)
    return gmrun.Run(o, config_file=cfg_path).run(num_steps, 200.0, root=root)


def compare(fa, fb, exact_tol=1e-9):
    for name in ("r.txt", "modes.txt", "config.toml"):
        with open(os.path.join(fa, name), "rb") as a, open(os.path.join(fb, name), "rb") as b:
            assert a.read() == b.read(), name
    for name in gmrun.DIAGNOSTIC_FILES:
        a, b = rows(os.path.join(fa, name)), rows(os.path.join(fb, name))
        assert a.shape == b.shape, name
        scale = max(np.max(np.abs(b)), 1e-300)
        assert np.max(np.abs(a - b)) <= exact_tol * scale, (name, np.max(np.abs(a - b)) / scale)
    for sub in ("psi", "psi_dt"):
        snaps = sorted(os.listdir(os.path.join(fb, sub)))
        assert snaps and snaps == sorted(os.listdir(os.path.join(fa, sub))), sub
        for snap in snaps:
            a, b = rows(os.path.join(fa, sub, snap)), rows(os.path.join(fb, sub, snap))
            assert np.allclose(a, b, rtol=2e-5, atol=1e-6 * np.max(np.abs(b))), (sub, snap)


def test_reference_driver_runs_and_resumes(tmp_path):
    cfg = tmp_path / "config.toml"
    cfg.write_text(config(30))
    modesbin = str(tmp_path / "initial_modes.bin")
    initial.main(["--config", str(cfg), "--output", modesbin])
    gdir = tmp_path / "gpu"
    gdir.mkdir()
    shutil.copy(cfg, gdir / "config.toml")
    fg = run_main(str(gdir), "config.toml", modesbin)
    fo = oracle_run(str(cfg), modesbin, str(tmp_path / "oracle"), 30)
    compare(fg, fo)
    # cut both folders back to snapshot 10 (snapshot 20 and the checkpoint
    # gone, as after a crash), extend them to 45 steps and resume: the C++
    # driver on the GPU, gm.run.Run on the oracle (main.cpp's resume path:
    # no initial state, flags, initialize, mask, resume)
    for folder in (fg, fo):
        os.remove(os.path.join(folder, "psi", "snapshot_000020.txt"))
        os.remove(os.path.join(folder, "checkpoint.bin"))
        with open(os.path.join(folder, "config.toml"), "w") as f:
            f.write(config(45))
    fr = run_main(str(gdir), os.path.basename(fg))
    assert os.path.samefile(fr, fg)
    o = new_oracle()
    o.set_flags(evolve_orbit=0, ramp_secondary_potential=1, steps_per_period=50.0, power_average_orbits=2.0,
                use_block_preconditioner=1)
    o.initialize()
    o.build_damping_mask(0.85, 1.0, 6)
    assert gmrun.Run(o).resume(fo, 45, 200.0) == 10
    # both resumed from the same six-digit snapshot 10 with the orbit rebuilt
    # from the diagnostic rows
    compare(fr, fo)
