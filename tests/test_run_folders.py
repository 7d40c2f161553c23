"""Run folders and resume (gm.run), after the reference's tests/restart.cpp.

Each case runs on the oracle (CPU) and on the HIP path (marked gpu) with the
reference's fixture: Nr = 65, L = 64, dt = 2, l_max = 2, mu = 0.2, M_c = 0.05,
q = 0.5, R = 20, absorption, save_every = 10, ramp of 0.5 periods, evolving
orbit with 32 steps per period and a one-orbit power average.

- exact resume from checkpoint.bin: every file identical to an uninterrupted run;
- a crash mid-write (partial row, snapshot ahead of the rows): identical too;
- no checkpoint, or an incomplete last snapshot: the orbit is rebuilt from the
  diagnostics and the rows agree to 2e-4 of each file's scale (snapshots keep
  6 significant digits);
- extending a finished run equals the longer run; a finished run is unchanged;
- a folder with another mode list, or without snapshots, is refused.
The text formats themselves are pinned against hand-written lines.
"""
import os
import shutil

import numpy as np
import pytest

from gm import run as gmrun

R_EX = 40.0
FIXTURE = dict(Nr=65, L=64.0, dt=2.0, epsilon=0.0, l_max=2, mu=0.2, M_c=0.05, q=0.5, a=0.0, R=20.0,
               counter=False, absorption=True, save_every=10, ramp_periods=0.5, epsilon_smear=4.0)


def _oracle_solver():
    import oracle
    return oracle.Oracle(**FIXTURE)


def _gpu_solver():
    import gm
    s = gm.CrankNicolsonModes(**FIXTURE)
    s.set_flags(apply_tile=32)  # reproducible reduction order across solvers
    return s


BACKENDS = [pytest.param(_oracle_solver, id="oracle"),
            pytest.param(_gpu_solver, id="gpu", marks=pytest.mark.gpu)]


def make_solver(factory):
    s = factory()
    s.set_flags(evolve_orbit=1, steps_per_period=32.0, power_average_orbits=1.0)
    return s


def initial_state(s):
    """tests/restart.cpp:28-35: Gaussian shells, phase 0.3 j, unit norm."""
    psi = np.zeros((s.n_modes, s.Nr), dtype=np.complex128)
    r = s.r
    for j in range(s.n_modes):
        i = np.arange(1, s.Nr - 1)
        psi[j, i] = np.exp(-((r[i] - 12.0 - 3 * j) / 6.0) ** 2) * np.exp(0.3j * j)
    return (psi / np.sqrt(np.vdot(psi, psi).real * s.dr)).ravel()


def new_run(factory, root, n_steps):
    os.makedirs(root, exist_ok=True)
    cfg = os.path.join(root, "config.toml")
    with open(cfg, "w") as f:
        f.write("# restart fixture\n")
    s = make_solver(factory)
    s.set_initial_state_full(initial_state(s))
    s.initialize()
    folder = gmrun.Run(s, config_file=cfg).run(n_steps, R_EX, root=root)
    s.close()
    return folder


def resume_run(factory, folder, n_steps):
    s = make_solver(factory)
    s.initialize()
    start = gmrun.Run(s).resume(folder, n_steps, R_EX)
    s.close()
    return start


def contents(path):
    with open(path, "rb") as f:
        return f.read()


def require_identical(expected, actual):
    compared = 0
    for base, _, files in os.walk(expected):
        for name in files:
            rel = os.path.relpath(os.path.join(base, name), expected)
            assert os.path.exists(os.path.join(actual, rel)), "missing " + rel
            assert contents(os.path.join(expected, rel)) == contents(os.path.join(actual, rel)), rel + " differs"
            compared += 1
    assert compared > 20


def read_rows(path):
    with open(path) as f:
        return [[float(x) for x in ln.split()] for ln in f.read().split("\n") if ln.strip()]


def max_difference(a, b):
    x, y = read_rows(a), read_rows(b)
    assert len(x) == len(y), os.path.basename(a)
    scale = max(abs(v) for row in x for v in row)
    worst = 0.0
    for rx, ry in zip(x, y):
        assert len(rx) == len(ry)
        worst = max(worst, max(abs(p - q) / scale for p, q in zip(rx, ry)))
    return worst


COMPARED = ("t.txt", "R_orb.txt", "theta_orb.txt", "Omega_orb.txt", "M_1.txt", "q.txt", "energy.txt",
            "p_bin.txt", "norms.txt", "flux.txt", "energy_flux.txt")


@pytest.mark.parametrize("factory", BACKENDS)
def test_exact_resume_from_checkpoint(factory, tmp_path):
    reference = new_run(factory, str(tmp_path / "reference"), 41)
    interrupted = new_run(factory, str(tmp_path / "interrupted"), 25)
    assert len(read_rows(os.path.join(interrupted, "t.txt"))) == 24
    assert resume_run(factory, interrupted, 41) == 20
    require_identical(reference, interrupted)


@pytest.mark.parametrize("factory", BACKENDS)
def test_resume_after_crash_mid_write(factory, tmp_path):
    reference = new_run(factory, str(tmp_path / "reference"), 41)
    crashed = new_run(factory, str(tmp_path / "crashed"), 25)
    with open(os.path.join(crashed, "flux.txt"), "a") as f:
        f.write("0.123 0.4")
    shutil.copyfile(os.path.join(crashed, "psi", "snapshot_000020.txt"),
                    os.path.join(crashed, "psi", "snapshot_000030.txt"))
    assert resume_run(factory, crashed, 41) == 20
    require_identical(reference, crashed)


@pytest.mark.parametrize("factory", BACKENDS)
def test_resume_without_checkpoint(factory, tmp_path):
    reference = new_run(factory, str(tmp_path / "reference"), 41)
    interrupted = new_run(factory, str(tmp_path / "interrupted"), 25)
    os.remove(os.path.join(interrupted, "checkpoint.bin"))
    resume_run(factory, interrupted, 41)
    for name in COMPARED:
        d = max_difference(os.path.join(reference, name), os.path.join(interrupted, name))
        assert d < 2e-4, (name, d)


@pytest.mark.parametrize("factory", BACKENDS)
def test_skips_incomplete_snapshot(factory, tmp_path):
    reference = new_run(factory, str(tmp_path / "reference"), 41)
    interrupted = new_run(factory, str(tmp_path / "interrupted"), 31)
    last = os.path.join(interrupted, "psi", "snapshot_000030.txt")
    text = contents(last)
    with open(last, "wb") as f:
        f.write(text[: len(text) // 2])
    assert resume_run(factory, interrupted, 41) == 20
    for name in COMPARED:
        d = max_difference(os.path.join(reference, name), os.path.join(interrupted, name))
        assert d < 2e-4, (name, d)


@pytest.mark.parametrize("factory", BACKENDS)
def test_extend_finished_run(factory, tmp_path):
    longer = new_run(factory, str(tmp_path / "long"), 61)
    shorter = new_run(factory, str(tmp_path / "short"), 41)
    resume_run(factory, shorter, 41)
    assert len(read_rows(os.path.join(shorter, "t.txt"))) == 40
    resume_run(factory, shorter, 61)
    require_identical(longer, shorter)


@pytest.mark.parametrize("factory", BACKENDS)
def test_rejects_mismatched_folders(factory, tmp_path):
    folder = new_run(factory, str(tmp_path / "run"), 12)
    with open(os.path.join(folder, "modes.txt"), "w") as f:
        f.write("0,0\n1,1\n")
    with pytest.raises(RuntimeError, match="modes.txt"):
        resume_run(factory, folder, 20)
    empty = new_run(factory, str(tmp_path / "empty"), 5)
    for name in os.listdir(os.path.join(empty, "psi")):
        os.remove(os.path.join(empty, "psi", name))
    with pytest.raises(RuntimeError, match="No usable snapshot"):
        resume_run(factory, empty, 20)


def test_run_folder_layout(tmp_path):
    """Files of a run, and the formats: r.txt and snapshots as the C++ stream
    default ('%g'), diagnostics as std::scientific with 12 digits."""
    folder = new_run(_oracle_solver, str(tmp_path / "layout"), 12)
    names = set(os.listdir(folder))
    assert set(gmrun.DIAGNOSTIC_FILES) <= names
    assert {"config.toml", "r.txt", "modes.txt", "checkpoint.bin", "psi", "psi_dt"} <= names
    assert sorted(os.listdir(os.path.join(folder, "psi"))) == ["snapshot_000000.txt", "snapshot_000010.txt"]
    assert os.listdir(os.path.join(folder, "psi_dt")) == ["snapshot_000000.txt"]
    with open(os.path.join(folder, "modes.txt")) as f:
        assert f.read() == "0,0\n1,-1\n1,1\n2,-2\n2,0\n2,2\n"
    with open(os.path.join(folder, "r.txt")) as f:
        r = f.read().split("\n")
    assert r[:3] == ["0", "1", "2"] and len(r) == 66 and r[-1] == ""
    with open(os.path.join(folder, "energy_modes.txt")) as f:
        rows = f.read().split("\n")
    assert len(rows) == 12 and rows[-1] == "" and len(rows[0].split(" ")) == 6
    for tok in rows[0].split(" "):
        mant, exp = tok.split("e")
        assert len(mant.lstrip("-")) == 14 and exp[0] in "+-" and len(exp) == 3
    assert gmrun.quantity_line(1.5) == "1.500000000000e+00\n"
    assert gmrun.quantity_line(-2.5e-300) == "-2.500000000000e-300\n"
    assert gmrun.quantity_line_modes([1.0, -0.0]) == "1.000000000000e+00 -0.000000000000e+00\n"
    assert gmrun.snapshot_text([1 + 2j, 1234567.0 - 1e-5j, 0.0001 + 0j]) == "1 2\n1.23457e+06 -1e-05\n0.0001 0\n"


def test_snapshot_loader_rejects_damaged_files(tmp_path):
    p = str(tmp_path / "s.txt")
    with open(p, "w") as f:
        f.write("1 2\n\n3 4\n")
    np.testing.assert_array_equal(gmrun.load_snapshot(p, 2), [1 + 2j, 3 + 4j])
    assert gmrun.load_snapshot(p, 3) is None
    with open(p, "w") as f:
        f.write("1 2\n3\n")
    assert gmrun.load_snapshot(p, 2) is None
    assert gmrun.load_snapshot(str(tmp_path / "missing.txt"), 2) is None
    with open(p, "w") as f:
        f.write("a\nb\nc")
    assert list(gmrun.line_ends(p)) == [2, 4]
