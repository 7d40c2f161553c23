"""Run folders: the text output, checkpoints and resume of a simulation, over
any solver handle with this package's `Solver` surface.

This is the host side of CrankNicolsonModes::run, ::evolve, ::step's output
and ::resume (src/CrankNicolsonModes.cpp:197-289,
src/CrankNicolsonModes_utils.cpp:10-100, src/CrankNicolsonModes_restart.cpp:22-399,
src/Utils.cpp:11-27). The solver computes; this module writes the same files,
byte for byte, and reads them back to resume.

Folder layout:
    Simulation_%Y_%m_%d_%H_%M_%S/     (local time, created in `root`)
        config.toml                   copy of the configuration file, when given
        r.txt                         one radius per line, C++ stream default ('%g')
        modes.txt                     'l,m' per line
        psi/snapshot_%06d.txt         're im' per line ('%g'), step 0 and every save_every
        psi_dt/snapshot_000000.txt    the Schrodinger right-hand side at t = 0
        checkpoint.bin                exact state at the last snapshot (written by the solver)
        <diagnostics>                 one line per step, std::scientific with 12 digits
                                      ('%.12e'), per-mode files space separated

Snapshots keep 6 significant digits, so a resume without a matching
checkpoint is accurate, not exact. With a checkpoint, a resumed run writes the
same bytes as an uninterrupted one, provided the solver's arithmetic is
reproducible: on the GPU pin `apply_tile` (the autotuner's choice depends on
timing, and the tile sets the order of the fused reductions).
"""
import os
import re
import shutil
import time

import numpy as np

# diagnostic_files() (src/CrankNicolsonModes_restart.cpp:109-127), in its order
DIAGNOSTIC_FILES = ("flux.txt", "energy_flux.txt", "energy.txt", "energy_modes.txt", "norms.txt", "t.txt",
                    "p_bin.txt", "p_bin_inst.txt", "dV_bin_dt.txt", "R_orb.txt", "Omega_orb.txt", "M_1.txt",
                    "q.txt", "theta_orb.txt")

# Columns of the diagnostic rows that rebuild the orbit without a checkpoint
# (reconstruct_state_from_diagnostics, src/CrankNicolsonModes_restart.cpp:238-250).
RESTORE_COLUMNS = ("t.txt", "R_orb.txt", "theta_orb.txt", "Omega_orb.txt", "M_1.txt", "q.txt", "p_bin.txt",
                   "p_bin_inst.txt")

_SNAPSHOT = re.compile(r"snapshot_(\d+)\.txt")


def quantity_line(value):
    """write_quantity_line (utils.cpp:85-88): std::scientific, precision 12."""
    return "%.12e\n" % value


def quantity_line_modes(values):
    """write_quantity_line_modes (utils.cpp:73-83): space separated."""
    return " ".join("%.12e" % v for v in values) + "\n"


def snapshot_text(psi):
    """save_snapshot (utils.cpp:52-71): 're im' per element, stream default
    format (printf '%g': 6 significant digits)."""
    a = np.asarray(psi, dtype=np.complex128).ravel()
    return "".join("%g %g\n" % v for v in zip(a.real.tolist(), a.imag.tolist()))


def snapshot_path(folder, name, step_id):
    return os.path.join(folder, name, "snapshot_%06d.txt" % step_id)


def load_snapshot(path, size):
    """load_snapshot (src/CrankNicolsonModes_restart.cpp:78-105): the state, or
    None when the file is missing, has a line that does not hold two numbers,
    or does not hold exactly `size` non-empty lines."""
    try:
        with open(path) as f:
            lines = [ln for ln in f.read().split("\n") if ln]
    except OSError:
        return None
    if len(lines) != size:
        return None
    try:
        v = np.array([ln.split(None, 2)[:2] for ln in lines], dtype=np.float64)
    except ValueError:
        return None
    if v.shape != (size, 2):
        return None
    return v[:, 0] + 1j * v[:, 1]


def line_ends(path):
    """line_ends (src/CrankNicolsonModes_restart.cpp:38-57): byte offsets just
    past each newline; a trailing partial line is not counted."""
    try:
        b = np.fromfile(path, dtype=np.uint8)
    except OSError as e:
        raise RuntimeError("Cannot open " + path) from e
    return np.flatnonzero(b == 10) + 1


def read_column(path, n_rows):
    """read_column (src/CrankNicolsonModes_restart.cpp:60-75): the leading
    number of each of the first n_rows lines."""
    try:
        with open(path) as f:
            lines = f.read().split("\n")
    except OSError as e:
        raise RuntimeError("Cannot open " + path) from e
    if len(lines) and lines[-1] == "":
        lines.pop()
    if len(lines) < n_rows:
        raise RuntimeError("Too few rows in " + path)
    out = np.empty(n_rows)
    for k in range(n_rows):
        tok = lines[k].split()
        try:
            out[k] = float(tok[0]) if tok else 0.0
        except ValueError:
            out[k] = 0.0
    return out


def create_output_folder(root="."):
    """create_output_folder (src/Utils.cpp:11-27)."""
    folder = os.path.join(root, time.strftime("Simulation_%Y_%m_%d_%H_%M_%S", time.localtime()))
    for d in (folder, os.path.join(folder, "psi"), os.path.join(folder, "psi_dt")):
        os.makedirs(d, exist_ok=True)
    return folder


class Run:
    """Output and resume of one solver (gm.CrankNicolsonModes or any object
    with the same methods). `save_every` defaults to the constructor's;
    `config_file` is copied to config.toml by run(); `log` receives the
    progress messages the reference prints."""

    def __init__(self, solver, save_every=None, config_file=None, log=None):
        self.solver = solver
        self.save_every = int(save_every if save_every is not None else solver.params.save_every)
        if self.save_every < 1:
            raise ValueError("save_every must be positive")
        self.config_file = config_file
        self.log = log or (lambda msg: None)
        self.folder = None
        self._files = {}

    # -- files --------------------------------------------------------------
    def _open_diagnostics(self, mode):
        self._close_diagnostics()
        for name in DIAGNOSTIC_FILES:
            try:
                self._files[name] = open(os.path.join(self.folder, name), mode)
            except OSError as e:
                raise RuntimeError("Cannot open " + os.path.join(self.folder, name)) from e

    def _close_diagnostics(self):
        for f in self._files.values():
            f.close()
        self._files = {}

    def flush_diagnostics(self):
        for f in self._files.values():
            f.flush()

    def close(self):
        self._close_diagnostics()

    def save_run_info(self):
        """save_run_info (utils.cpp:10-45)."""
        if self.config_file is not None:
            shutil.copyfile(self.config_file, os.path.join(self.folder, "config.toml"))
        with open(os.path.join(self.folder, "r.txt"), "w") as f:
            f.write("".join("%g\n" % x for x in np.asarray(self.solver.r).tolist()))
        with open(os.path.join(self.folder, "modes.txt"), "w") as f:
            f.write("".join("%d,%d\n" % lm for lm in self.solver.modes))

    def save_snapshot(self, step_id, psi, name):
        with open(snapshot_path(self.folder, name, step_id), "w") as f:
            f.write(snapshot_text(psi))

    def save_checkpoint(self, step_id):
        self.solver.save_checkpoint(os.path.join(self.folder, "checkpoint.bin"), step_id)

    # -- run ----------------------------------------------------------------
    def step(self, step_id):
        """One step and its output (src/CrankNicolsonModes.cpp:145-234)."""
        s = self.solver
        d = s.step(step_id, diag=True)
        sc = s.scalars()
        w = {name: f.write for name, f in self._files.items()}
        w["R_orb.txt"](quantity_line(sc["R"]))
        w["theta_orb.txt"](quantity_line(sc["theta"]))
        w["Omega_orb.txt"](quantity_line(sc["Omega"]))
        w["M_1.txt"](quantity_line(sc["M1"]))
        w["q.txt"](quantity_line(sc["q"]))
        w["t.txt"](quantity_line(sc["t"]))
        w["energy_flux.txt"](quantity_line_modes(d["energy_flux"]))
        w["flux.txt"](quantity_line_modes(d["flux"]))
        w["energy.txt"](quantity_line(d["energy"]))
        w["p_bin.txt"](quantity_line(d["p_bin"]))
        w["p_bin_inst.txt"](quantity_line(d["p_bin_inst"]))
        w["dV_bin_dt.txt"](quantity_line(d["dV_bin_dt"]))
        w["norms.txt"](quantity_line_modes(d["norms"]))
        w["energy_modes.txt"](quantity_line_modes(d["energies"]))
        if step_id % self.save_every == 0:
            self.save_snapshot(step_id, s.psi, "psi")
            self.flush_diagnostics()
            self.save_checkpoint(step_id)
        return d

    def _stopped(self):
        s = self.solver
        return (not getattr(s, "isolated", False)) and bool(s.get_flags().evolve_orbit) and s.scalars()["R"] <= 6.0

    def evolve(self, first_step, n_steps):
        """evolve (src/CrankNicolsonModes.cpp:272-289): steps first_step ..
        n_steps-1, stopping early once an evolving orbit reaches R <= 6."""
        n = first_step
        for n in range(first_step, n_steps):
            if self._stopped():
                break
            self.step(n)
        self.flush_diagnostics()
        self.log("Simulation completed!")

    def run(self, n_steps, r_ex, root="."):
        """run (src/CrankNicolsonModes.cpp:236-259): a new folder under root,
        the run information, snapshot and checkpoint 0, then steps 1 ..
        n_steps-1. Returns the folder."""
        s = self.solver
        s.begin_run(r_ex, n_steps)
        self.folder = create_output_folder(root)
        self.log("Output folder: " + self.folder)
        self._open_diagnostics("w")
        self.save_run_info()
        self.save_snapshot(0, s.psi, "psi")
        self.save_snapshot(0, s.schrodinger_rhs(s.psi), "psi_dt")
        self.save_checkpoint(0)
        self.evolve(1, n_steps)
        return self.folder

    # -- resume -------------------------------------------------------------
    def _check_modes(self, folder):
        path = os.path.join(folder, "modes.txt")
        if not os.path.exists(path):
            return
        modes = self.solver.modes
        count = 0
        with open(path) as f:
            for line in f.read().split("\n"):
                if not line:
                    continue
                try:
                    l, m = (int(x) for x in line.split(",")[:2])
                except ValueError:
                    l, m = None, None
                if count >= len(modes) or tuple(modes[count]) != (l, m):
                    raise RuntimeError("modes.txt does not match the modes built from config.toml")
                count += 1
        if count != len(modes):
            raise RuntimeError("modes.txt does not match the modes built from config.toml")

    def resume(self, folder, n_steps, r_ex):
        """resume (src/CrankNicolsonModes_restart.cpp:287-399): restart from the
        latest complete snapshot that has all its diagnostic rows, exactly
        from checkpoint.bin when it belongs to that snapshot, otherwise from
        the snapshot and the orbit rebuilt from the diagnostics; drop what was
        written after it and step on to n_steps - 1."""
        s = self.solver
        self.log("Resuming simulation in " + folder)
        s.begin_run(r_ex, n_steps)
        self.folder = folder
        self._check_modes(folder)
        ends = {name: line_ends(os.path.join(folder, name)) for name in DIAGNOSTIC_FILES}
        complete_rows = min(len(e) for e in ends.values())
        psi_dir = os.path.join(folder, "psi")
        try:
            names = os.listdir(psi_dir)
        except OSError as e:
            raise RuntimeError("Cannot read " + psi_dir) from e
        steps = sorted((int(m.group(1)) for m in map(_SNAPSHOT.fullmatch, names) if m), reverse=True)
        start, psi = -1, None
        for cand in steps:
            if cand > complete_rows:
                continue
            path = snapshot_path(folder, "psi", cand)
            psi = load_snapshot(path, s.Nr * s.n_modes)
            if psi is not None:
                start = cand
                break
            self.log("Skipping incomplete snapshot " + path)
        if start < 0:
            raise RuntimeError("No usable snapshot found in " + psi_dir)
        self.log("Restarting from snapshot %d" % start)
        s.psi = psi
        if s.load_checkpoint(os.path.join(folder, "checkpoint.bin"), start):
            self.log("Exact state restored from checkpoint.bin.")
        else:
            self.log("No checkpoint for this snapshot: rebuilding the orbit state from the diagnostics.")
            cols = [read_column(os.path.join(folder, name), start) for name in RESTORE_COLUMNS]
            s.restore_from_diagnostics(np.stack(cols, axis=1) if start else np.zeros((0, 8)))
        for name in DIAGNOSTIC_FILES:
            keep = 0 if start == 0 else int(ends[name][start - 1])
            os.truncate(os.path.join(folder, name), keep)
        self._open_diagnostics("a")
        if start + 1 >= n_steps:
            self.log("The run had already reached its last step (%d)." % (n_steps - 1))
            self._close_diagnostics()
            return start
        self.evolve(start + 1, n_steps)
        return start
