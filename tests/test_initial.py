"""MODESBIN reader/writer and the translated-atom generator (CPU).

The reader cases follow tests/config_and_io.cpp:130-156 of the reference. The
generator is pinned by mathematics, not by a second copy of its own formulas:
Y_lm against scipy's sph_harm_y, the untranslated atom against the textbook
closed forms of R_nl, the translated field against the analytic orbital
(field reconstruction from the modes), norm conservation under translation,
and the rotation that carries the primary onto the secondary.
"""
import math
import os
import struct

import numpy as np
import pytest
from scipy import special

from gm import initial


def profile(nr, scale):
    i = np.arange(nr)
    return scale * i - 0.5j * scale * i * i


def test_modesbin_reader(tmp_path):
    nr = 6
    path = str(tmp_path / "state.bin")
    initial.write_modesbin(path, nr, [((1, 1), profile(nr, 2)), ((0, 0), profile(nr, 1)), ((7, 7), profile(nr, 9))])
    psi = initial.load_modesbin(path, [(0, 0), (1, -1), (1, 1)])
    assert psi.size == 3 * nr
    assert np.array_equal(psi[:nr], profile(nr, 1))
    assert not psi[nr:2 * nr].any()
    assert np.array_equal(psi[2 * nr:], profile(nr, 2))

    data = open(path, "rb").read()
    bad_magic = str(tmp_path / "bad_magic.bin")
    open(bad_magic, "wb").write(b"NOTMODES" + data[8:])
    bad_version = str(tmp_path / "bad_version.bin")
    open(bad_version, "wb").write(data[:8] + struct.pack("<i", 2) + data[12:])
    with pytest.raises(RuntimeError, match="is not a MODESBIN file"):
        initial.load_modesbin(bad_magic, [(0, 0)])
    with pytest.raises(RuntimeError, match="Unsupported MODESBIN version"):
        initial.load_modesbin(bad_version, [(0, 0)])
    with pytest.raises(RuntimeError, match="Cannot open initial-state file"):
        initial.load_modesbin(str(tmp_path / "missing.bin"), [(0, 0)])
    truncated = str(tmp_path / "truncated.bin")
    open(truncated, "wb").write(data[:-8])
    with pytest.raises(RuntimeError, match="truncated"):
        initial.load_modesbin(truncated, [(0, 0)])


def test_modesbin_layout_bytes(tmp_path):
    """Header and record layout, byte by byte (include/InitialData.hpp:18-34)."""
    path = str(tmp_path / "one.bin")
    initial.write_modesbin(path, 2, [((3, -1), np.array([1.5 - 2j, -0.25 + 4j]))])
    data = open(path, "rb").read()
    expect = (b"MODESBIN" + struct.pack("<i", 1) + struct.pack("<q", 2) + struct.pack("<q", 1) +
              struct.pack("<ii", 3, -1) + struct.pack("<dd", 1.5, -0.25) + struct.pack("<dd", -2.0, 4.0))
    assert data == expect


def test_solid_ylm_matches_scipy():
    rng = np.random.default_rng(3)
    x, y, z = rng.standard_normal((3, 400))
    # points on the z axis, both poles
    x[:4] = y[:4] = 0.0
    z[:4] = [1.0, -2.0, 0.5, -0.1]
    rho = np.sqrt(x * x + y * y + z * z)
    theta = np.arccos(z / rho)
    phi = np.arctan2(y, x)
    for l in range(13):
        for m in range(-l, l + 1):
            got = initial.solid_ylm(l, m, x, y, z, rho)
            want = special.sph_harm_y(l, m, theta, phi)
            assert np.max(np.abs(got - want)) < 1e-12 * max(1.0, np.max(np.abs(want))), (l, m)


def radial_closed_form(n, l, r, a):
    """Textbook hydrogen radial functions, written out (not from Laguerre)."""
    s = r / a
    pref = a ** -1.5
    table = {
        (1, 0): 2.0 * np.exp(-s),
        (2, 1): s * np.exp(-s / 2) / (2.0 * math.sqrt(6.0)),
        (3, 2): 4.0 * s * s * np.exp(-s / 3) / (81.0 * math.sqrt(30.0)),
        (3, 1): 8.0 * s * (1 - s / 6) * np.exp(-s / 3) / (27.0 * math.sqrt(6.0)),
        (4, 3): s ** 3 * np.exp(-s / 4) / (768.0 * math.sqrt(35.0)),
    }
    return pref * table[(n, l)]


@pytest.mark.parametrize("nlm", [(1, 0, 0), (2, 1, 1), (3, 2, 0), (3, 2, -2), (3, 1, -1), (4, 3, 1)])
def test_untranslated_atom_is_one_mode(nlm):
    n, l, m = nlm
    mu, M = 0.3, 1.7
    a = 1.0 / (mu * mu * M)
    r = np.linspace(0.0, 40 * n * n * a / 10, 300)
    modes = initial.all_modes(6)
    U = initial.translated_atom(r, 6, mu, 0.0, n=n, l=l, m=m, M=M, modes=modes).reshape(len(modes), -1)
    want = r * radial_closed_form(n, l, r, a)
    j = modes.index((l, m))
    assert np.max(np.abs(U[j] - want)) < 1e-12 * np.max(np.abs(want))
    others = np.delete(U, j, axis=0)
    assert np.max(np.abs(others)) < 1e-12 * np.max(np.abs(want))


def analytic_211(x, y, z, a):
    rho = np.sqrt(x * x + y * y + z * z)
    return radial_closed_form(2, 1, rho, a) * (-math.sqrt(3.0 / (8.0 * math.pi))) * np.where(rho > 0, (x + 1j * y) / np.where(rho > 0, rho, 1), 0)


def test_translated_field_reconstruction():
    """Sum_lm U_lm(r)/r Y_lm(theta, phi) rebuilds psi_211(x - d e_x) at points
    off the quadrature grid. Away from the atom's centre the l_max 24 series is
    converged to ~1e-11 of the orbital's peak; on the sphere r = d it passes
    through the cusp of exp(-rho/2a), where the series converges only
    algebraically in l (9.5e-6, 7.7e-7, 2.7e-7 at l_max 12, 24, 36)."""
    mu, d = 0.15, 50.0 / 3.0
    a = 1.0 / (mu * mu)
    l_max = 24
    r = np.array([2.0, 9.0, d, 25.0, 60.0, 150.0, 400.0])
    modes = initial.all_modes(l_max)
    U = initial.translated_atom(r, l_max, mu, d, modes=modes).reshape(len(modes), -1)
    rng = np.random.default_rng(7)
    theta = np.arccos(rng.uniform(-1, 1, 25))
    phi = rng.uniform(0, 2 * np.pi, 25)
    Y = np.array([special.sph_harm_y(l, m, theta, phi) for (l, m) in modes])
    for i, ri in enumerate(r):
        field = (U[:, i] / ri) @ Y
        x = ri * np.sin(theta) * np.cos(phi)
        y = ri * np.sin(theta) * np.sin(phi)
        z = ri * np.cos(theta)
        want = analytic_211(x - d, y, z, a)
        peak = np.max(np.abs(radial_closed_form(2, 1, np.array([2 * a]), a)))
        tol = 2e-6 if ri == d else 1e-8
        assert np.max(np.abs(field - want)) < tol * peak, ri


def test_translation_conserves_norm_and_parity():
    """The translated atom keeps unit norm; with the offset on x (y -> -y maps
    the atom onto its conjugate) every U_lm is real, and the z -> -z parity of
    psi_211 kills every l + m odd mode."""
    mu, d = 0.15, 50.0 / 3.0
    a = 1.0 / (mu * mu)
    r = np.linspace(0.0, 60 * a, 6001)
    modes = initial.all_modes(16)
    U = initial.translated_atom(r, 16, mu, d, modes=modes).reshape(len(modes), -1)
    assert abs(initial.modes_norm(U, r) - 1.0) < 1e-6
    assert np.max(np.abs(U.imag)) < 1e-13 * np.max(np.abs(U))
    odd = [j for j, (l, m) in enumerate(modes) if (l + m) % 2]
    assert np.max(np.abs(U[odd])) < 1e-13 * np.max(np.abs(U))


def write_config(path, **over):
    cfg = {"N": 257, "L": 400.0, "q": 1.0, "R": 30.0, "mu": 0.2, "epsilon": 0.0, "l_max": 8, "n": 2, "l": 1, "m": 1}
    cfg.update(over)
    text = (f"[grid]\nN = {cfg['N']}\nL = {cfg['L']}\n"
            f"[binary]\nq = {cfg['q']}\nR = {cfg['R']}\n"
            f"[atom]\nmu = {cfg['mu']}\nn = {cfg['n']}\nl = {cfg['l']}\nm = {cfg['m']}\n"
            f"[simulation]\nepsilon = {cfg['epsilon']}\nl_max = {cfg['l_max']}\n")
    with open(path, "w") as f:
        f.write(text)
    return cfg


def test_generator_command_line(tmp_path):
    cfg_path = str(tmp_path / "config.toml")
    cfg = write_config(cfg_path, epsilon=0.01)
    out = str(tmp_path / "initial_modes.bin")
    initial.main(["--config", cfg_path, "--output", out])
    data = open(out, "rb").read()
    Nr, count = struct.unpack_from("<qq", data, 12)
    assert (Nr, count) == (cfg["N"], (cfg["l_max"] + 1) ** 2)
    off, order = 28, []
    for _ in range(count):
        order.append(struct.unpack_from("<ii", data, off))
        off += 8 + 16 * Nr
    assert order == initial.all_modes(cfg["l_max"])
    modes = initial.all_modes(cfg["l_max"])
    U = initial.load_modesbin(out, modes).reshape(len(modes), Nr)
    r = np.linspace(0.0, cfg["L"], Nr)
    r[0] = cfg["epsilon"]
    assert abs(initial.modes_norm(U, r) - 1.0) < 1e-14
    # the solver's mode subset reads from the same file
    sub = initial.load_modesbin(out, initial.modes_for(cfg["l_max"]))
    assert sub.size == Nr * len(initial.modes_for(cfg["l_max"]))


def test_secondary_is_the_primary_rotated_by_pi():
    """With equal masses the secondary at -R/2 is the primary at +R/2 turned by
    pi about z: U_lm(secondary) = (-1)^(m + m_atom) U_lm(primary)."""
    cfg = {"grid": {"N": 129, "L": 300.0}, "binary": {"q": 1.0, "R": 24.0},
           "atom": {"mu": 0.25, "n": 3, "l": 2, "m": -1}, "simulation": {"epsilon": 0.0, "l_max": 10}}
    r, modes, Up = initial.generate(cfg, "primary")
    _, _, Us = initial.generate(cfg, "secondary")
    sign = np.array([(-1.0) ** (m - 1) for (_, m) in modes])[:, None]
    assert np.max(np.abs(Us - sign * Up)) < 1e-12 * np.max(np.abs(Up))


def test_secondary_uses_its_own_mass():
    """R = 0 leaves both atoms at the origin; the secondary's Bohr radius is
    1/(mu^2 q M1) and its amplitude key scales it before normalisation."""
    cfg = {"grid": {"N": 2001, "L": 4000.0}, "binary": {"q": 0.5, "R": 0.0, "M1": 1.0},
           "atom": {"mu": 0.2, "n": 2, "l": 1, "m": 1},
           "atom_secondary": {"n": 3, "l": 2, "m": 2, "amplitude": 3.0},
           "simulation": {"epsilon": 0.0, "l_max": 4, "normalize_initial_state": False}}
    r, modes, U = initial.generate(cfg, "secondary")
    a2 = 1.0 / (0.2 * 0.2 * 0.5)
    want = 3.0 * r * radial_closed_form(3, 2, r, a2)
    j = modes.index((2, 2))
    assert np.max(np.abs(U[j] - want)) < 1e-12 * np.max(np.abs(want))


def test_generator_rejects_a_mode_above_l_max():
    cfg = {"grid": {"N": 9, "L": 10.0}, "binary": {"q": 1.0, "R": 2.0},
           "atom": {"mu": 0.2, "n": 4, "l": 3, "m": 1}, "simulation": {"epsilon": 0.0, "l_max": 2}}
    with pytest.raises(ValueError, match="not included in l_max"):
        initial.generate(cfg)
