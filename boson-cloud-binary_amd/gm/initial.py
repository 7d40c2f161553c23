"""MODESBIN initial states.

The format is the reference's (include/InitialData.hpp:18-34,
src/InitialData.cpp:7-79): the 8 bytes "MODESBIN", int32 version 1, int64 Nr,
int64 number of stored modes, then per mode int32 l, int32 m, Nr doubles of
Re U_lm and Nr doubles of Im U_lm (little-endian). Stored modes that the
solver does not list are ignored; listed modes that are not stored are zero.

`translated_atom` builds the state the reference's initial-state generator
(initial_states/get_initial_state_COM_fast.py) writes for a hydrogenic
(n, l, m) atom translated off the centre of mass along x:
U_lm(r) = r * integral Y_lm^* psi_nlm(x - offset e_x) dOmega. The reference
rotates the atom so that the offset lies on z, translates along z with a 1-D
quadrature and rotates back with Wigner-D matrices; here the same integral is
taken directly, with a Gauss-Legendre (cos theta) x FFT (phi) product
quadrature of the translated field, evaluated without angles through the
Cartesian form of Y_lm.

`generate` and the command line (`python -m gm.initial --config config.toml
--output initial_modes.bin [--atom secondary] [--n-quad Q]`) mirror that
script: same configuration keys, every (l, m) with l <= l_max written in
(l, m) order, the grid linspace(0, L, N) with r[0] = epsilon, the amplitude
factor and the trapezoid-rule normalisation.
"""
import math
import os
import struct
from concurrent.futures import ThreadPoolExecutor

import numpy as np
from scipy import special

MAGIC = b"MODESBIN"


def write_modesbin(path, Nr, profiles):
    """profiles: iterable of ((l, m), complex array of length Nr)."""
    profiles = list(profiles)
    with open(path, "wb") as f:
        f.write(MAGIC)
        f.write(struct.pack("<i", 1))
        f.write(struct.pack("<qq", int(Nr), len(profiles)))
        for (l, m), u in profiles:
            u = np.asarray(u, dtype=np.complex128)
            assert u.size == Nr
            f.write(struct.pack("<ii", int(l), int(m)))
            f.write(np.ascontiguousarray(u.real, dtype="<f8").tobytes())
            f.write(np.ascontiguousarray(u.imag, dtype="<f8").tobytes())


def load_modesbin(path, modes):
    """load_initial_modes_binary (src/InitialData.cpp:7-79): the stacked state
    (mode-major) for the solver's mode list. A later duplicate of a stored
    (l, m) replaces the earlier one, as in the reference's map."""
    try:
        with open(path, "rb") as f:
            data = f.read()
    except OSError:
        raise RuntimeError(f"Cannot open initial-state file {path}") from None
    if len(data) < 28:
        raise RuntimeError(f"{path} is not a MODESBIN file")
    if data[:8] != MAGIC:
        raise RuntimeError(f"{path} is not a MODESBIN file")
    (version,) = struct.unpack_from("<i", data, 8)
    if version != 1:
        raise RuntimeError(f"Unsupported MODESBIN version in {path}")
    Nr, count = struct.unpack_from("<qq", data, 12)
    off = 28
    stored = {}
    if Nr < 0 or count < 0 or len(data) < off + count * (8 + 16 * Nr):
        raise RuntimeError(f"{path} is truncated")
    for _ in range(count):
        l, m = struct.unpack_from("<ii", data, off)
        off += 8
        re = np.frombuffer(data, dtype="<f8", count=Nr, offset=off)
        off += 8 * Nr
        im = np.frombuffer(data, dtype="<f8", count=Nr, offset=off)
        off += 8 * Nr
        stored[(l, m)] = re + 1j * im
    psi = np.zeros(Nr * len(modes), dtype=np.complex128)
    for j, lm in enumerate(modes):
        if tuple(lm) in stored:
            psi[j * Nr:(j + 1) * Nr] = stored[tuple(lm)]
    return psi


def modes_for(l_max):
    """build_modes (src/CrankNicolsonModes.cpp:107-128): l - m even."""
    return [(l, m) for l in range(l_max + 1) for m in range(-l, l + 1) if (l - m) % 2 == 0]


def hydrogen_radial(n, l, rho, a0):
    """Normalised hydrogenic radial function R_nl(rho) for Bohr radius a0."""
    x = 2.0 * rho / (n * a0)
    norm = math.sqrt((2.0 / (n * a0)) ** 3 * math.factorial(n - l - 1) / (2 * n * math.factorial(n + l)))
    return norm * np.exp(-x / 2) * x ** l * special.eval_genlaguerre(n - l - 1, 2 * l + 1, x)


def solid_ylm(l, m, x, y, z, rho):
    """Y_lm (Condon-Shortley phase) at the points (x, y, z) with rho = |(x, y, z)|,
    without angles: Y_lm = N_lm (-1)^m (2m-1)!! Q_l^m(z/rho) ((x + iy)/rho)^m for
    m >= 0, where Q_l^m = P_l^m / sin^m theta follows the associated Legendre
    recurrence in l, and Y_l,-m = (-1)^m conj(Y_lm). At rho = 0 the direction is
    taken as +z."""
    am = abs(m)
    with np.errstate(invalid="ignore", divide="ignore"):
        inv = np.where(rho > 0, 1.0 / rho, 0.0)
    c = np.where(rho > 0, z * inv, 1.0)
    q = np.ones_like(c)
    if l > am:
        q0, q = q, (2 * am + 1) * c
        for k in range(am + 2, l + 1):
            q0, q = q, ((2 * k - 1) * c * q - (k + am - 1) * q0) / (k - am)
    dfact = float(np.prod(np.arange(2 * am - 1, 0, -2))) if am else 1.0
    norm = math.sqrt((2 * l + 1) / (4 * math.pi) * math.factorial(l - am) / math.factorial(l + am))
    val = ((-1) ** am * norm * dfact) * q * ((x + 1j * y) * inv) ** am
    if m < 0:
        val = (-1) ** am * np.conj(val)
    return val


def translated_atom(r, l_max, mu, offset, n=2, l=1, m=1, n_theta=None, n_phi=None, chunk=128, workers=None,
                    M=1.0, modes=None):
    """The modes (default: the solver's, l - m even, l <= l_max) of psi_nlm
    centred at (offset, 0, 0) around a mass M, Bohr radius 1/(mu^2 M); returns
    the stacked state of shape (len(modes) * Nr,).

    The field is sampled on a Gauss-Legendre (theta) x uniform (phi) grid per
    chunk of radii, transformed over phi by FFT, and projected on the
    normalised Legendre factors of every m at once with one matrix product per
    m. Chunks run on a thread pool (numpy releases the GIL in its kernels);
    each writes its own columns, so the result does not depend on `workers`."""
    r = np.asarray(r, dtype=float)
    Nr = r.size
    modes = modes_for(l_max) if modes is None else [tuple(lm) for lm in modes]
    if max(abs(mm) for _, mm in modes) >= (n_phi or (4 * l_max + 32)) // 2:
        raise ValueError("n_phi too small for the largest |m|")
    a0 = 1.0 / (mu * mu * M)
    nt = n_theta or (2 * l_max + 24)
    nph = n_phi or (4 * l_max + 32)
    ct, wt = np.polynomial.legendre.leggauss(nt)
    theta = np.arccos(ct)
    st = np.sqrt(1.0 - ct * ct)
    phi = 2 * np.pi * np.arange(nph) / nph
    A = st[:, None] * np.cos(phi)[None, :]
    B = st[:, None] * np.sin(phi)[None, :]
    C = np.broadcast_to(ct[:, None], A.shape)
    # per m: the stacked modes and the (theta x mode) table of w_theta N_lm P_l^m
    # (Condon-Shortley; the negative-m factor carries (-1)^m)
    tables = []
    for mm in sorted({mm for _, mm in modes}):
        js = [j for j, (_, m2) in enumerate(modes) if m2 == mm]
        am = abs(mm)
        sign = -1.0 if (mm < 0 and am % 2) else 1.0
        T = np.array([sign * wt * np.real(special.sph_harm_y(modes[j][0], am, theta, 0.0)) for j in js]).T
        tables.append((mm, np.array(js), T))
    out = np.zeros((len(modes), Nr), dtype=np.complex128)

    def one(s):
        rs = r[s:s + chunk]
        rr = rs[:, None, None]
        x = rr * A - offset
        y = rr * B
        z = rr * C
        rho = np.sqrt(x * x + y * y + z * z)
        field = hydrogen_radial(n, l, rho, a0) * solid_ylm(l, m, x, y, z, rho)
        # sum_phi e^{-i m phi} field * 2 pi / nph
        F = np.fft.fft(field, axis=2) * (2 * np.pi / nph)
        for mm, js, T in tables:
            out[js, s:s + chunk] = (F[:, :, mm % nph] @ T).T * rs

    nw = workers or min(16, os.cpu_count() or 1)
    with ThreadPoolExecutor(max_workers=nw) as ex:
        list(ex.map(one, range(0, Nr, chunk)))
    return out.ravel()


def all_modes(l_max):
    """Every (l, m) with l <= l_max, in (l, m) order: the generator's mode list
    (get_initial_state_COM_fast.py:151 builds it, save_modes_binary sorts it)."""
    return [(l, m) for l in range(l_max + 1) for m in range(-l, l + 1)]


def modes_norm(U, r):
    """Sum over modes of the trapezoid rule of |U_lm|^2 on r
    (get_initial_state_COM_fast.py:160-161); U is (modes, Nr)."""
    return float(np.sum(np.trapezoid(np.abs(U) ** 2, x=r, axis=1)))


def _load_toml(path):
    try:
        import tomllib
    except ModuleNotFoundError:
        import tomli as tomllib
    with open(path, "rb") as f:
        return tomllib.load(f)


def generate(config, atom="primary", n_quad=None, workers=None):
    """The state get_initial_state_COM_fast.py writes for `config` (a parsed
    config.toml): returns (r, modes, U) with U of shape (len(modes), N).

    Keys read (get_initial_state_COM_fast.py:56-67,121-127,571-624): grid.N,
    grid.L, binary.q, binary.R, binary.M1 (1), atom.mu, simulation.epsilon,
    simulation.l_max, atom.n/l/m, atom_secondary.n/l/m (default: the primary's),
    atom.amplitude / atom_secondary.amplitude (1),
    simulation.normalize_initial_state (true), simulation.initial_norm (1),
    simulation.translation_n_quad (64). The primary (mass M1) sits at +R1 and
    the secondary (mass M2 = q M1) at -R2, R1 = M2 R/(M1 + M2),
    R2 = M1 R/(M1 + M2). n_quad is the Gauss-Legendre order in cos theta, raised
    to l_max + 1 where it is smaller, so that every Y_lm projection of a
    degree-<= 2 l_max polynomial stays exact."""
    if atom not in ("primary", "secondary"):
        raise ValueError("atom must be 'primary' or 'secondary'")
    N = int(config["grid"]["N"])
    L = float(config["grid"]["L"])
    q = float(config["binary"]["q"])
    R = float(config["binary"]["R"])
    M1 = float(config["binary"].get("M1", 1.0))
    M2 = q * M1
    mu = float(config["atom"]["mu"])
    r0 = float(config["simulation"]["epsilon"])
    l_max = int(config["simulation"]["l_max"])
    sim = config["simulation"]
    if n_quad is None:
        n_quad = int(sim.get("translation_n_quad", 64))
    n1, l1, m1 = int(config["atom"]["n"]), int(config["atom"]["l"]), int(config["atom"]["m"])
    sec = config.get("atom_secondary", {})
    if atom == "primary":
        nlm = (n1, l1, m1)
        M, offset = M1, M2 / (M1 + M2) * R
        amplitude = float(config["atom"].get("amplitude", 1.0))
    else:
        nlm = (int(sec.get("n", n1)), int(sec.get("l", l1)), int(sec.get("m", m1)))
        M, offset = M2, -M1 / (M1 + M2) * R
        amplitude = float(sec.get("amplitude", 1.0))
    if nlm[1] > l_max:
        raise ValueError(f"Atomic mode ({nlm[1]},{nlm[2]}) is not included in l_max={l_max}.")
    r = np.linspace(0.0, L, N)
    r[0] = r0
    modes = all_modes(l_max)
    U = translated_atom(r, l_max, mu, offset, n=nlm[0], l=nlm[1], m=nlm[2], n_theta=max(n_quad, l_max + 1),
                        workers=workers, M=M, modes=modes).reshape(len(modes), N)
    if amplitude != 1.0:
        U *= amplitude
    if bool(sim.get("normalize_initial_state", True)):
        norm = modes_norm(U, r)
        if norm <= 0.0:
            raise ValueError("Cannot normalize a state with zero norm.")
        U *= math.sqrt(float(sim.get("initial_norm", 1.0)) / norm)
    return r, modes, U


def main(argv=None):
    """Command line of get_initial_state_COM_fast.py:38-46."""
    import argparse
    ap = argparse.ArgumentParser(description="Write a translated hydrogenic atom as a MODESBIN file.")
    ap.add_argument("--config", required=True)
    ap.add_argument("--output", required=True)
    ap.add_argument("--atom", choices=["primary", "secondary"], default="primary")
    ap.add_argument("--n-quad", type=int, default=None)
    ap.add_argument("--workers", type=int, default=None)
    args = ap.parse_args(argv)
    r, modes, U = generate(_load_toml(args.config), atom=args.atom, n_quad=args.n_quad, workers=args.workers)
    write_modesbin(args.output, r.size, zip(modes, U))
    print(f"Saved {len(modes)} modes x {r.size} radial points to {args.output} (norm {modes_norm(U, r):.16e})")


def perturbed_state(base, Nr, K, rel=1e-3, seed=0, dr=None):
    """Add i.i.d. complex normal noise on interior nodes, scaled to `rel` of the
    state's RMS, then normalise to dr * |U|^2 = 1."""
    rng = np.random.default_rng(seed)
    U = np.array(base, dtype=np.complex128).reshape(K, Nr)
    rms = math.sqrt(np.mean(np.abs(U) ** 2))
    noise = rng.standard_normal((K, Nr)) + 1j * rng.standard_normal((K, Nr))
    noise[:, 0] = noise[:, -1] = 0
    U = U + rel * rms / math.sqrt(2) * noise
    U[:, 0] = U[:, -1] = 0
    if dr is not None:
        U /= math.sqrt(np.vdot(U, U).real * dr)
    return U.ravel()


if __name__ == "__main__":
    main()
