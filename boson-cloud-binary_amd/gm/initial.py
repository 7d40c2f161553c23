"""MODESBIN initial states.

The format is the reference's (include/InitialData.hpp:18-34,
src/InitialData.cpp:7-79): the 8 bytes "MODESBIN", int32 version 1, int64 Nr,
int64 number of stored modes, then per mode int32 l, int32 m, Nr doubles of
Re U_lm and Nr doubles of Im U_lm (little-endian). Stored modes that the
solver does not list are ignored; listed modes that are not stored are zero.

`translated_atom` builds the state the reference's initial_states scripts
write for a hydrogenic (n, l, m) atom on the primary, translated to the centre
of mass: U_lm(r) = r * integral Y_lm^* psi_nlm(x - x_1) dOmega, computed here
with a Gauss-Legendre x FFT product quadrature over the sphere.
"""
import math
import struct

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
    (mode-major) for the solver's mode list."""
    with open(path, "rb") as f:
        data = f.read()
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


def translated_atom(r, l_max, mu, offset, n=2, l=1, m=1, n_theta=None, n_phi=None, chunk=512):
    """Modes (l - m even, l <= l_max) of psi_nlm centred at (offset, 0, 0),
    Bohr radius 1/mu^2; returns the stacked state of shape (K * Nr,)."""
    r = np.asarray(r, dtype=float)
    Nr = r.size
    modes = modes_for(l_max)
    a0 = 1.0 / (mu * mu)
    nt = n_theta or (2 * l_max + 24)
    nph = n_phi or (4 * l_max + 32)
    ct, wt = np.polynomial.legendre.leggauss(nt)
    theta = np.arccos(ct)
    st = np.sqrt(1.0 - ct * ct)
    phi = 2 * np.pi * np.arange(nph) / nph
    # normalised Legendre factors N_lm P_l^m(cos theta) (Condon-Shortley), m >= 0
    ylm = {}
    for (ll, mm) in modes:
        am = abs(mm)
        if (ll, am) not in ylm:
            ylm[(ll, am)] = np.real(special.sph_harm_y(ll, am, theta, 0.0))
    out = np.zeros((len(modes), Nr), dtype=np.complex128)
    cp, sp = np.cos(phi), np.sin(phi)
    for s in range(0, Nr, chunk):
        rr = r[s:s + chunk][:, None, None]
        x = rr * st[None, :, None] * cp[None, None, :] - offset
        y = rr * st[None, :, None] * sp[None, None, :]
        z = rr * ct[None, :, None] * np.ones_like(cp)[None, None, :]
        rho = np.sqrt(x * x + y * y + z * z)
        with np.errstate(invalid="ignore", divide="ignore"):
            cos_t = np.where(rho > 0, z / rho, 1.0)
            ph = np.arctan2(y, x)
        ang = special.sph_harm_y(l, m, np.arccos(np.clip(cos_t, -1, 1)), ph)
        field = hydrogen_radial(n, l, rho, a0) * ang
        # sum_phi e^{-i m phi} field * 2 pi / nph
        F = np.fft.fft(field, axis=2) * (2 * np.pi / nph)
        for j, (ll, mm) in enumerate(modes):
            col = F[:, :, mm % nph]
            yl = ylm[(ll, abs(mm))]
            if mm < 0 and abs(mm) % 2:
                yl = -yl
            out[j, s:s + chunk] = r[s:s + chunk] * np.sum(col * (wt * yl)[None, :], axis=1)
    return out.ravel()


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
