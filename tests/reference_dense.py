"""Independent dense restatement of the reference's own test oracles.

Follows tests/parallel_correctness.cpp:49-107 (dense H0, V1, V2, dVdt and the
CN step solved by LU) and tests/SphericalAbsorptionReference.hpp:19-118 (the
s-wave sink through the addition theorem) of rdellamonica/boson-cloud-binary.
The angular factors come from sympy's exact Wigner 3j symbols and spherical
harmonics, and the quadrature from numpy's Gauss-Legendre, so nothing here
shares code with the oracle or with the product.
"""
import functools
import math

import numpy as np
from sympy import N as sym_N
from sympy import Ynm, pi as sym_pi
from sympy.physics.wigner import wigner_3j


@functools.lru_cache(maxsize=None)
def w3j(a, b, c, d, e, f):
    return float(sym_N(wigner_3j(a, b, c, d, e, f), 30))


@functools.lru_cache(maxsize=None)
def sph_plm0(l, m):
    """Y_l^m(pi/2, 0) with the Condon-Shortley phase, m >= 0."""
    return float(sym_N(Ynm(l, m, sym_pi / 2, 0).expand(func=True), 30))


def modes_for(l_max):
    return [(l, m) for l in range(l_max + 1) for m in range(-l, l + 1) if (l - m) % 2 == 0]


def angular(l, m, l1, m1, ls):
    ms = m1 - m
    if abs(ms) > ls:
        return 0.0
    y = sph_plm0(ls, abs(ms))
    if ms < 0 and abs(ms) % 2:
        y = -y
    pref = (1.0 if m1 % 2 == 0 else -1.0) * math.sqrt((2 * l + 1) * (2 * l1 + 1) * (2 * ls + 1) / (4 * math.pi))
    return pref * w3j(ls, l1, l, 0, 0, 0) * w3j(ls, l1, l, -ms, m1, -m) * y


def operators(Nr, dr, r, l_max, mu, M1, M2, q, R1, R2, theta, Omega, counter, ramp):
    """Dense h0, v1, v2, dvdt at one orbit state (tests/parallel_correctness.cpp:53-107)."""
    modes = modes_for(l_max)
    K = len(modes)
    n = Nr * K
    h0 = np.zeros((n, n), dtype=complex)
    v1 = np.zeros_like(h0)
    v2 = np.zeros_like(h0)
    dvdt = np.zeros_like(h0)
    for j, (l, m) in enumerate(modes):
        for i in range(1, Nr - 1):
            row = j * Nr + i
            h0[row, row] = (2 / (dr * dr) + l * (l + 1) / (r[i] * r[i])) / (2 * mu)
            h0[row, row - 1] = h0[row, row + 1] = -1 / (2 * mu * dr * dr)
        for k, (l1, m1) in enumerate(modes):
            ms = m1 - m
            for ls in range(abs(l1 - l), l1 + l + 1):
                if abs(ms) > ls:
                    continue
                ang = angular(l, m, l1, m1, ls)
                if ang == 0.0:
                    continue
                phase = np.exp(1j * ms * theta)
                rows = j * Nr + np.arange(1, Nr - 1)
                cols = k * Nr + np.arange(1, Nr - 1)
                ri = r[1:Nr - 1]

                def radial(radius):
                    return np.minimum(ri, radius) ** ls / np.maximum(ri, radius) ** (ls + 1)

                common = -mu * (M1 + M2) * 4 * math.pi / (2 * ls + 1) * ang * phase
                a1 = common / (1 + q) * radial(R1)
                a2 = common * (1.0 if ms % 2 == 0 else -1.0) * q / (1 + q) * radial(R2)
                v1[rows, cols] += a1
                v2[rows, cols] += a2
                dvdt[rows, cols] += (a1 + ramp * a2) * (-1.0 if counter else 1.0) * 1j * ms * Omega
    return h0, v1, v2, dvdt


def absorption_moments(Nr, dr, r, l_max, width, d):
    """tests/SphericalAbsorptionReference.hpp:48-74, integrating in the distance s."""
    mom = np.zeros((2 * l_max + 1, Nr))
    if width < dr or width <= 0:
        return mom
    xg, wg = np.polynomial.legendre.leggauss(256)
    for i in range(1, Nr - 1):
        ri = r[i]
        low, high = abs(ri - d), min(width, ri + d)
        if d < 1e-8 * width:
            if ri < width:
                mom[0, i] = 4 * math.pi * math.exp(-1 / (1 - ri * ri / (width * width)))
        elif low < high:
            s = 0.5 * (high - low) * xg + 0.5 * (high + low)
            w = 0.5 * (high - low) * wg
            u = np.clip((ri * ri + d * d - s * s) / (2 * ri * d), -1.0, 1.0)
            ker = 2 * math.pi / (ri * d) * w * s * np.exp(-1 / (1 - s * s / (width * width)))
            for l in range(2 * l_max + 1):
                mom[l, i] += np.sum(ker * np.polynomial.legendre.Legendre.basis(l)(u))
    norm = np.sum(dr * r[1:Nr - 1] ** 2 * mom[0, 1:Nr - 1])
    if norm > 0:
        mom /= norm
    return mom


def swave_projector(Nr, dr, r, l_max, width, d, theta, body):
    """tests/SphericalAbsorptionReference.hpp:94-106."""
    mom = absorption_moments(Nr, dr, r, l_max, width, d)
    phi = theta + (0.0 if body == 1 else math.pi)
    modes = modes_for(l_max)
    v = np.zeros(Nr * len(modes), dtype=complex)
    for j, (l, m) in enumerate(modes):
        y = sph_plm0(l, abs(m))
        if m < 0 and abs(m) % 2:
            y = -y
        direction = np.conj(y * np.exp(1j * m * phi))
        v[j * Nr + 1:j * Nr + Nr - 1] = r[1:Nr - 1] * mom[l, 1:Nr - 1] * direction
    return v


def sink(Nr, dr, r, l_max, smear, M1, M2, R1, R2, theta, ramp):
    """Dense Gamma = sum_a 16 pi M_a^2 ramp dr v_a v_a^H (SphericalAbsorptionReference.hpp:109-118)."""
    gamma = np.zeros((Nr * len(modes_for(l_max)),) * 2, dtype=complex)
    for body, mass, d in ((1, M1, R1), (2, M2, R2)):
        v = swave_projector(Nr, dr, r, l_max, smear * mass, d, theta, body)
        gamma += 16 * math.pi * mass * mass * ramp * dr * np.outer(v, v.conj())
    return gamma


def smoothstep(x):
    return 3 * x * x - 2 * x * x * x
