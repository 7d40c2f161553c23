"""The polar-quadrature coupling's host tables on the CPU (tests/cpp/theta_tables.cpp).

The program emulates the GPU kernel's arithmetic lane by lane (the 4x4x4
matrix-core blocks, the l_max-class layout, the rule of a multipole cut, the U
sums, the banded Toeplitz product) on the tables the library builds, and
compares the coupling with the direct sum over Gaunt terms from Racah's 3j
formula (the routine the Gaunt-form tables use; pinned against sympy in
test_oracle.py). The Gauss-Legendre rule of 8 ng points is exact for
polynomials of degree 16 ng - 1, so a tile with cut lc needs 16 ng - 1 >=
2 l_max + lc: one rule smaller must fail.
"""
import json
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
PROG = os.path.join(HERE, "cpp", "_bin", "theta_tables")


def run(*args):
    if not os.path.exists(PROG):
        pytest.fail("tests/cpp/_bin/theta_tables is not built: run boson-cloud-binary_amd/build.py")
    res = subprocess.run([PROG] + [str(a) for a in args], capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stdout + res.stderr
    return json.loads(res.stdout.strip().splitlines()[-1])


@pytest.mark.parametrize("lmax,cls,lc", [(0, 4, 0), (1, 4, 2), (2, 4, 4), (3, 4, 6), (4, 4, 8), (5, 8, 10),
                                         (6, 8, 12), (6, 8, 3), (8, 8, 5), (11, 12, 22), (12, 12, 24),
                                         (16, 16, 9), (17, 20, 34), (20, 20, 11), (20, 20, 23), (20, 20, 24),
                                         (20, 20, 40)])
def test_polar_coupling_equals_gaunt_sum(lmax, cls, lc):
    r = run(lmax, cls, lc)
    assert 16 * r["rule"] - 1 >= 2 * lmax + lc
    assert r["rel_err"] < 5e-15, r


@pytest.mark.parametrize("lmax,cls,lc", [(6, 8, 12), (12, 12, 24), (20, 20, 23), (20, 20, 40)])
def test_one_rule_smaller_is_not_exact(lmax, cls, lc):
    r = run(lmax, cls, lc)
    if r["rule"] == 1:
        pytest.skip("no smaller rule")
    bad = run(lmax, cls, lc, r["rule"] - 1)
    assert bad["rel_err"] > 1e-8, bad
