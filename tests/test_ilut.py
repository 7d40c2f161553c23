"""Host ILUT + BiCGSTAB (gm_ilut_bicgstab), the solve behind the "ilut"
preconditioner key and the fallback of a failed radial solve
(src/CrankNicolsonModes_matrices.cpp:18-24, 173-185, 203-214).

Eigen's IncompleteLUT is not in the reference tree, so the factors are not
compared with Eigen's (parity unpinned for the factors, as for Eigen's own
iteration counts). What is pinned: the solution against scipy's sparse direct
solve on CN-shaped systems, an exact factor (no dropping, unlimited fill)
converging in one iteration, the zero-pivot replacement, and the errors.
"""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import gm


def cn_system(K=6, N=200, dt=2.0, dr=1.0, mu=0.2, seed=0, node_major=True):
    """A = I + i dt/2 H with H = radial Laplacian per mode (centrifugal term
    l(l+1)/2mu r^2) plus a random Hermitian coupling between modes at each node."""
    rng = np.random.default_rng(seed)
    r = dr * np.arange(1, N + 1)
    ls = np.repeat(np.arange(K), 1)
    rows, cols, vals = [], [], []

    def idx(j, i):
        return i * K + j if node_major else j * N + i

    for i in range(N):
        C = rng.standard_normal((K, K)) + 1j * rng.standard_normal((K, K))
        C = 0.05 * (C + C.conj().T) / np.exp(r[i] / 50.0)
        for j in range(K):
            rows.append(idx(j, i)); cols.append(idx(j, i))
            vals.append(1.0 / (mu * dr * dr) + ls[j] * (ls[j] + 1) / (2 * mu * r[i] ** 2) + C[j, j].real)
            for k in range(K):
                if k != j:
                    rows.append(idx(j, i)); cols.append(idx(k, i)); vals.append(C[j, k])
            for i2 in (i - 1, i + 1):
                if 0 <= i2 < N:
                    rows.append(idx(j, i)); cols.append(idx(j, i2)); vals.append(-1.0 / (2 * mu * dr * dr))
    n = K * N
    H = sp.csr_matrix((np.array(vals, dtype=np.complex128), (rows, cols)), shape=(n, n))
    A = (sp.identity(n, dtype=np.complex128, format="csr") + 0.5j * dt * H).tocsr()
    A.sort_indices()
    b = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    return A, b


def solve(A, b, **kw):
    return gm.ilut_bicgstab(A.indptr, A.indices, A.data, b, **kw)


def relres(A, x, b):
    return np.linalg.norm(A @ x - b) / np.linalg.norm(b)


@pytest.mark.parametrize("eigen_mode", [True, False])
def test_ilut_solves_cn_system(eigen_mode):
    A, b = cn_system()
    x, its = solve(A, b, eigen_mode=eigen_mode)
    ref = spla.spsolve(A.tocsc(), b)
    assert relres(A, x, b) <= 1e-10
    assert np.linalg.norm(x - ref) / np.linalg.norm(ref) < 1e-9
    assert 1 <= its <= 5


def test_exact_factor_converges_in_one_iteration():
    """No dropping and unlimited fill: ILUT is the exact LU of the (banded,
    node-major) matrix, so the preconditioned operator is the identity."""
    A, b = cn_system(K=4, N=60, seed=3)
    x, its = solve(A, b, droptol=0.0, fill_factor=10 ** 6)
    ref = spla.spsolve(A.tocsc(), b)
    assert its == 1
    assert np.linalg.norm(x - ref) / np.linalg.norm(ref) < 1e-12


def test_dropping_costs_iterations_not_accuracy():
    A, b = cn_system(K=4, N=80, seed=5, dt=20.0)
    _, its_full = solve(A, b, droptol=0.0, fill_factor=10 ** 6)
    x, its = solve(A, b, droptol=0.2, fill_factor=1)
    assert its > its_full
    assert relres(A, x, b) <= 1e-10


def test_mode_major_order_also_converges():
    """The step hands the factor the node-major order; any order solves."""
    A, b = cn_system(K=3, N=40, seed=7, node_major=False)
    x, _ = solve(A, b)
    assert relres(A, x, b) <= 1e-10


def test_guess_is_used():
    A, b = cn_system(K=3, N=40, seed=8)
    ref = spla.spsolve(A.tocsc(), b)
    x, its = solve(A, b, x0=ref)
    assert its == 0 and np.array_equal(x, ref)


def test_zero_pivot_is_replaced():
    """A zero diagonal becomes (1e-4 + droptol) |row|_2 (Saad's ILUT remedy)."""
    A = sp.csr_matrix(np.array([[0.0, 1.0], [1.0, 0.0]], dtype=np.complex128))
    A.sort_indices()
    b = np.array([2.0 + 1j, -3.0])
    x, _ = solve(A, b)
    np.testing.assert_allclose(x, [-3.0, 2.0 + 1j], rtol=0, atol=1e-9)


def test_zero_rhs_gives_zero():
    A, _ = cn_system(K=2, N=10)
    x, its = solve(A, np.zeros(A.shape[0], dtype=np.complex128), x0=np.ones(A.shape[0]))
    assert its == 0 and not x.any()


def test_empty_row_fails_the_factorization():
    A = sp.csr_matrix((np.array([1.0 + 0j]), ([0], [0])), shape=(2, 2))
    with pytest.raises(gm.GMError, match="ILUT factorization failed"):
        solve(A, np.ones(2, dtype=np.complex128))


def test_malformed_csr_is_rejected():
    ptr = np.array([0, 2, 1])
    with pytest.raises(gm.InvalidArgument):
        gm.ilut_bicgstab(ptr, np.array([0, 1]), np.ones(2), np.ones(2))
    with pytest.raises(gm.InvalidArgument):  # columns not increasing
        gm.ilut_bicgstab(np.array([0, 2, 3]), np.array([1, 0, 1]), np.ones(3), np.ones(2))


def test_unreached_tolerance_reports_iterations():
    A, b = cn_system(K=4, N=80, seed=5, dt=20.0)
    with pytest.raises(gm.GMError, match="above the tolerance"):
        solve(A, b, droptol=0.5, fill_factor=1, max_iterations=1)
