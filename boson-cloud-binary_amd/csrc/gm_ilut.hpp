// gm_ilut.hpp - host ILUT preconditioner and BiCGSTAB for the full-ILUT solve
// of solve_CN (src/CrankNicolsonModes_matrices.cpp:18-24, 173-185, 203-214).
//
// The reference uses Eigen 3.4's IncompleteLUT (fill factor 50, drop tolerance
// 1e-9) under Eigen's BiCGSTAB (tolerance 1e-10, 5000 iterations), or under its
// own BiCGSTAB with the Woodbury sink correction when sinks are active. Eigen
// is not part of the reference tree; what is implemented here is the published
// algorithm: Saad's dual-threshold ILUT (Y. Saad, "ILUT: a dual threshold
// incomplete LU factorization", Numer. Linear Algebra Appl. 1 (1994) 387-402),
// with the row-wise IKJ elimination, entries below droptol * |row|_2 dropped,
// the `fill` largest kept in each of L and U, and a zero pivot replaced by
// (1e-4 + droptol) * |row|_2. `fill` follows Eigen's fill-factor parameter:
// nnz(A) * fill_factor / n + 1 entries per row, half in L and half in U.
//
// Ordering: the caller hands the matrix in node-major order (row i*K + j for
// mode j at node i). There the CN operator is block tridiagonal with K x K
// blocks, so the factor stays inside a band of about 3K per row (Eigen orders
// by AMD instead; the factors differ, the solution does not).
#pragma once

#include <complex>
#include <cstddef>
#include <functional>
#include <string>
#include <vector>

namespace gmh {

using cplx = std::complex<double>;

// Compressed sparse rows, columns sorted within each row.
struct CsrZ {
    int n = 0;
    std::vector<int> ptr, col;
    std::vector<cplx> val;
    size_t nnz() const { return val.size(); }
    void multiply(const cplx *x, cplx *y) const;  // y = A x
};

class Ilut {
public:
    // false (and a reason) for an empty row; the factor is then unusable.
    bool compute(const CsrZ &A, double droptol, int fill_factor, std::string *why);
    // x = (L U)^{-1} b; x and b may alias.
    void solve(const cplx *b, cplx *x) const;
    size_t nnz() const { return lval.size() + uval.size(); }

private:
    int n = 0;
    std::vector<int> lptr, lcol, uptr, ucol;  // L strict lower (unit diagonal); U diagonal first
    std::vector<cplx> lval, uval;
};

using LinOp = std::function<void(const cplx *in, cplx *out)>;

// Right-preconditioned BiCGSTAB on n unknowns, from the guess in x.
// eigen_mode = true: Eigen 3.4's recurrence (restart threshold eps^2 |r0|^2,
// the iteration counter reset at the first restart); false: the reference's
// own (src/CrankNicolsonModes_matrices.cpp:47-91, restart threshold
// 1e-32 |r0|^2). Returns whether |r|^2 <= tol^2 |b|^2; iterations are counted.
bool bicgstab(const LinOp &A, const LinOp &M, const cplx *b, cplx *x, size_t n, double tol, int max_iterations,
              bool eigen_mode, int *iterations);

}  // namespace gmh
