// gm_ilut.cpp - host ILUT(fill, droptol) and BiCGSTAB (see gm_ilut.hpp).
#include "gm_ilut.hpp"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <queue>

namespace gmh {

void CsrZ::multiply(const cplx *x, cplx *y) const
{
    for (int i = 0; i < n; ++i) {
        cplx acc = 0.0;
        for (int q = ptr[i]; q < ptr[i + 1]; ++q) acc += val[q] * x[col[q]];
        y[i] = acc;
    }
}

namespace {

// Keep the `keep` entries of largest magnitude among cols (values in w).
void keep_largest(std::vector<int> &cols, const std::vector<cplx> &w, int keep)
{
    if (static_cast<int>(cols.size()) <= keep) return;
    std::nth_element(cols.begin(), cols.begin() + keep, cols.end(),
                     [&](int a, int b) { return std::abs(w[a]) > std::abs(w[b]); });
    cols.resize(static_cast<size_t>(keep));
}

}  // namespace

bool Ilut::compute(const CsrZ &A, double droptol, int fill_factor, std::string *why)
{
    n = A.n;
    const long long per_row = std::min<long long>(n, static_cast<long long>(A.nnz()) * fill_factor / std::max(n, 1) + 1);
    const int keepL = static_cast<int>(per_row / 2), keepU = static_cast<int>(per_row / 2);
    lptr.assign(static_cast<size_t>(n) + 1, 0);
    uptr.assign(static_cast<size_t>(n) + 1, 0);
    lcol.clear(), ucol.clear(), lval.clear(), uval.clear();
    std::vector<cplx> w(static_cast<size_t>(n), 0.0);
    std::vector<char> present(static_cast<size_t>(n), 0);
    std::vector<int> touched, lkept, upart;
    std::priority_queue<int, std::vector<int>, std::greater<int>> pending;  // L columns, ascending
    for (int i = 0; i < n; ++i) {
        touched.clear();
        lkept.clear();
        upart.clear();
        double norm2 = 0.0;
        auto add = [&](int c, cplx v) {
            present[c] = 1;
            w[c] = v;
            touched.push_back(c);
            if (c < i) pending.push(c);
            else if (c > i) upart.push_back(c);
        };
        add(i, 0.0);
        for (int q = A.ptr[i]; q < A.ptr[i + 1]; ++q) {
            const int c = A.col[q];
            norm2 += std::norm(A.val[q]);
            if (present[c]) w[c] += A.val[q];
            else add(c, A.val[q]);
        }
        const double rnorm = std::sqrt(norm2);
        if (!(rnorm > 0.0)) {
            if (why) *why = "ILUT: row " + std::to_string(i) + " of the matrix is empty";
            for (int c : touched) present[c] = 0, w[c] = 0.0;
            return false;
        }
        const double tau = droptol * rnorm;
        // row-wise elimination with the finished rows of U, in column order
        while (!pending.empty()) {
            const int k = pending.top();
            pending.pop();
            const cplx f = w[k] / uval[static_cast<size_t>(uptr[k])];
            if (std::abs(f) <= tau) {
                w[k] = 0.0;
                continue;
            }
            w[k] = f;
            lkept.push_back(k);
            for (int q = uptr[k] + 1; q < uptr[k + 1]; ++q) {
                const int c = ucol[q];
                if (present[c]) w[c] -= f * uval[q];
                else add(c, -f * uval[q]);
            }
        }
        // second dropping rule: below tau, then the largest `keep` of L and of U
        upart.erase(std::remove_if(upart.begin(), upart.end(), [&](int c) { return std::abs(w[c]) <= tau; }),
                    upart.end());
        keep_largest(lkept, w, keepL);
        keep_largest(upart, w, keepU);
        std::sort(lkept.begin(), lkept.end());
        std::sort(upart.begin(), upart.end());
        for (int c : lkept) {
            lcol.push_back(c);
            lval.push_back(w[c]);
        }
        lptr[i + 1] = static_cast<int>(lval.size());
        cplx d = w[i];
        if (d == cplx(0.0)) d = (1e-4 + droptol) * rnorm;
        ucol.push_back(i);
        uval.push_back(d);
        for (int c : upart) {
            ucol.push_back(c);
            uval.push_back(w[c]);
        }
        uptr[i + 1] = static_cast<int>(uval.size());
        for (int c : touched) present[c] = 0, w[c] = 0.0;
    }
    return true;
}

void Ilut::solve(const cplx *b, cplx *x) const
{
    if (x != b) std::copy(b, b + n, x);
    for (int i = 0; i < n; ++i) {
        cplx acc = x[i];
        for (int q = lptr[i]; q < lptr[i + 1]; ++q) acc -= lval[q] * x[lcol[q]];
        x[i] = acc;
    }
    for (int i = n - 1; i >= 0; --i) {
        cplx acc = x[i];
        for (int q = uptr[i] + 1; q < uptr[i + 1]; ++q) acc -= uval[q] * x[ucol[q]];
        x[i] = acc / uval[static_cast<size_t>(uptr[i])];
    }
}

namespace {

cplx dotc(const cplx *a, const cplx *b, size_t n)  // sum conj(a) b
{
    cplx s = 0.0;
    for (size_t q = 0; q < n; ++q) s += std::conj(a[q]) * b[q];
    return s;
}

double norm2(const cplx *a, size_t n)
{
    double s = 0.0;
    for (size_t q = 0; q < n; ++q) s += std::norm(a[q]);
    return s;
}

}  // namespace

bool bicgstab(const LinOp &A, const LinOp &M, const cplx *b, cplx *x, size_t n, double tol, int max_iterations,
              bool eigen_mode, int *iterations)
{
    *iterations = 0;
    const double b2 = norm2(b, n);
    if (b2 == 0.0) {
        std::fill(x, x + n, cplx(0.0));
        return true;
    }
    const double threshold = tol * tol * b2;
    std::vector<cplx> r(n), r0(n), v(n, 0.0), p(n, 0.0), y(n), z(n), s(n), t(n);
    auto residual = [&] {
        A(x, r.data());
        for (size_t q = 0; q < n; ++q) r[q] = b[q] - r[q];
    };
    residual();
    r0 = r;
    double r0_norm2 = norm2(r0.data(), n);
    const double fixed_restart = r0_norm2 * 1e-32;
    const double eps2 = DBL_EPSILON * DBL_EPSILON;
    cplx rho = 1.0, alpha = 1.0, w = 1.0;
    int i = 0, restarts = 0;
    while (norm2(r.data(), n) > threshold && i < max_iterations) {
        const cplx rho_old = rho;
        rho = dotc(r0.data(), r.data(), n);
        if (std::abs(rho) < (eigen_mode ? eps2 * r0_norm2 : fixed_restart)) {
            // lost bi-orthogonality: restart from the current residual
            residual();
            r0 = r;
            r0_norm2 = norm2(r.data(), n);
            rho = r0_norm2;
            if (eigen_mode && restarts++ == 0) i = 0;
        }
        const cplx beta = (rho / rho_old) * (alpha / w);
        for (size_t q = 0; q < n; ++q) p[q] = r[q] + beta * (p[q] - w * v[q]);
        M(p.data(), y.data());
        A(y.data(), v.data());
        alpha = rho / dotc(r0.data(), v.data(), n);
        for (size_t q = 0; q < n; ++q) s[q] = r[q] - alpha * v[q];
        M(s.data(), z.data());
        A(z.data(), t.data());
        const double tn = norm2(t.data(), n);
        w = tn > 0.0 ? dotc(t.data(), s.data(), n) / tn : cplx(0.0);
        for (size_t q = 0; q < n; ++q) {
            x[q] += alpha * y[q] + w * z[q];
            r[q] = s[q] - w * t[q];
        }
        ++i;
    }
    *iterations = i;
    return norm2(r.data(), n) <= threshold;
}

}  // namespace gmh
