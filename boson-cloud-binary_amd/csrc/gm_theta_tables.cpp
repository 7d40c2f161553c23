// gm_theta_tables.cpp - host tables of the polar-quadrature coupling (gm_theta.hpp).
#include "gm_theta.hpp"

#include <cmath>

namespace gmt {

void gauss_legendre_half(int n, std::vector<long double> &x, std::vector<long double> &w)
{
    const int N = 2 * n;
    const long double pi = 3.141592653589793238462643383279502884L;
    x.assign(n, 0.0L);
    w.assign(n, 0.0L);
    for (int k = 0; k < n; ++k) {
        // k-th positive root of P_N, from the largest down (Tricomi's start)
        long double z = std::cos(pi * (k + 0.75L) / (N + 0.5L));
        long double dp = 1.0L;
        for (int it = 0; it < 100; ++it) {
            long double p0 = 1.0L, p1 = z;
            for (int l = 2; l <= N; ++l) {
                const long double p2 = ((2 * l - 1) * z * p1 - (l - 1) * p0) / l;
                p0 = p1;
                p1 = p2;
            }
            dp = N * (z * p1 - p0) / (z * z - 1.0L);
            const long double dz = p1 / dp;
            z -= dz;
            if (std::fabs(dz) < 1e-21L) break;
        }
        long double p0 = 1.0L, p1 = z;
        for (int l = 2; l <= N; ++l) {
            const long double p2 = ((2 * l - 1) * z * p1 - (l - 1) * p0) / l;
            p0 = p1;
            p1 = p2;
        }
        dp = N * (z * p1 - p0) / (z * z - 1.0L);
        x[n - 1 - k] = z;
        w[n - 1 - k] = 2.0L / ((1.0L - z * z) * dp * dp);
    }
}

void assoc_legendre(int L, long double x, std::vector<long double> &out)
{
    out.assign(static_cast<size_t>(L + 1) * (L + 1), 0.0L);
    const long double s = std::sqrt(std::fmax(0.0L, 1.0L - x * x));
    long double pmm = std::sqrt(0.5L);
    for (int m = 0; m <= L; ++m) {
        if (m > 0) pmm = -pmm * std::sqrt((2.0L * m + 1.0L) / (2.0L * m)) * s;
        out[static_cast<size_t>(m) * (L + 1) + m] = pmm;
        if (m + 1 <= L) out[static_cast<size_t>(m + 1) * (L + 1) + m] = std::sqrt(2.0L * m + 3.0L) * x * pmm;
        for (int l = m + 2; l <= L; ++l) {
            const long double a = std::sqrt((4.0L * l * l - 1.0L) / (static_cast<long double>(l) * l - m * m));
            const long double b = std::sqrt((static_cast<long double>(l - 1) * (l - 1) - m * m) /
                                            (4.0L * (l - 1) * (l - 1) - 1.0L));
            out[static_cast<size_t>(l) * (L + 1) + m] =
                a * (x * out[static_cast<size_t>(l - 1) * (L + 1) + m] - b * out[static_cast<size_t>(l - 2) * (L + 1) + m]);
        }
    }
}

ThetaTables build_theta_tables(int lmax, int LM)
{
    ThetaTables T;
    T.lmax = lmax;
    T.LM = LM;
    const int L2 = 2 * lmax;
    // block layout of the class LM; entries of modes l > l_max (or l* > 2 l_max) are zero
    auto n_m = [&](int m) { return (LM - std::abs(m)) / 2 + 1; };
    auto n_l = [&](int d) { return (2 * LM - d) / 2 + 1; };
    for (int m = -LM; m <= LM; ++m) T.ks += (n_m(m) + 3) / 4;
    for (int d = 0; d <= 2 * LM; ++d) T.kd += (n_l(d) + 3) / 4;
    T.ngmax = theta_rule(lmax, L2);
    const size_t tri = static_cast<size_t>(T.ngmax) * (T.ngmax + 1) / 2;
    T.A.assign(tri * T.ks * 16, 0.0);
    T.C.assign(tri * T.ks * 16, 0.0);
    T.U.assign(tri * T.kd * 16, 0.0);
    const long double inv2pi = 0.159154943091895335768883763372514362L;
    std::vector<long double> p0;
    assoc_legendre(L2, 0.0L, p0);
    const int W = L2 + 1;
    // signed p_l^m(x) from the table of m >= 0
    auto pf = [&](const std::vector<long double> &P, int l, int m) {
        const long double v = P[static_cast<size_t>(l) * W + std::abs(m)];
        return (m < 0 && (std::abs(m) & 1)) ? -v : v;
    };
    for (int ng = 1; ng <= T.ngmax; ++ng) {
        const int n = 4 * ng;
        std::vector<long double> x, w;
        gauss_legendre_half(n, x, w);
        std::vector<std::vector<long double>> P(n);
        for (int i = 0; i < n; ++i) assoc_legendre(L2, x[i], P[i]);
        const size_t baseA = static_cast<size_t>(T.ks) * 16 * ng * (ng - 1) / 2;
        const size_t baseU = static_cast<size_t>(T.kd) * 16 * ng * (ng - 1) / 2;
        for (int t = 0; t < ng; ++t) {
            int ks = 0;
            for (int m = -LM; m <= LM; ++m) {
                const int ns = (n_m(m) + 3) / 4;
                for (int s = 0; s < ns; ++s, ++ks) {
                    for (int a = 0; a < 4; ++a)
                        for (int b = 0; b < 4; ++b) {
                            const size_t at = baseA + (static_cast<size_t>(t) * T.ks + ks) * 16 + 4 * a + b;
                            // phase A: row b = point 4 t + b, k a = mode (m, 4 s + a)
                            const int la = std::abs(m) + 2 * (4 * s + a);
                            if (la <= lmax) T.A[at] = static_cast<double>(pf(P[4 * t + b], la, m));
                            // phase C: row b = mode (m, 4 s + b), k a = point 4 t + a; the
                            // half rule's weight is twice the full rule's
                            const int lb = std::abs(m) + 2 * (4 * s + b);
                            if (lb <= lmax)
                                T.C[at] = static_cast<double>(2.0L * w[4 * t + a] * pf(P[4 * t + a], lb, m));
                        }
                }
            }
            int kd = 0;
            for (int d = 0; d <= 2 * LM; ++d) {
                const int ns = (n_l(d) + 3) / 4;
                for (int s = 0; s < ns; ++s, ++kd)
                    for (int a = 0; a < 4; ++a)
                        for (int b = 0; b < 4; ++b) {
                            const int ls = d + 2 * (4 * s + a);
                            if (ls > L2) continue;
                            T.U[baseU + (static_cast<size_t>(t) * T.kd + kd) * 16 + 4 * a + b] =
                                static_cast<double>(inv2pi * pf(p0, ls, d) * pf(P[4 * t + b], ls, d));
                        }
            }
        }
    }
    return T;
}

}  // namespace gmt
