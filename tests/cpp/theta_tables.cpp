// The polar-quadrature coupling's host tables (boson-cloud-binary_amd/csrc/
// gm_theta_tables.cpp) against the direct Gaunt sum, on the CPU.
//
// Emulates the GPU kernel's arithmetic lane by lane: v_mfma_f64_4x4x4f64
// (lane 16 a + 4 g + b holds A[g][row b][k a], B[g][k a][col b] and
// D[g][row a][col b]), the class layout of the tables, the rule of a cut,
// the U sums masked above the cut, phase A, the banded Toeplitz phase B and
// phase C; then compares (V x)_j = sum_k sum_l* G_l*[j, k] E_l* x_k with G
// from Racah's 3j formula (gm_math.cpp, the routine the Gaunt tables use),
// the convention of src/CrankNicolsonModes_hamiltonian.cpp:68-104.
//
//   theta_tables <l_max> <class> <lc> [rule]   prints {"lmax":..,"lc":..,"rule":..,"rel_err":..}
// (rule forces a quadrature rule, e.g. one too small for the cut)
#include "gm_math.hpp"
#include "gm_theta.hpp"

#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

typedef std::complex<double> cd;

static void mfma(const double *A, const double *B, double *D)
{
    double out[64];
    for (int l = 0; l < 64; ++l) {
        const int i = l >> 4, g = (l >> 2) & 3, j = l & 3;
        double s = D[l];
        for (int k = 0; k < 4; ++k) s += A[16 * k + 4 * g + i] * B[16 * k + 4 * g + j];
        out[l] = s;
    }
    for (int l = 0; l < 64; ++l) D[l] = out[l];
}

int main(int argc, char **argv)
{
    if (argc < 4) {
        std::fprintf(stderr, "usage: theta_tables <l_max> <class> <lc>\n");
        return 2;
    }
    const int lmax = std::atoi(argv[1]), LM = std::atoi(argv[2]), lc = std::atoi(argv[3]);
    if (lmax < 0 || LM < lmax || lc < 0 || lc > 2 * lmax) return 2;
    const int K = (lmax + 1) * (lmax + 2) / 2, Ls = 2 * lmax + 1, NM = 2 * LM + 1;
    const gmt::ThetaTables T = gmt::build_theta_tables(lmax, LM);
    auto jof = [](int l, int m) { return l * (l + 1) / 2 + (m + l) / 2; };
    auto nsm = [&](int m) { return ((LM - std::abs(m)) / 2 + 1 + 3) / 4; };
    auto nsd = [&](int d) { return ((2 * LM - d) / 2 + 1 + 3) / 4; };
    std::mt19937_64 rng(1);
    std::normal_distribution<double> nd;
    // one tile of 8 nodes: x[j][n], profiles E[l*][n] (zero above the cut)
    std::vector<cd> x(K * 8);
    std::vector<double> E(Ls * 8);
    for (auto &v : x) v = cd(nd(rng), nd(rng));
    for (int ls = 0; ls < Ls; ++ls)
        for (int n = 0; n < 8; ++n) E[ls * 8 + n] = ls <= lc ? nd(rng) * std::exp(-0.2 * ls) : 0.0;
    // direct sum over the Gaunt terms
    std::vector<cd> yd(K * 8, 0.0);
    for (int l = 0; l <= lmax; ++l)
        for (int m = -l; m <= l; m += 2)
            for (int l1 = 0; l1 <= lmax; ++l1)
                for (int m1 = -l1; m1 <= l1; m1 += 2) {
                    const int ms = m1 - m;
                    for (int ls = std::abs(l - l1); ls <= std::min(l + l1, lc); ls += 2) {
                        if (std::abs(ms) > ls) continue;
                        const double w1 = gmh::wigner3j(ls, l1, l, 0, 0, 0);
                        if (w1 == 0.0) continue;
                        const double norm = std::sqrt((2.0 * l + 1) * (2.0 * l1 + 1) * (2.0 * ls + 1) / (4.0 * M_PI));
                        const double w2 = gmh::wigner3j(ls, l1, l, -ms, m1, -m);
                        double yv = gmh::sph_plm0(ls, std::abs(ms));
                        if (ms < 0 && std::abs(ms) % 2 != 0) yv = -yv;
                        const double G = (m1 % 2 == 0 ? 1.0 : -1.0) * norm * w1 * w2 * yv;
                        for (int n = 0; n < 8; ++n) yd[jof(l, m) * 8 + n] += G * E[ls * 8 + n] * x[jof(l1, m1) * 8 + n];
                    }
                }
    // the kernel, lane by lane: column c = 4 g + b is node c >> 1, part c & 1
    const int ng = argc > 4 ? std::atoi(argv[4]) : gmt::theta_rule(lmax, lc);
    if (ng < 1 || ng > T.ngmax) return 2;
    const size_t tri = static_cast<size_t>(ng) * (ng - 1) / 2;
    std::vector<double> sX(K * 16);
    for (int j = 0; j < K; ++j)
        for (int c = 0; c < 16; ++c) sX[j * 16 + c] = (c & 1) ? x[j * 8 + (c >> 1)].imag() : x[j * 8 + (c >> 1)].real();
    std::vector<double> yacc(static_cast<size_t>(T.ks) * 64, 0.0);
    for (int t = 0; t < ng; ++t) {
        const double *TA = T.A.data() + (tri + t) * T.ks * 16;
        const double *TC = T.C.data() + (tri + t) * T.ks * 16;
        const double *TU = T.U.data() + (tri + t) * T.kd * 16;
        std::vector<double> U(static_cast<size_t>(NM) * 64, 0.0), X(static_cast<size_t>(NM) * 64, 0.0);
        int kd = 0;
        for (int d = 0; d <= 2 * LM; ++d) {
            double D[64] = {0};
            if (d <= lc)
                for (int s = 0; s < nsd(d); ++s) {
                    double A[64], B[64];
                    for (int lane = 0; lane < 64; ++lane) {
                        const int a = lane >> 4, g = (lane >> 2) & 3, b = lane & 3;
                        A[lane] = TU[(kd + s) * 16 + 4 * a + b];
                        const int ls = d + 2 * (4 * s + a);
                        B[lane] = ls <= lc ? E[ls * 8 + ((4 * g + b) >> 1)] : 0.0;
                    }
                    mfma(A, B, D);
                }
            for (int lane = 0; lane < 64; ++lane) U[d * 64 + lane] = D[lane];
            kd += nsd(d);
        }
        int ks = 0;
        for (int m = -LM; m <= LM; ++m) {
            double D[64] = {0};
            for (int s = 0; s < nsm(m); ++s) {
                double A[64], B[64];
                for (int lane = 0; lane < 64; ++lane) {
                    const int a = lane >> 4, g = (lane >> 2) & 3, b = lane & 3;
                    A[lane] = TA[(ks + s) * 16 + 4 * a + b];
                    const int l = std::abs(m) + 2 * (4 * s + a);
                    B[lane] = l <= lmax ? sX[jof(l, m) * 16 + 4 * g + b] : sX[4 * g + b];
                }
                mfma(A, B, D);
            }
            for (int lane = 0; lane < 64; ++lane) X[(m + LM) * 64 + lane] = D[lane];
            ks += nsm(m);
        }
        ks = 0;
        for (int m = -LM; m <= LM; ++m) {
            const int mi = m + LM;
            double Y[64];
            for (int lane = 0; lane < 64; ++lane) {
                double y = U[lane] * X[mi * 64 + lane];
                for (int d = 1; d <= 2 * LM; ++d) {
                    if (mi + d < NM) y += U[d * 64 + lane] * X[(mi + d) * 64 + lane];
                    if (mi - d >= 0) y += U[d * 64 + lane] * X[(mi - d) * 64 + lane];
                }
                Y[lane] = y;
            }
            for (int s = 0; s < nsm(m); ++s) {
                double A[64];
                for (int lane = 0; lane < 64; ++lane) A[lane] = TC[(ks + s) * 16 + 4 * (lane >> 4) + (lane & 3)];
                mfma(A, Y, &yacc[static_cast<size_t>(ks + s) * 64]);
            }
            ks += nsm(m);
        }
    }
    // accumulator (m, s), lane (a, g, b): mode (m, 4 s + a) at column 4 g + b
    std::vector<cd> ye(K * 8, 0.0);
    int ks = 0;
    for (int m = -LM; m <= LM; ++m) {
        for (int s = 0; s < nsm(m); ++s)
            for (int lane = 0; lane < 64; ++lane) {
                const int a = lane >> 4, g = (lane >> 2) & 3, b = lane & 3;
                const int l = std::abs(m) + 2 * (4 * s + a), c = 4 * g + b;
                if (l > lmax) continue;
                const double v = yacc[static_cast<size_t>(ks + s) * 64 + lane];
                ye[jof(l, m) * 8 + (c >> 1)] += (c & 1) ? cd(0, v) : cd(v, 0);
            }
        ks += nsm(m);
    }
    double num = 0, den = 0;
    for (int i = 0; i < K * 8; ++i) {
        num += std::norm(ye[i] - yd[i]);
        den += std::norm(yd[i]);
    }
    std::printf("{\"lmax\": %d, \"class\": %d, \"lc\": %d, \"rule\": %d, \"points\": %d, \"rel_err\": %.6e}\n", lmax, LM,
                lc, ng, 4 * ng, den > 0 ? std::sqrt(num / den) : std::sqrt(num));
    return 0;
}
