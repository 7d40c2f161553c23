// gm_theta.hpp - tables of the polar-quadrature form of the multipole coupling.
//
// The coupling of the reference (src/CrankNicolsonModes_hamiltonian.cpp:68-104,
// 176-222) is, per radial node, V_jk = sum_l* E_l*(r) G_l*[j, k] with the
// Gaunt coefficient
//   G_l*[j, k] = (1 / 2 pi) p_l*^m*(0) int_-1^1 p_lj^mj p_lk^mk p_l*^m* dx,
// m* = m_k - m_j, p_l^m the associated Legendre function normalised to
// int_-1^1 p^2 dx = 1 (Condon-Shortley phase, p_l^-m = (-1)^m p_l^m). The
// integrand is a polynomial of degree l_j + l_k + l* and even in x, so a
// Gauss-Legendre rule of 2 n points (n on x > 0, weights doubled) gives it
// exactly when l_j + l_k + l* <= 4 n - 1. With U_d(x, r) = (1 / 2 pi)
// sum_l* E_l*(r) p_l*^d(0) p_l*^d(x) (l* = d, d + 2, ...), U_-d = U_d:
//   X_i[m']  = sum_{k: m_k = m'} p_k(x_i) x_k             (phase A)
//   Y_i[m]   = sum_m' U_|m - m'|(x_i) X_i[m']               (phase B)
//   (V x)_j  = sum_i w_i p_j(x_i) Y_i[m_j]                   (phase C)
// equals sum_k V_jk x_k to rounding. A tile whose multipoles stop at lc needs
// l_j + l_k + l* <= 2 l_max + lc: rule ng has n = 4 ng points, enough when
// 16 ng - 1 >= 2 l_max + lc.
//
// Phases A, C and the U sums run on v_mfma_f64_4x4x4f64 (four 4x4x4 blocks).
// Lane = 16 a + 4 g + b holds A-operand [block g][row b][k a], B-operand
// [g][k a][col b] and result [g][row a][col b]; so a result lane holds the
// B-operand position of the next product, and phase B works on one
// (quadrature point 4 t + a, column 4 g + b) per lane. Every table holds, per
// 4x4 A block, the 16 values at index 4 a + b (the blocks g share them).
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

namespace gmt {

// The blocks follow the layout of the kernel's l_max class LM >= l_max (so
// that its loops have compile-time bounds); entries of modes l > l_max and of
// multipoles l* > 2 l_max are zero.
struct ThetaTables {
    int lmax = 0, LM = 0;
    int ks = 0;     // (m, s) blocks of phases A and C: m = -LM..LM, s < ceil(n_m / 4), n_m = (LM - |m|) / 2 + 1
    int kd = 0;     // (d, s) blocks of the U sums: d = 0..2 LM, s < ceil(n_l*(d) / 4), n_l* = (2 LM - d) / 2 + 1
    int ngmax = 0;  // rules ng = 1..ngmax (4 ng points on x > 0)
    // TA[rule][t][ks][16]: p_(m, 4 s + a)(x_(4 t + b)), rule ng at ks * 16 * ng (ng - 1) / 2
    std::vector<double> A;
    // TC[rule][t][ks][16]: w_(4 t + a) p_(m, 4 s + b)(x_(4 t + a))
    std::vector<double> C;
    // TU[rule][t][kd][16]: (1 / 2 pi) p_l*^d(0) p_l*^d(x_(4 t + b)), l* = d + 2 (4 s + a), rule at kd * 16 * ng (ng - 1) / 2
    std::vector<double> U;
};

// Quadrature rule ng (4 ng points) of a tile whose multipole cut is lc: the
// smallest with 16 ng - 1 >= 2 l_max + lc.
__host__ __device__ inline int theta_rule(int lmax, int lc) { return (2 * lmax + lc + 16) / 16; }

ThetaTables build_theta_tables(int lmax, int LM);

// Gauss-Legendre rule of 2 n points on [-1, 1]: the n positive nodes
// (ascending) and their weights, in long double (Newton on P_2n).
void gauss_legendre_half(int n, std::vector<long double> &x, std::vector<long double> &w);
// p_l^m(x) for 0 <= m <= l <= L at one x: out[l * (L + 1) + m].
void assoc_legendre(int L, long double x, std::vector<long double> &out);

}  // namespace gmt
