// gm_theta.hip - the operator apply with the coupling in polar-quadrature form.
//
// The coupling sum_k V_jk(r) x_k of the reference (hamiltonian.cpp:68-104,
// 176-222) is evaluated through a Gauss-Legendre rule in x = cos(theta)
// (gm_theta.hpp has the identity and the tables):
//   U   U_d(x_i, r) = (1/2pi) sum_l* E_l*(r) p_l*^d(0) p_l*^d(x_i)   (d <= lc)
//   A   X_i[m']     = sum_(k: m_k = m') p_k(x_i) xt_k                (xt = e^(i m theta) x)
//   B   Y_i[m]      = sum_m' U_|m - m'|(x_i) X_i[m']                 (|m - m'| <= lc)
//   C   (V x)_j     = e^(-i m_j theta) sum_i w_i p_j(x_i) Y_i[m_j]
// A workgroup is one wavefront and one tile of 8 radial nodes, whose 16
// columns are (node, real or imaginary part): every table is real, so the two
// parts are transformed alike. U, A and C run on v_mfma_f64_4x4x4f64 in
// quadrature groups of four points (four 4x4x4 blocks over four columns
// each); B is a banded Toeplitz product on the VALU, one (point, column) per
// lane, in the apply's arithmetic (FP64, or FP32 for the mixed-precision
// Krylov applies). Work per node: about 2 (2 l_max + lc) n_pts K real
// multiply-adds on the matrix core and n_pts (2 l_max + 1)(2 lc + 1) on the
// VALU, against the 2 K^2 + 2 K (number of Gaunt terms) of the direct sum.
// The epilogue is the Gaunt-form apply's (stencil, CN form, sinks, fused
// reductions: gm_devfn.hpp).
#include "gm_devfn.hpp"
#include "gm_theta.hpp"

#include <type_traits>

namespace gmk {

namespace {

constexpr int TT = 8;  // radial nodes of a tile

// Accumulator layout at l_max = LM: m = -LM .. LM in order, ceil(n_m / 4)
// blocks each (n_m = (LM - |m|) / 2 + 1 modes); a compile-time table, so that
// the indices of the register arrays fold to constants.
__host__ __device__ constexpr int ns_of(int LM, int m)
{
    return ((LM - (m < 0 ? -m : m)) / 2 + 1 + 3) / 4;
}
__host__ __device__ constexpr int kcount(int LM)
{
    int o = 0;
    for (int m = -LM; m <= LM; ++m) o += ns_of(LM, m);
    return o;
}
template <int LM>
struct TLay {
    int ko[2 * LM + 2];    // first accumulator of m = i - LM
    int me[kcount(LM)];    // m of accumulator e
    int se[kcount(LM)];    // block s of accumulator e
    constexpr TLay() : ko(), me(), se()
    {
        int o = 0;
        for (int i = 0; i <= 2 * LM; ++i) {
            ko[i] = o;
            for (int s = 0; s < ns_of(LM, i - LM); ++s, ++o) {
                me[o] = i - LM;
                se[o] = s;
            }
        }
        ko[2 * LM + 1] = o;
    }
};

__device__ __forceinline__ double mf(double a, double b, double c)
{
    return __builtin_amdgcn_mfma_f64_4x4x4f64(a, b, c, 0, 0, 0);
}

// the value of lane ^ 1 (quad_perm [1, 0, 3, 2])
__device__ __forceinline__ double swap1(double v)
{
    const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), 0xB1, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), 0xB1, 0xF, 0xF, false);
    return __hiloint2double(hi, lo);
}

__device__ __forceinline__ int iabs(int v) { return v < 0 ? -v : v; }

}  // namespace

template <int LM, typename T>
__device__ __forceinline__ void k_tapply_body(const ApplyArgs &a, Blk B)
{
    constexpr int NM = 2 * LM + 1;
    constexpr int KL = kcount(LM);
    constexpr TLay<LM> lay{};
    constexpr int KMAX = (LM + 1) * (LM + 2) / 2;
    __shared__ double sX[KMAX * 16];        // phased x of the tile: [mode][column]
    __shared__ double sE[(2 * LM + 1) * TT];  // profile E_l* of the tile's nodes (0 above the cut)
    const gmd::Geometry &g = a.g;
    gmd::Krylov *Kr = a.Kr;
    if ((a.mode == AM_V || a.mode == AM_T) && Kr->done) return;
    const gmd::StepParams *P = a.P;
    const int lane = threadIdx.x;
    const int a4 = lane >> 4, b4 = lane & 3;
    const int la = 4 * a4 + b4;                       // index in a table's 4x4 block
    const int col = 4 * ((lane >> 2) & 3) + b4;       // column: node col >> 1, part col & 1
    const int nd = col >> 1, part = col & 1;
    const int N = g.N, lmax = g.lmax, K = g.K, Ls = g.Ls;
    const int t0 = B.id * TT;
    const int lc = g.lcut ? g.lcut[t0 / gmd::TILE] : 2 * lmax;
    for (int e = lane; e < K * TT; e += 64) {
        const int j = e >> 3, n = e & 7, i = t0 + n;
        double2 v = make_double2(0.0, 0.0);
        if (i < N) v = cmul(P->ph[mode_m(j) + lmax], a.X[(size_t)j * N + i]);
        sX[j * 16 + 2 * n] = v.x;
        sX[j * 16 + 2 * n + 1] = v.y;
    }
    for (int e = lane; e < (2 * lmax + 1) * TT; e += 64) {
        const int ls = e >> 3, i = t0 + (e & 7);
        sE[e] = (ls <= lc && i < N) ? g.E[(size_t)((ls & 1) * Ls + ls) * N + i] : 0.0;
    }
    __syncthreads();
    const int ng = gmt::theta_rule(lmax, lc);
    const size_t tri = (size_t)ng * (ng - 1) / 2;
    const double *TA = g.TA + tri * g.tks * 16;
    const double *TC = g.TC + tri * g.tks * 16;
    const double *TU = g.TU + tri * g.tkd * 16;
    double yacc[KL];
#pragma unroll
    for (int e = 0; e < KL; ++e) yacc[e] = 0.0;
    for (int t = 0; t < ng; ++t) {
        // U: U_d at (point 4 t + a4, node nd) for d <= lc
        T U[2 * LM + 1];
#pragma unroll
        for (int d = 0; d <= 2 * LM; ++d) U[d] = T(0);
        {
            const double *tu = TU + (size_t)t * g.tkd * 16 + la;
            int kd = 0;
#pragma unroll
            for (int d = 0; d <= 2 * LM; ++d) {
                if (d > lc) break;
                const int nsc = ((lc - d) / 2 + 1 + 3) / 4;
                double acc = 0.0;
                for (int s = 0; s < nsc; ++s) {
                    const int ls = d + 2 * (4 * s + a4);
                    acc = mf(tu[(size_t)(kd + s) * 16], ls <= lc ? sE[ls * TT + nd] : 0.0, acc);
                }
                U[d] = T(acc);
                kd += ((2 * lmax - d) / 2 + 1 + 3) / 4;
            }
        }
        // A: X[m'] at (point 4 t + a4, column col)
        T X[NM];
        {
            const double *ta = TA + (size_t)t * g.tks * 16 + la;
            int ks = 0;
#pragma unroll
            for (int mi = 0; mi < NM; ++mi) {
                const int m = mi - LM;
                X[mi] = T(0);
                if (iabs(m) > lmax) continue;
                const int nm = (lmax - iabs(m)) / 2 + 1, ns = (nm + 3) / 4;
                double acc = 0.0;
                for (int s = 0; s < ns; ++s) {
                    int l = iabs(m) + 2 * (4 * s + a4);
                    if (l > lmax) l = iabs(m);  // zero in the table: any finite row
                    acc = mf(ta[(size_t)(ks + s) * 16], sX[(l * (l + 1) / 2 + (m + l) / 2) * 16 + col], acc);
                }
                X[mi] = T(acc);
                ks += ns;
            }
        }
        // B then C: Y[m] on the VALU, then its blocks of (V x) on the matrix core
        {
            const double *tc = TC + (size_t)t * g.tks * 16 + la;
            int ks = 0;
#pragma unroll
            for (int mi = 0; mi < NM; ++mi) {
                const int m = mi - LM;
                if (iabs(m) > lmax) continue;
                T y = U[0] * X[mi];
#pragma unroll
                for (int d = 1; d <= 2 * LM; d += 2) {
                    if (d > lc) break;
                    // U[d + 1] is zero above the cut
                    T xs = T(0), xs2 = T(0);
                    if (mi + d < NM) xs = X[mi + d];
                    if (mi - d >= 0) xs = xs + X[mi - d];
                    y = fma(U[d], xs, y);
                    if (d + 1 <= 2 * LM) {
                        if (mi + d + 1 < NM) xs2 = X[mi + d + 1];
                        if (mi - d - 1 >= 0) xs2 = xs2 + X[mi - d - 1];
                        y = fma(U[d + 1], xs2, y);
                    }
                }
                const double yd = (double)y;
                const int ns = ((lmax - iabs(m)) / 2 + 1 + 3) / 4;
#pragma unroll
                for (int s = 0; s < lay.ko[mi + 1] - lay.ko[mi]; ++s) {
                    if (s >= ns) break;
                    yacc[lay.ko[mi] + s] = mf(tc[(size_t)(ks + s) * 16], yd, yacc[lay.ko[mi] + s]);
                }
                ks += ns;
            }
        }
    }
    // Epilogue: accumulator e holds (V x) of mode (m_e, 4 s_e + a4) at column
    // col. Lanes col and col ^ 1 hold the real and imaginary parts of the same
    // element; part 0 takes the element of e and part 1 that of e + 1.
    const ApplyEpi ep = apply_epi(a);
    double red[3] = {0.0, 0.0, 0.0};
    const int i = t0 + nd;
#pragma unroll
    for (int e = 0; e < KL; e += 2) {
        const double v0 = yacc[e];
        const double v1 = e + 1 < KL ? yacc[e + 1] : 0.0;
        const double recv = swap1(part ? v0 : v1);
        const double2 Cv = part ? make_double2(recv, v1) : make_double2(v0, recv);
        const int e1 = e + 1 < KL ? e + 1 : e;
        const int m = part ? lay.me[e1] : lay.me[e];
        const int s = part ? lay.se[e1] : lay.se[e];
        if (part && e + 1 >= KL) continue;
        if (iabs(m) > lmax || i >= N) continue;
        if (4 * s + a4 > (lmax - iabs(m)) / 2) continue;
        const int l = iabs(m) + 2 * (4 * s + a4);
        const int j = l * (l + 1) / 2 + (m + l) / 2;
        apply_element(a, ep, j, l, i, cdotm(P->ph[m + lmax], Cv), red);
    }
    apply_reduce(a, red, B);
}

template <int LM, typename T>
__global__ __launch_bounds__(64) void k_tapply(ApplyArgs a)
{
    k_tapply_body<LM, T>(a, grid_blk());
}
// One launch over many solvers: member m owns blocks [off[m], off[m + 1]).
template <int LM, typename T>
__global__ __launch_bounds__(64) void k_tapply_batch(const ApplyArgs *__restrict__ args, const int *__restrict__ off,
                                                     int M)
{
    const int m = batch_member(off, M);
    const ApplyArgs a = args[m];
    k_tapply_body<LM, T>(a, Blk{(int)blockIdx.x - __builtin_amdgcn_readfirstlane(off[m]),
                              __builtin_amdgcn_readfirstlane(off[m + 1] - off[m])});
}

int tapply_class(int lmax, int mixed)
{
    const int maxlm = mixed ? 32 : 24;
    if (lmax < 0 || lmax > maxlm) return 0;
    return lmax <= 4 ? 4 : (lmax + 3) / 4 * 4;
}

int tapply_grid(int N) { return (N + TT - 1) / TT; }

template <int LM, typename T>
static hipError_t launch_t(const ApplyArgs *a, const ApplyArgs *args, const int *off, int M, int total, hipStream_t st)
{
    if (args)
        hipLaunchKernelGGL((k_tapply_batch<LM, T>), dim3(total), dim3(64), 0, st, args, off, M);
    else
        hipLaunchKernelGGL((k_tapply<LM, T>), dim3(tapply_grid(a->g.N)), dim3(64), 0, st, *a);
    return hipGetLastError();
}

template <typename T>
static hipError_t launch_tt(int cls, const ApplyArgs *a, const ApplyArgs *args, const int *off, int M, int total,
                            hipStream_t st)
{
    switch (cls) {
    case 4: return launch_t<4, T>(a, args, off, M, total, st);
    case 8: return launch_t<8, T>(a, args, off, M, total, st);
    case 12: return launch_t<12, T>(a, args, off, M, total, st);
    case 16: return launch_t<16, T>(a, args, off, M, total, st);
    case 20: return launch_t<20, T>(a, args, off, M, total, st);
    case 24: return launch_t<24, T>(a, args, off, M, total, st);
    default: break;
    }
    if constexpr (std::is_same<T, float>::value) {
        if (cls == 28) return launch_t<28, T>(a, args, off, M, total, st);
        if (cls == 32) return launch_t<32, T>(a, args, off, M, total, st);
    }
    return hipErrorInvalidValue;
}

hipError_t launch_tapply(const ApplyArgs &a, hipStream_t st)
{
    const int cls = tapply_class(a.g.lmax, a.mixed);
    if (!cls || !a.g.TA) return hipErrorInvalidValue;
    return a.mixed ? launch_tt<float>(cls, &a, nullptr, nullptr, 0, 0, st)
                   : launch_tt<double>(cls, &a, nullptr, nullptr, 0, 0, st);
}

hipError_t launch_tapply_batch(const ApplyArgs *args, const int *off, int M, int total, int cls, int mixed,
                               hipStream_t st)
{
    return mixed ? launch_tt<float>(cls, nullptr, args, off, M, total, st)
                 : launch_tt<double>(cls, nullptr, args, off, M, total, st);
}

}  // namespace gmk
