// gm_theta_apply.hpp - the operator apply with the coupling in polar-quadrature form
// (kernel templates; gm_theta_f64.hip instantiates them).
//
// The coupling sum_k V_jk(r) x_k of the reference (hamiltonian.cpp:68-104,
// 176-222) is evaluated through a Gauss-Legendre rule in x = cos(theta)
// (gm_theta.hpp has the identity and the tables):
//   U   U_d(x_i, r) = (1/2pi) sum_l* E_l*(r) p_l*^d(0) p_l*^d(x_i)   (d <= lc)
//   A   X_i[m']     = sum_(k: m_k = m') p_k(x_i) xt_k                (xt = e^(i m theta) x)
//   B   Y_i[m]      = sum_m' U_|m - m'|(x_i) X_i[m']                 (|m - m'| <= lc)
//   C   (V x)_j     = e^(-i m_j theta) sum_i w_i p_j(x_i) Y_i[m_j]
// A workgroup is one tile of 8 radial nodes, whose 16 columns are (node,
// real or imaginary part): every table is real, so the two parts are
// transformed alike. Its four wavefronts split the output m. U, A and C run
// on v_mfma_f64_4x4x4f64 in quadrature groups of four points (four 4x4x4
// blocks over four columns each); B is a banded Toeplitz product on the VALU
// in FP64, one (point, column) per lane. Work per node: about 2 (2 l_max + lc) n_pts K real
// multiply-adds on the matrix core and n_pts (2 l_max + 1)(2 lc + 1) on the
// VALU, against the 2 K^2 + 2 K (number of Gaunt terms) of the direct sum.
// The epilogue is the Gaunt-form apply's (stencil, CN form, sinks, fused
// reductions: gm_devfn.hpp).
#pragma once
#include "gm_devfn.hpp"
#include "gm_theta.hpp"

#include <type_traits>

namespace gmk {

namespace {

constexpr int TT = 8;  // radial nodes of a wavefront's tile
constexpr int TW = 4;  // wavefronts of a tile's workgroup (each takes a quarter of the output m)
// epilogue elements per thread with their loads in flight together (4: the
// apply 1.5 % faster than 2 at l_max 20, N 16384, still two waves per SIMD)
#ifndef GM_TEB
#define GM_TEB 4
#endif
#ifndef GM_TWAVES
#define GM_TWAVES 2
#endif

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

// Block layout of the U sums at class LM: d = 0 .. 2 LM, ceil(n_l*(d) / 4)
// blocks each (n_l* = (2 LM - d) / 2 + 1 multipoles l* = d, d + 2, ...).
template <int LM>
struct DLay {
    int ko[2 * LM + 2];  // first block of d
    int ns[2 * LM + 1];  // blocks of d
    constexpr DLay() : ko(), ns()
    {
        int o = 0;
        for (int d = 0; d <= 2 * LM; ++d) {
            ko[d] = o;
            ns[d] = ((2 * LM - d) / 2 + 1 + 3) / 4;
            o += ns[d];
        }
        ko[2 * LM + 1] = o;
    }
};

__device__ __forceinline__ double mf(double a, double b, double c)
{
    return __builtin_amdgcn_mfma_f64_4x4x4f64(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ int iabs(int v) { return v < 0 ? -v : v; }

// f(integral_constant<I>) for I = B .. E - 1 (unrolled at compile time)
template <int I, int E, class F>
__device__ __forceinline__ void static_for(F &&f)
{
    if constexpr (I < E) {
        f(std::integral_constant<int, I>{});
        static_for<I + 1, E>(f);
    }
}
// the same while f returns true
template <int I, int E, class F>
__device__ __forceinline__ void static_while(F &&f)
{
    if constexpr (I < E) {
        if (f(std::integral_constant<int, I>{})) static_while<I + 1, E>(f);
    }
}

}  // namespace

// Doubles per tile of the U-sum table at class LM: every point group of the
// largest rule and every diagonal d <= 2 LM, 4 points x TT nodes each.
template <int LM>
__host__ __device__ constexpr size_t ut_stride(int lmax)
{
    return (size_t)gmt::theta_rule(lmax, 2 * lmax) * (2 * LM + 1) * 4 * TT;
}

// The U sums of one tile: U_d(x_i, r) = sum_l* E_l*(r) TU[i][d][l*] for the
// tile's rule and every point group, d <= lc (zero above), on the matrix
// core, into Ut. They depend on the orbit's radial profiles and the cut only,
// so they are rebuilt with those, not per apply. Wavefront W of the tile's
// workgroup takes d = W, W + TW, ...
template <int LM, int W>
__device__ __forceinline__ void tapply_u_wave(const gmd::Geometry &g, const double *sE, double *out, int lc, int ng,
                                              int lane)
{
    constexpr DLay<LM> dl{};
    constexpr int KD = dl.ko[2 * LM + 1];
    const int a4 = lane >> 4, b4 = lane & 3;
    const int la = 4 * a4 + b4;
    const int col = 4 * ((lane >> 2) & 3) + b4;
    const int nd = col >> 1;
    const size_t tri = (size_t)ng * (ng - 1) / 2;
    for (int t = 0; t < ng; ++t) {
        const double *tu = g.TU + (tri + t) * (size_t)KD * 16 + la;
        static_for<0, (2 * LM + 1 - W + TW - 1) / TW>([&](auto kc) {
            constexpr int d = W + TW * kc();
            double acc = 0.0;
            if (d <= lc) {
                static_for<0, dl.ns[d]>([&](auto sc) {
                    constexpr int s = sc();
                    const int ls = d + 2 * (4 * s + a4);
                    const double e = sE[(ls < 2 * LM ? ls : 2 * LM) * TT + nd];
                    acc = mf(tu[(dl.ko[d] + s) * 16], ls <= lc ? e : 0.0, acc);
                });
            }
            // the two parts of a node hold the same value
            if ((col & 1) == 0) out[(t * (2 * LM + 1) + d) * 4 * TT] = acc;
        });
    }
}

template <int LM>
__device__ __forceinline__ void tapply_u_body(const gmd::Geometry g, double *Ut, int tile)
{
    __shared__ double sE[(2 * LM + 1) * TT];
    const int lane = threadIdx.x & 63, wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int N = g.N, lmax = g.lmax, Ls = g.Ls;
    const int t0 = tile * TT;
    const int lc = g.lcut ? g.lcut[t0 / gmd::TILE] : 2 * lmax;
    for (int e = threadIdx.x; e < (2 * LM + 1) * TT; e += 64 * TW) {
        const int ls = e >> 3, i = t0 + (e & 7);
        sE[e] = (ls <= lc && ls <= 2 * lmax && i < N) ? g.E[(size_t)((ls & 1) * Ls + ls) * N + i] : 0.0;
    }
    __syncthreads();
    const int ng = gmt::theta_rule(lmax, lc);
    const int a4 = lane >> 4, nd = (4 * ((lane >> 2) & 3) + (lane & 3)) >> 1;
    double *out = Ut + (size_t)tile * ut_stride<LM>(lmax) + a4 * TT + nd;
    static_assert(TW == 4, "four wavefronts per tile");
    if (wv == 0) tapply_u_wave<LM, 0>(g, sE, out, lc, ng, lane);
    else if (wv == 1) tapply_u_wave<LM, 1>(g, sE, out, lc, ng, lane);
    else if (wv == 2) tapply_u_wave<LM, 2>(g, sE, out, lc, ng, lane);
    else tapply_u_wave<LM, 3>(g, sE, out, lc, ng, lane);
}

template <int LM>
__global__ __launch_bounds__(64 * TW) void k_tapply_u(gmd::Geometry g, double *Ut)
{
    tapply_u_body<LM>(g, Ut, blockIdx.x);
}

// Output m of wavefront W, interleaved: m + LM = W, W + TW, W + 2 TW, ...
// (every wavefront takes near and far m alike, so the banded product of
// phase B, whose cost per m depends on the cut, stays balanced); n m, their
// m + LM and the first of their accumulators in the wavefront's local
// layout.
template <int LM, int W>
struct WList {
    static constexpr int n = (2 * LM + 1 - W + TW - 1) / TW;
    int mi[n > 0 ? n : 1];
    int ko[n + 1];
    constexpr WList() : mi(), ko()
    {
        int o = 0;
        for (int k = 0; k < n; ++k) {
            mi[k] = W + TW * k;
            ko[k] = o;
            o += ns_of(LM, mi[k] - LM);
        }
        ko[n] = o;
    }
};

// LDS of one workgroup (one tile of TT nodes).
template <int LM>
struct TShared {
    static constexpr int KMAX = (LM + 1) * (LM + 2) / 2;
    double x[KMAX * 16];               // x of the tile [mode][node][re, im]
    double2 ph[2 * LM + 1];            // e^(i m theta), m = -LM .. LM (zero above l_max)
    union {
        struct {
            double X[(2 * LM + 1) * 64];  // X[m'] of the current point group, per lane
            double A[kcount(LM) * 16];    // A operands of phase A, then of phase C
        };
        double C[KMAX * 16];           // after the point groups: (V xt) [mode][node][re, im]
    };
};

// The work of wavefront W: phase A for its m', phase B for its m (all X from
// LDS, the U sums from their table), phase C for its m, then its
// accumulators into the coupling tile. (Interleaving the m of the four
// wavefronts instead of giving each a contiguous range made the apply 3 %
// faster at l_max 20, N 16384.)
template <int LM, int W>
__device__ __forceinline__ void tapply_wave(const ApplyArgs &a, TShared<LM> &sh, int lc, int ng, int lane, int tile)
{
    constexpr int NM = 2 * LM + 1;
    constexpr TLay<LM> lay{};
    constexpr int KL = kcount(LM);
    constexpr DLay<LM> dl{};
    constexpr int KD = dl.ko[2 * LM + 1];
    constexpr WList<LM, W> wl{};
    constexpr int NW = WList<LM, W>::n, NACC = wl.ko[NW];
    const gmd::Geometry &g = a.g;
    const int a4 = lane >> 4, b4 = lane & 3;
    const int la = 4 * a4 + b4;                  // index in a table's 4x4 block
    const int col = 4 * ((lane >> 2) & 3) + b4;  // column: node col >> 1, part col & 1
    const int nd = col >> 1, part = col & 1;
    const int lmax = g.lmax;
    const size_t tri = (size_t)ng * (ng - 1) / 2;
    // U_d of this tile at (point 4 t + a4, node nd): Ut[((t (2 LM + 1) + d) 4 + a4) TT + nd]
    const double *ut = g.Ut + (size_t)tile * ut_stride<LM>(g.lmax) + a4 * TT + nd;
    double yacc[NACC > 0 ? NACC : 1];
    static_for<0, NACC>([&](auto e) { yacc[e()] = 0.0; });
    // Software pipeline: a table's global loads are issued a phase before its
    // copy into LDS, and a diagonal group's U sums a group before their use.
    constexpr int NT = (KL * 8 + 64 * TW - 1) / (64 * TW);
    double2 rt[NT];
    const auto load_table = [&](const double *base) {
        static_for<0, NT>([&](auto k) {
            const int e = threadIdx.x + 64 * TW * k();
            rt[k()] = e < KL * 8 ? ((const double2 *)base)[e] : make_double2(0.0, 0.0);
        });
    };
    const auto store_table = [&]() {
        static_for<0, NT>([&](auto k) {
            const int e = threadIdx.x + 64 * TW * k();
            if (e < KL * 8) ((double2 *)sh.A)[e] = rt[k()];
        });
    };
    double un[4];
    const auto load_u = [&](int t, auto d0c) {
        constexpr int d0 = d0c();
        static_for<0, 4>([&](auto qc) {
            constexpr int d = d0 + qc();
            if constexpr (d <= 2 * LM) un[qc()] = ut[(t * (2 * LM + 1) + d) * 4 * TT];
            else un[qc()] = 0.0;
        });
    };
    load_table(g.TA + tri * (size_t)KL * 16);
    for (int t = 0; t < ng; ++t) {
        __syncthreads();
        store_table();  // TA of point group t
        load_u(t, std::integral_constant<int, 0>{});
        load_table(g.TC + (tri + t) * (size_t)KL * 16);
        __syncthreads();
        // A: X[m'] at (point 4 t + a4, column col) for this wavefront's m'
        static_for<0, NW>([&](auto kc) {
            constexpr int mi = wl.mi[kc()], m = mi - LM, am = m < 0 ? -m : m;
            constexpr int nb = lay.ko[mi + 1] - lay.ko[mi];
            // the column's part of e^(i m theta) x: c x_re - s x_im or s x_re + c x_im
            const double2 ph = sh.ph[mi];
            const double pr = part ? ph.y : ph.x, pim = part ? ph.x : -ph.y;
            double av[nb], bv[nb];
            static_for<0, nb>([&](auto sc) {
                constexpr int s = sc();
                const int l = am + 2 * (4 * s + a4);
                const int j = l <= lmax ? l * (l + 1) / 2 + (m + l) / 2 : 0;  // zero in the table above l_max
                av[s] = sh.A[(lay.ko[mi] + s) * 16 + la];
                bv[s] = pr * sh.x[j * 16 + 2 * nd] + pim * sh.x[j * 16 + 2 * nd + 1];
            });
            double acc = 0.0;
            static_for<0, nb>([&](auto sc) { acc = mf(av[sc()], bv[sc()], acc); });
            sh.X[mi * 64 + lane] = acc;
        });
        __syncthreads();
        store_table();  // TC of point group t
        if (t + 1 < ng) load_table(g.TA + (tri + t + 1) * (size_t)KL * 16);
        double X[NM];
        static_for<0, NM>([&](auto mic) { X[mic()] = sh.X[mic() * 64 + lane]; });
        // B: Y[m] = sum_d U_d (X[m + d] + X[m - d]) for this wavefront's m, the
        // diagonals in groups of four (the table holds zeros above the cut)
        double Y[NW > 0 ? NW : 1];
        static_for<0, NW>([&](auto kc) { Y[kc()] = 0.0; });
        static_while<0, (2 * LM + 4) / 4>([&](auto gc) {
            constexpr int d0 = 4 * gc();
            if (d0 > lc) return false;
            double u[4];
            static_for<0, 4>([&](auto qc) { u[qc()] = un[qc()]; });
            if constexpr (d0 + 4 <= 2 * LM)
                if (d0 + 4 <= lc) load_u(t, std::integral_constant<int, d0 + 4>{});
            static_for<0, 4>([&](auto qc) {
                constexpr int d = d0 + qc();
                if constexpr (d <= 2 * LM) {
                    static_for<0, NW>([&](auto kc) {
                        constexpr int k = kc(), mi = wl.mi[k];
                        if constexpr (d == 0) {
                            Y[k] = u[0] * X[mi];
                        } else {
                            if constexpr (mi + d < NM) Y[k] = fma(u[qc()], X[mi + d], Y[k]);
                            if constexpr (mi - d >= 0) Y[k] = fma(u[qc()], X[mi - d], Y[k]);
                        }
                    });
                }
            });
            return true;
        });
        __syncthreads();
        // C: the blocks of (V x) of this wavefront's m on the matrix core
        static_for<0, NW>([&](auto kc) {
            constexpr int k = kc(), mi = wl.mi[k];
            constexpr int nb = lay.ko[mi + 1] - lay.ko[mi];
            const double yd = Y[k];
            double av[nb];
            static_for<0, nb>([&](auto sc) { av[sc()] = sh.A[(lay.ko[mi] + sc()) * 16 + la]; });
            static_for<0, nb>([&](auto sc) {
                constexpr int e = wl.ko[k] + sc();
                yacc[e] = mf(av[sc()], yd, yacc[e]);
            });
        });
    }
    // accumulator e: (V xt) of mode (m_e, 4 s_e + a4) at column col, into
    // the coupling tile over X and the tables (read for the last time before
    // the barrier)
    __syncthreads();
    static_for<0, NW>([&](auto kc) {
        constexpr int k = kc(), mi = wl.mi[k], m = mi - LM, am = m < 0 ? -m : m;
        static_for<0, wl.ko[k + 1] - wl.ko[k]>([&](auto sc) {
            constexpr int s = sc();
            const int l = am + 2 * (4 * s + a4);
            if (l <= lmax) sh.C[(l * (l + 1) / 2 + (m + l) / 2) * 16 + col] = yacc[wl.ko[k] + s];
        });
    });
}

template <int LM>
__device__ __forceinline__ void k_tapply_body(const ApplyArgs &a, Blk B)
{
    constexpr int KMAX = (LM + 1) * (LM + 2) / 2;
    __shared__ TShared<LM> sh;
    const gmd::Geometry &g = a.g;
    gmd::Krylov *Kr = a.Kr;
    if ((a.mode == AM_V || a.mode == AM_T) && Kr->done) return;
    const gmd::StepParams *P = a.P;
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    const int N = g.N, lmax = g.lmax, K = g.K;
    const int t0 = B.id * TT;
    const int lc = g.lcut ? g.lcut[t0 / gmd::TILE] : 2 * lmax;
    // the tile's x, all loads in flight before the first use (element e:
    // mode e >> 3, node e & 7; rows above K and nodes past N load node t0
    // and store zeros)
    static_assert(sizeof(sh.C) <= sizeof(sh.X) + sizeof(sh.A), "the coupling tile fits over X and A");
    for (int e = threadIdx.x; e < 2 * LM + 1; e += 64 * TW) {
        const int m = e - LM;
        sh.ph[e] = (m >= -lmax && m <= lmax) ? P->ph[m + lmax] : make_double2(0.0, 0.0);
    }
    {
        constexpr int NE = (KMAX * TT + 64 * TW - 1) / (64 * TW);
        double2 v[NE];
        static_for<0, NE>([&](auto ec) {
            const int e = threadIdx.x + 64 * TW * ec(), j = e >> 3, i = t0 + (e & 7);
            v[ec()] = a.X[(size_t)(j < K ? j : 0) * N + (i < N ? i : t0)];
        });
        static_for<0, NE>([&](auto ec) {
            const int e = threadIdx.x + 64 * TW * ec(), j = e >> 3, n = e & 7;
            if (j < K) {
                const double2 w = t0 + n < N ? v[ec()] : make_double2(0.0, 0.0);
                sh.x[j * 16 + 2 * n] = w.x;
                sh.x[j * 16 + 2 * n + 1] = w.y;
            }
        });
    }
    const int ng = gmt::theta_rule(lmax, lc);
    // Every register array is indexed by template constants only (static_for,
    // one code path per wavefront); the point-group tables are shared in LDS.
    static_assert(TW == 4, "four wavefronts per tile");
    if (wv == 0) tapply_wave<LM, 0>(a, sh, lc, ng, lane, B.id);
    else if (wv == 1) tapply_wave<LM, 1>(a, sh, lc, ng, lane, B.id);
    else if (wv == 2) tapply_wave<LM, 2>(a, sh, lc, ng, lane, B.id);
    else tapply_wave<LM, 3>(a, sh, lc, ng, lane, B.id);
    __syncthreads();
    // Epilogue over the tile's (mode, node), eight consecutive nodes of a mode
    // per eight threads, in batches of EB whose loads are issued before their
    // arithmetic and stores: C = e^(-i m theta) (V xt), then the stencil, CN
    // form, sinks and the fused reductions.
    const ApplyEpi ep = apply_epi(a);
    double red[3] = {0.0, 0.0, 0.0};
    // every element of a thread lies at node t0 + (threadIdx.x & 7)
    const int ie = t0 + (threadIdx.x & (TT - 1));
    const double cr = ie < N ? centrifugal(g, ie) : 0.0;
    constexpr int EB = GM_TEB;
    for (int e0 = threadIdx.x; e0 < K * TT; e0 += 64 * TW * EB) {
        ElemIn in[EB];
        static_for<0, EB>([&](auto bc) {
            const int e = e0 + 64 * TW * bc(), j = e < K * TT ? e >> 3 : 0, n = e & 7;
            const int i = t0 + n < N ? t0 + n : t0;
            const double *xr = sh.x + j * 16;
            // x and its interior neighbours from the tile, the neighbours
            // across the tile's edges (n = 0, TT - 1) from memory
            in[bc()] = apply_element_load_r0(a, j, i);
            in[bc()].xi = make_double2(xr[2 * n], xr[2 * n + 1]);
            in[bc()].xm = n > 0 ? make_double2(xr[2 * n - 2], xr[2 * n - 1]) : a.X[(size_t)j * N + (i > 0 ? i - 1 : 0)];
            in[bc()].xp = n < TT - 1 ? make_double2(xr[2 * n + 2], xr[2 * n + 3])
                                     : a.X[(size_t)j * N + (i < N - 1 ? i + 1 : N - 1)];
        });
        static_for<0, EB>([&](auto bc) {
            const int e = e0 + 64 * TW * bc(), j = e >> 3, n = e & 7, i = t0 + n;
            if (e >= K * TT || i >= N) return;
            // l from j = l (l + 1) / 2 + (m + l) / 2 (single-precision root, corrected)
            int l = (int)((sqrtf(8.0f * j + 1.0f) - 1.0f) * 0.5f);
            if ((l + 1) * (l + 2) / 2 <= j) ++l;
            if (l * (l + 1) / 2 > j) --l;
            const int m = 2 * (j - l * (l + 1) / 2) - l;
            const double2 Cv = make_double2(sh.C[j * 16 + 2 * n], sh.C[j * 16 + 2 * n + 1]);
            apply_element(a, ep, j, l, i, cdotm(sh.ph[m + LM], Cv), in[bc()], cr, red);
        });
    }
    apply_reduce(a, red, B);
}

template <int LM>
__global__ __launch_bounds__(64 * TW, GM_TWAVES) void k_tapply(ApplyArgs a)
{
    k_tapply_body<LM>(a, grid_blk());
}
// One launch over many solvers: member m owns blocks [off[m], off[m + 1]).
template <int LM>
__global__ __launch_bounds__(64 * TW, GM_TWAVES) void k_tapply_batch(const ApplyArgs *__restrict__ args, const int *__restrict__ off,
                                                     int M)
{
    const int m = batch_member(off, M);
    const ApplyArgs a = args[m];
    k_tapply_body<LM>(a, Blk{(int)blockIdx.x - __builtin_amdgcn_readfirstlane(off[m]),
                              __builtin_amdgcn_readfirstlane(off[m + 1] - off[m])});
}

template <int LM>
hipError_t launch_t(const ApplyArgs *a, const ApplyArgs *args, const int *off, int M, int total, hipStream_t st)
{
    if (args)
        hipLaunchKernelGGL((k_tapply_batch<LM>), dim3(total), dim3(64 * TW), 0, st, args, off, M);
    else
        hipLaunchKernelGGL((k_tapply<LM>), dim3(tapply_grid(a->g.N)), dim3(64 * TW), 0, st, *a);
    return hipGetLastError();
}

template <int LM>
hipError_t launch_u(const gmd::Geometry &g, double *Ut, hipStream_t st)
{
    hipLaunchKernelGGL((k_tapply_u<LM>), dim3((g.N + TT - 1) / TT), dim3(64 * TW), 0, st, g, Ut);
    return hipGetLastError();
}

// Launch of the l_max class cls (gm_theta_f64.hip).
hipError_t launch_tt(int cls, const ApplyArgs *a, const ApplyArgs *args, const int *off, int M, int total,
                     hipStream_t st);

}  // namespace gmk
