// gm_kernels.hip - gfx950 kernels of the Crank-Nicolson step.
//
// Kernels (all FP64, complex values as double2):
//   k_build_E   radial multipole profiles (src/CrankNicolsonModes_hamiltonian.cpp:12-42,179-185)
//   k_apply     H, A = 1 + i dt/2 H and their fused reductions: three-point
//               radial stencil plus the Gaunt coupling formed per node from the
//               E and G tables (hamiltonian.cpp:44-66,176-243, matrices.cpp:27-40)
//   k_factor    Thomas pivots of each mode's radial tridiagonal
//               (include/RadialBlockPreconditioner.hpp:43-76)
//   k_psweep    forward/backward substitution as block scans of affine maps
//               over segments of each mode's row, fused with the BiCGSTAB
//               p / s updates and the Woodbury sink correction
//               (include/RadialBlockPreconditioner.hpp:79-97, matrices.cpp:128-157)
//               (and the Woodbury columns P^-1 v_a)
//   k_update    x += alpha y + w z, r = s - w t with fused dot products and the
//               loop control of matrices.cpp:65-89 / Eigen BiCGSTAB
//   k_sink      projections v_a^H x over the absorber supports (absorption.cpp:100-125)
//   k_diag_final diagnostics of extraction.cpp:8-129
//
// Reductions are deterministic: every block writes one partial per quantity
// and the last block to finish sums the partials in block order.
#include "gm_kernels.hpp"
// device helpers (complex arithmetic, block ranges, reductions, apply elements)
#include "gm_devfn.hpp"

#include <cmath>
#include <mutex>
#include <type_traits>

namespace gmk {

using gmd::MBlock;
using gmd::MChunk;
using gmd::Geometry;
using gmd::Krylov;
using gmd::SinkDev;
using gmd::StepParams;


// ---------------------------------------------------------------- E table

__global__ void k_build_E(int N, int Ls, double dr, double eps, double mu, const StepParams *P, double *E, float *E32)
{
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= Ls * N) return;
    const double R1 = P->R1, R2 = P->R2, q = P->q, Mt = P->Mt, prf = P->prf, w1 = P->w1;
    const int ls = idx / N, i = idx - ls * N;
    double e0 = 0.0, e1 = 0.0;
    if (i > 0 && i < N - 1) {
        const double r = eps + i * dr;
        double f1, f2;
        if (r >= R1) f1 = 1 / (1 + q) * pow(R1, (double)ls) / pow(r, (double)(ls + 1));
        else f1 = 1 / (1 + q) * pow(r, (double)ls) / pow(R1, (double)(ls + 1));
        if (r >= R2) f2 = q / (1 + q) * pow(R2, (double)ls) / pow(r, (double)(ls + 1));
        else f2 = q / (1 + q) * pow(r, (double)ls) / pow(R2, (double)(ls + 1));
        const double c = -mu * Mt * 4.0 * M_PI / (2.0 * ls + 1.0);
        e0 = c * (w1 * f1 + prf * f2);
        e1 = c * (w1 * f1 - prf * f2);
    }
    E[(size_t)ls * N + i] = e0;
    E[(size_t)(Ls + ls) * N + i] = e1;
    E32[(size_t)ls * N + i] = (float)e0;
    E32[(size_t)(Ls + ls) * N + i] = (float)e1;
}

// Largest multipole l* of each node tile whose profile exceeds tau times the
// tile's leading profile (one wavefront per tile). tau = 0 keeps every l*.
// Each lane first takes its own node's maximum over l* and its own highest
// l* above the threshold; a wave maximum of each gives the tile's values
// (max over nodes of the per-node highest l* = highest l* whose maximum over
// nodes passes).
__global__ __launch_bounds__(64) void k_lcut(int N, int Ls, double tau, const double *E, int *lcut)
{
    const int tile = blockIdx.x;
    const int i = tile * gmd::TILE + threadIdx.x;
    auto prof = [&](int ls) {
        return i < N ? fmax(fabs(E[(size_t)ls * N + i]), fabs(E[(size_t)(Ls + ls) * N + i])) : 0.0;
    };
    double gmax = 0.0;
    for (int ls = 0; ls < Ls; ++ls) gmax = fmax(gmax, prof(ls));
    for (int o = 32; o > 0; o >>= 1) gmax = fmax(gmax, __shfl_xor(gmax, o));
    int keep = tau > 0.0 ? 0 : Ls - 1;
    if (tau > 0.0) {
        for (int ls = 0; ls < Ls; ++ls)
            if (prof(ls) > tau * gmax) keep = ls;
        for (int o = 32; o > 0; o >>= 1) keep = max(keep, __shfl_xor(keep, o));
    }
    if (threadIdx.x == 0) lcut[tile] = keep;
}

// ---------------------------------------------------------------- operator apply

typedef double d4 __attribute__((ext_vector_type(4)));


// Coupling of the 16 rows of a chunk at the NG x 16 nodes of a tile, on the
// matrix cores. The chunk's pass table lists its (l1, slot group) blocks; each
// pass runs the columns m1 = 0 .. l1 of degree l1, and each (pass, m1) step is
//   W[row][node] = sum_t G[m1][t][row] E_t(node)   (one MFMA per 16 nodes)
//   acc[row][node] += W[row][node] xt_(l1,m1)(node)   (VALU)
// for the slots t = 0 .. 3 of the pass, l* = ls0 + 2 t, with xt = exp(i m1
// theta) x. Lane layout (gfx950): A[row = lane%16][k = lane/16],
// B[k = lane/16][node = lane%16], D rows per Arith::row, node = lane%16.
// Arithmetic of the coupling: FP64 (matrix-core products and accumulation in
// double) or FP32 (products on v_mfma_f32_16x16x4f32 and accumulation in
// float, for the mixed-precision Krylov applies).
typedef float f4 __attribute__((ext_vector_type(4)));
template <typename T> struct Arith;
template <> struct Arith<double> {
    typedef double2 C;
    typedef d4 V4;
    static __device__ __forceinline__ V4 mfma(double a, double b, const V4 &c)
    {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ C from(double2 v) { return v; }
    static __device__ __forceinline__ double2 to_d(C v) { return v; }
    // D row held in accumulator register r by lane group q (measured layout)
    static __device__ __forceinline__ int row(int q, int r) { return q + 4 * r; }
};
template <> struct Arith<float> {
    typedef float2 C;
    typedef f4 V4;
    static __device__ __forceinline__ V4 mfma(float a, float b, const V4 &c)
    {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ C from(double2 v) { return make_float2((float)v.x, (float)v.y); }
    static __device__ __forceinline__ double2 to_d(C v) { return make_double2(v.x, v.y); }
    static __device__ __forceinline__ int row(int q, int r) { return 4 * q + r; }
};

template <int NG, typename T>
__device__ __forceinline__ void acc_update(typename Arith<T>::C (&acc)[NG][4], int n, const typename Arith<T>::V4 &D,
                                           typename Arith<T>::C xv)
{
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        acc[n][r].x = fma(D[r], xv.x, acc[n][r].x);
        acc[n][r].y = fma(D[r], xv.y, acc[n][r].y);
    }
}

// Raw buffer loads: a scalar resource over a whole table, a fixed per-lane
// byte offset and a per-step scalar byte offset (reads past num_records
// return zeros). Word 3 of the resource is the gfx9 default format.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buffer_rsrc(const void *p, unsigned bytes)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), (short)0, (int)bytes, 0x00020000);
}
template <typename V> __device__ __forceinline__ V buffer_ld(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff);
template <> __device__ __forceinline__ float buffer_ld<float>(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, (int)voff, (int)soff, 0));
}
template <> __device__ __forceinline__ double buffer_ld<double>(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff)
{
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, (int)voff, (int)soff, 0));
}
template <> __device__ __forceinline__ float2 buffer_ld<float2>(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff)
{
    return __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(r, (int)voff, (int)soff, 0));
}
template <> __device__ __forceinline__ double2 buffer_ld<double2>(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff)
{
    return __builtin_bit_cast(double2, __builtin_amdgcn_raw_buffer_load_b128(r, (int)voff, (int)soff, 0));
}

// The coupling of one chunk at one tile, one stream of steps from one of the
// chunk's step lists (each sorted by the first slot's l*, so that a tile with
// multipole cut lc runs a prefix or a range of it). The A
// fragment (the step's [slot][row] Gaunt values) and the phased column vectors of the next RING
// steps are in flight, as buffer loads with fixed per-lane offsets and the
// step's scalar offsets; the B fragment E_{l*}(node) is read from LDS one step
// ahead. No lane masking is needed: G holds zeros for the padding slots of a
// block and the LDS rows above lc hold zeros. After the last step the ring
// repeats the last step's offsets (unused loads), so the loads stay
// unconditional.
//
// GRP = 2 runs pairs: each step holds two consecutive slot groups of one
// column (fragments at g and g + 64, LDS rows from e and e + 8), chained on
// the matrix core (the second MFMA accumulates onto the first), so the column
// is loaded and multiplied in once for eight slots. GRP = 1 runs singles, or
// pairs as their first group alone. The stream is the `total` steps from
// `st`, rounded up to whole ring rounds; the caller makes sure that the
// entries past `total` add nothing or belong to no other stream.
template <int NG, typename T, int GRP>
__device__ __forceinline__ void mfma_stream(const Geometry &g, const T *GM, const typename Arith<T>::C *Xt, int t0,
                                            const gmd::MStep *st, int total, const T *sE,
                                            typename Arith<T>::C (&acc)[NG][4])
{
    typedef Arith<T> AR;
    typedef typename AR::C C;
    constexpr int TN = 16 * NG;
    constexpr int RING = gmd::RING;
    const int lane = threadIdx.x & 63, q = lane >> 4, col = lane & 15;
    const int N = g.N;
    if (total == 0) return;
    const __amdgpu_buffer_rsrc_t rg = buffer_rsrc(GM, g.gm_n * (unsigned)sizeof(T));
    const __amdgpu_buffer_rsrc_t rx = buffer_rsrc(Xt, (unsigned)g.K * (unsigned)N * (unsigned)sizeof(C));
    // per-lane byte offsets; nodes past the grid repeat node N-1 (they only
    // feed accumulators never stored)
    const uint32_t g_lane = (uint32_t)(q * 16 + col) * (uint32_t)sizeof(T);
    uint32_t x_lane[NG];
#pragma unroll
    for (int n = 0; n < NG; ++n) x_lane[n] = (uint32_t)min(t0 + 16 * n + col, N - 1) * (uint32_t)sizeof(C);
    const uint32_t e_lane = (uint32_t)(2 * q * TN + col) * (uint32_t)sizeof(T);
    const uint32_t xcol = (uint32_t)N * (uint32_t)sizeof(C);
    // The stream runs a multiple of RING steps (one loop exit per ring
    // round). The steps past the tile's `total` are the chunk's steps whose
    // first slot lies above lc, or its padding steps: both read LDS rows above
    // lc, which hold zeros, so they add nothing. Step entries are read RING
    // steps before their loads are issued.
    const int padded = (total + RING - 1) / RING * RING;
    auto entry = [&](const int step) { return st[min(step, padded - 1)]; };
    gmd::MStep ent[RING];
#pragma unroll
    for (int k = 0; k < RING; ++k) ent[k] = entry(k);
    T fr[RING][GRP];
    C xv[RING][NG];
    uint32_t eoff[RING];
    // issue the loads of `step` into slot k (ent[k] holds its entry) and read
    // the entry of the step RING later
    auto issue = [&](const int k, const int step) {
        const gmd::MStep e = ent[k];
#pragma unroll
        for (int h = 0; h < GRP; ++h)
            fr[k][h] = buffer_ld<T>(rg, g_lane + (uint32_t)(64 * h * sizeof(T)), (uint32_t)e.g * (uint32_t)sizeof(T));
#pragma unroll
        for (int n = 0; n < NG; ++n) xv[k][n] = buffer_ld<C>(rx, x_lane[n], (uint32_t)e.x * xcol);
        eoff[k] = (uint32_t)e.e * (uint32_t)(TN * sizeof(T));
        ent[k] = entry(step + RING);
    };
    auto bf_load = [&](const int k, T (&bf)[GRP][NG]) {
        const T *eb = (const T *)((const char *)sE + eoff[k] + e_lane);
#pragma unroll
        for (int h = 0; h < GRP; ++h)
#pragma unroll
            for (int n = 0; n < NG; ++n) bf[h][n] = eb[8 * h * TN + 16 * n];
    };
    const typename AR::V4 z = {T(0), T(0), T(0), T(0)};
    auto consume = [&](const int k, const T (&bf)[GRP][NG]) {
        if constexpr (NG == 1) {
            typename AR::V4 D0 = AR::mfma(fr[k][0], bf[0][0], z);
            if constexpr (GRP == 2) D0 = AR::mfma(fr[k][1], bf[1][0], D0);
            acc_update<NG, T>(acc, 0, D0, xv[k][0]);
        } else {
            // node groups in pairs: the accumulation of group n runs while
            // group n + 1 is on the matrix core
#pragma unroll
            for (int n = 0; n < NG; n += 2) {
                typename AR::V4 D0 = AR::mfma(fr[k][0], bf[0][n], z);
                typename AR::V4 D1 = AR::mfma(fr[k][0], bf[0][n + 1], z);
                if constexpr (GRP == 2) {
                    D0 = AR::mfma(fr[k][1], bf[1][n], D0);
                    D1 = AR::mfma(fr[k][1], bf[1][n + 1], D1);
                }
                __builtin_amdgcn_sched_barrier(0);
                acc_update<NG, T>(acc, n, D0, xv[k][n]);
                __builtin_amdgcn_sched_barrier(0);
                acc_update<NG, T>(acc, n + 1, D1, xv[k][n + 1]);
                __builtin_amdgcn_sched_barrier(0);
            }
        }
    };
    // prologue in ring order (kept by the barriers), so that the loop is
    // entered with the same loads in flight as at its back edge
#pragma unroll
    for (int k = 0; k < RING; ++k) {
        issue(k, k);
        __builtin_amdgcn_sched_barrier(0);
    }
    T bfa[GRP][NG], bfb[GRP][NG];
    bf_load(0, bfa);
#pragma unroll 1
    for (int s = 0; s < padded; s += RING) {
#pragma unroll
        for (int k = 0; k < RING; ++k) {
            // slot k + 1 was issued before slot k is consumed: its B fragment
            // is read now and used at the next step
            if (k & 1) {
                bf_load((k + 1) % RING, bfa);
                consume(k, bfb);
            } else {
                bf_load((k + 1) % RING, bfb);
                consume(k, bfa);
            }
            // unconditional refill (a conditional one would make the
            // compiler copy registers with loads in flight)
            issue(k, s + k + RING);
        }
    }
}

// Operator apply. A workgroup: one tile of TN = 16 NG radial nodes and a group
// of AW row chunks (16 modes of equal l parity each), one chunk per wavefront;
// the tile's profiles E_{l*} are staged in LDS. Many small workgroups per CU
// let the dispatcher balance tiles near the bodies (many multipoles) against
// far ones. The epilogue adds the three-point stencil, the centrifugal term
// and the CAP (hamiltonian.cpp:44-66,224-243), forms A = 1 + i dt/2 H
// (matrices.cpp:27-40) with the rank-one sinks (matrices.cpp:111-116), and
// fuses the reductions.
template <int NG, typename T>
__device__ __forceinline__ void k_apply_body(const ApplyArgs &a, Blk B)
{
    typedef Arith<T> AR;
    constexpr int TN = 16 * NG;
    // rows l* = 0 .. lc, then zeros (single-group padding steps read up to
    // ZROW + 6, paired ones up to ZROW + 14)
    constexpr int ER = gmd::EROWS;
    __shared__ T sE[ER * TN];
    // the workgroup's chunk rows (mode, l, m) and the phases exp(i m theta),
    // staged with the profiles for the epilogue
    __shared__ int sJ[gmd::AW][gmd::MR], sL[gmd::AW][gmd::MR], sM[gmd::AW][gmd::MR];
    __shared__ double2 sPh[2 * gmd::MAXLM + 1];
    const int MODE = a.mode;
    const Geometry &g = a.g;
    Krylov *Kr = a.Kr;
    if ((MODE == AM_V || MODE == AM_T) && Kr->done) return;
    const StepParams *P = a.P;
    const int lane = threadIdx.x & 63, q = lane >> 4, col = lane & 15;
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int N = g.N, lmax = g.lmax;
    // block b runs on XCD b % 8 (round-robin dispatch): the chunk groups of a
    // tile share one XCD's L2, tiles are spread over the XCDs and, within an
    // XCD, dispatched in order of increasing radius (decreasing multipole work)
    const int b = B.id;
    const int ng = a.ngroups;
    const int local = b >> 3;
    const int grp = local % ng;
    const int tile = (local / ng) * 8 + (b & 7);
    const int ntiles = (N + TN - 1) / TN;
    const bool item_ok = tile < ntiles;
    const int chn = grp * gmd::AW + wv;
    const bool wave_ok = item_ok && chn < g.nchunks;
    const int t0 = tile * TN;
    const int nsink = P->nsink;
    const bool gq_mode = MODE == AM_RHS;
    const double2 coef0 = gq_mode ? Kr->gq[0] : Kr->coefA[0];
    const double2 coef1 = gq_mode ? Kr->gq[1] : Kr->coefA[1];
    const double rhs_nrm = (MODE == AM_RES || MODE == AM_RST) ? Kr->rhs_nrm : 0.0;
    double red[3] = {0.0, 0.0, 0.0};
    typename AR::C acc[NG][4];
#pragma unroll
    for (int n = 0; n < NG; ++n)
#pragma unroll
        for (int r = 0; r < 4; ++r) acc[n][r] = AR::from(make_double2(0.0, 0.0));
    const T *E = std::is_same<T, double>::value ? (const T *)g.E : (const T *)g.E32;
    const T *GM = std::is_same<T, double>::value ? (const T *)g.GM : (const T *)g.GM32;
    if (item_ok) {
        // multipoles above lcut are below 2^-64 of the tile's leading one
        const int lc = g.lcut ? g.lcut[t0 / gmd::TILE] : 2 * lmax;
        for (int e = threadIdx.x; e < (lc + 1) * TN; e += 64 * gmd::AW) {
            const int ls = e / TN, node = t0 + (e & (TN - 1));
            sE[e] = node < N ? E[(size_t)((ls & 1) * g.Ls + ls) * N + node] : T(0);
        }
        for (int e = (lc + 1) * TN + threadIdx.x; e < ER * TN; e += 64 * gmd::AW) sE[e] = T(0);
        for (int e = threadIdx.x; e < gmd::AW * gmd::MR; e += 64 * gmd::AW) {
            const int w = e / gmd::MR, r = e % gmd::MR, c = grp * gmd::AW + w;
            const bool ok = c < g.nchunks;
            sJ[w][r] = ok ? g.mchunks[c].row_j[r] : -1;
            sL[w][r] = ok ? g.mchunks[c].row_l[r] : 0;
            sM[w][r] = ok ? g.mchunks[c].row_m[r] : 0;
        }
        for (int e = threadIdx.x; e < 2 * lmax + 1; e += 64 * gmd::AW) sPh[e] = P->ph[e];
        __syncthreads();
        if (wave_ok) {
            const typename AR::C *Xt = static_cast<const typename AR::C *>(a.Xt);
            // Whole pairs, then pairs cut after their first group, then
            // singles. The pair stream runs its count rounded up to whole
            // ring rounds; the pairs it runs past the whole ones are exact as
            // pairs too (a second group above the cut reads zero rows), so
            // the cut pairs start after them.
            const int cix = chn * (2 * lmax + 1) + lc;
            const gmd::MStep *pl = g.msteps + g.pair_start[chn];
            const int pf = g.pair_full[cix], pe = g.pair_end[cix];
            const int hb = min((pf + gmd::RING - 1) / gmd::RING * gmd::RING, pe);
            mfma_stream<NG, T, 2>(g, GM, Xt, t0, pl, pf, sE, acc);
            mfma_stream<NG, T, 1>(g, GM, Xt, t0, pl + hb, pe - hb, sE, acc);
            mfma_stream<NG, T, 1>(g, GM, Xt, t0, g.msteps + g.step_start[chn], g.step_count[cix], sE, acc);
        }
    }
    if (wave_ok) {
        // epilogue: lane owns rows q + 4r and nodes 16 n + col
        const double alpha2 = 1.0 / (g.dr * g.dr);
        const double hdt = P->half_dt;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            // two rows at a time: bounds the loads the compiler keeps in flight
            if ((r & 1) == 0) asm volatile("" ::: "memory");
            const int rl = AR::row(q, r);
            const int j = sJ[wv][rl];
            const int l = sL[wv][rl], m = sM[wv][rl];
#pragma unroll
            for (int n = 0; n < NG; ++n) {
                const int i = t0 + 16 * n + col;
                if (j < 0 || i >= N) continue;
                const size_t row = (size_t)j * N + i;
                const double2 Cc = cdotm(sPh[m + lmax], AR::to_d(acc[n][r]));
                if (a.Cout) a.Cout[row] = Cc;
                const double2 xi = load_x(a, j, i);
                const double2 capv = g.cap ? g.cap[i] : make_double2(0.0, 0.0);
                double2 Hx;
                if (i > 0 && i < N - 1) {
                    const double ri = g.eps + i * g.dr;
                    const double h0d = -(-2.0 * alpha2 - (double)(l * (l + 1)) / (ri * ri)) / (2 * g.mu);
                    const double2 xm = load_x(a, j, i - 1);
                    const double2 xp = load_x(a, j, i + 1);
                    Hx = cadd(cscale(g.off, cadd(xm, xp)), cadd(cmul(make_double2(h0d + capv.x, capv.y), xi), Cc));
                } else {
                    Hx = cmul(capv, xi);
                }
                if (MODE == AM_H) {
                    a.Y[row] = Hx;
                } else if (MODE == AM_V || MODE == AM_T || MODE == AM_RES || MODE == AM_A || MODE == AM_RST) {
                    double2 out = cadd(xi, imul(hdt, Hx));
                    if (nsink) out = cadd(out, sink_term(a, nsink, coef0, coef1, j, i));
                    if (MODE == AM_V) {
                        a.Y[row] = out;
                        const double2 d = cdotm(a.R0[row], out);
                        red[0] += d.x;
                        red[1] += d.y;
                    } else if (MODE == AM_T) {
                        a.Y[row] = out;
                        const double2 d = cdotm(out, a.R0[row]);
                        red[0] += d.x;
                        red[1] += d.y;
                        red[2] += cnorm2(out);
                    } else if (MODE == AM_A) {
                        a.Y[row] = out;
                    } else if (MODE == AM_RST) {
                        // r = b - A x of the normalised system, r0' = |rhs| r
                        const double2 bq = make_double2(a.R0[row].x / rhs_nrm, a.R0[row].y / rhs_nrm);
                        const double2 rq = csub(bq, out);
                        a.Y[row] = rq;
                        a.Y2[row] = cscale(rhs_nrm, rq);
                        red[0] += cnorm2(rq);
                    } else {
                        double2 bq = a.R0[row];
                        if (a.r0_raw && rhs_nrm != 0.0) bq = make_double2(bq.x / rhs_nrm, bq.y / rhs_nrm);
                        red[0] += cnorm2(csub(out, bq));
                    }
                } else if (MODE == AM_RHS) {
                    const double2 sHx = imul(hdt, Hx);
                    double2 st = make_double2(0.0, 0.0);
                    if (nsink) st = cscale(0.5 * hdt, sink_term(a, nsink, coef0, coef1, j, i));  // (dt/4) Gamma psi
                    const double2 rhs = csub(csub(xi, sHx), st);
                    const double2 ax = cadd(cadd(xi, sHx), st);
                    const double2 r0 = make_double2(rhs.x - ax.x, rhs.y - ax.y);
                    a.Y[row] = rhs;
                    a.Y2[row] = r0;
                    red[0] += cnorm2(rhs);
                    red[1] += cnorm2(r0);
                }
            }
        }
    }
    if (MODE == AM_V) {
        if (block_partials<2>({red[0], red[1]}, a.partial, &Kr->counter, B)) {
            const double dr_ = sum_partials(a.partial, 0, B), di = sum_partials(a.partial, 1, B);
            if (threadIdx.x == 0) {
                Kr->alpha = cdiv(Kr->rho, make_double2(dr_, di));
                Kr->counter = 0;
            }
        }
    } else if (MODE == AM_T) {
        if (block_partials<3>({red[0], red[1], red[2]}, a.partial, &Kr->counter, B)) {
            const double tr = sum_partials(a.partial, 0, B), ti = sum_partials(a.partial, 1, B);
            const double tt = sum_partials(a.partial, 2, B);
            if (threadIdx.x == 0) {
                Kr->w = tt > 0.0 ? make_double2(tr / tt, ti / tt) : make_double2(0.0, 0.0);
                Kr->counter = 0;
            }
        }
    } else if (MODE == AM_RHS) {
        if (block_partials<2>({red[0], red[1]}, a.partial, &Kr->counter, B)) {
            const double s = sum_partials(a.partial, 0, B), s0 = sum_partials(a.partial, 1, B);
            if (threadIdx.x == 0) {
                Kr->rhs_nrm = sqrt(s);
                Kr->r0_raw2 = s0;
                Kr->counter = 0;
            }
        }
    } else if (MODE == AM_RES || MODE == AM_RST) {
        // (one reduction site for both: a second copy of the block reduction
        // raised the register count of every apply instance by a fifth)
        if (block_partials<1>({red[0]}, a.partial, &Kr->counter, B)) {
            const double s = sum_partials(a.partial, 0, B);
            if (threadIdx.x == 0) {
                if (MODE == AM_RES) {
                    Kr->res2 = s;
                } else {
                    // Lost bi-orthogonality (matrices.cpp:67-74, and Eigen's BiCGSTAB):
                    // restart from the true residual, r0 = r, rho = |r|^2 (scaled by
                    // |rhs| with the unnormalised shadow residual); Eigen also takes the
                    // new |r0|^2 as its threshold base and resets the iteration counter
                    // on the first restart. The iteration then goes on from beta with
                    // the old p, v, alpha and w.
                    const double rho = Kr->rhs_nrm * s;
                    Kr->rho = make_double2(rho, 0.0);
                    if (Kr->eigen_mode) {
                        Kr->r0_sq = rho;
                        if (Kr->restarts == 0) Kr->iter = 0;
                    }
                    Kr->restarts += 1;
                    Kr->r2 = s;
                    Kr->beta = cmul(cdiv(Kr->rho, Kr->rho_old), cdiv(Kr->alpha, Kr->w));
                    Kr->restart = 0;
                    Kr->done = 0;
                }
                Kr->counter = 0;
            }
        }
    }
}
template <int NG, typename T>
__global__ __launch_bounds__(64 * gmd::AW) void k_apply(ApplyArgs a) { k_apply_body<NG, T>(a, grid_blk()); }
// One launch over many solvers: member m owns blocks [off[m], off[m + 1]).
template <int NG, typename T>
__global__ __launch_bounds__(64 * gmd::AW) void k_apply_batch(const ApplyArgs *__restrict__ args, const int *__restrict__ off, int M)
{
    const int m = batch_member(off, M);
    const ApplyArgs a = args[m];
    k_apply_body<NG, T>(a, Blk{(int)blockIdx.x - __builtin_amdgcn_readfirstlane(off[m]),
                             __builtin_amdgcn_readfirstlane(off[m + 1] - off[m])});
}
// The same launch with the members' arguments at apply_pool[first + m]: from
// constant memory the compiler reloads fields where they are used, and the
// kernel keeps the register count of a plain launch (a global argument table
// cost about 30 more VGPRs, and with them waves per SIMD).
__constant__ ApplyArgs apply_pool[APPLY_POOL];
template <int NG, typename T>
__global__ __launch_bounds__(64 * gmd::AW) void k_apply_pooled(int first, const int *__restrict__ off, int M)
{
    const int m = batch_member(off, M);
    k_apply_body<NG, T>(apply_pool[first + m], Blk{(int)blockIdx.x - __builtin_amdgcn_readfirstlane(off[m]),
                                                   __builtin_amdgcn_readfirstlane(off[m + 1] - off[m])});
}

// Diagnostics apply on the post-mask state (extraction.cpp:8-129). The
// coupling is node-local and the mask is one radial profile for every mode, so
// V psi = |rhs| mask (V x) with V x kept by the acceptance apply; only the
// stencil is evaluated here. One wavefront per (mode, tile of DIAG_NODES nodes) writes
// the tile sums of |U|^2 and Re<U|H U> below idx_ex, Re<U|H U>, Im(m <U|V U>)
// and Re<dU/dt|V U> up to idx_ex inclusive, and |U|^2 over the grid.
__device__ __forceinline__ void k_diag_apply_body(const ApplyArgs &a, Blk B)
{
    const Geometry &g = a.g;
    const StepParams *P = a.P;
    const Krylov *Kr = a.Kr;
    const int lane = threadIdx.x & 63;
    const int N = g.N, nt = (N + gmd::DIAG_NODES - 1) / gmd::DIAG_NODES;
    const int item = B.id * 4 + (threadIdx.x >> 6);
    if (item >= g.K * nt) return;
    const int j = item / nt, tile = item - j * nt;
    const int kk = (int)(sqrt(8.0 * j + 1.0) - 1.0) / 2;
    const int l = (kk + 1) * (kk + 2) / 2 <= j ? kk + 1 : kk;
    const int m = 2 * (j - l * (l + 1) / 2) - l;
    const double nrm = Kr->rhs_nrm;
    const int ix = P->idx_ex;
    double rp[6] = {0, 0, 0, 0, 0, 0};
    // four nodes per lane, 64 apart (coalesced per pass), summed in that order
#pragma unroll
    for (int u = 0; u < gmd::DIAG_NODES / 64; ++u) {
        const int i = tile * gmd::DIAG_NODES + 64 * u + lane;
        if (i >= N) continue;
        const size_t row = (size_t)j * N + i;
        const double2 Cc = cscale(nrm * (a.mask ? a.mask[i] : 1.0), a.Cin[row]);
        const double2 xi = load_x(a, j, i);
        const double2 capv = g.cap ? g.cap[i] : make_double2(0.0, 0.0);
        double2 Hx;
        if (i > 0 && i < N - 1) {
            const double ri = g.eps + i * g.dr;
            const double h0d = -(-2.0 / (g.dr * g.dr) - (double)(l * (l + 1)) / (ri * ri)) / (2 * g.mu);
            const double2 xm = load_x(a, j, i - 1);
            const double2 xp = load_x(a, j, i + 1);
            Hx = cadd(cscale(g.off, cadd(xm, xp)), cadd(cmul(make_double2(h0d + capv.x, capv.y), xi), Cc));
        } else {
            Hx = cmul(capv, xi);
        }
        double2 dpsi = make_double2(Hx.y, -Hx.x);  // -i H psi
        if (P->nsink) dpsi = csub(dpsi, cscale(0.5, sink_term(a, P->nsink, Kr->gq[0], Kr->gq[1], j, i)));
        const double2 eh = cdotm(xi, Hx);
        const double n2 = cnorm2(xi);
        if (i < ix) {
            rp[0] += n2;
            rp[1] += eh.x;
        }
        if (i <= ix) {
            rp[2] += eh.x;
            rp[3] += m * cdotm(xi, Cc).y;
            rp[4] += cdotm(dpsi, Cc).x;
        }
        rp[5] += n2;
        if (i == ix) a.dpsi_ix[j] = dpsi;
    }
#pragma unroll
    for (int qq = 0; qq < 6; ++qq) {
        const double s = wave_sum(rp[qq]);
        if (lane == 0) a.rowpart[((size_t)tile * g.K + j) * 6 + qq] = s;
    }
}
__global__ __launch_bounds__(256) void k_diag_apply(ApplyArgs a) { k_diag_apply_body(a, grid_blk()); }
// One launch over many solvers: member m owns blocks [off[m], off[m + 1]).
__global__ __launch_bounds__(256) void k_diag_apply_batch(const ApplyArgs *__restrict__ args, const int *__restrict__ off, int M)
{
    const int m = batch_member(off, M);
    const ApplyArgs a = args[m];
    k_diag_apply_body(a, Blk{(int)blockIdx.x - __builtin_amdgcn_readfirstlane(off[m]),
                             __builtin_amdgcn_readfirstlane(off[m + 1] - off[m])});
}

// ---------------------------------------------------------------- Thomas factor

// Diagonal of mode j's CN matrix at node i: 1 + i dt/2 (H0_ii + V_jj(r_i) + CAP_i)
// (hamiltonian.cpp:224-243, matrices.cpp:27-40); boundary rows keep only the CAP.
__device__ __forceinline__ double2 cn_diag(const FactorArgs &f, const double *Gjj, int l, int i)
{
    const int N = f.N;
    double2 h = make_double2(0.0, 0.0);
    if (i > 0 && i < N - 1) {
        const double ri = f.eps + i * f.dr;
        const double alpha2 = 1.0 / (f.dr * f.dr);
        // sum_s G_jj[s] E_2s(r_i) in order of s, the profile loads issued
        // eight at a time (one memory latency per eight terms)
        double v = 0.0;
        for (int s0 = 0; s0 <= l; s0 += 8) {
            double ev[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) ev[u] = s0 + u <= l ? f.E[(size_t)(2 * (s0 + u)) * N + i] : 0.0;
#pragma unroll
            for (int u = 0; u < 8; ++u)
                if (s0 + u <= l) v += Gjj[s0 + u] * ev[u];
        }
        h.x = -(-2.0 * alpha2 - (double)(l * (l + 1)) / (ri * ri)) / (2 * f.mu) + v;
    }
    if (f.cap) h = cadd(h, f.cap[i]);
    const double hdt = f.P->half_dt;
    return make_double2(1.0 - hdt * h.y, hdt * h.x);
}

struct Mob {
    double2 a, b, c, d;  // [[a, b], [c, d]] acting on (u, v), pivot = u / v
};

// Divide the map by its largest entry (the pivot is a ratio).
__device__ __forceinline__ void mob_rescale(Mob &r)
{
    double m = fmax(fmax(fabs(r.a.x), fabs(r.a.y)), fmax(fabs(r.b.x), fabs(r.b.y)));
    m = fmax(m, fmax(fmax(fabs(r.c.x), fabs(r.c.y)), fmax(fabs(r.d.x), fabs(r.d.y))));
    if (m > 0.0 && isfinite(m)) {
        const double sc = 1.0 / m;
        r.a = cscale(sc, r.a);
        r.b = cscale(sc, r.b);
        r.c = cscale(sc, r.c);
        r.d = cscale(sc, r.d);
    }
}

// first, then second: second * first, rescaled
__device__ __forceinline__ Mob mob_compose(const Mob &x, const Mob &y)
{
    Mob r;
    r.a = cadd(cmul(y.a, x.a), cmul(y.b, x.c));
    r.b = cadd(cmul(y.a, x.b), cmul(y.b, x.d));
    r.c = cadd(cmul(y.c, x.a), cmul(y.d, x.c));
    r.d = cadd(cmul(y.c, x.b), cmul(y.d, x.d));
    mob_rescale(r);
    return r;
}

// x, then an interior node's map (u, v) -> (d u + nl2 v, u), unscaled
__device__ __forceinline__ Mob mob_interior(const Mob &x, double2 d, double2 nl2)
{
    Mob r;
    r.a = cadd(cmul(d, x.a), cmul(nl2, x.c));
    r.b = cadd(cmul(d, x.b), cmul(nl2, x.d));
    r.c = x.a;
    r.d = x.b;
    return r;
}

// x, then a boundary node's reset (u, v) -> (d v, v), unscaled
__device__ __forceinline__ Mob mob_reset(const Mob &x, double2 d)
{
    Mob r;
    r.a = cmul(d, x.c);
    r.b = cmul(d, x.d);
    r.c = x.c;
    r.d = x.d;
    return r;
}

__device__ __forceinline__ Mob shfl_mob(const Mob &v, int src)
{
    Mob r;
    r.a.x = __shfl(v.a.x, src); r.a.y = __shfl(v.a.y, src);
    r.b.x = __shfl(v.b.x, src); r.b.y = __shfl(v.b.y, src);
    r.c.x = __shfl(v.c.x, src); r.c.y = __shfl(v.c.y, src);
    r.d.x = __shfl(v.d.x, src); r.d.y = __shfl(v.d.y, src);
    return r;
}

// Block exclusive scan of Moebius maps in increasing thread order (T
// threads); returns the thread's prefix and leaves the block total in `tot`.
template <int T>
__device__ __forceinline__ Mob mob_block_scan(const Mob &loc, Mob &tot)
{
    __shared__ Mob wt[T / 64];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    Mob inc = loc;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const Mob o = shfl_mob(inc, lane >= d ? lane - d : 0);
        if (lane >= d) inc = mob_compose(o, inc);
    }
    Mob ex = shfl_mob(inc, lane > 0 ? lane - 1 : 0);
    if (lane == 0) {
        ex.a = make_double2(1.0, 0.0);
        ex.b = make_double2(0.0, 0.0);
        ex.c = make_double2(0.0, 0.0);
        ex.d = make_double2(1.0, 0.0);
    }
    if (lane == 63) wt[wv] = inc;
    __syncthreads();
    Mob pre;
    pre.a = make_double2(1.0, 0.0);
    pre.b = make_double2(0.0, 0.0);
    pre.c = make_double2(0.0, 0.0);
    pre.d = make_double2(1.0, 0.0);
    tot = pre;
    for (int w = 0; w < T / 64; ++w) {
        if (w == wv) pre = tot;
        tot = mob_compose(tot, wt[w]);
    }
    __syncthreads();
    return mob_compose(pre, ex);
}

// Thomas pivots of every mode (include/RadialBlockPreconditioner.hpp:43-76):
// pivot_i = d_i - lower_i inv_{i-1} upper_{i-1}, with lower = upper = l (the
// stencil entry of A) on interior rows, so pivot_i = d_i - l^2 / pivot_{i-1}
// for 2 <= i <= N-2 and pivot_i = d_i otherwise. The recurrence is a Moebius
// map: the workgroup scans the maps of PE consecutive nodes per thread (2x2
// products, rescaled), then every thread replays its nodes with the
// reference's arithmetic from its incoming pivot.
//
// k_factor_win: block j nseg + g computes the pivots of segment g of mode j
// over the window of the preconditioner sweeps (the segment and PHALO nodes
// on either side) with the pivot at the window's first node reset to its
// diagonal. A pivot's error is multiplied by |l / pivot|^2 per node, so over
// the left halo by the square of the product that bounds the sweeps' dropped
// carries. From the window's pivots on [b - PHALO + 1, b + PHALO - 1] (the
// segment's own and its right halo's, which follow from the segment with an
// exact carry) the block also checks boundary b = g1 for the sweeps: flag
// pseq[j nseg + g] unless both halo products are below 2^-56. FT threads
// take the window's PT x PE nodes, FE consecutive ones each (fewer threads
// with more nodes each spend less of the block on the scan, whose cost is
// per thread).
constexpr int FT = 256, FE = gmd::PT * gmd::PE / FT;
static_assert(FT * FE == gmd::PT * gmd::PE, "the factor window covers the sweeps' window");
__global__ __launch_bounds__(FT) void k_factor_win(FactorArgs f)
{
    constexpr int T = FT, E = FE;
    const int N = f.N, nseg = gmd::precond_nseg(N);
    const int j = blockIdx.x / nseg, g = blockIdx.x - j * nseg;
    const int tid = threadIdx.x;
    int l = 0;
    while ((l + 1) * (l + 2) / 2 <= j) ++l;
    const double *Gjj = f.Gdiag + (size_t)j * (f.lmax + 1);
    const double2 ell = make_double2(0.0, f.P->half_dt * f.off);
    const double2 ell2 = cmul(ell, ell);
    const double2 nl2 = make_double2(-ell2.x, -ell2.y);
    const int g0 = g * gmd::PSEG, g1 = nseg == 1 ? N : min(N, g0 + gmd::PSEG);
    const int w0 = max(0, g0 - gmd::PHALO), w1 = nseg == 1 ? N : min(N, g1 + gmd::PHALO);
    const int i0 = w0 + tid * E;
    // interior rows 2 .. N - 2 apply the recurrence; the window's first node
    // and the boundary rows reset the pivot to the diagonal
    auto interior = [&](int i) { return i >= 2 && i <= N - 2 && i != w0; };
    // the window's diagonals, computed at consecutive nodes per thread index
    // (coalesced reads of the radial profiles) and handed over through LDS
    // (window node q at sdv[q + q / E], so that the E consecutive nodes of a
    // thread sit at a stride of E + 1 entries)
    __shared__ double2 sdv[T * E + T];
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const int q = e * T + tid, i = w0 + q;
        if (i < w1) sdv[q + q / E] = cn_diag(f, Gjj, l, i);
    }
    __syncthreads();
    double2 dv[E];
    Mob loc;
    loc.a = make_double2(1.0, 0.0);
    loc.b = make_double2(0.0, 0.0);
    loc.c = make_double2(0.0, 0.0);
    loc.d = make_double2(1.0, 0.0);
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const int i = i0 + e;
        if (i >= w1) {
            dv[e] = make_double2(1.0, 0.0);
        } else {
            dv[e] = sdv[tid * (E + 1) + e];
            loc = interior(i) ? mob_interior(loc, dv[e], nl2) : mob_reset(loc, dv[e]);
        }
    }
    mob_rescale(loc);
    Mob tot;
    const Mob P = mob_block_scan<T>(loc, tot);
    // incoming pivot: P applied to (1, 1) (the window starts with a reset, so
    // the value entering it is arbitrary)
    const double2 u0 = cadd(P.a, P.b), v0 = cadd(P.c, P.d);
    double2 prev_inv = cdiv(make_double2(1.0, 0.0), cdiv(u0, v0));
    int failed = 0;
    double2 iv[E];
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const int i = i0 + e;
        iv[e] = make_double2(0.0, 0.0);
        if (i >= w1) continue;
        double2 piv = dv[e];
        if (interior(i)) piv = csub(piv, cmul(cmul(ell, prev_inv), ell));
        // |piv|^2 is zero or not finite exactly when |piv| is
        const double ap2 = cnorm2(piv);
        if (ap2 == 0.0 || !isfinite(ap2)) {
            failed = 1;
            iv[e] = make_double2(0.0, 0.0);
        } else {
            iv[e] = cdiv(make_double2(1.0, 0.0), piv);
        }
        if (i >= g0 && i < g1) {
            f.inv[(size_t)j * N + i] = iv[e];
            if (f.inv32) f.inv32[(size_t)j * N + i] = make_float2((float)iv[e].x, (float)iv[e].y);
        }
        prev_inv = iv[e];
    }
    if (failed) atomicOr(f.failed, 1);
    if (!f.pseq || nseg == 1) return;
    // the halo check of boundary b = g1 (none after the last segment)
    const int b = g1;
    const double lell = 0.5 * log(cnorm2(ell)), lim = -56.0 * 0.6931471805599453;
    double sl = 0.0, sr = 0.0;
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const int i = i0 + e;
        // only the nodes within PHALO of b enter the products
        if (i >= w1 || i < b - gmd::PHALO + 1 || i > b + gmd::PHALO - 1) continue;
        // a_i = 0 on the boundary rows (the product is then zero)
        const double t = (i >= 1 && i <= N - 2) ? lell + 0.5 * log(cnorm2(iv[e])) : -INFINITY;
        if (i <= b) sl += t;
        if (i >= b) sr += t;
    }
    __shared__ double hs[T / 64][2];
    sl = wave_sum(sl);
    sr = wave_sum(sr);
    if ((tid & 63) == 0) {
        hs[tid >> 6][0] = sl;
        hs[tid >> 6][1] = sr;
    }
    __syncthreads();
    if (tid == 0) {
        sl = sr = 0.0;
        for (int w = 0; w < T / 64; ++w) {
            sl += hs[w][0];
            sr += hs[w][1];
        }
        // a right halo that reaches node N - 1 drops nothing
        const bool bad = g < nseg - 1 && (!(sl < lim) || (b + gmd::PHALO < N && !(sr < lim)));
        f.pseq[j * nseg + g] = bad;
    }
}

// k_factor: one workgroup per mode over the whole row, its tiles in order
// with exact carries, for the modes whose segments k_factor_win flagged (or
// every mode when given no segment flags); it sets all of the mode's flags to
// its own check, and the mode's verdict pseq[K nseg + j] that the sweeps read.
__global__ __launch_bounds__(gmd::PT) void k_factor(FactorArgs f)
{
    constexpr int T = gmd::PT, E = gmd::PE, TE = T * E;
    const int j = blockIdx.x, tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int nseg = gmd::precond_nseg(f.N);
    if (f.pseq) {
        // k_factor_win's pivots stand unless a boundary of this mode failed
        // its check (one window per row is exact)
        int any = 0;
        for (int g = 0; g + 1 < nseg; ++g) any |= f.pseq[j * nseg + g];
        if (!any) {
            if (tid == 0) f.pseq[f.K * nseg + j] = 0;
            return;
        }
    }
    int l = 0;
    while ((l + 1) * (l + 2) / 2 <= j) ++l;
    const int mi = j - l * (l + 1) / 2;
    (void)mi;
    const double *Gjj = f.Gdiag + (size_t)j * (f.lmax + 1);
    const double2 ell = make_double2(0.0, f.P->half_dt * f.off);
    const double2 ell2 = cmul(ell, ell);
    const double2 nl2 = make_double2(-ell2.x, -ell2.y);
    const int N = f.N;
    const int ntile = (N + TE - 1) / TE;
    double2 carry = make_double2(1.0, 0.0);  // pivot before node 0 (unused: node 0 resets)
    int failed = 0;
    for (int t = 0; t < ntile; ++t) {
        const int i0 = t * TE + tid * E;
        double2 dv[E];
        Mob loc;
        loc.a = make_double2(1.0, 0.0);
        loc.b = make_double2(0.0, 0.0);
        loc.c = make_double2(0.0, 0.0);
        loc.d = make_double2(1.0, 0.0);
        // the thread's E maps (identity beyond the grid), rescaled once: the
        // entries of a product of E maps stay far inside the double range
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const int i = i0 + e;
            if (i >= N) {
                dv[e] = make_double2(1.0, 0.0);
            } else {
                dv[e] = cn_diag(f, Gjj, l, i);
                loc = i >= 2 && i <= N - 2 ? mob_interior(loc, dv[e], nl2) : mob_reset(loc, dv[e]);
            }
        }
        mob_rescale(loc);
        // block exclusive scan of the thread maps (increasing node order)
        Mob tot;
        const Mob P = mob_block_scan<T>(loc, tot);
        // incoming pivot of this thread: P applied to (carry, 1)
        const double2 u = cadd(cmul(P.a, carry), P.b), v = cadd(cmul(P.c, carry), P.d);
        double2 prev_inv = cdiv(make_double2(1.0, 0.0), cdiv(u, v));
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const int i = i0 + e;
            if (i >= N) break;
            double2 piv = dv[e];
            if (i >= 2 && i <= N - 2) piv = csub(piv, cmul(cmul(ell, prev_inv), ell));
            const double ap = sqrt(cnorm2(piv));
            double2 inv;
            if (ap == 0.0 || !isfinite(ap)) {
                failed = 1;
                inv = make_double2(0.0, 0.0);
            } else {
                inv = cdiv(make_double2(1.0, 0.0), piv);
            }
            f.inv[(size_t)j * N + i] = inv;
            if (f.inv32) f.inv32[(size_t)j * N + i] = make_float2((float)inv.x, (float)inv.y);
            prev_inv = inv;
        }
        // carry: the last pivot of the tile, from the block total
        {
            const double2 uu = cadd(cmul(tot.a, carry), tot.b), vv = cadd(cmul(tot.c, carry), tot.d);
            carry = cdiv(uu, vv);
        }
    }
    if (failed) atomicOr(f.failed, 1);
    if (!f.pseq) return;
    // The segmented sweeps drop, at each end of a window, a carry that the
    // recurrence then multiplies by a_i = -l inv_i over the halo: the forward
    // carry into segment s + 1 by the product over its left halo, nodes
    // [b - PHALO + 1, b], the backward carry into segment s by the product
    // over its right halo, [b, b + PHALO - 1], b = (s + 1) PSEG. The mode is
    // solved whole unless every such product is below 2^-56.
    __shared__ int bad;
    if (tid == 0) bad = 0;
    __syncthreads();  // (also makes this block's pivots visible to its threads)
    const double lell = 0.5 * log(cnorm2(ell)), lim = -56.0 * 0.6931471805599453;
    for (int w = wv; w < 2 * (nseg - 1); w += T / 64) {
        const int b = (w / 2 + 1) * gmd::PSEG;
        const bool right = w & 1;
        // a right halo that reaches node N - 1 drops nothing
        if (right && b + gmd::PHALO >= N) continue;
        const int lo = right ? b : b - gmd::PHALO + 1;
        double sum = 0.0;
        for (int i = lo + lane; i < lo + gmd::PHALO; i += 64) {
            // a_i = 0 on the boundary rows (the product is then zero)
            const double2 iv = f.inv[(size_t)j * N + i];
            sum += (i >= 1 && i <= N - 2) ? lell + 0.5 * log(cnorm2(iv)) : -INFINITY;
        }
        for (int o = 32; o > 0; o >>= 1) sum += __shfl_xor(sum, o);
        if (lane == 0 && !(sum < lim)) atomicOr(&bad, 1);
    }
    __syncthreads();
    if (tid < nseg) f.pseq[j * nseg + tid] = tid + 1 < nseg ? bad : 0;
    if (tid == 0) f.pseq[f.K * nseg + j] = bad;
}

// ---------------------------------------------------------------- Thomas solve

// Workgroup barrier that orders LDS accesses only: global loads and stores in
// flight stay in flight (__syncthreads also fences global memory, which waits
// for every outstanding load and store).
__device__ __forceinline__ void lds_barrier()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

template <typename C>
struct Aff {
    C A, F;
};

// first, then second
template <typename C>
__device__ __forceinline__ Aff<C> compose(const Aff<C> &first, const Aff<C> &second)
{
    Aff<C> r;
    r.A = cmul(second.A, first.A);
    r.F = cadd(cmul(second.A, first.F), second.F);
    return r;
}

template <typename C>
__device__ __forceinline__ Aff<C> shfl_aff(const Aff<C> &v, int src)
{
    Aff<C> r;
    r.A.x = __shfl(v.A.x, src);
    r.A.y = __shfl(v.A.y, src);
    r.F.x = __shfl(v.F.x, src);
    r.F.y = __shfl(v.F.y, src);
    return r;
}

// Exclusive scan of affine maps over the PT threads of a block in the order
// of increasing (REV = false) or decreasing (REV = true) thread index.
// Returns the thread's exclusive prefix; `total` receives the block total.
template <bool REV, typename C>
__device__ Aff<C> block_scan(Aff<C> v, Aff<C> &total)
{
    __shared__ Aff<C> wt[gmd::PT / 64];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int vl = REV ? 63 - lane : lane;  // position in scan order within the wave
    Aff<C> inc = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const int src = REV ? lane + d : lane - d;
        const Aff<C> o = shfl_aff(inc, src < 0 ? 0 : (src > 63 ? 63 : src));
        if (vl >= d) inc = compose(o, inc);
    }
    Aff<C> ex = shfl_aff(inc, REV ? (lane < 63 ? lane + 1 : 63) : (lane > 0 ? lane - 1 : 0));
    if (vl == 0) {
        ex.A = cone<C>();
        ex.F = czero<C>();
    }
    if (vl == 63) wt[wv] = inc;
    lds_barrier();
    const int nw = gmd::PT / 64;
    Aff<C> pre;
    pre.A = cone<C>();
    pre.F = czero<C>();
    Aff<C> tot = pre;
    for (int k = 0; k < nw; ++k) {
        const int w = REV ? nw - 1 - k : k;
        if (w == wv) pre = tot;
        tot = compose(tot, wt[w]);
    }
    lds_barrier();
    total = tot;
    return compose(pre, ex);
}

__device__ __forceinline__ int lds_slot(int pos) { return pos + pos / gmd::PE; }

// exp(i m theta) y into the coupling scratch of the next apply, in its arithmetic
__device__ __forceinline__ void store_phased(void *Xt, int f32, size_t q, double2 ph, double2 y)
{
    const double2 v = cmul(ph, y);
    if (f32) ((float2 *)Xt)[q] = make_float2((float)v.x, (float)v.y);
    else ((double2 *)Xt)[q] = v;
}


// The p / s sweeps of the preconditioner: p = r + beta (p - w v) (PV_P) or
// s = r - alpha v (PV_S), then y = P^-1 p or z = P^-1 s by the Thomas
// recurrence (RadialBlockPreconditioner.hpp:78-97) as two first-order affine
// recurrences, y_i = a_i y_(i-1) + inv_i in_i and x_i = a_i x_(i+1) + y_i with
// a_i = -l inv_i (zero on the boundary rows).
//
// A block solves segment g of mode j: the window [w0, w1) of at most one
// PT x PE tile, the segment [g0, g1) and up to PHALO nodes on either
// side, is loaded once (each thread PE consecutive nodes), both recurrences
// run as block scans with the carries into the window's ends dropped, and the
// segment's nodes are stored. A dropped carry is multiplied over a halo by
// the product of |a_i|, which k_factor_win checks to stay below 2^-56 at
// every segment boundary (then the segment's output is the whole-row solve's
// to rounding). A mode that fails the check is solved whole by its segment-0
// block: the same recurrences over the row's tiles in order, with exact
// carries, the forward result kept in the output vector. RT = float sweeps
// in FP32 on the FP32 pivots; the p / s update and the output stay FP64.
// PV_WCOL solves the Woodbury columns P^-1 v_a the same way, with the compact
// sink vector as the right-hand side.
template <int VAR, typename RT>
__device__ __forceinline__ void k_psweep_body(const PrecArgs &a, Blk B)
{
    static_assert(VAR == PV_P || VAR == PV_S || (VAR == PV_WCOL && std::is_same<RT, double>::value),
                  "p or s sweep, or the FP64 Woodbury columns");
    constexpr bool WCOL = VAR == PV_WCOL;
    typedef typename Arith<RT>::C C;
    constexpr bool F32 = std::is_same<RT, float>::value;
    constexpr int T = gmd::PT, E = gmd::PE, TE = T * E;
    Krylov *Kr = a.Kr;
    if (!WCOL && Kr->done) return;
    const int nseg = a.nseg;
    const int N = a.N, tid = threadIdx.x;
    // PV_WCOL: the launch is ncol column grids of K nseg blocks; column col
    // solves P^-1 v_col into the sink's Woodbury column
    const int col = WCOL ? B.id / (a.K * nseg) : 0;
    const int bid = B.id - col * a.K * nseg;
    // Woodbury (a.wood): the segments holding sink-support nodes, and segment
    // 0 of every mode (a mode solved whole stores its row there), store
    // partial projections v_a^H y for k_wcoef. Their blocks come first in the
    // grid (j = b / count); the partial of holding block b is slot b.
    const int ns_w = a.wood ? a.Ps->nsink : 0;
    const int s0lo = ns_w > 0 ? a.sk[0].lo : 0, s0hi = ns_w > 0 ? a.sk[0].lo + a.sk[0].n : 0;
    const int s1lo = ns_w > 1 ? a.sk[1].lo : 0, s1hi = ns_w > 1 ? a.sk[1].lo + a.sk[1].n : 0;
    auto holds = [&](int gg) {
        const int lo = gg * gmd::PSEG, hi = nseg == 1 ? N : min(N, lo + gmd::PSEG);
        return gg == 0 || (lo < s0hi && s0lo < hi) || (lo < s1hi && s1lo < hi);
    };
    int count = 0;
    for (int gg = 0; gg < nseg; ++gg) count += holds(gg);
    int j, g = 0, rank;
    {
        const bool holder = bid < a.K * count;
        const int per = holder ? count : nseg - count, b = holder ? bid : bid - a.K * count;
        j = b / per;
        rank = b - j * per;
        // the rank-th segment (in order) that holds / does not hold
        for (int gg = 0, k = 0; gg < nseg; ++gg)
            if (holds(gg) == holder && k++ == rank) g = gg;
    }
    // a mode is solved whole when any of its segment boundaries failed the
    // halo check (k_factor_win; k_factor stores the mode's verdict after the
    // boundary flags, one load here)
    const bool whole = nseg > 1 && a.pseq[a.K * nseg + j] != 0;
    const size_t base_j = (size_t)j * N;
    const SinkDev &skc = a.sk[col];
    // (the sink columns, read-only to every other kernel, are PV_WCOL's output)
    double2 *const Y = WCOL ? const_cast<double2 *>(skc.W) : a.Y;
    const C ell = Arith<RT>::from(make_double2(0.0, a.Ps->half_dt * a.off));
    // the first iteration reads r = r0' / |rhs| from the unnormalised shadow
    // residual and takes p = r
    const bool it0 = !WCOL && Kr->fresh != 0;
    const bool first = VAR == PV_P && it0;
    const double nrm0 = it0 ? Kr->rhs_nrm : 1.0;
    const double2 *Rin = it0 ? a.R0 : a.R;
    double2 beta = make_double2(0, 0), wv = make_double2(0, 0), alpha = make_double2(0, 0);
    if (VAR == PV_P) {
        beta = Kr->beta;
        wv = Kr->w;
    }
    if (VAR == PV_S) alpha = Kr->alpha;
    // PV_P reads p from one buffer and writes the new p into the other (a
    // window's halo reads p that the neighbouring segment updates)
    const double2 *Pin = a.Pv;
    double2 *Pout = a.Pv;
    if (VAR == PV_P) {
        const bool second = Kr->pb != 0;
        Pin = second ? a.Pv2 : a.Pv;
        Pout = second ? a.Pv : a.Pv2;
    }
    const double2 phj = a.Xt ? a.Ps->ph[mode_m(j) + a.lmax] : make_double2(1.0, 0.0);
    // the thread's share of the projections v_a^H y over the sink supports
    double wp[4] = {0.0, 0.0, 0.0, 0.0};
    int nz_lo = N, nz_hi = 0;  // PV_WCOL: the column's non-zero node range
    auto store_out = [&](const int idx, const double2 yv) __attribute__((always_inline)) {
        ntst(Y + base_j + idx, yv);
        if (WCOL && (yv.x != 0.0 || yv.y != 0.0)) {
            nz_lo = min(nz_lo, idx);
            nz_hi = max(nz_hi, idx + 1);
        }
        if (a.Xt) store_phased(a.Xt, a.xt_f32, base_j + idx, phj, yv);
        if (idx >= s0lo && idx < s0hi) {
            const double2 d = cdotm(a.sk[0].v[(size_t)j * a.sk[0].n + (idx - s0lo)], yv);
            wp[0] += d.x;
            wp[1] += d.y;
        }
        if (idx >= s1lo && idx < s1hi) {
            const double2 d = cdotm(a.sk[1].v[(size_t)j * a.sk[1].n + (idx - s1lo)], yv);
            wp[2] += d.x;
            wp[3] += d.y;
        }
    };
    auto coef = [&](const C iv) __attribute__((always_inline)) {
        return C{-(ell.x * iv.x - ell.y * iv.y), -(ell.x * iv.y + ell.y * iv.x)};
    };
    // PV_WCOL's right-hand side: the sink vector v_col, zero outside its support
    auto vcol = [&](const int idx) __attribute__((always_inline)) {
        const int d = idx - skc.lo;
        return d >= 0 && d < skc.n ? skc.v[(size_t)j * skc.n + d] : make_double2(0.0, 0.0);
    };
    // the p / s update (PV_WCOL: the column's right-hand side as it is)
    auto update = [&](const double2 r, const double2 v, const double2 p) __attribute__((always_inline)) {
        if constexpr (WCOL) return r;
        const double2 rq = it0 ? make_double2(r.x / nrm0, r.y / nrm0) : r;
        if constexpr (VAR == PV_P) return first ? rq : cadd(rq, cmul(beta, csub(p, cmul(wv, v))));
        return csub(rq, cmul(alpha, v));
    };
    // Forward over [w0, w1) (positions tid E + e): in_i, its store on [p0, p1),
    // a_i and f_i = inv_i in_i; a_w0 = 0 when drop (the carry into the window
    // is dropped), and the identity past w1. Returns y in fe, a in ae.
    auto forward = [&](const int w0, const int w1, const int p0, const int p1, const bool drop, C &carry, C (&ae)[E],
                       C (&fe)[E]) __attribute__((always_inline)) {
        double2 r[E] = {}, v[E] = {}, p[E] = {};
        C iv[E];
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const int idx = min(w0 + tid * E + e, w1 - 1);
            const size_t q = base_j + idx;
            if constexpr (WCOL) {
                r[e] = vcol(idx);
            } else {
                r[e] = Rin[q];
                v[e] = a.V[q];
                if (VAR == PV_P && !first) p[e] = Pin[q];
            }
            if constexpr (F32) iv[e] = a.inv32[q];
            else iv[e] = a.inv[q];
        }
        Aff<C> loc;
        loc.A = cone<C>();
        loc.F = czero<C>();
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const int idx = w0 + tid * E + e;
            ae[e] = cone<C>();
            fe[e] = czero<C>();
            if (idx < w1) {
                const double2 in = update(r[e], v[e], p[e]);
                if (!WCOL && idx >= p0 && idx < p1) Pout[base_j + idx] = in;
                fe[e] = cmul(iv[e], Arith<RT>::from(in));
                ae[e] = (idx >= 1 && idx <= N - 2 && !(drop && idx == w0)) ? coef(iv[e]) : czero<C>();
            }
            loc.A = cmul(ae[e], loc.A);
            loc.F = cadd(cmul(ae[e], loc.F), fe[e]);
        }
        Aff<C> tot;
        const Aff<C> ex = block_scan<false>(loc, tot);
        C y = cadd(cmul(ex.A, carry), ex.F);
#pragma unroll
        for (int e = 0; e < E; ++e) {
            y = cadd(cmul(ae[e], y), fe[e]);
            fe[e] = y;
        }
        carry = cadd(cmul(tot.A, carry), tot.F);
    };
    // Backward over [w0, w1) from y in fe and a in ae: x_i = a_i x_(i+1) + y_i,
    // a_(w1 - 1) = 0 when drop, the identity past w1. Returns x in fe.
    auto backward = [&](const int w0, const int w1, const bool drop, C &carry, C (&ae)[E], C (&fe)[E])
                        __attribute__((always_inline)) {
        Aff<C> loc;
        loc.A = cone<C>();
        loc.F = czero<C>();
#pragma unroll
        for (int e = E - 1; e >= 0; --e) {
            const int idx = w0 + tid * E + e;
            if (idx >= w1) {
                ae[e] = cone<C>();
                fe[e] = czero<C>();
            } else if (drop && idx == w1 - 1) {
                ae[e] = czero<C>();
            }
            loc.A = cmul(ae[e], loc.A);
            loc.F = cadd(cmul(ae[e], loc.F), fe[e]);
        }
        Aff<C> tot;
        const Aff<C> ex = block_scan<true>(loc, tot);
        C x = cadd(cmul(ex.A, carry), ex.F);
#pragma unroll
        for (int e = E - 1; e >= 0; --e) {
            x = cadd(cmul(ae[e], x), fe[e]);
            fe[e] = x;
        }
        carry = cadd(cmul(tot.A, carry), tot.F);
    };
    C ae[E], fe[E];
    // The window's r, v, p (PV_WCOL: v_col) and pivots are loaded coalesced
    // (position tid + T e), before the mode's verdict is known (the indices
    // are valid either way, and the loads overlap its latency); f = inv in
    // goes through LDS to the thread's consecutive positions, where the
    // thread loads its pivots once more for a; the output returns through
    // LDS to coalesced stores.
    const int g0 = g * gmd::PSEG, g1 = nseg == 1 ? N : min(N, g0 + gmd::PSEG);
    const int w0 = max(0, g0 - gmd::PHALO), w1 = nseg == 1 ? N : min(N, g1 + gmd::PHALO);
    double2 r[E] = {}, v[E] = {}, p[E] = {};
    C ivw[E];
#pragma unroll
    for (int e = 0; e < E; ++e) {
        const int idx = min(w0 + tid + T * e, w1 - 1);
        const size_t q = base_j + idx;
        // (the pivots, read by every sweep, stay cacheable)
        if constexpr (WCOL) {
            r[e] = vcol(idx);
        } else {
            r[e] = ntld(Rin + q);
            v[e] = ntld(a.V + q);
            if (VAR == PV_P && !first) p[e] = ntld(Pin + q);
        }
        if constexpr (F32) ivw[e] = a.inv32[q];
        else ivw[e] = a.inv[q];
    }
    if (!whole) {
        __shared__ C sf[TE + TE / E];
        {
            const C (&iv)[E] = ivw;
#pragma unroll
            for (int e = 0; e < E; ++e) {
                const int pos = tid + T * e, idx = w0 + pos;
                C fv = czero<C>();
                if (idx < w1) {
                    const double2 in = update(r[e], v[e], p[e]);
                    if (!WCOL && idx >= g0 && idx < g1) ntst(Pout + base_j + idx, in);
                    fv = cmul(iv[e], Arith<RT>::from(in));
                }
                sf[lds_slot(pos)] = fv;
            }
        }
        C iv[E];
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const size_t q = base_j + min(w0 + tid * E + e, w1 - 1);
            if constexpr (F32) iv[e] = a.inv32[q];
            else iv[e] = a.inv[q];
        }
        lds_barrier();
        // forward, a_w0 = 0 (the carry into the window is dropped), the
        // identity past w1
        Aff<C> loc;
        loc.A = cone<C>();
        loc.F = czero<C>();
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const int idx = w0 + tid * E + e;
            fe[e] = sf[lds_slot(tid * E + e)];
            ae[e] = idx >= w1 ? cone<C>() : (idx >= 1 && idx <= N - 2 && idx != w0) ? coef(iv[e]) : czero<C>();
            loc.A = cmul(ae[e], loc.A);
            loc.F = cadd(cmul(ae[e], loc.F), fe[e]);
        }
        Aff<C> tot;
        C y = block_scan<false>(loc, tot).F;
#pragma unroll
        for (int e = 0; e < E; ++e) {
            y = cadd(cmul(ae[e], y), fe[e]);
            fe[e] = y;
        }
        C carry = czero<C>();
        backward(w0, w1, true, carry, ae, fe);
        // (every thread read its slots before the scans' barriers)
#pragma unroll
        for (int e = 0; e < E; ++e) sf[lds_slot(tid * E + e)] = fe[e];
        lds_barrier();
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const int idx = w0 + tid + T * e;
            if (idx >= g0 && idx < g1) store_out(idx, Arith<RT>::to_d(sf[lds_slot(tid + T * e)]));
        }
    } else if (g == 0) {
        // the whole row, tile by tile with exact carries; the forward result
        // goes to the output vector and is read back by the backward pass
        const int ntile = (N + TE - 1) / TE;
        C carry = czero<C>();
        for (int t = 0; t < ntile; ++t) {
            const int w0 = t * TE, w1 = min(N, w0 + TE);
            forward(w0, w1, 0, N, false, carry, ae, fe);
#pragma unroll
            for (int e = 0; e < E; ++e) {
                const int idx = w0 + tid * E + e;
                if (idx < w1) Y[base_j + idx] = Arith<RT>::to_d(fe[e]);
            }
        }
        __syncthreads();  // (with its global fence: the block's forward results are visible)
        carry = czero<C>();
        for (int t = ntile - 1; t >= 0; --t) {
            const int w0 = t * TE, w1 = min(N, w0 + TE);
#pragma unroll
            for (int e = 0; e < E; ++e) {
                const int idx = w0 + tid * E + e;
                const size_t q = base_j + min(idx, w1 - 1);
                C iv;
                if constexpr (F32) iv = a.inv32[q];
                else iv = a.inv[q];
                fe[e] = Arith<RT>::from(Y[q]);
                ae[e] = (idx >= 1 && idx <= N - 2) ? coef(iv) : czero<C>();
            }
            backward(w0, w1, false, carry, ae, fe);
            // (every thread of the block read its forward values before the scan's barriers)
#pragma unroll
            for (int e = 0; e < E; ++e) {
                const int idx = w0 + tid * E + e;
                if (idx < w1) store_out(idx, Arith<RT>::to_d(fe[e]));
            }
        }
    }
    if constexpr (WCOL) {
        // the column's exact non-zero node range over all modes: one atomic
        // pair per block that stored a non-zero (far segments store zeros)
        __shared__ int nzw[T / 64][2];
        for (int o = 32; o > 0; o >>= 1) {
            nz_lo = min(nz_lo, __shfl_xor(nz_lo, o));
            nz_hi = max(nz_hi, __shfl_xor(nz_hi, o));
        }
        if ((tid & 63) == 0) {
            nzw[tid >> 6][0] = nz_lo;
            nzw[tid >> 6][1] = nz_hi;
        }
        __syncthreads();
        if (tid == 0) {
            for (int w = 1; w < T / 64; ++w) {
                nz_lo = min(nz_lo, nzw[w][0]);
                nz_hi = max(nz_hi, nzw[w][1]);
            }
            if (nz_lo < nz_hi) {
                atomicMin(a.range + 2 * col, nz_lo);
                atomicMax(a.range + 2 * col + 1, nz_hi);
            }
        }
        return;
    }
    // Woodbury: the holding blocks (the first K count of the grid) store
    // their partial projections v_a^H y in a fixed order; the correction
    // kernel that follows sums them (no block of this launch waits for
    // another)
    if (!a.wood || B.id >= a.K * count) return;
    __shared__ double ws[T / 64][4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const double v = wave_sum(wp[q]);
        if ((tid & 63) == 0) ws[tid >> 6][q] = v;
    }
    __syncthreads();
    if (tid == 0) {
        for (int q = 0; q < 4; ++q) {
            double v = 0.0;
            for (int w = 0; w < T / 64; ++w) v += ws[w][q];
            a.wpart[(size_t)q * a.K * count + B.id] = v;
        }
        if (B.id == 0) Kr->wparts = a.K * count;
    }
}
template <int VAR, typename RT>
__global__ __launch_bounds__(gmd::PT) void k_psweep(PrecArgs a) { k_psweep_body<VAR, RT>(a, grid_blk()); }
// One launch over many solvers: member m owns blocks [off[m], off[m + 1]).
template <int VAR, typename RT>
__global__ __launch_bounds__(gmd::PT) void k_psweep_batch(const PrecArgs *__restrict__ args, const int *__restrict__ off, int M)
{
    const int m = batch_member(off, M);
    const PrecArgs a = args[m];
    k_psweep_body<VAR, RT>(a, Blk{(int)blockIdx.x - __builtin_amdgcn_readfirstlane(off[m]),
                                  __builtin_amdgcn_readfirstlane(off[m + 1] - off[m])});
}

// ---------------------------------------------------------------- BiCGSTAB vector update

__device__ void loop_top(Krylov *Kr, double r2, double2 rho_new)
{
    // while (r^2 > tol^2 |b|^2 && i < maxit): rho_old = rho; rho = r0.r; restart check; beta
    Kr->r2 = r2;
    if (!(r2 > Kr->tol2 && Kr->iter < Kr->maxit)) {
        Kr->done = 1;
        return;
    }
    Kr->rho_old = Kr->rho;
    Kr->rho = rho_new;
    const double thr = Kr->eigen_mode ? Kr->rst_eps2 * Kr->r0_sq : Kr->restart_thr;
    if (sqrt(cnorm2(rho_new)) < thr) {
        Kr->restart = 1;
        Kr->done = 1;
        return;
    }
    Kr->beta = cmul(cdiv(Kr->rho, Kr->rho_old), cdiv(Kr->alpha, Kr->w));
}

__device__ __forceinline__ void k_update_body(const UpdateArgs &a, Blk B)
{
    Krylov *Kr = a.Kr;
    if (Kr->done) return;
    const double2 al = Kr->alpha, w = Kr->w;
    double v[3] = {0, 0, 0};
    for (size_t q = (size_t)B.id * blockDim.x + threadIdx.x; q < a.n; q += (size_t)B.n * blockDim.x) {
        const double2 x = cadd(ntld(a.X + q), cadd(cmul(al, ntld(a.Yc + q)), cmul(w, ntld(a.Zc + q))));
        ntst(a.X + q, x);
        const double2 r = csub(ntld(a.S + q), cmul(w, ntld(a.T + q)));
        ntst(a.R + q, r);
        const double2 d = cdotm(ntld(a.R0 + q), r);
        v[0] += d.x;
        v[1] += d.y;
        v[2] += cnorm2(r);
    }
    if (block_partials<3>(v, a.partial, &Kr->counter, B)) {
        const double rr = sum_partials(a.partial, 0, B), ri = sum_partials(a.partial, 1, B);
        const double r2 = sum_partials(a.partial, 2, B);
        if (threadIdx.x == 0) {
            Kr->iter += 1;
            Kr->fresh = 0;
            Kr->pb ^= 1;  // the p update of this iteration wrote the other buffer
            loop_top(Kr, r2, make_double2(rr, ri));
            Kr->counter = 0;
        }
    }
}
__global__ __launch_bounds__(256) void k_update(UpdateArgs a) { k_update_body(a, grid_blk()); }
// One launch over many solvers: member m owns blocks [off[m], off[m + 1]).
__global__ __launch_bounds__(256) void k_update_batch(const UpdateArgs *__restrict__ args, const int *__restrict__ off, int M)
{
    const int m = batch_member(off, M);
    const UpdateArgs a = args[m];
    k_update_body(a, Blk{(int)blockIdx.x - __builtin_amdgcn_readfirstlane(off[m]),
                             __builtin_amdgcn_readfirstlane(off[m + 1] - off[m])});
}

// Start of the normalised solve (Eigen BiCGSTAB on A x = b, b = rhs/|rhs|,
// guess x = psi/|rhs|): the rhs apply left |rhs|, the raw shadow residual
// r0' = rhs - A psi in R0 and |r0'|^2. The shadow residual enters BiCGSTAB
// only through ratios of inner products, so it stays unnormalised (r0' =
// |rhs| r0): rho and r0.v carry the factor |rhs|, which cancels in alpha and
// beta, and the restart thresholds are scaled by it. r = r0'/|rhs| is formed
// where the first iteration reads it; p = v = 0 are never read.
__device__ __forceinline__ void k_init_body(const InitArgs &a, Blk B)
{
    Krylov *Kr = a.Kr;
    const double nrm = Kr->rhs_nrm;
    if (nrm != 0.0) {
        for (size_t q = (size_t)B.id * blockDim.x + threadIdx.x; q < a.n; q += (size_t)B.n * blockDim.x) {
            const double2 x = a.Psi[q];
            a.X[q] = make_double2(x.x / nrm, x.y / nrm);
        }
    }
    if (B.id == 0 && threadIdx.x == 0) {
        // |r|^2 and |b|^2 of the normalised system
        const double r2 = nrm != 0.0 ? Kr->r0_raw2 / (nrm * nrm) : 0.0;
        const double b2 = nrm != 0.0 ? 1.0 : 0.0;
        const double rho0 = nrm != 0.0 ? Kr->r0_raw2 / nrm : 0.0;  // r0'.r = |rhs| |r|^2
        Kr->counter = 0;
        Kr->iter = 0;
        Kr->maxit = a.maxit;
        Kr->eigen_mode = a.eigen_mode;
        Kr->restarts = 0;
        Kr->restart = 0;
        Kr->fresh = 1;
        Kr->pb = 0;
        Kr->rst_eps2 = a.rst_eps2;
        Kr->rst_ref = a.rst_ref;
        Kr->accepted = 0;
        Kr->done = 0;
        Kr->zero_rhs = nrm == 0.0;
        Kr->rho = make_double2(1.0, 0.0);
        Kr->alpha = make_double2(1.0, 0.0);
        Kr->w = make_double2(1.0, 0.0);
        Kr->r0_sq = rho0;               // |rhs| |r0|^2: Eigen's eps^2 |r0|^2 test on the scaled rho
        Kr->restart_thr = rho0 * a.rst_ref;  // fixed at the start (matrices.cpp:56)
        Kr->tol2 = a.tol * a.tol * b2;
        if (nrm == 0.0) {
            Kr->done = 1;
        } else {
            loop_top(Kr, r2, make_double2(rho0, 0.0));
        }
    }
}
__global__ __launch_bounds__(256) void k_init(InitArgs a) { k_init_body(a, grid_blk()); }
// One launch over many solvers: member m owns blocks [off[m], off[m + 1]).
__global__ __launch_bounds__(256) void k_init_batch(const InitArgs *__restrict__ args, const int *__restrict__ off, int M)
{
    const int m = batch_member(off, M);
    const InitArgs a = args[m];
    k_init_body(a, Blk{(int)blockIdx.x - __builtin_amdgcn_readfirstlane(off[m]),
                             __builtin_amdgcn_readfirstlane(off[m + 1] - off[m])});
}

// Acceptance (matrices.cpp:118-121,166-170): psi = mask * |rhs| x when
// |A x - b| <= 1e-9, or 0 for a zero right-hand side.
__device__ __forceinline__ void k_accept_body(const AcceptArgs &a, Blk B)
{
    const Krylov *Kr = a.Kr;
    const double nrm = Kr->rhs_nrm;
    const double res = sqrt(Kr->res2);
    const bool zero = Kr->zero_rhs;
    const bool ok = zero || res <= 1e-9;
    if (B.id == 0 && threadIdx.x == 0) {
        a.status[0] = ok ? 1 : 0;
        a.resid[0] = zero ? 0.0 : res;
    }
    if (!ok) return;
    for (size_t q = (size_t)B.id * blockDim.x + threadIdx.x; q < a.n; q += (size_t)B.n * blockDim.x) {
        double2 x = zero ? make_double2(0.0, 0.0) : cscale(nrm, a.X[q]);
        if (a.mask) x = cscale(a.mask[q % a.N], x);
        a.Psi[q] = x;
    }
}
__global__ __launch_bounds__(256) void k_accept(AcceptArgs a) { k_accept_body(a, grid_blk()); }
// One launch over many solvers: member m owns blocks [off[m], off[m + 1]).
__global__ __launch_bounds__(256) void k_accept_batch(const AcceptArgs *__restrict__ args, const int *__restrict__ off, int M)
{
    const int m = batch_member(off, M);
    const AcceptArgs a = args[m];
    k_accept_body(a, Blk{(int)blockIdx.x - __builtin_amdgcn_readfirstlane(off[m]),
                             __builtin_amdgcn_readfirstlane(off[m + 1] - off[m])});
}

// ---------------------------------------------------------------- sinks

// One workgroup. mode SK_PROJ: out[a] = scale * v_a^H X (stored per the dest
// selector); SK_SETUP: Woodbury capacitance G = diag(1/((dt/4) g)) + V^H W and its inverse.
__device__ __forceinline__ void k_sink_body(const SinkArgs &a, Blk B)
{
    Krylov *Kr = a.Kr;
    const StepParams *P = a.P;
    const int ns = P->nsink;
    const int tid = threadIdx.x;
    // loops over the at most two sinks have constant bounds, so the
    // accumulators stay in registers
    double v[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    const int N = a.N;
#pragma unroll
    for (int s = 0; s < 2; ++s) {
        if (s >= ns) break;
        const SinkDev &sk = a.sk[s];
        const int tot = a.K * sk.n;
        for (int q = tid; q < tot; q += blockDim.x) {
            const int k = q / sk.n, d = q - k * sk.n;
            const size_t gi = (size_t)k * N + sk.lo + d;
            if (a.mode == SK_PROJ) {
                const double2 d2 = cdotm(sk.v[q], a.X[gi]);
                v[2 * s] += d2.x;
                v[2 * s + 1] += d2.y;
            } else {
                const double2 rot = P->wrot[mode_m(k) + a.lmax];
#pragma unroll
                for (int u = 0; u < 2; ++u) {
                    if (u >= ns) break;
                    const double2 d2 = cdotm(sk.v[q], cmul(rot, a.sk[u].W[gi]));
                    v[4 * s + 2 * u] += d2.x;
                    v[4 * s + 2 * u + 1] += d2.y;
                }
            }
        }
    }
    __shared__ double ws[16][8];
    __shared__ double red[8];
    for (int q = 0; q < 8; ++q) {
        const double s = wave_sum(v[q]);
        if ((tid & 63) == 0) ws[tid >> 6][q] = s;
    }
    __syncthreads();
    if (tid != 0) return;
    for (int q = 0; q < 8; ++q) {
        double s = 0;
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) s += ws[w][q];
        red[q] = s;
    }
    if (a.mode == SK_PROJ) {
        for (int s = 0; s < ns; ++s) {
            const double sc = a.scale * (a.scale_by_nrm ? Kr->rhs_nrm : 1.0);
            const double2 pr = cscale(sc, make_double2(red[2 * s], red[2 * s + 1]));
            if (a.dest == SD_START) {
                Kr->qold[s] = pr;
                Kr->gq[s] = cscale(P->g[s], pr);
            } else if (a.dest == SD_QNEW) {
                Kr->qnew[s] = pr;
            } else if (a.dest == SD_POST) {
                Kr->qpost[s] = pr;
                Kr->gq[s] = cscale(P->g[s], pr);
            } else if (a.dest == SD_COEFA_QNEW) {
                // one projection of the accepted iterate x: the rank-one
                // coefficients of the residual apply and |rhs| v^H x
                const double2 raw = make_double2(red[2 * s], red[2 * s + 1]);
                Kr->coefA[s] = cscale(0.5 * P->half_dt * P->g[s], raw);
                Kr->qnew[s] = cscale(Kr->rhs_nrm, raw);
            } else {
                Kr->coefA[s] = cscale(0.5 * P->half_dt * P->g[s], pr);
            }
        }
    } else {
        double2 G[4];
#pragma unroll
        for (int s = 0; s < 2; ++s)
#pragma unroll
            for (int u = 0; u < 2; ++u) {
                G[s * 2 + u] = make_double2(red[4 * s + 2 * u], red[4 * s + 2 * u + 1]);
                if (s == u && s < ns) G[s * 2 + u].x += 1.0 / (0.5 * P->half_dt * P->g[s]);
            }
        for (int q = 0; q < 4; ++q) Kr->G[q] = G[q];
        if (ns == 1) {
            Kr->Ginv[0] = cdiv(make_double2(1.0, 0.0), G[0]);
        } else if (ns == 2) {
            const double2 det = csub(cmul(G[0], G[3]), cmul(G[2], G[1]));
            const double2 id = cdiv(make_double2(1.0, 0.0), det);
            Kr->Ginv[0] = cmul(G[3], id);
            Kr->Ginv[1] = cmul(make_double2(-G[1].x, -G[1].y), id);
            Kr->Ginv[2] = cmul(make_double2(-G[2].x, -G[2].y), id);
            Kr->Ginv[3] = cmul(G[0], id);
        }
    }
}
__global__ __launch_bounds__(1024) void k_sink(SinkArgs a) { k_sink_body(a, grid_blk()); }
// One launch over many solvers: member m owns blocks [off[m], off[m + 1]).
__global__ __launch_bounds__(1024) void k_sink_batch(const SinkArgs *__restrict__ args, const int *__restrict__ off, int M)
{
    const int m = batch_member(off, M);
    const SinkArgs a = args[m];
    k_sink_body(a, Blk{(int)blockIdx.x - __builtin_amdgcn_readfirstlane(off[m]),
                             __builtin_amdgcn_readfirstlane(off[m + 1] - off[m])});
}

// Woodbury correction of a preconditioned vector, in place over the node
// range where the columns W_a can be non-zero (matrices.cpp:151):
// y -= sum_a c_a W_a.
// Woodbury coefficients (matrices.cpp:146-153), one block of 256 threads
// between a preconditioner apply and its correction: c = Ginv (V^H y) from the
// preconditioner's partial projections, summed in a fixed order. The rank-one
// coefficients of A applied to the corrected y - W c, (dt/4) g_a v_a^H
// (y - W c), equal c: with G = D + V^H W and D = diag(1/((dt/4) g)),
// V^H (y - W c) = V^H y - (G - D) Ginv V^H y = D c.
__device__ __forceinline__ void k_wcoef_body(const WcorrArgs &a)
{
    Krylov *Kr = a.Kr;
    if (Kr->done) return;
    const int ns = a.nsink;
    // a contiguous slice of the partials per thread, then a fixed-order
    // combine across the block
    __shared__ double ws[4][4];
    const int nb = Kr->wparts, per = (nb + 255) / 256, b0 = threadIdx.x * per;
    double pr[4] = {0.0, 0.0, 0.0, 0.0};
    for (int b = b0; b < b0 + per && b < nb; ++b)
#pragma unroll
        for (int q = 0; q < 4; ++q) pr[q] += a.wpart[(size_t)q * nb + b];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const double v = wave_sum(pr[q]);
        if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6][q] = v;
    }
    __syncthreads();
    if (threadIdx.x != 0) return;
    for (int q = 0; q < 4; ++q) pr[q] = ws[0][q] + ws[1][q] + ws[2][q] + ws[3][q];
    const double2 proj[2] = {make_double2(pr[0], pr[1]), make_double2(pr[2], pr[3])};
    for (int s = 0; s < ns; ++s) {
        double2 c = make_double2(0.0, 0.0);
        for (int u = 0; u < ns; ++u) c = cadd(c, cmul(Kr->Ginv[s * 2 + u], proj[u]));
        Kr->c[s] = c;
        Kr->coefA[s] = c;
    }
}
__global__ __launch_bounds__(256) void k_wcoef(WcorrArgs a) { k_wcoef_body(a); }
// One launch over many solvers: member m owns block off[m] (one block each).
__global__ __launch_bounds__(256) void k_wcoef_batch(const WcorrArgs *__restrict__ args, const int *__restrict__ off, int M)
{
    k_wcoef_body(args[batch_member(off, M)]);
}

// y -= W c (the stored columns rotated to the step's angle), and the phased copy
__device__ __forceinline__ void k_wcorr_body(const WcorrArgs &a, Blk B)
{
    const Krylov *Kr = a.Kr;
    if (Kr->done) return;
    const int ns = a.nsink;
    const int lo = min(Kr->wrange[0], ns > 1 ? Kr->wrange[2] : a.N);
    const int hi = max(Kr->wrange[1], ns > 1 ? Kr->wrange[3] : 0);
    if (hi <= lo) return;
    const int width = hi - lo;
    const double2 c0 = Kr->c[0], c1 = Kr->c[1];
    const size_t n = (size_t)width * a.K;
    for (size_t q = (size_t)B.id * blockDim.x + threadIdx.x; q < n; q += (size_t)B.n * blockDim.x) {
        const int k = (int)(q / width), i = lo + (int)(q - (size_t)k * width);
        const size_t gi = (size_t)k * a.N + i;
        // stored columns W(theta_W); W(theta) = exp(-i m (theta - theta_W)) W(theta_W)
        double2 w = cmul(c0, a.W0[gi]);
        if (ns > 1) w = cadd(w, cmul(c1, a.W1[gi]));
        const int mk = mode_m(k);
        const double2 y = csub(a.Y[gi], cmul(a.P->wrot[mk + a.lmax], w));
        a.Y[gi] = y;
        if (a.Xt) store_phased(a.Xt, a.xt_f32, gi, a.P->ph[mk + a.lmax], y);
    }
}
__global__ __launch_bounds__(256) void k_wcorr(WcorrArgs a) { k_wcorr_body(a, grid_blk()); }
// One launch over many solvers: member m owns blocks [off[m], off[m + 1]).
__global__ __launch_bounds__(256) void k_wcorr_batch(const WcorrArgs *__restrict__ args, const int *__restrict__ off, int M)
{
    const int m = batch_member(off, M);
    const WcorrArgs a = args[m];
    k_wcorr_body(a, Blk{(int)blockIdx.x - __builtin_amdgcn_readfirstlane(off[m]),
                             __builtin_amdgcn_readfirstlane(off[m + 1] - off[m])});
}

// ---------------------------------------------------------------- apply_sink on device vectors

// One wavefront per mode k: the projections v_a^H X over the mode's part of
// each support, stored per mode; the last wavefront to finish sums the modes
// in order (fixed order, so the result does not depend on timing) and forms
// coef_a = scale_a (v_a^H X).
__global__ __launch_bounds__(64) void k_sink_project(SinkApplyArgs a)
{
    const int k = blockIdx.x, lane = threadIdx.x;
    SinkScratch *scr = a.scr;
    double v[4] = {0, 0, 0, 0};
#pragma unroll
    for (int s = 0; s < 2; ++s) {
        if (s >= a.nsink) break;
        const SinkDev &sk = a.sk[s];
        for (int d = lane; d < sk.n; d += 64) {
            const double2 t = cdotm(sk.v[(size_t)k * sk.n + d], a.X[(size_t)k * a.N + sk.lo + d]);
            v[2 * s] += t.x;
            v[2 * s + 1] += t.y;
        }
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) v[q] = wave_sum(v[q]);
    __shared__ bool last;
    if (lane == 0) {
#pragma unroll
        for (int q = 0; q < 4; ++q)
            __hip_atomic_store(&scr->part[k][q], v[q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        // the same ordering argument as block_partials
        __atomic_signal_fence(__ATOMIC_SEQ_CST);
        __builtin_amdgcn_s_waitcnt(0);
        __atomic_signal_fence(__ATOMIC_SEQ_CST);
        last = __hip_atomic_fetch_add(&scr->counter, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) ==
               gridDim.x - 1;
    }
    __syncthreads();
    if (!last) return;
    __threadfence();
    double t[4] = {0, 0, 0, 0};
    // contiguous slice of modes per lane, then a fixed-order wave sum
    const int per = (a.K + 63) / 64, k0 = lane * per;
    for (int kk = k0; kk < k0 + per && kk < a.K; ++kk)
#pragma unroll
        for (int q = 0; q < 4; ++q) t[q] += vload(&scr->part[kk][q]);
#pragma unroll
    for (int q = 0; q < 4; ++q) t[q] = wave_sum(t[q]);
    if (lane == 0) {
        for (int s = 0; s < 2; ++s)
            scr->coef[s] = s < a.nsink ? cscale(a.scale[s], make_double2(t[2 * s], t[2 * s + 1]))
                                       : make_double2(0.0, 0.0);
        scr->counter = 0;
    }
}

// One workgroup per mode: Y (or zero) + coef_0 v_0 + coef_1 v_1, the second
// support after the first (they may share nodes).
__global__ __launch_bounds__(256) void k_sink_expand(SinkApplyArgs a)
{
    const int k = blockIdx.x;
    double2 *Yk = a.Y + (size_t)k * a.N;
    if (a.zero_base)
        for (int i = threadIdx.x; i < a.N; i += blockDim.x) Yk[i] = make_double2(0.0, 0.0);
    __syncthreads();
    for (int s = 0; s < a.nsink; ++s) {
        const SinkDev &sk = a.sk[s];
        const double2 c = vload2(&a.scr->coef[s]);
        for (int d = threadIdx.x; d < sk.n; d += blockDim.x) {
            const int i = sk.lo + d;
            Yk[i] = cadd(Yk[i], cmul(c, sk.v[(size_t)k * sk.n + d]));
        }
        __syncthreads();
    }
}

__global__ void k_cscale(double2 *y, size_t n, double2 c)
{
    for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x)
        y[q] = cmul(c, y[q]);
}

// ---------------------------------------------------------------- absorber profiles and projectors

// update_absorption_profile (absorption.cpp:12-81) for one support node per
// block (the first nb0 blocks: body 0): the cap of half-width `width` around
// the body seen from radius r, integrated over z = 1 - cos(gamma) with the
// Gauss-Legendre rule, multipole by multipole,
//   mult_l(r) = sum_k 2 pi capw w_k exp(-1 / (1 - dist^2 / width^2)) P_l(1 - z_k).
// Thread k evaluates quadrature point k and its Legendre recurrence; the
// terms of 32 multipoles at a time meet in LDS and thread l adds its row in
// quadrature order, so each sum has the order of the host restatement.
__global__ __launch_bounds__(ABS_MAX_NQ) void k_abs_profile(AbsorberArgs a, int nb0)
{
#pragma clang fp contract(off)
    const int b = (int)blockIdx.x < nb0 ? 0 : 1;
    const int d = (int)blockIdx.x - (b ? nb0 : 0);
    const AbsorberBody &B = a.b[b];
    const int n = B.n, degree = 2 * a.lmax, k = threadIdx.x;
    const double rv = a.r[B.lo + d], radius = B.radius, width = B.width;
    const double delta = rv - radius;
    double *mult = B.mult;
    if (fabs(delta) >= width || radius == 0.0) {
        // outside the cap: no weight; the body at the origin: the s-wave only
        for (int l = k; l <= degree; l += blockDim.x) {
            double v = 0.0;
            if (l == 0 && fabs(delta) < width) {
                const double xx = rv / width;
                v = 4.0 * M_PI * exp(-1.0 / (1.0 - xx * xx));
            }
            mult[(size_t)l * n + d] = v;
        }
        return;
    }
    const double product = 2.0 * rv * radius;
    const double capw = fmin(2.0, (width * width - delta * delta) / product);
    double weight = 0.0, u = 0.0;
    if (k < a.nq) {
        const double z = capw * a.glx[k];
        u = 1.0 - z;
        const double dist2 = delta * delta + product * z;
        weight = 2.0 * M_PI * capw * a.glw[k] * exp(-1.0 / (1.0 - dist2 / (width * width)));
    }
    constexpr int LG = 32;
    __shared__ double term[LG][ABS_MAX_NQ];
    double previous = 1.0, current = u;
    for (int l0 = 0; l0 <= degree; l0 += LG) {
        for (int q = 0; q < LG && l0 + q <= degree; ++q) {
            const int l = l0 + q;
            double pl;
            if (l == 0) {
                pl = 1.0;
            } else if (l == 1) {
                pl = current;
            } else {
                const double next = ((2 * l - 1) * u * current - (l - 1) * previous) / l;
                previous = current;
                current = next;
                pl = next;
            }
            term[q][k] = l == 0 ? weight : weight * pl;
        }
        __syncthreads();
        if (k < LG && l0 + k <= degree) {
            double acc = 0.0;
            for (int q = 0; q < a.nq; ++q) acc += term[k][q];
            mult[(size_t)(l0 + k) * n + d] = acc;
        }
        __syncthreads();
    }
}

// build_swave_projector (absorption.cpp:127-151): block (j, body) writes row
// j of the body's projector,
//   v_j(r) = sqrt(4 pi) r (mult_{l_j}(r) / norm) c_j exp(-i m_j theta) (x (-1)^m_j for the secondary),
// with norm = sum_d dr r^2 mult_0(r) (each block adds it in node order) and
// c_j the angular coefficient of the (j, 00) pair.
__global__ __launch_bounds__(64) void k_abs_project(AbsorberArgs a)
{
#pragma clang fp contract(off)
    const AbsorberBody &B = a.b[blockIdx.y];
    const int j = blockIdx.x, n = B.n;
    if (n == 0) return;
    double norm = 0.0;
    for (int d = 0; d < n; ++d) norm += a.dr * a.r[B.lo + d] * a.r[B.lo + d] * B.mult[d];
    const int l = mode_l(j), ms = -mode_m(j);
    const double ang = a.ang[j];
    const double2 ph = a.P->ph[ms + a.lmax];
    const double parity = B.secondary && ms % 2 != 0 ? -1.0 : 1.0;
    const double sqrt4pi = sqrt(4.0 * M_PI);
    for (int d = threadIdx.x; d < n; d += blockDim.x) {
        double2 val = make_double2(0.0, 0.0);
        if (ang != 0.0) {
            double mv = B.mult[(size_t)l * n + d];
            if (norm > 0.0) mv /= norm;
            const double m = mv * ang;
            const double sc = a.r[B.lo + d] * sqrt4pi * parity;
            val = make_double2(sc * (m * ph.x), sc * (m * ph.y));
        }
        B.v[(size_t)j * n + d] = val;
    }
}

__global__ void k_sink_scratch_clear(SinkScratch *scr)
{
    if (threadIdx.x == 0) scr->counter = 0;
}

// ---------------------------------------------------------------- diagnostics

// One wavefront per mode: sums of the per-tile partials of the diagnostics
// apply, the flux at r_ex, and the mode's share of the global sums.
__device__ __forceinline__ void k_diag_final_body(const DiagArgs &a, Blk B)
{
    const int j = B.id, lane = threadIdx.x;
    const int K = a.K, N = a.N, ix = a.ix, nt = a.ntiles;
    const double dr = a.dr, rix = a.eps + ix * dr;
    double v[6] = {0, 0, 0, 0, 0, 0};
    // contiguous tile ranges per lane keep the summation order fixed
    const int per = (nt + 63) / 64;
    for (int t = lane * per; t < lane * per + per && t < nt; ++t) {
        const double *rp = a.rowpart + ((size_t)t * K + j) * 6;
        for (int q = 0; q < 6; ++q) v[q] += rp[q];
    }
    for (int q = 0; q < 6; ++q) v[q] = wave_sum(v[q]);
    if (lane == 0) {
        const double2 um = a.Psi[(size_t)j * N + ix - 1], u0 = a.Psi[(size_t)j * N + ix],
                      up = a.Psi[(size_t)j * N + ix + 1];
        const double2 dU = csub(cscale(1.0 / (2.0 * dr), csub(up, um)), make_double2(u0.x / rix, u0.y / rix));
        a.out[j] = cdotm(u0, dU).y;
        a.out[K + j] = cdotm(a.dpsi_ix[j], dU).x / a.mu;
        a.out[2 * K + j] = v[0] * dr;
        a.out[3 * K + j] = v[1] * dr;
        for (int q = 0; q < 4; ++q) a.out[(4 + q) * K + j] = v[2 + q];
    }
}
__global__ __launch_bounds__(64) void k_diag_final(DiagArgs a) { k_diag_final_body(a, grid_blk()); }
// One launch over many solvers: member m owns blocks [off[m], off[m + 1]).
__global__ __launch_bounds__(64) void k_diag_final_batch(const DiagArgs *__restrict__ args, const int *__restrict__ off, int M)
{
    const int m = batch_member(off, M);
    const DiagArgs a = args[m];
    k_diag_final_body(a, Blk{(int)blockIdx.x - __builtin_amdgcn_readfirstlane(off[m]),
                             __builtin_amdgcn_readfirstlane(off[m + 1] - off[m])});
}

// ---------------------------------------------------------------- launchers

int apply_grid(int N, int nchunks, int ng, int *ngroups)
{
    const int tn = 16 * ng;
    const int ntiles = (N + tn - 1) / tn;
    *ngroups = (nchunks + gmd::AW - 1) / gmd::AW;
    return (ntiles + 7) / 8 * 8 * *ngroups;
}

// Column vectors of the coupling, exp(i m theta) x in the arithmetic of the
// apply (FP64, or FP32 for the mixed Krylov applies): nbx blocks per mode.
__device__ __forceinline__ void k_phase_body(const PhaseArgs &a, Blk B)
{
    const int j = B.id / a.nbx, bx = B.id - j * a.nbx;
    const int kk = (int)(sqrt(8.0 * j + 1.0) - 1.0) / 2;
    const int l = (kk + 1) * (kk + 2) / 2 <= j ? kk + 1 : kk;
    const int m = 2 * (j - l * (l + 1) / 2) - l;
    const double2 ph = a.P->ph[m + a.lmax];
    const size_t base = (size_t)j * a.N;
    for (int i = bx * blockDim.x + threadIdx.x; i < a.N; i += a.nbx * blockDim.x) {
        const double2 v = cmul(ph, a.X[base + i]);
        if (a.f32) static_cast<float2 *>(a.Xt)[base + i] = make_float2((float)v.x, (float)v.y);
        else static_cast<double2 *>(a.Xt)[base + i] = v;
    }
}
__global__ __launch_bounds__(256) void k_phase(PhaseArgs a) { k_phase_body(a, grid_blk()); }
__global__ __launch_bounds__(256) void k_phase_batch(const PhaseArgs *__restrict__ args, const int *__restrict__ off, int M)
{
    const int m = batch_member(off, M);
    const PhaseArgs a = args[m];
    k_phase_body(a, Blk{(int)blockIdx.x - __builtin_amdgcn_readfirstlane(off[m]),
                             __builtin_amdgcn_readfirstlane(off[m + 1] - off[m])});
}

template <int NG>
static hipError_t launch_apply_ng(const ApplyArgs &a, int grid, hipStream_t st)
{
    if (a.mixed)
        hipLaunchKernelGGL((k_apply<NG, float>), dim3(grid), dim3(64 * gmd::AW), 0, st, a);
    else
        hipLaunchKernelGGL((k_apply<NG, double>), dim3(grid), dim3(64 * gmd::AW), 0, st, a);
    return hipGetLastError();
}

PhaseArgs phase_args(const ApplyArgs &a)
{
    PhaseArgs p{};
    p.X = a.X;
    p.Xt = a.Xt;
    p.N = a.g.N;
    p.lmax = a.g.lmax;
    p.nbx = (a.g.N + 255) / 256 < 16 ? (a.g.N + 255) / 256 : 16;
    p.f32 = a.mixed;
    p.P = a.P;
    return p;
}

hipError_t launch_apply(ApplyArgs a, int mode, hipStream_t st)
{
    if (a.ng == TAPPLY_NG) {
        // the polar form reads X and applies the phases itself
        a.mode = mode;
        return launch_tapply(a, st);
    }
    if (a.ng != 1 && a.ng != 2 && a.ng != 4) return hipErrorInvalidValue;
    if (!a.Xt) return hipErrorInvalidValue;
    const int grid = apply_grid(a.g.N, a.g.nchunks, a.ng, &a.ngroups);
    a.mode = mode;
    if (!a.xt_ready) {
        const PhaseArgs pa = phase_args(a);
        hipLaunchKernelGGL(k_phase, dim3(pa.nbx * a.g.K), dim3(256), 0, st, pa);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    if (a.ng == 1) return launch_apply_ng<1>(a, grid, st);
    if (a.ng == 2) return launch_apply_ng<2>(a, grid, st);
    return launch_apply_ng<4>(a, grid, st);
}

hipError_t launch_diag_apply(const ApplyArgs &a, hipStream_t st)
{
    const int items = a.g.K * ((a.g.N + gmd::DIAG_NODES - 1) / gmd::DIAG_NODES);
    hipLaunchKernelGGL(k_diag_apply, dim3((items + 3) / 4), dim3(256), 0, st, a);
    return hipGetLastError();
}

hipError_t launch_build_E(int N, int Ls, double dr, double eps, double mu, const StepParams *P, double *E, float *E32,
                          hipStream_t st)
{
    const int n = Ls * N;
    hipLaunchKernelGGL(k_build_E, dim3((n + 255) / 256), dim3(256), 0, st, N, Ls, dr, eps, mu, P, E, E32);
    return hipGetLastError();
}

hipError_t launch_lcut(int N, int Ls, double tau, const double *E, int *lcut, hipStream_t st)
{
    const int ntiles = (N + gmd::TILE - 1) / gmd::TILE;
    hipLaunchKernelGGL(k_lcut, dim3(ntiles), dim3(64), 0, st, N, Ls, tau, E, lcut);
    return hipGetLastError();
}

hipError_t launch_factor(const FactorArgs &f, hipStream_t st)
{
    // the segments' pivots and checks, then the whole-row factor of the
    // flagged modes (without segment flags: every mode whole)
    if (f.pseq) hipLaunchKernelGGL(k_factor_win, dim3(f.K * gmd::precond_nseg(f.N)), dim3(FT), 0, st, f);
    hipLaunchKernelGGL(k_factor, dim3(f.K), dim3(gmd::PT), 0, st, f);
    return hipGetLastError();
}

hipError_t launch_precond(const PrecArgs &a, int var, hipStream_t st)
{
    const bool f32 = a.inv32 != nullptr && var != PV_WCOL;
    if (a.nseg != gmd::precond_nseg(a.N)) return hipErrorInvalidValue;
    const dim3 g(precond_blocks(a.K, a.N));
    if (var == PV_P && f32) hipLaunchKernelGGL((k_psweep<PV_P, float>), g, dim3(gmd::PT), 0, st, a);
    else if (var == PV_S && f32) hipLaunchKernelGGL((k_psweep<PV_S, float>), g, dim3(gmd::PT), 0, st, a);
    else if (var == PV_P) hipLaunchKernelGGL((k_psweep<PV_P, double>), g, dim3(gmd::PT), 0, st, a);
    else if (var == PV_S) hipLaunchKernelGGL((k_psweep<PV_S, double>), g, dim3(gmd::PT), 0, st, a);
    else hipLaunchKernelGGL((k_psweep<PV_WCOL, double>), dim3(precond_blocks(a.K, a.N) * a.ncol), dim3(gmd::PT), 0, st, a);
    return hipGetLastError();
}

static int stream_grid(size_t n)
{
    size_t g = (n + 255) / 256;
    if (g > (size_t)gmd::RED_BLOCKS) g = gmd::RED_BLOCKS;
    return (int)g;
}

int precond_blocks(int K, int N) { return K * gmd::precond_nseg(N); }
int stream_blocks(size_t n) { return stream_grid(n); }
int phase_blocks(int N, int K) { return ((N + 255) / 256 < 16 ? (N + 255) / 256 : 16) * K; }
int diag_apply_blocks(int N, int K) { return (K * ((N + gmd::DIAG_NODES - 1) / gmd::DIAG_NODES) + 3) / 4; }

__global__ void k_krylov_gather(gmd::Krylov *const *kr, int *out, int M)
{
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= M) return;
    const gmd::Krylov *k = kr[m];
    out[3 * m] = k->done;
    out[3 * m + 1] = k->restart;
    out[3 * m + 2] = k->iter;
}

hipError_t launch_krylov_gather(gmd::Krylov *const *kr, int *out, int M, hipStream_t st)
{
    hipLaunchKernelGGL(k_krylov_gather, dim3((M + 63) / 64), dim3(64), 0, st, kr, out, M);
    return hipGetLastError();
}

template <int NG>
static void launch_apply_batch_ng(const BatchLaunch &b, hipStream_t st)
{
    const ApplyArgs *a = static_cast<const ApplyArgs *>(b.args);
    if (b.pool >= 0 && b.mixed)
        hipLaunchKernelGGL((k_apply_pooled<NG, float>), dim3(b.total), dim3(64 * gmd::AW), 0, st, b.pool, b.off, b.M);
    else if (b.pool >= 0)
        hipLaunchKernelGGL((k_apply_pooled<NG, double>), dim3(b.total), dim3(64 * gmd::AW), 0, st, b.pool, b.off, b.M);
    else if (b.mixed)
        hipLaunchKernelGGL((k_apply_batch<NG, float>), dim3(b.total), dim3(64 * gmd::AW), 0, st, a, b.off, b.M);
    else
        hipLaunchKernelGGL((k_apply_batch<NG, double>), dim3(b.total), dim3(64 * gmd::AW), 0, st, a, b.off, b.M);
}

hipError_t launch_batch(const BatchLaunch &b, hipStream_t st)
{
    if (b.total <= 0 || b.M < 1) return hipSuccess;
    const dim3 g(b.total);
    switch (b.kernel) {
    case BK_APPLY:
        if (b.ng == TAPPLY_NG)
            return launch_tapply_batch(static_cast<const ApplyArgs *>(b.args), b.off, b.M, b.total, b.var, st);
        if (b.ng == 1) launch_apply_batch_ng<1>(b, st);
        else if (b.ng == 2) launch_apply_batch_ng<2>(b, st);
        else if (b.ng == 4) launch_apply_batch_ng<4>(b, st);
        else return hipErrorInvalidValue;
        break;
    case BK_PRECOND: {
        const PrecArgs *a = static_cast<const PrecArgs *>(b.args);
        if (b.var == PV_P && b.f32) hipLaunchKernelGGL((k_psweep_batch<PV_P, float>), g, dim3(gmd::PT), 0, st, a, b.off, b.M);
        else if (b.var == PV_S && b.f32) hipLaunchKernelGGL((k_psweep_batch<PV_S, float>), g, dim3(gmd::PT), 0, st, a, b.off, b.M);
        else if (b.var == PV_P) hipLaunchKernelGGL((k_psweep_batch<PV_P, double>), g, dim3(gmd::PT), 0, st, a, b.off, b.M);
        else if (b.var == PV_S) hipLaunchKernelGGL((k_psweep_batch<PV_S, double>), g, dim3(gmd::PT), 0, st, a, b.off, b.M);
        else return hipErrorInvalidValue;
        break;
    }
    case BK_UPDATE:
        hipLaunchKernelGGL(k_update_batch, g, dim3(256), 0, st, static_cast<const UpdateArgs *>(b.args), b.off, b.M);
        break;
    case BK_INIT:
        hipLaunchKernelGGL(k_init_batch, g, dim3(256), 0, st, static_cast<const InitArgs *>(b.args), b.off, b.M);
        break;
    case BK_ACCEPT:
        hipLaunchKernelGGL(k_accept_batch, g, dim3(256), 0, st, static_cast<const AcceptArgs *>(b.args), b.off, b.M);
        break;
    case BK_SINK:
        hipLaunchKernelGGL(k_sink_batch, g, dim3(1024), 0, st, static_cast<const SinkArgs *>(b.args), b.off, b.M);
        break;
    case BK_WCOEF:
        hipLaunchKernelGGL(k_wcoef_batch, g, dim3(256), 0, st, static_cast<const WcorrArgs *>(b.args), b.off, b.M);
        break;
    case BK_WCORR:
        hipLaunchKernelGGL(k_wcorr_batch, g, dim3(256), 0, st, static_cast<const WcorrArgs *>(b.args), b.off, b.M);
        break;
    case BK_DIAG_APPLY:
        hipLaunchKernelGGL(k_diag_apply_batch, g, dim3(256), 0, st, static_cast<const ApplyArgs *>(b.args), b.off, b.M);
        break;
    case BK_DIAG_FINAL:
        hipLaunchKernelGGL(k_diag_final_batch, g, dim3(64), 0, st, static_cast<const DiagArgs *>(b.args), b.off, b.M);
        break;
    case BK_PHASE:
        hipLaunchKernelGGL(k_phase_batch, g, dim3(256), 0, st, static_cast<const PhaseArgs *>(b.args), b.off, b.M);
        break;
    default:
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_update(const UpdateArgs &a, hipStream_t st)
{
    hipLaunchKernelGGL(k_update, dim3(stream_grid(a.n)), dim3(256), 0, st, a);
    return hipGetLastError();
}

hipError_t launch_init(const InitArgs &a, hipStream_t st)
{
    hipLaunchKernelGGL(k_init, dim3(stream_grid(a.n)), dim3(256), 0, st, a);
    return hipGetLastError();
}

hipError_t launch_accept(const AcceptArgs &a, hipStream_t st)
{
    hipLaunchKernelGGL(k_accept, dim3(stream_grid(a.n)), dim3(256), 0, st, a);
    return hipGetLastError();
}

hipError_t launch_sink(const SinkArgs &a, hipStream_t st)
{
    hipLaunchKernelGGL(k_sink, dim3(1), dim3(1024), 0, st, a);
    return hipGetLastError();
}

hipError_t launch_diag_final(const DiagArgs &a, hipStream_t st)
{
    hipLaunchKernelGGL(k_diag_final, dim3(a.K), dim3(64), 0, st, a);
    return hipGetLastError();
}

// Static reduction scratch of the allocation-free apply_sink
// (gm_apply_sink_device): one SinkScratch per GPU, held for the life of the
// process. Host code never passes `scratch` itself to a kernel: its device
// address on each GPU comes from hipGetSymbolAddress, looked up once per
// device with that device current.
static __device__ __align__(256) unsigned char scratch[sizeof(SinkScratch)];

namespace {
constexpr int kMaxDevices = 64;
std::mutex scratch_mu;
SinkScratch *scratch_ptr[kMaxDevices] = {};
}  // namespace

hipError_t sink_scratch_address(SinkScratch **out)
{
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(scratch_mu);
    if (!scratch_ptr[dev]) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        scratch_ptr[dev] = static_cast<SinkScratch *>(p);
    }
    *out = scratch_ptr[dev];
    return hipSuccess;
}

namespace {
std::mutex pool_mu;
std::vector<char> pool_used[kMaxDevices];
ApplyArgs *pool_addr[kMaxDevices] = {};
}  // namespace

hipError_t apply_pool_acquire(int count, int *first)
{
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
    if (count < 1 || count > APPLY_POOL) return hipErrorOutOfMemory;
    std::lock_guard<std::mutex> lk(pool_mu);
    std::vector<char> &used = pool_used[dev];
    if (used.empty()) used.assign(APPLY_POOL, 0);
    // first fit
    int run = 0;
    for (int i = 0; i < APPLY_POOL; ++i) {
        run = used[i] ? 0 : run + 1;
        if (run == count) {
            const int f = i + 1 - count;
            for (int k = f; k <= i; ++k) used[k] = 1;
            *first = f;
            return hipSuccess;
        }
    }
    return hipErrorOutOfMemory;
}

void apply_pool_release(int device, int first, int count)
{
    if (device < 0 || device >= kMaxDevices || first < 0) return;
    std::lock_guard<std::mutex> lk(pool_mu);
    std::vector<char> &used = pool_used[device];
    for (int k = first; k < first + count && k < (int)used.size(); ++k) used[k] = 0;
}

hipError_t apply_pool_write(int first, const ApplyArgs *src, int count)
{
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
    if (first < 0 || count < 0 || first + count > APPLY_POOL) return hipErrorInvalidValue;
    ApplyArgs *base = nullptr;
    {
        std::lock_guard<std::mutex> lk(pool_mu);
        if (!pool_addr[dev]) {
            void *p = nullptr;
            e = hipGetSymbolAddress(&p, HIP_SYMBOL(apply_pool));
            if (e != hipSuccess) return e;
            pool_addr[dev] = static_cast<ApplyArgs *>(p);
        }
        base = pool_addr[dev];
    }
    return hipMemcpy(base + first, src, sizeof(ApplyArgs) * count, hipMemcpyHostToDevice);
}

hipError_t launch_sink_apply(const SinkApplyArgs &a, hipStream_t st)
{
    if (a.K < 1 || a.K > SINK_MAX_MODES || !a.scr || !a.X || !a.Y || a.nsink < 0 || a.nsink > 2)
        return hipErrorInvalidValue;
    // the counter starts at zero (a fill kernel, so the sequence may be captured)
    hipLaunchKernelGGL(k_sink_scratch_clear, dim3(1), dim3(64), 0, st, a.scr);
    if (a.nsink > 0) hipLaunchKernelGGL(k_sink_project, dim3(a.K), dim3(64), 0, st, a);
    hipLaunchKernelGGL(k_sink_expand, dim3(a.K), dim3(256), 0, st, a);
    return hipGetLastError();
}

hipError_t launch_absorbers(const AbsorberArgs &a, hipStream_t st)
{
    if (a.nq < 1 || a.nq > ABS_MAX_NQ || 2 * a.lmax + 1 > ABS_MAX_NQ) return hipErrorInvalidValue;
    const int nb0 = a.b[0].build ? a.b[0].n : 0, nb1 = a.b[1].build ? a.b[1].n : 0;
    if (nb0 + nb1 > 0) hipLaunchKernelGGL(k_abs_profile, dim3(nb0 + nb1), dim3(ABS_MAX_NQ), 0, st, a, nb0);
    if (a.b[0].n + a.b[1].n > 0) hipLaunchKernelGGL(k_abs_project, dim3(a.K, 2), dim3(64), 0, st, a);
    return hipGetLastError();
}

hipError_t launch_cscale(double2 *y, size_t n, double2 c, hipStream_t st)
{
    hipLaunchKernelGGL(k_cscale, dim3(stream_grid(n)), dim3(256), 0, st, y, n, c);
    return hipGetLastError();
}

hipError_t launch_wcorr(const WcorrArgs &a, hipStream_t st)
{
    hipLaunchKernelGGL(k_wcoef, dim3(1), dim3(256), 0, st, a);
    hipLaunchKernelGGL(k_wcorr, dim3(stream_grid((size_t)a.N * a.K)), dim3(256), 0, st, a);
    return hipGetLastError();
}

}  // namespace gmk
