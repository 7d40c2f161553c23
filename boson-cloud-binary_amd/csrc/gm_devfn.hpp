// gm_devfn.hpp - device helpers shared by the kernel files (gm_kernels.hip,
// gm_theta.hip): complex arithmetic, mode numbers, the block ranges of
// plain and batched launches, deterministic block reductions and the
// operator apply's element loads.
#pragma once
#include "gm_kernels.hpp"

namespace gmk {

// ---------------------------------------------------------------- helpers

__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double2 cmul(double2 a, double2 b)
{
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// conj(a) * b
__device__ __forceinline__ double2 cdotm(double2 a, double2 b)
{
    return make_double2(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ double2 cscale(double s, double2 a) { return make_double2(s * a.x, s * a.y); }
__device__ __forceinline__ double2 cdiv(double2 a, double2 b)
{
    const double d = b.x * b.x + b.y * b.y;
    return make_double2((a.x * b.x + a.y * b.y) / d, (a.y * b.x - a.x * b.y) / d);
}
__device__ __forceinline__ double cnorm2(double2 a) { return a.x * a.x + a.y * a.y; }
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b)
{
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
template <typename C> __device__ __forceinline__ C cone();
template <> __device__ __forceinline__ double2 cone<double2>() { return make_double2(1.0, 0.0); }
template <> __device__ __forceinline__ float2 cone<float2>() { return make_float2(1.0f, 0.0f); }
template <typename C> __device__ __forceinline__ C czero();
template <> __device__ __forceinline__ double2 czero<double2>() { return make_double2(0.0, 0.0); }
template <> __device__ __forceinline__ float2 czero<float2>() { return make_float2(0.0f, 0.0f); }
// i * h * a for real h
__device__ __forceinline__ double2 imul(double h, double2 a) { return make_double2(-h * a.y, h * a.x); }

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

__device__ __forceinline__ double vload(const double *p) { return *(const volatile double *)p; }

__device__ __forceinline__ double2 vload2(const double2 *p)
{
    const volatile double *q = (const volatile double *)p;
    return make_double2(q[0], q[1]);
}

// Azimuthal number m of mode j (modes ordered by l, then m = -l, -l+2, ..., l).
__device__ __forceinline__ int mode_l(int j)
{
    const int kk = (int)(sqrt(8.0 * j + 1.0) - 1.0) / 2;
    return (kk + 1) * (kk + 2) / 2 <= j ? kk + 1 : kk;
}

__device__ __forceinline__ int mode_m(int j)
{
    const int l = mode_l(j);
    return 2 * (j - l * (l + 1) / 2) - l;
}

// The blocks of one launch that belong to one solver: this block's index
// among them and their count. A plain launch gives each solver the whole
// grid; a batched ragged launch (one grid over many solvers) gives each its
// own contiguous range of blocks.
struct Blk {
    int id, n;
};
__device__ __forceinline__ Blk grid_blk() { return Blk{(int)blockIdx.x, (int)gridDim.x}; }

// The member of a batched launch that owns this block: the largest m with
// off[m] <= blockIdx.x (off holds M + 1 non-decreasing block offsets).
__device__ __forceinline__ int batch_member(const int *off, int M)
{
    const int b = blockIdx.x;
    int lo = 0, hi = M;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (__builtin_amdgcn_readfirstlane(off[mid]) <= b) lo = mid;
        else hi = mid;
    }
    // uniform over the block: the member's arguments load into scalar registers
    return __builtin_amdgcn_readfirstlane(lo);
}

// Sum NQ per-thread values over the block; every block stores its totals in
// partial[q * nblocks + block]. Returns true in the last block to finish, whose
// threads then see every partial. blockDim.x must be a multiple of 64, <= 1024.
template <int NQ>
__device__ bool block_partials(const double (&v)[NQ], double *partial, unsigned *counter, Blk B)
{
    __shared__ double ws[16][NQ];
    __shared__ bool last;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nw = blockDim.x >> 6;
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        const double s = wave_sum(v[q]);
        if (lane == 0) ws[wv][q] = s;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int q = 0; q < NQ; ++q) {
            double s = 0.0;
            for (int w = 0; w < nw; ++w) s += ws[w][q];
            // agent-scope store: written through to the device coherence
            // point, so no release fence (a write-back of this XCD's whole
            // L2, which every block of a streaming kernel would wait on) is
            // needed for the partials
            __hip_atomic_store(partial + (size_t)q * B.n + B.id, s, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
        }
        // The partial stores have completed before this block is counted.
        // The ordering rests on the AMDGPU memory model for gfx950: an
        // agent-scope store is performed at the device coherence point once
        // vmcnt reaches zero, so after s_waitcnt vmcnt(0) it is visible to
        // any agent-scope load issued later, and the relaxed counter RMW
        // cannot be issued before the wait. The signal fences are compiler
        // barriers: they keep the compiler from moving the stores below the
        // wait or the RMW above it (a release fence would instead write back
        // this XCD's whole L2).
        __atomic_signal_fence(__ATOMIC_SEQ_CST);
        __builtin_amdgcn_s_waitcnt(0);
        __atomic_signal_fence(__ATOMIC_SEQ_CST);
        const unsigned t = __hip_atomic_fetch_add(counter, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        last = (t == (unsigned)B.n - 1);
    }
    __syncthreads();
    // acquire in the one block that reads every partial (coherent loads)
    if (last) __threadfence();
    return last;
}

// In the last block: sum the partials of quantity q in block order.
__device__ inline double sum_partials(const double *partial, int q, Blk B)
{
    __shared__ double ws[16];
    const int nb = B.n;
    double s = 0.0;
    // contiguous slice per thread keeps the order fixed for a fixed grid
    const int per = (nb + blockDim.x - 1) / blockDim.x;
    const int b0 = threadIdx.x * per;
    for (int b = b0; b < b0 + per && b < nb; ++b) s += vload(partial + (size_t)q * nb + b);
    // fixed-order combine across threads
    s = wave_sum(s);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) ws[wv] = s;
    __syncthreads();
    double t = 0.0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t += ws[w];
    __syncthreads();
    return t;
}

// ---------------------------------------------------------------- apply elements

// Non-temporal loads and stores for the streaming passes (each element read
// or written once per launch: the BiCGSTAB update, the sweeps' vectors, the
// apply's shadow-residual reads and its outputs): they keep those streams
// from displacing what the operator apply re-reads every launch (its U-sum
// table, the quadrature tables, the pivots) from the caches. Measured at
// l_max 20, N 16384: the apply 146 -> 131 us, k_update 84 -> 79 us, the
// sweeps 3-5 us faster each.
typedef double nt_d2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ double2 ntld(const double2 *p)
{
    const nt_d2 v = __builtin_nontemporal_load(reinterpret_cast<const nt_d2 *>(p));
    return make_double2(v.x, v.y);
}
__device__ __forceinline__ void ntst(double2 *p, double2 v)
{
    nt_d2 w;
    w.x = v.x;
    w.y = v.y;
    __builtin_nontemporal_store(w, reinterpret_cast<nt_d2 *>(p));
}

__device__ __forceinline__ double2 load_x(const ApplyArgs &a, int k, int i)
{
    return a.X[(size_t)k * a.g.N + i];
}

// Rank-one sink term sum_a coef_a v_a at (mode j, node i).
__device__ __forceinline__ double2 sink_term(const ApplyArgs &a, int nsink, double2 coef0, double2 coef1, int j, int i)
{
    double2 t = make_double2(0.0, 0.0);
    if (nsink > 0) {
        const int d = i - a.sk[0].lo;
        if (d >= 0 && d < a.sk[0].n) t = cadd(t, cmul(coef0, a.sk[0].v[(size_t)j * a.sk[0].n + d]));
    }
    if (nsink > 1) {
        const int d = i - a.sk[1].lo;
        if (d >= 0 && d < a.sk[1].n) t = cadd(t, cmul(coef1, a.sk[1].v[(size_t)j * a.sk[1].n + d]));
    }
    return t;
}

}  // namespace gmk

namespace gmk {

// Per-launch values of the operator apply's epilogue.
struct ApplyEpi {
    double2 coef0, coef1;  // rank-one sink coefficients (dt/4) g_a (v_a^H x), or g_a (v_a^H psi) for AM_RHS
    double rhs_nrm;        // |rhs| (AM_RES, AM_RST)
    double alpha2, hdt;    // 1 / dr^2, dt / 2
    double c0;             // alpha2 / mu: the stencil's part of H0_ii
    int nsink;
};

__device__ __forceinline__ ApplyEpi apply_epi(const ApplyArgs &a)
{
    ApplyEpi e;
    const bool gq_mode = a.mode == AM_RHS;
    e.coef0 = gq_mode ? a.Kr->gq[0] : a.Kr->coefA[0];
    e.coef1 = gq_mode ? a.Kr->gq[1] : a.Kr->coefA[1];
    e.rhs_nrm = (a.mode == AM_RES || a.mode == AM_RST) ? a.Kr->rhs_nrm : 0.0;
    e.alpha2 = 1.0 / (a.g.dr * a.g.dr);
    e.hdt = a.P->half_dt;
    e.c0 = e.alpha2 / a.g.mu;
    e.nsink = a.P->nsink;
    return e;
}

// The loads of one element (mode j, node i) of the apply: x at i and its two
// neighbours (clamped to the grid; used only on interior rows), the CAP and
// the reduction partner R0 where the mode has one. Split from the arithmetic
// so that a caller can issue the loads of several elements before using any.
struct ElemIn {
    double2 xi, xm, xp, cap, r0;
};

// The element's loads other than x (the CAP and R0), for callers holding x.
__device__ __forceinline__ ElemIn apply_element_load_r0(const ApplyArgs &a, int j, int i)
{
    const gmd::Geometry &g = a.g;
    const int MODE = a.mode;
    ElemIn in;
    in.cap = g.cap ? g.cap[i] : make_double2(0.0, 0.0);
    const bool uses_r0 = MODE == AM_V || MODE == AM_T || MODE == AM_RES || MODE == AM_RST;
    in.r0 = uses_r0 ? ntld(a.R0 + (size_t)j * g.N + i) : make_double2(0.0, 0.0);
    return in;
}

__device__ __forceinline__ ElemIn apply_element_load(const ApplyArgs &a, int j, int i)
{
    const gmd::Geometry &g = a.g;
    const int N = g.N, MODE = a.mode;
    const size_t row = (size_t)j * N + i;
    ElemIn in;
    in.xi = a.X[row];
    in.xm = a.X[row - (i > 0 ? 1 : 0)];
    in.xp = a.X[row + (i < N - 1 ? 1 : 0)];
    in.cap = g.cap ? g.cap[i] : make_double2(0.0, 0.0);
    const bool uses_r0 = MODE == AM_V || MODE == AM_T || MODE == AM_RES || MODE == AM_RST;
    in.r0 = uses_r0 ? ntld(a.R0 + row) : make_double2(0.0, 0.0);
    return in;
}

// 1 / (2 mu r_i^2), the node's factor of the centrifugal term l (l + 1) / (2 mu r^2)
__device__ __forceinline__ double centrifugal(const gmd::Geometry &g, int i)
{
    const double ri = g.eps + i * g.dr;
    return i > 0 ? 1.0 / (2 * g.mu * ri * ri) : 0.0;
}

// One element (mode j = (l, m), node i) of the apply, given the coupling Cc
// = (V x)_j at node i, the element's loads and the node's centrifugal factor
// cr: the three-point stencil, the centrifugal term and the CAP
// (hamiltonian.cpp:44-66,224-243), A = 1 + i dt/2 H (matrices.cpp:27-40)
// with the rank-one sinks (matrices.cpp:111-116), the output of the mode and
// its terms of the fused reductions.
__device__ __forceinline__ void apply_element(const ApplyArgs &a, const ApplyEpi &ep, int j, int l, int i, double2 Cc,
                                              const ElemIn &in, double cr, double (&red)[3])
{
    const gmd::Geometry &g = a.g;
    const int N = g.N, MODE = a.mode;
    const size_t row = (size_t)j * N + i;
    if (a.Cout) a.Cout[row] = Cc;
    const double2 xi = in.xi, capv = in.cap;
    double2 Hx;
    if (i > 0 && i < N - 1) {
        const double h0d = ep.c0 + (double)(l * (l + 1)) * cr;
        Hx = cadd(cscale(g.off, cadd(in.xm, in.xp)), cadd(cmul(make_double2(h0d + capv.x, capv.y), xi), Cc));
    } else {
        Hx = cmul(capv, xi);
    }
    if (MODE == AM_H) {
        ntst(a.Y + row, Hx);
    } else if (MODE == AM_RHS) {
        const double2 sHx = imul(ep.hdt, Hx);
        double2 st = make_double2(0.0, 0.0);
        if (ep.nsink) st = cscale(0.5 * ep.hdt, sink_term(a, ep.nsink, ep.coef0, ep.coef1, j, i));  // (dt/4) Gamma psi
        const double2 rhs = csub(csub(xi, sHx), st);
        const double2 ax = cadd(cadd(xi, sHx), st);
        const double2 r0 = make_double2(rhs.x - ax.x, rhs.y - ax.y);
        ntst(a.Y + row, rhs);
        ntst(a.Y2 + row, r0);
        red[0] += cnorm2(rhs);
        red[1] += cnorm2(r0);
    } else {
        double2 out = cadd(xi, imul(ep.hdt, Hx));
        if (ep.nsink) out = cadd(out, sink_term(a, ep.nsink, ep.coef0, ep.coef1, j, i));
        if (MODE == AM_V) {
            ntst(a.Y + row, out);
            const double2 d = cdotm(in.r0, out);
            red[0] += d.x;
            red[1] += d.y;
        } else if (MODE == AM_T) {
            ntst(a.Y + row, out);
            const double2 d = cdotm(out, in.r0);
            red[0] += d.x;
            red[1] += d.y;
            red[2] += cnorm2(out);
        } else if (MODE == AM_A) {
            ntst(a.Y + row, out);
        } else if (MODE == AM_RST) {
            // r = b - A x of the normalised system, r0' = |rhs| r
            const double2 bq = make_double2(in.r0.x / ep.rhs_nrm, in.r0.y / ep.rhs_nrm);
            const double2 rq = csub(bq, out);
            ntst(a.Y + row, rq);
            ntst(a.Y2 + row, cscale(ep.rhs_nrm, rq));
            red[0] += cnorm2(rq);
        } else {  // AM_RES
            double2 bq = in.r0;
            if (a.r0_raw && ep.rhs_nrm != 0.0) bq = make_double2(bq.x / ep.rhs_nrm, bq.y / ep.rhs_nrm);
            red[0] += cnorm2(csub(out, bq));
        }
    }
}

// The fused reductions of the apply and the BiCGSTAB scalars the last block
// forms from them (matrices.cpp:47-91, Eigen's BiCGSTAB).
__device__ __forceinline__ void apply_reduce(const ApplyArgs &a, const double (&red)[3], Blk B)
{
    gmd::Krylov *Kr = a.Kr;
    const int MODE = a.mode;
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

}  // namespace gmk
