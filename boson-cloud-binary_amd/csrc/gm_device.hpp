// gm_device.hpp - data shared between the host solver and the gfx950 kernels.
//
// Layout in HBM (one solver):
//   complex vectors   K*Nr double2, mode-major (mode j at node i is j*Nr + i),
//                     the reference's stacking (include/CrankNicolsonModes.hpp:10-13);
//   E table           2 * (2 l_max + 1) * Nr doubles: the radial multipole
//                     profiles c_l* (f1 + sign prf f2) for even (index 0) and
//                     odd (index 1) m*-parity pairs;
//   GM table          the MFMA A operands in coupling-stream order: per step
//                     of a chunk's pair or single list, [slot][row] doubles
//                     (8 or 4 slots x 16 rows), the Gaunt coefficient of the
//                     row's mode with the step's column at each slot's l*
//                     (host-side they come from per-(chunk, l1) blocks, MBlock);
//   preconditioner    K*Nr double2 inverse pivots of each mode's radial tridiagonal.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace gmd {

constexpr int TILE = 64;   // radial nodes of a multipole-cut tile
constexpr int DIAG_NODES = 256;  // radial nodes per diagnostics wavefront (4 per lane)
#ifndef GM_AW
#define GM_AW 4
#endif
constexpr int AW = GM_AW;  // row chunks (wavefronts) per operator-apply workgroup
constexpr int MAXLM = 40;  // largest supported l_max + 1
// 256 threads of 8 nodes: the sweeps 6-10 us faster each than 512 of 4 at
// l_max 20, N 16384 (the affine scan's cost is per thread)
#ifndef GM_PT
#define GM_PT 256
#endif
#ifndef GM_PE
#define GM_PE 8
#endif
constexpr int PT = GM_PT;  // threads of a preconditioner or factor workgroup
constexpr int PE = GM_PE;  // consecutive nodes per preconditioner thread
constexpr int RED_BLOCKS = 1024;  // max partial-sum slots per reduction
#ifndef GM_PHALO
#define GM_PHALO 192
#endif
// The p / s sweeps of the preconditioner split a mode's row into segments of
// PSEG nodes, one workgroup each, solved over a window of one PT x PE tile:
// the segment and PHALO nodes on either side, whose dropped carries decay
// below rounding (k_factor checks it per mode).
constexpr int PHALO = GM_PHALO;
constexpr int PSEG = PT * PE - 2 * PHALO;
static_assert(PSEG > PHALO, "preconditioner segments longer than their halos");
// segments per mode: one (the whole row in one window) up to PT x PE nodes
__host__ __device__ inline int precond_nseg(int N) { return N <= PT * PE ? 1 : (N + PSEG - 1) / PSEG; }

constexpr int MR = 16;     // rows of an MFMA chunk

// Sixteen rows (modes of equal l parity) handled by one wavefront of the
// operator apply; padding rows have row_j = -1.
struct MChunk {
    int row_j[MR], row_l[MR], row_m[MR];
    int par, pad[3];
};

// Gaunt coefficients of one (chunk, l1) block: G[goff + (m1 * nTp + t) * 16 + row]
// is the coefficient of the chunk row with mode (l1, m1) at multipole
// l* = 2 (tlo + t) + p, p = (l + l1) mod 2; nTp is nT rounded up to a
// multiple of 4, the slots t >= nT holding zeros.
struct MBlock {
    int goff, tlo, nT, pad;
};

// Depth of the operand ring of the coupling stream; each of a chunk's step
// lists ends with RING - 1 padding steps that add nothing.
#ifndef GM_RING
#define GM_RING 4
#endif
constexpr int RING = GM_RING;
// LDS row read by padding steps (every row above a tile's cut holds zeros).
constexpr int ZROW = 2 * MAXLM - 1;
// LDS rows of the staged profiles: a paired padding step reads ZROW .. ZROW + 14.
constexpr int EROWS = ZROW + 15;

// One step of a chunk's coupling stream: four consecutive multipole slots of
// the (chunk, l1) block at one column m1 (a single), or two such groups of the
// same column (a pair: eight slots, l* from e to e + 14, 128 fragment values).
// A chunk's pairs and singles are each sorted by the l* of their first slot,
// so a tile with multipole cut lc runs prefixes and ranges of the lists.
struct MStep {
    int g;    // fragment of the step's slots x 16 rows in the stream-ordered G table
    int x;    // column mode kb + m1
    int e;    // l* of the group's first slot (its LDS row)
    int pad;
};

// Constant geometry and tables of one solver (device pointers).
struct Geometry {
    int N, K, lmax, Ls;        // Ls = 2 l_max + 1
    int ntiles, nchunks;       // node tiles of TILE, MFMA row chunks
    double dr, eps, mu, off;   // off = -1/(2 mu dr^2), the stencil off-diagonal of H0
    const MChunk *mchunks;
    const MBlock *mblocks;     // nchunks * (l_max + 1)
    const MStep *msteps;       // steps of every chunk's pair and single lists
    const int *step_start;     // nchunks offsets into msteps of the single lists
    const int *step_count;     // nchunks x (2 l_max + 1): singles a tile with cut lc runs
    const int *pair_start;     // nchunks offsets into msteps of the pair lists
    const int *pair_full;      // nchunks x (2 l_max + 1): pairs run whole at cut lc (a prefix)
    const int *pair_end;       // nchunks x (2 l_max + 1): end of the pairs run as their first group
    unsigned gm_n;             // elements of GM / GM32
    const double *GM;          // Gaunt coefficients in MFMA block layout
    const double *E;           // 2 * Ls * N
    const float *GM32;         // FP32 copies for the mixed-precision Krylov applies
    const float *E32;
    const int *lcut;           // per node tile: largest multipole l* evaluated
    const double2 *cap;        // Nr diagonal CAP or nullptr
    // polar-quadrature form of the coupling (gm_theta.hpp): MFMA operand
    // tables of phases A, C and the U sums, all rules; tks / tkd blocks per point group
    const double *TA, *TC, *TU;
    int tks, tkd;
    const double *Ut;          // U sums of every 8-node tile (rebuilt with E and the cut)
};

// One s-wave sink restricted to its radial support [lo, lo+n).
struct SinkDev {
    int lo, n;
    const double2 *v;  // K * n (mode-major over the support)
    const double2 *W;  // full-length P^-1 v, or nullptr
    int wlo, whi;      // nodes where W can be non-zero
};

// Per-step scalars, resident in device memory so that a captured graph can
// replay a step with new values.
struct StepParams {
    double2 ph[2 * MAXLM + 1];  // exp(i m theta), index m + l_max
    double2 wrot[2 * MAXLM + 1];  // exp(-i m (theta - theta_W)): Woodbury columns computed at theta_W
    double half_dt;             // dt / 2
    double dt;
    double g[2];                // sink strengths g_a (0 when inactive)
    int nsink;                  // active sinks in the solve (0..2), in order
    int sink_body[2];           // body (0 or 1) of each active sink
    int idx_ex;                 // extraction index
    int pad;
    // orbit of the radial profiles: positions of the bodies, mass ratio,
    // total mass and the companion's ramp weight
    double R1, R2, q, Mt, prf;
    double w1;                  // weight of the primary's profile (1; 0 only to read off V2 alone)
};

// BiCGSTAB state, device resident (src/CrankNicolsonModes_matrices.cpp:47-91).
struct Krylov {
    double2 rho, rho_old, alpha, w, beta;
    double r2, tol2, r0_sq, restart_thr;
    double rhs_nrm;             // |rhs| of the CN system
    double r0_raw2;             // |rhs - A psi|^2 (the unnormalised shadow residual)
    double res2;                // |A x - b|^2 of the acceptance check
    double2 c[2];               // Woodbury coefficients of the last preconditioner apply
    double2 coefA[2];           // rank-one coefficients (dt/4) g_a (v_a^H x) for the next apply
    double2 gq[2];              // g_a (v_a^H psi) for Gamma psi
    double2 qold[2];            // v_a^H psi at the start of the step
    double2 qnew[2];            // v_a^H psi after the solve, before the mask
    double2 qpost[2];           // v_a^H psi after the mask
    double2 Ginv[4];            // inverse of the k x k Woodbury capacitance
    double2 G[4];
    int wrange[4];              // [lo, hi) of each Woodbury column W_a
    double rst_eps2, rst_ref;   // restart thresholds as fractions of |r0|^2 (Eigen's eps^2; 1e-32 with sinks)
    int iter, maxit, done, restart, eigen_mode, restarts, zero_rhs, accepted;
    unsigned counter;
    int fresh;                  // no iteration yet since k_init: r is read from r0' / |rhs| and p = r
    int pb;                     // p is held in the second p buffer (the p update writes the other one)
    int wparts;                 // slots of the preconditioner's partial Woodbury projections
};

// Output of the diagnostics reduction.
struct DiagOut {
    double energy, pbin_sum, dv, mass;
};

}  // namespace gmd
