// gm_kernels.hpp - launch interface of the gfx950 kernels (gm_kernels.hip).
#pragma once
#include "gm_device.hpp"

namespace gmk {

// AM_RST: the BiCGSTAB restart (r = b - A x into Y, r0' = |rhs| r into Y2, |r|^2)
enum ApplyMode { AM_H = 0, AM_A = 1, AM_V = 2, AM_T = 3, AM_RHS = 4, AM_RES = 5, AM_RST = 6 };
// PV_WCOL: the Woodbury columns W_s = P^-1 v_s of the first `ncol` sinks in
// one launch (sk[s].W written, range[2 s, 2 s + 1] widened); PV_P / PV_S: the
// two preconditioner applies of a BiCGSTAB iteration.
enum PrecVar { PV_WCOL = 0, PV_P = 1, PV_S = 2 };
enum SinkMode { SK_PROJ = 0, SK_SETUP = 1 };
enum SinkDest { SD_START = 0, SD_QNEW = 1, SD_POST = 2, SD_COEFA = 3, SD_COEFA_QNEW = 4 };

struct ApplyArgs {
    gmd::Geometry g;
    const gmd::StepParams *P;
    gmd::Krylov *Kr;
    const double2 *X;    // input vector
    double2 *Y;          // output
    double2 *Y2;         // second output (AM_RHS: rhs - A psi, the unnormalised shadow residual; AM_RST: r0')
    const double2 *R0;   // unnormalised shadow residual (AM_V), s (AM_T), b or the raw rhs (AM_RES, AM_RST)
    int r0_raw;          // AM_RES, AM_RST: R0 holds the raw rhs; b = rhs / |rhs| is formed here
    gmd::SinkDev sk[2];
    double *partial;     // reduction partials (3 * grid doubles)
    double *rowpart;     // diagnostics apply: 6 doubles per (tile, mode)
    double2 *dpsi_ix;    // diagnostics apply: dpsi/dt at the extraction node per mode
    double2 *Cout;       // AM_RES: coupling V x written per row (or nullptr)
    const double2 *Cin;  // diagnostics apply: V x of the acceptance apply; V psi = |rhs| mask V x
    const double *mask;  // diagnostics apply: damping mask, or nullptr
    int ngroups;         // chunk groups per tile, set by the launcher
    int mixed;           // coupling in FP32 (E32, GM32) with FP32 accumulation
    int ng;              // 16-node groups per tile (1, 2 or 4), or 8: the polar-quadrature form (8-node tiles)
    void *Xt;            // scratch: exp(i m theta) X, double2 (FP64) or float2 (mixed), K * N
    int xt_ready;        // Xt already holds the phased X (written by its producer): no phase pass
    int mode;            // ApplyMode, set by the launcher
};

// exp(i m theta) X into Xt (double2, or float2 when f32), nbx blocks per mode
struct PhaseArgs {
    const double2 *X;
    void *Xt;
    int N, lmax, nbx, f32;
    const gmd::StepParams *P;
};

struct FactorArgs {
    int N, K, lmax;
    double dr, eps, mu, off;
    const gmd::StepParams *P;  // dt / 2
    const double *Gdiag;  // K * (l_max + 1): coefficient of (j, j) at l* = 2 s
    const double *E;
    const double2 *cap;
    double2 *inv;
    float2 *inv32;        // FP32 copy of the pivots, or nullptr
    int *failed;
    int *pseq;            // per mode and segment boundary (K nseg): 1 when a halo is too short; then per mode (K): solve whole
};

struct PrecArgs {
    int N, K;
    double off;          // the stencil off-diagonal; dt / 2 comes from Ps
    const double2 *inv;
    gmd::Krylov *Kr;
    const double2 *R;
    const double2 *R0;   // unnormalised shadow residual: r = R0 / |rhs| at the first iteration
    double2 *Pv;         // PV_P: p in the first buffer (Krylov pb = 0); PV_S: s (written)
    double2 *Pv2;        // PV_P: the second p buffer (the update reads one buffer and writes the other)
    const int *pseq;     // K nseg boundary flags, then per mode (K): solve the row whole
    int nseg;            // PV_P / PV_S: segments per mode (precond_nseg)
    double *wpart;       // PV_P / PV_S with wood: the partial projections v_a^H y (4 per holding block)
    const double2 *V;
    int ncol;            // PV_WCOL: number of sink columns
    double2 *Y;          // output (PV_P / PV_S)
    int wood;            // run the Woodbury epilogue
    gmd::SinkDev sk[2];
    const gmd::StepParams *Ps;
    int *range;          // PV_WCOL: per column, atomically widened [lo, hi) of non-zero output nodes
    int lmax;            // Woodbury epilogue: mode -> m for the column rotation
    const float2 *inv32; // FP32 pivots: the p / s variants sweep in FP32 on them (mixed precision), or nullptr
    void *Xt;            // also write exp(i m theta) y here for the next apply (or nullptr)
    int xt_f32;          // Xt holds float2 (FP32 coupling) instead of double2
};

struct WcorrArgs {
    int N, K, nsink, lmax;
    const gmd::StepParams *P;  // wrot: rotation of the stored Woodbury columns
    gmd::Krylov *Kr;
    const double *wpart;  // the preconditioner's partial projections v_a^H y (Krylov wparts slots)
    const double2 *W0, *W1;
    double2 *Y;
    void *Xt;            // phased copy of Y to correct as well (or nullptr)
    int xt_f32;
};

struct UpdateArgs {
    size_t n;
    gmd::Krylov *Kr;
    double2 *X;
    const double2 *Yc, *Zc, *S, *T, *R0;
    double2 *R;
    double *partial;
};

struct InitArgs {
    size_t n;
    gmd::Krylov *Kr;
    const double2 *Psi;
    double2 *X;
    double tol;
    int maxit, eigen_mode;
    double rst_eps2, rst_ref;  // restart thresholds relative to |r0|^2
};

struct AcceptArgs {
    size_t n;
    int N;
    const gmd::Krylov *Kr;
    const double2 *X;
    const double *mask;
    double2 *Psi;
    int *status;
    double *resid;
};

struct SinkArgs {
    int N, K, lmax;
    int mode, dest;
    double scale;
    int scale_by_nrm;    // multiply by |rhs| (device value)
    const double2 *X;
    gmd::SinkDev sk[2];
    gmd::Krylov *Kr;
    const gmd::StepParams *P;
};

struct DiagArgs {
    int N, K, ix, ntiles;
    double dr, eps, mu;
    const double2 *Psi;
    const double2 *dpsi_ix;
    const double *rowpart;
    double *out;         // flux, energy_flux, norms, energies, then E, p_bin, dV, mass shares (K each)
};

// apply_sink (absorption.cpp:100-109) on device vectors: Y = base + sum_a
// scale_a (v_a^H X) v_a over the supports, base being Y itself (accumulate) or
// zero. The projections reduce through a scratch block (SinkScratch).
constexpr int SINK_MAX_MODES = gmd::MAXLM * (gmd::MAXLM + 1) / 2;  // l_max <= MAXLM - 1
struct SinkScratch {
    unsigned counter;
    unsigned pad;
    double2 coef[2];
    double part[SINK_MAX_MODES][4];  // per-mode projections of both sinks (re, im)
};
struct SinkApplyArgs {
    int N, K, nsink;
    gmd::SinkDev sk[2];
    double scale[2];     // factor g_a of each sink
    const double2 *X;
    double2 *Y;
    int zero_base;       // Y = sum ... (1) or Y += sum ... (0)
    SinkScratch *scr;    // device scratch of at least sizeof(SinkScratch)
};
hipError_t launch_sink_apply(const SinkApplyArgs &a, hipStream_t st);
// Device address of the library's static SinkScratch on the current device,
// looked up with hipGetSymbolAddress on the first call per device.
hipError_t sink_scratch_address(SinkScratch **out);
// y = c y over n complex values
hipError_t launch_cscale(double2 *y, size_t n, double2 c, hipStream_t st);

// Absorber of one body (absorption.cpp:12-81,127-151): the smeared cap's
// multipole profile over the interior support [lo, lo + n) and its s-wave
// projector v (K x n, at the step's angle).
struct AbsorberBody {
    int lo, n;            // support nodes (n = 0: no absorber)
    int build;            // rebuild the profile: the body moved or its width changed
    int secondary;        // the secondary's sign on odd m
    double radius, width;
    double *mult;         // (2 l_max + 1) x n multipoles (unnormalised)
    double2 *v;           // K x n projector
};
constexpr int ABS_MAX_NQ = 128;  // quadrature points of one profile node (one per thread)
struct AbsorberArgs {
    int K, lmax, nq;
    double dr;
    const double *r;      // radial nodes
    const double *glx, *glw;  // Gauss-Legendre rule on [0, 1], nq points
    const double *ang;    // per mode j: angular coefficient of the (j, 00) pair with l* = l_j
    const gmd::StepParams *P;  // exp(i m theta)
    AbsorberBody b[2];
};
hipError_t launch_absorbers(const AbsorberArgs &a, hipStream_t st);

// One launch over many solvers (a batched ragged launch): member m's
// arguments are args[m] (device memory) and it owns the blocks
// [off[m], off[m + 1]) of a grid of `total` blocks. Kernels and their block
// counts are those of the plain launches; a member with no blocks sits out.
enum BatchKernel {
    BK_APPLY = 0, BK_PRECOND, BK_UPDATE, BK_INIT, BK_ACCEPT, BK_SINK, BK_WCORR, BK_DIAG_APPLY, BK_DIAG_FINAL, BK_PHASE,
    BK_WCOEF
};
struct BatchLaunch {
    int kernel;        // BatchKernel
    const void *args;  // device array of the kernel's argument struct, one per member
    const int *off;    // device array of M + 1 block offsets
    int M, total;
    int ng, mixed;     // BK_APPLY: tile and precision of every member of the launch
    int var, f32;      // BK_PRECOND: variant and sweep precision; BK_APPLY with ng = TAPPLY_NG: var = l_max class
    int pool;          // BK_APPLY: first of the M argument entries in the constant pool, or -1 (args)
};
// Argument entries of batched applies in constant memory, per device: the
// kernel reads its member's fields with scalar, invariant loads where it
// needs them instead of holding them in registers.
constexpr int APPLY_POOL = 2048;
hipError_t apply_pool_acquire(int count, int *first);  // current device; hipErrorOutOfMemory when full
void apply_pool_release(int device, int first, int count);
hipError_t apply_pool_write(int first, const ApplyArgs *src, int count);  // synchronous, current device
hipError_t launch_batch(const BatchLaunch &b, hipStream_t st);
// Blocks of the plain launches, for laying out batched ones.
int precond_blocks(int K, int N);
int stream_blocks(size_t n);
int phase_blocks(int N, int K);
int diag_apply_blocks(int N, int K);
// kr[m]->done, ->restart and ->iter into out[3 m .. 3 m + 2]
hipError_t launch_krylov_gather(gmd::Krylov *const *kr, int *out, int M, hipStream_t st);

// grid of the operator apply for tiles of 16 ng nodes (ng = 1, 2 or 4)
int apply_grid(int N, int nchunks, int ng, int *ngroups);
// The apply with the coupling in polar-quadrature form (gm_theta.hip), chosen
// by ng = TAPPLY_NG: one 64-thread block per 8 radial nodes, compiled per
// l_max class (a multiple of 4, up to 24).
constexpr int TAPPLY_NG = 8;
int tapply_class(int lmax, int mixed);  // 0: not available
int tapply_grid(int N);
hipError_t launch_tapply(const ApplyArgs &a, hipStream_t st);
// The polar form's U-sum table (doubles per solver; 0 without a class) and
// its build from the radial profiles and the cut (g.Ut unused here).
size_t tapply_u_doubles(int N, int lmax);
hipError_t launch_tapply_u(const gmd::Geometry &g, double *Ut, hipStream_t st);
hipError_t launch_tapply_batch(const ApplyArgs *args, const int *off, int M, int total, int cls, hipStream_t st);
hipError_t launch_apply(ApplyArgs a, int mode, hipStream_t st);
PhaseArgs phase_args(const ApplyArgs &a);
// radial profiles from the orbit in P (R1, R2, q, Mt, prf)
hipError_t launch_build_E(int N, int Ls, double dr, double eps, double mu, const gmd::StepParams *P, double *E,
                          float *E32, hipStream_t st);
hipError_t launch_lcut(int N, int Ls, double tau, const double *E, int *lcut, hipStream_t st);
hipError_t launch_factor(const FactorArgs &f, hipStream_t st);
hipError_t launch_precond(const PrecArgs &a, int var, hipStream_t st);
hipError_t launch_update(const UpdateArgs &a, hipStream_t st);
hipError_t launch_init(const InitArgs &a, hipStream_t st);
hipError_t launch_accept(const AcceptArgs &a, hipStream_t st);
hipError_t launch_sink(const SinkArgs &a, hipStream_t st);
hipError_t launch_wcorr(const WcorrArgs &a, hipStream_t st);
hipError_t launch_diag_apply(const ApplyArgs &a, hipStream_t st);
hipError_t launch_diag_final(const DiagArgs &a, hipStream_t st);

}  // namespace gmk
