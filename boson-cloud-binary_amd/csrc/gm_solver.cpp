// gm_solver.cpp - the solver handle behind include/gm.h.
//
// The host side restates the orbit bookkeeping of the reference's
// CrankNicolsonModes::step (src/CrankNicolsonModes.cpp:145-234) and keeps the
// Gaunt and absorber tables; all per-node work runs in the gfx950 kernels of
// gm_kernels.hip on one HIP stream. psi and every Krylov vector stay resident
// in HBM; a step returns to the host only the solver status and the
// diagnostics that step() writes (O(n_modes) doubles).
#include "gm.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <thread>
#include <tuple>
#include <vector>

#include "gm_kernels.hpp"
#include "gm_ilut.hpp"
#include "gm_math.hpp"
#include "gm_theta.hpp"

namespace {

struct DeviceError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

void hip_check(hipError_t e, const char *what)
{
    if (e != hipSuccess) throw DeviceError(std::string(what) + ": " + hipGetErrorString(e));
}

#define HIPCK(x) hip_check((x), #x)

template <class T>
struct DevBuf {
    T *p = nullptr;
    size_t n = 0;
    void alloc(size_t count)
    {
        release();
        if (count == 0) return;
        HIPCK(hipMalloc(reinterpret_cast<void **>(&p), count * sizeof(T)));
        n = count;
    }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    ~DevBuf() { release(); }
};

template <class T>
struct PinBuf {
    T *p = nullptr;
    void alloc(size_t count)
    {
        release();
        HIPCK(hipHostMalloc(reinterpret_cast<void **>(&p), count * sizeof(T), hipHostMallocDefault));
    }
    void release()
    {
        if (p) (void)hipHostFree(p);
        p = nullptr;
    }
    ~PinBuf() { release(); }
};

// A sparse operator in CSR, rows sorted by column (gm_operator_csr).
struct OpCsr {
    std::vector<int64_t> ptr;
    std::vector<int32_t> col;
    std::vector<std::complex<double>> val;
};

struct Profile {
    double radius = -1.0, width = -1.0;
    bool valid = false;
    bool build = false;          // the device multipoles are to be rebuilt
    int lo = 0, n = 0;           // interior support nodes [lo, lo+n)
};

struct Sink {
    int lo = 0, n = 0;
    double strength = 0.0;
    DevBuf<double> dmult;        // (2 l_max + 1) * n profile multipoles of the support
    DevBuf<double2> dv;          // K * n projector over the support (device only)
    DevBuf<double2> W;           // P^-1 v, full length
    // host copy of v and its squared norm, read back on demand (host ILUT,
    // sink strength queries)
    std::vector<double2> v;
    double norm2 = 0.0;
    bool v_host = false;
};

}  // namespace

struct gm_solver {
    // ---- constructor arguments and public members (include/CrankNicolsonModes.hpp:88-164)
    int Nr = 0, save_every = 0, l_max = 0, K = 0;
    double epsilon = 0, mu = 0, q = 0, R = 0, theta = 0, phase = 0, Omega = 0, Omega_0 = 0, P = 0;
    double R1 = 0, R2 = 0, M1 = 0, M2 = 0, a = 0, L = 0, dr = 0, dt = 0;
    bool counter = false;
    double M_c = 0;
    bool isolated = false;
    double ramp_periods = 0, initial_period = 0;
    bool absorption = false;
    double epsilon_smear = 0;
    gm_flags flags{};
    bool initialized = false;
    std::vector<int> ml, mm;
    std::vector<double> r;
    // ---- time-dependent state
    double t_i = 0, ramp_factor = 0, prf = 0, Omega_V = 0, p_bin = 0;
    double Ah1 = 0, Ah2 = 0;
    bool sink_active = false;
    int mode00 = -1;
    double r_ex = 0;
    int idx_ex = 1;
    double last_residual = 0;
    int last_iterations = 0;
    std::unique_ptr<gmh::RollingPhaseAverage> avg;
    Profile prof[2];
    Sink sink[2];
    // ---- mask and CAP
    bool mask = false, cap = false;
    std::vector<double> damping;
    std::vector<double2> capv;
    int cap_version = 0;
    // ---- device
    hipStream_t st = nullptr;
    std::vector<int> goff_h;
    std::vector<double> G_h;
    std::vector<gmd::MChunk> mchunks_h;
    std::vector<gmd::MBlock> mblocks_h;
    std::vector<double> GM_h, Gdiag_h;
    DevBuf<gmd::MChunk> mchunks;
    DevBuf<gmd::MBlock> mblocks;
    DevBuf<int> step_count, pair_full, pair_end, pair_start;
    unsigned gm_n = 0;  // elements of GM / GM32 (Gaunt fragments in stream order)
    DevBuf<gmd::MStep> msteps;
    DevBuf<int> mstep_start;
    DevBuf<double> GM, Gdiag, E, dmask;
    DevBuf<float> GM32, E32;  // FP32 copies for the mixed-precision Krylov applies
    DevBuf<int> lcut;
    // polar-quadrature form of the coupling (gm_theta.hpp)
    std::shared_ptr<const gmt::ThetaTables> theta_h;
    DevBuf<double> TA, TC, TU;
    DevBuf<double> Ut;  // U sums of every tile, rebuilt with the radial profiles and the cut
    // absorbers: radial nodes, the cap quadrature and the (j, 00) angular coefficients
    DevBuf<double> dr_nodes, dglx, dglw, dang;
    int abs_nq = 0;
    double coupling_tol = 5.421010862427522e-20;  // 2^-64
    DevBuf<double2> dcap;
    DevBuf<double2> psi, b, x, rr, r0, p, p2, v, y, z, s, t, inv, tmp;
    DevBuf<int> pseq;  // per mode and segment boundary: the halo check failed; then per mode: solve the row whole
    DevBuf<double> wpart;  // partial Woodbury projections of the p / s sweeps (4 per segment block)
    DevBuf<float2> inv32;  // FP32 pivots (mixed-precision preconditioner sweeps)
    DevBuf<double2> xt;         // phased column vectors of the coupling (FP64; FP32 uses the first half)
    DevBuf<double> partial, rowpart, diagout;
    DevBuf<gmk::SinkScratch> sinkscr;  // reduction scratch of the device apply_sink
    gmk::SinkScratch *static_scr = nullptr;  // this GPU's static scratch (gm_apply_sink_device)
    DevBuf<double2> dpsi_ix;
    DevBuf<gmd::StepParams> dP;
    DevBuf<gmd::Krylov> dK;
    DevBuf<int> dflags;  // [0] factor failure, [1] accept status
    DevBuf<double> dresid;
    PinBuf<gmd::StepParams> hP;
    PinBuf<gmd::Krylov> hK;
    PinBuf<double> hdiag;
    PinBuf<int> hflags;
    PinBuf<double> hresid;
    PinBuf<int> hwr;
    int ntiles = 0, nchunks = 0;
    // caches
    bool E_valid = false;
    bool E_pending = false;  // profiles of a new orbit not yet enqueued
    double E_key[6] = {0, 0, 0, 0, 0, 0};
    bool fac_valid = false;
    double fac_key[3] = {0, 0, 0};
    int fac_cap_version = -1;
    // Woodbury columns W_a = P^-1 v_a stored at angle w_theta (see build_perturbation)
    bool w_valid = false, w_dirty = true;
    double w_key[14] = {};
    double w_theta = 0.0;
    int prev_iters = 4;
    // BiCGSTAB restart thresholds relative to |r0|^2: Eigen's eps^2 without
    // sinks, the reference's 1e-32 with them (gm_set_restart_thresholds)
    double rst_eps2 = 4.930380657631324e-32, rst_ref = 1e-32;
    int last_restarts = 0;
    // gm_operator_csr: the operator of a size query, kept for the fill call
    OpCsr op_cache;
    int op_which = -1;
    // progress of the step in flight (step_begin / solve_advance / step_finish)
    double step_start = 0.0;
    bool step_wood = false;
    bool step_E = false, step_fac = false;  // profiles / pivots rebuilt at the solve start
    bool step_ilut = false;  // this step's solve is the host ILUT one
    bool step_fp64 = false;  // a mixed-precision step solved again with the FP64 operator
    bool retry_mixed = false;  // flags.precision switched from mixed for that retry (restored on acceptance)
    int step_enqueued = 0;
    // captured parts of the step, keyed by every kernel argument that can change
    struct GraphEntry {
        int kind = 0;
        std::vector<uint64_t> key;
        hipGraphExec_t exec = nullptr;
    };
    std::vector<GraphEntry> graphs;
    // last key requested per kind and precision (capture once it repeats; a
    // mixed step retried in FP64 alternates the two)
    std::vector<uint64_t> last_key[3][2];
    // operator-apply form per precision, 0 until tuned: Gaunt-form tiles of
    // 16 ng nodes (ng = 1, 2, 4) or the polar form (gmk::TAPPLY_NG), and the
    // measured time per candidate (ms; 16, 32, 64 nodes, polar; 0 if not timed)
    int apply_ng[2] = {0, 0};
    double apply_ms[2][4] = {{0, 0, 0, 0}, {0, 0, 0, 0}};
    std::string err;
    // profiling (gm_profile_enable)
    bool prof_on = false;
    std::vector<hipEvent_t> ev_pool;
    struct Pending {
        int cat;
        hipEvent_t a, b;
    };
    std::vector<Pending> pending;
    double prof_ms[GM_PROF_NCAT] = {};
    int64_t prof_n[GM_PROF_NCAT] = {};

    size_t nvec() const { return static_cast<size_t>(Nr) * K; }
    double ramp_end_time() const { return ramp_periods * initial_period; }
};

namespace {

thread_local std::string g_create_err;

void ensure_device()
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n < 1) throw DeviceError("no HIP device is available");
}


// ------------------------------------------------------------------ profiling

hipEvent_t take_event(gm_solver &S)
{
    if (S.ev_pool.empty()) {
        hipEvent_t e;
        HIPCK(hipEventCreate(&e));
        return e;
    }
    hipEvent_t e = S.ev_pool.back();
    S.ev_pool.pop_back();
    return e;
}

// Launch `f` (returning hipError_t) on the solver stream, timed when profiling.
template <class F>
void launch(gm_solver &S, int cat, F &&f)
{
    if (!S.prof_on) {
        HIPCK(f());
        return;
    }
    hipEvent_t a = take_event(S), b = take_event(S);
    HIPCK(hipEventRecord(a, S.st));
    HIPCK(f());
    HIPCK(hipEventRecord(b, S.st));
    S.pending.push_back({cat, a, b});
}

// After a stream synchronisation: fold the recorded intervals into the totals.
void prof_collect(gm_solver &S)
{
    for (auto &p : S.pending) {
        float ms = 0.f;
        HIPCK(hipEventElapsedTime(&ms, p.a, p.b));
        S.prof_ms[p.cat] += ms;
        S.prof_n[p.cat] += 1;
        S.ev_pool.push_back(p.a);
        S.ev_pool.push_back(p.b);
    }
    S.pending.clear();
}

// ------------------------------------------------------------------ tables

double angular(const gm_solver &S, int j, int k, int ls);

// Gaunt blocks: for each (l, l1), [m1][s][m] with l* = |l - l1| + 2 s, the
// angular coefficient of build_coupling_cache (hamiltonian.cpp:77-95).
void build_gaunt(gm_solver &S)
{
    // blocks (l, l1) of [m1][s][m] coefficients; offsets first, then the
    // blocks in parallel on the host cores (binary128 3j symbols dominate)
    const int Lm = S.l_max + 1;
    S.goff_h.assign(Lm * Lm, 0);
    size_t total = 0;
    for (int l = 0; l <= S.l_max; ++l)
        for (int l1 = 0; l1 <= S.l_max; ++l1) {
            S.goff_h[l * Lm + l1] = static_cast<int>(total);
            total += static_cast<size_t>(l1 + 1) * (std::min(l, l1) + 1) * (l + 1);
        }
    S.G_h.assign(total, 0.0);
    auto block = [&](int l, int l1) {
        double *G = S.G_h.data() + S.goff_h[l * Lm + l1];
        const int Sl = std::min(l, l1) + 1, lo = std::abs(l - l1);
        for (int s = 0; s < Sl; ++s) {
            const int ls = lo + 2 * s;
            const double w1 = gmh::wigner3j(ls, l1, l, 0, 0, 0);
            const double norm = std::sqrt((2.0 * l + 1) * (2.0 * l1 + 1) * (2.0 * ls + 1) / (4.0 * M_PI));
            for (int m1i = 0; m1i <= l1; ++m1i)
                for (int mi = 0; mi <= l; ++mi) {
                    const int m = -l + 2 * mi, m1 = -l1 + 2 * m1i;
                    const int ms = m1 - m;
                    double ang = 0.0;
                    if (std::abs(ms) <= ls && w1 != 0.0) {
                        const double w2 = gmh::wigner3j(ls, l1, l, -ms, m1, -m);
                        double yv = gmh::sph_plm0(ls, std::abs(ms));
                        if (ms < 0 && std::abs(ms) % 2 != 0) yv = -yv;
                        ang = (m1 % 2 == 0 ? 1.0 : -1.0) * norm * w1 * w2 * yv;
                    }
                    G[(m1i * Sl + s) * (l + 1) + mi] = ang;
                }
        }
    };
    const int nb = Lm * Lm;
    const int nt = std::max(1, std::min(nb, static_cast<int>(std::thread::hardware_concurrency())));
    std::vector<std::thread> pool;
    for (int t = 0; t < nt; ++t)
        pool.emplace_back([&, t] {
            for (int b = t; b < nb; b += nt) block(b / Lm, b % Lm);
        });
    for (auto &th : pool) th.join();
}

// MFMA tables: 16-row chunks of modes with equal l parity (largest l first),
// and for every (chunk, l1) the Gaunt block [m1][t][row] with
// l* = 2 (tlo + t) + p, p = (l + l1) mod 2 (zero where the row has no term).
void build_mfma_tables(gm_solver &S)
{
    const int Lm = S.l_max + 1;
    S.mchunks_h.clear();
    for (int par = 0; par < 2; ++par) {
        // rows of one l parity ordered by m, then by decreasing l: a chunk of
        // 16 rows spans a narrow band of m, so columns far from it in m1 (and
        // the low multipoles of the columns near it) carry no coupling
        std::vector<int> rows;
        for (int l = S.l_max; l >= 0; --l)
            if (l % 2 == par)
                for (int mi = 0; mi <= l; ++mi) rows.push_back(l * (l + 1) / 2 + mi);
        std::stable_sort(rows.begin(), rows.end(), [&](int x, int y) {
            return S.mm[x] != S.mm[y] ? S.mm[x] < S.mm[y] : S.ml[x] > S.ml[y];
        });
        for (size_t r0 = 0; r0 < rows.size(); r0 += gmd::MR) {
            gmd::MChunk c{};
            c.par = par;
            for (int r = 0; r < gmd::MR; ++r) {
                const size_t k = r0 + r;
                c.row_j[r] = k < rows.size() ? rows[k] : -1;
                c.row_l[r] = k < rows.size() ? S.ml[rows[k]] : 0;
                c.row_m[r] = k < rows.size() ? S.mm[rows[k]] : 0;
            }
            S.mchunks_h.push_back(c);
        }
    }
    // blocks and the chunk work estimate (MFMA slot groups x columns)
    auto block_range = [&](const gmd::MChunk &c, int l1, int &tlo, int &nT) {
        const int p = (c.par + l1) & 1;
        int lo = 1 << 30, hi = -1;
        for (int r = 0; r < gmd::MR; ++r) {
            if (c.row_j[r] < 0) continue;
            const int l = c.row_l[r];
            lo = std::min(lo, (std::abs(l - l1) - p) / 2);
            hi = std::max(hi, (l + l1 - p) / 2);
        }
        tlo = lo;
        nT = hi - lo + 1;
    };
    std::vector<std::pair<long, int>> order;
    for (size_t c = 0; c < S.mchunks_h.size(); ++c) {
        long work = 0;
        for (int l1 = 0; l1 <= S.l_max; ++l1) {
            int tlo, nT;
            block_range(S.mchunks_h[c], l1, tlo, nT);
            work += static_cast<long>(l1 + 1) * ((nT + 3) / 4);
        }
        order.push_back({-work, static_cast<int>(c)});
    }
    std::stable_sort(order.begin(), order.end());
    std::vector<gmd::MChunk> sorted;
    for (auto &o : order) sorted.push_back(S.mchunks_h[o.second]);
    S.mchunks_h.swap(sorted);
    S.mblocks_h.assign(S.mchunks_h.size() * Lm, gmd::MBlock{});
    S.GM_h.clear();
    for (size_t c = 0; c < S.mchunks_h.size(); ++c) {
        const gmd::MChunk &ch = S.mchunks_h[c];
        for (int l1 = 0; l1 <= S.l_max; ++l1) {
            gmd::MBlock &B = S.mblocks_h[c * Lm + l1];
            block_range(ch, l1, B.tlo, B.nT);
            B.goff = static_cast<int>(S.GM_h.size());
            const int p = (ch.par + l1) & 1;
            // slots padded with zeros to a multiple of four (one MFMA k-group)
            const int nTp = (B.nT + 3) / 4 * 4;
            for (int m1i = 0; m1i <= l1; ++m1i)
                for (int t = 0; t < nTp; ++t)
                    for (int r = 0; r < gmd::MR; ++r) {
                        double v = 0.0;
                        if (ch.row_j[r] >= 0 && t < B.nT)
                            v = angular(S, ch.row_j[r], l1 * (l1 + 1) / 2 + m1i, 2 * (B.tlo + t) + p);
                        S.GM_h.push_back(v);
                    }
        }
    }
    // fragments read past the last block stay in bounds
    S.GM_h.insert(S.GM_h.end(), 4 * gmd::MR, 0.0);
    S.nchunks = static_cast<int>(S.mchunks_h.size());
    // diagonal coefficients (j, j) at l* = 2 s for the Thomas pivots
    S.Gdiag_h.assign(static_cast<size_t>(S.K) * Lm, 0.0);
    for (int j = 0; j < S.K; ++j)
        for (int s = 0; s <= S.ml[j]; ++s) S.Gdiag_h[static_cast<size_t>(j) * Lm + s] = angular(S, j, j, 2 * s);
    S.ntiles = (S.Nr + gmd::TILE - 1) / gmd::TILE;
}

// The Gaunt and MFMA tables depend on l_max only: built once per process and
// l_max, then copied into every solver of that l_max.
struct HostTables {
    std::vector<int> goff;
    std::vector<double> G, GM, Gdiag;
    std::vector<gmd::MChunk> mchunks;
    std::vector<gmd::MBlock> mblocks;
    int nchunks = 0;
    std::shared_ptr<const gmt::ThetaTables> theta;
};
std::mutex tables_mu;
std::map<int, std::shared_ptr<const HostTables>> tables_cache;

void host_tables(gm_solver &S)
{
    std::shared_ptr<const HostTables> T;
    {
        std::lock_guard<std::mutex> lk(tables_mu);
        auto it = tables_cache.find(S.l_max);
        if (it != tables_cache.end()) T = it->second;
    }
    if (!T) {
        build_gaunt(S);
        build_mfma_tables(S);
        auto H = std::make_shared<HostTables>();
        H->goff = S.goff_h;
        H->G = S.G_h;
        H->GM = S.GM_h;
        H->Gdiag = S.Gdiag_h;
        H->mchunks = S.mchunks_h;
        H->mblocks = S.mblocks_h;
        H->nchunks = S.nchunks;
        if (const int cls = gmk::tapply_class(S.l_max, 0))
            H->theta = std::make_shared<const gmt::ThetaTables>(gmt::build_theta_tables(S.l_max, cls));
        S.theta_h = H->theta;
        std::lock_guard<std::mutex> lk(tables_mu);
        tables_cache.emplace(S.l_max, H);
        return;
    }
    S.goff_h = T->goff;
    S.G_h = T->G;
    S.GM_h = T->GM;
    S.Gdiag_h = T->Gdiag;
    S.mchunks_h = T->mchunks;
    S.mblocks_h = T->mblocks;
    S.nchunks = T->nchunks;
    S.theta_h = T->theta;
    S.ntiles = (S.Nr + gmd::TILE - 1) / gmd::TILE;
}

double angular(const gm_solver &S, int j, int k, int ls)
{
    const int l = S.ml[j], l1 = S.ml[k];
    const int mi = (S.mm[j] + l) / 2, m1i = (S.mm[k] + l1) / 2;
    const int Sl = std::min(l, l1) + 1, lo = std::abs(l - l1);
    const int s = (ls - lo) / 2;
    if (ls < lo || s >= Sl || (ls - lo) % 2) return 0.0;
    return S.G_h[S.goff_h[l * (S.l_max + 1) + l1] + (m1i * Sl + s) * (l + 1) + mi];
}

gmd::Geometry geometry(const gm_solver &S)
{
    gmd::Geometry g{};
    g.N = S.Nr;
    g.K = S.K;
    g.lmax = S.l_max;
    g.Ls = 2 * S.l_max + 1;
    g.ntiles = S.ntiles;
    g.nchunks = S.nchunks;
    g.dr = S.dr;
    g.eps = S.epsilon;
    g.mu = S.mu;
    g.off = -(1.0 / (S.dr * S.dr)) / (2 * S.mu);
    g.mchunks = S.mchunks.p;
    g.mblocks = S.mblocks.p;
    g.step_count = S.step_count.p;
    g.msteps = S.msteps.p;
    g.step_start = S.mstep_start.p;
    g.pair_start = S.pair_start.p;
    g.pair_full = S.pair_full.p;
    g.pair_end = S.pair_end.p;
    g.gm_n = S.gm_n;
    g.GM = S.GM.p;
    g.GM32 = S.GM32.p;
    g.E32 = S.E32.p;
    g.E = S.E.p;
    g.lcut = S.lcut.p;
    g.cap = S.cap ? S.dcap.p : nullptr;
    g.TA = S.TA.p;
    g.TC = S.TC.p;
    g.TU = S.TU.p;
    g.tks = S.theta_h ? S.theta_h->ks : 0;
    g.tkd = S.theta_h ? S.theta_h->kd : 0;
    g.Ut = S.Ut.p;
    return g;
}

// ------------------------------------------------------------------ absorbers

// update_absorption_profile (src/CrankNicolsonModes_absorption.cpp:12-81):
// the interior nodes within `width` of the body; the multipoles over them are
// rebuilt on the device (k_abs_profile) when the body or its width changed.
void update_profile(gm_solver &S, Profile &pr, double radius, double width)
{
    if (pr.valid && pr.radius == radius && pr.width == width) return;
    pr.valid = true;
    pr.radius = radius;
    pr.width = width;
    pr.lo = 0;
    pr.n = 0;
    pr.build = false;
    if (width < S.dr || width <= 0.0 || radius < 0.0) return;
    // r_i = epsilon + i dr is monotone: scan a window around the body
    const int N = S.Nr;
    const int c = static_cast<int>(std::floor((radius - S.epsilon) / S.dr));
    const int span = static_cast<int>(std::ceil(width / S.dr)) + 2;
    const int a = std::max(1, c - span), b = std::min(N - 1, c + span + 1);
    int lo = -1, hi = -1;
    for (int i = a; i < b; ++i)
        if (std::abs(S.r[i] - radius) < width) {
            if (lo < 0) lo = i;
            hi = i + 1;
        }
    if (lo < 0) return;
    pr.lo = lo;
    pr.n = hi - lo;
    pr.build = true;
}

// The cap quadrature (nq = max(128, 2 l_max + 16) Gauss-Legendre points on
// [0, 1], absorption.cpp:27-34), the radial nodes and the angular
// coefficients of the (j, 00) pairs on the device, once per solver.
std::mutex gl_mu;
std::map<int, std::pair<std::vector<double>, std::vector<double>>> gl_cache;

void absorber_tables(gm_solver &S)
{
    const int nq = std::max(128, 2 * S.l_max + 16);
    if (nq > gmk::ABS_MAX_NQ) throw std::invalid_argument("l_max too large for the absorber quadrature");
    std::pair<std::vector<double>, std::vector<double>> gl;
    {
        std::lock_guard<std::mutex> lk(gl_mu);
        auto it = gl_cache.find(nq);
        if (it == gl_cache.end()) {
            gmh::gauss_legendre01(nq, gl.first, gl.second);
            gl_cache.emplace(nq, gl);
        } else {
            gl = it->second;
        }
    }
    std::vector<double> ang(S.K, 0.0);
    if (S.mode00 >= 0)
        for (int j = 0; j < S.K; ++j) ang[j] = angular(S, j, S.mode00, S.ml[j]);
    if (S.dr_nodes.n != static_cast<size_t>(S.Nr)) S.dr_nodes.alloc(S.Nr);
    if (S.dglx.n != static_cast<size_t>(nq)) {
        S.dglx.alloc(nq);
        S.dglw.alloc(nq);
    }
    if (S.dang.n != static_cast<size_t>(S.K)) S.dang.alloc(S.K);
    // pageable sources: synchronous copies
    HIPCK(hipMemcpy(S.dr_nodes.p, S.r.data(), S.Nr * sizeof(double), hipMemcpyHostToDevice));
    HIPCK(hipMemcpy(S.dglx.p, gl.first.data(), nq * sizeof(double), hipMemcpyHostToDevice));
    HIPCK(hipMemcpy(S.dglw.p, gl.second.data(), nq * sizeof(double), hipMemcpyHostToDevice));
    HIPCK(hipMemcpy(S.dang.p, ang.data(), S.K * sizeof(double), hipMemcpyHostToDevice));
    S.abs_nq = nq;
}

// Device storage of a body's absorber over its support.
void size_sink(gm_solver &S, Sink &sk, int n)
{
    const size_t nm = static_cast<size_t>(2 * S.l_max + 1) * std::max(n, 1);
    const size_t nv = static_cast<size_t>(S.K) * std::max(n, 1);
    if (sk.dmult.n < nm) sk.dmult.alloc(nm);
    if (sk.dv.n < nv) sk.dv.alloc(nv);
}

// The two bodies' profiles (when rebuilt) and projectors at the step's angle
// (build_swave_projector, absorption.cpp:127-151), after the step parameters
// (exp(i m theta)) are on the device.
void launch_absorbers(gm_solver &S)
{
    gmk::AbsorberArgs a{};
    a.K = S.K;
    a.lmax = S.l_max;
    a.nq = S.abs_nq;
    a.dr = S.dr;
    a.r = S.dr_nodes.p;
    a.glx = S.dglx.p;
    a.glw = S.dglw.p;
    a.ang = S.dang.p;
    a.P = S.dP.p;
    for (int b = 0; b < 2; ++b) {
        Profile &pr = S.prof[b];
        Sink &sk = S.sink[b];
        gmk::AbsorberBody &B = a.b[b];
        B.lo = pr.lo;
        B.n = pr.n;
        B.build = pr.build;
        B.secondary = b == 1;
        B.radius = pr.radius;
        B.width = pr.width;
        B.mult = sk.dmult.p;
        B.v = sk.dv.p;
        pr.build = false;
        sk.v_host = false;
    }
    launch(S, GM_PROF_SMALL, [&] { return gmk::launch_absorbers(a, S.st); });
}

// The host copy of a sink's projector (and its squared norm), read back from
// the device after the stream has drained.
const Sink &sink_host(const gm_solver &S, int body)
{
    Sink &sk = const_cast<Sink &>(S.sink[body]);
    if (!sk.v_host) {
        sk.v.assign(static_cast<size_t>(S.K) * sk.n, make_double2(0.0, 0.0));
        sk.norm2 = 0.0;
        if (!sk.v.empty()) {
            HIPCK(hipStreamSynchronize(S.st));
            HIPCK(hipMemcpy(sk.v.data(), sk.dv.p, sk.v.size() * sizeof(double2), hipMemcpyDeviceToHost));
            for (const double2 &val : sk.v) sk.norm2 += val.x * val.x + val.y * val.y;
        }
        sk.v_host = true;
    }
    return sk;
}

// build_spherical_absorption (absorption.cpp:153-182).
void build_absorption(gm_solver &S)
{
    update_profile(S, S.prof[0], S.R1, S.epsilon_smear * S.M1);
    update_profile(S, S.prof[1], S.R2, S.epsilon_smear * S.M2);
    double ramp = 1.0;
    if (!S.isolated && S.t_i < S.ramp_end_time()) ramp = 0.0;
    else if (!S.isolated && S.t_i < S.ramp_end_time() + S.initial_period) {
        const double xx = (S.t_i - S.ramp_end_time()) / S.initial_period;
        ramp = 3.0 * xx * xx - 2.0 * xx * xx * xx;
    }
    S.Ah1 = 16.0 * M_PI * S.M1 * S.M1 * ramp;
    S.Ah2 = 16.0 * M_PI * S.M2 * S.M2 * ramp;
    for (int b = 0; b < 2; ++b) {
        Sink &sk = S.sink[b];
        sk.lo = S.prof[b].lo;
        sk.n = S.prof[b].n;
        size_sink(S, sk, sk.n);
    }
    S.sink[0].strength = S.Ah1 * S.dr;
    S.sink[1].strength = S.Ah2 * S.dr;
}

// ------------------------------------------------------------------ per-step operator

// The reference drops a sink whose projector has |v|^2 = 0 (absorption.cpp:
// 153-182). A body off the origin with a non-empty support has |v|^2 > 0: the
// (00, 00) row holds sqrt(4 pi) r mult_0(r) c_00 with c_00 != 0, and at the
// support node nearest the body (within dr / 2 <= width / 2) the cap weight
// exp(-1 / (1 - dist^2 / width^2)) is at least exp(-4/3) near the cap's
// centre. At the origin the weights of a support just inside `width` can
// underflow to zero; such a sink is kept with v = 0, where its Woodbury terms
// and absorbed mass vanish, the same step as without it.
bool sink_nonzero(const gm_solver &S, int body)
{
    return S.sink[body].n > 0 && S.mode00 >= 0;
}

// Radial profiles E of the orbit in the device step parameters, and the
// multipole cut of each tile.
void launch_profiles(gm_solver &S)
{
    launch(S, GM_PROF_FACTOR, [&] {
        return gmk::launch_build_E(S.Nr, 2 * S.l_max + 1, S.dr, S.epsilon, S.mu, S.dP.p, S.E.p, S.E32.p, S.st);
    });
    launch(S, GM_PROF_FACTOR, [&] {
        return gmk::launch_lcut(S.Nr, 2 * S.l_max + 1, S.coupling_tol, S.E.p, S.lcut.p, S.st);
    });
    // the polar form's U sums follow the profiles and the cut (skipped once
    // both precisions settled on the Gaunt form)
    const auto polar = [&](int p) { return S.apply_ng[p] == 0 || S.apply_ng[p] == gmk::TAPPLY_NG; };
    if (S.Ut.p && (polar(0) || polar(1)))
        launch(S, GM_PROF_FACTOR, [&] { return gmk::launch_tapply_u(geometry(S), S.Ut.p, S.st); });
}

// build_time_dependent_perturbation (hamiltonian.cpp:245-268) on the device
// tables. The step scalars go to the device first; the radial profiles, when
// the orbit moved, are rebuilt at once, or with `defer` inside the step's
// captured start (E_pending).
void build_perturbation(gm_solver &S, double time, bool defer = false)
{
    S.ramp_factor = 1.0;
    if (S.isolated) {
        S.Omega = 0.0;
    } else if (time < S.ramp_end_time()) {
        const double xx = time / S.ramp_end_time();
        S.ramp_factor = 3.0 * xx * xx - 2.0 * xx * xx * xx;
        S.Omega = S.Omega_0 * S.ramp_factor;
    } else {
        S.Omega = std::sqrt((S.M1 + S.M2) / std::pow(S.R, 3));
    }
    S.prf = S.flags.ramp_secondary_potential ? S.ramp_factor : 1.0;
    // radial profiles: rebuilt only when the orbit or the companion weight changed
    const double key[6] = {S.R1, S.R2, S.q, S.M1 + S.M2, S.prf, S.mu};
    if (!S.E_valid || std::memcmp(key, S.E_key, sizeof(key)) != 0) {
        S.E_pending = true;
        std::memcpy(S.E_key, key, sizeof(key));
        S.E_valid = true;
        S.fac_valid = false;
    }
    gmd::StepParams &P = *S.hP.p;
    P.R1 = S.R1;
    P.R2 = S.R2;
    P.q = S.q;
    P.Mt = S.M1 + S.M2;
    P.prf = S.prf;
    P.w1 = 1.0;
    for (int m = -S.l_max; m <= S.l_max; ++m)
        P.ph[m + S.l_max] = make_double2(std::cos(m * S.theta), std::sin(m * S.theta));
    S.Omega_V = S.Omega;
    if (S.absorption) build_absorption(S);
    S.sink_active = S.absorption && time >= S.ramp_end_time();
    P.half_dt = S.dt * 0.5;
    P.dt = S.dt;
    P.idx_ex = S.idx_ex;
    P.nsink = 0;
    P.g[0] = P.g[1] = 0.0;
    if (S.sink_active)
        for (int bdy = 0; bdy < 2; ++bdy)
            if (S.sink[bdy].strength != 0.0 && sink_nonzero(S, bdy)) {
                P.sink_body[P.nsink] = bdy;
                P.g[P.nsink] = S.sink[bdy].strength;
                P.nsink++;
            }
    // Woodbury columns: P is block-diagonal in the modes and v_a(theta) =
    // D(theta) v_a(0) with D = diag(exp(-i m_j theta)) (absorption.cpp:127-151),
    // so W_a(theta) = D(theta - theta_W) W_a(theta_W) and V^H W does not depend
    // on theta. The columns are solved again only when P (dt, profiles, CAP) or
    // an absorber support changes; otherwise the kernels rotate the stored ones.
    {
        double wk[14] = {S.dt, S.E_key[0], S.E_key[1], S.E_key[2], S.E_key[3], S.E_key[4], S.E_key[5],
                         static_cast<double>(S.cap_version), static_cast<double>(P.nsink),
                         static_cast<double>(P.sink_body[0] + 2 * P.sink_body[1]), 0, 0, 0, 0};
        for (int s = 0; s < P.nsink; ++s) {
            const Sink &sk = S.sink[P.sink_body[s]];
            wk[10 + 2 * s] = sk.lo;
            wk[11 + 2 * s] = sk.n;
        }
        S.w_dirty = P.nsink > 0 && (!S.w_valid || std::memcmp(wk, S.w_key, sizeof(wk)) != 0);
        if (S.w_dirty) {
            std::memcpy(S.w_key, wk, sizeof(wk));
            S.w_theta = S.theta;
        }
        for (int m = -S.l_max; m <= S.l_max; ++m)
            P.wrot[m + S.l_max] = make_double2(std::cos(m * (S.theta - S.w_theta)), -std::sin(m * (S.theta - S.w_theta)));
    }
    HIPCK(hipMemcpyAsync(S.dP.p, S.hP.p, sizeof(gmd::StepParams), hipMemcpyHostToDevice, S.st));
    if (S.absorption) launch_absorbers(S);
    if (!defer && S.E_pending) {
        launch_profiles(S);
        S.E_pending = false;
    }
}

void fill_sinks(const gm_solver &S, gmd::SinkDev (&sk)[2])
{
    const gmd::StepParams &P = *S.hP.p;
    for (int s = 0; s < 2; ++s) {
        sk[s] = gmd::SinkDev{};
        if (s < P.nsink) {
            const Sink &src = S.sink[P.sink_body[s]];
            sk[s].lo = src.lo;
            sk[s].n = src.n;
            sk[s].v = src.dv.p;
            sk[s].W = src.W.p;
        }
    }
}

gmk::ApplyArgs apply_args(gm_solver &S, const double2 *X, double2 *Y)
{
    gmk::ApplyArgs a{};
    a.g = geometry(S);
    a.P = S.dP.p;
    a.Kr = S.dK.p;
    a.X = X;
    a.Y = Y;
    fill_sinks(S, a.sk);
    a.partial = S.partial.p;
    a.rowpart = S.rowpart.p;
    a.dpsi_ix = S.dpsi_ix.p;
    a.ng = S.apply_ng[0] ? S.apply_ng[0] : 2;
    a.Xt = S.xt.p;
    return a;
}

gmk::PrecArgs prec_args(gm_solver &S)
{
    gmk::PrecArgs a{};
    a.N = S.Nr;
    a.K = S.K;
    a.off = -(1.0 / (S.dr * S.dr)) / (2 * S.mu);
    a.inv = S.inv.p;
    a.Kr = S.dK.p;
    fill_sinks(S, a.sk);
    a.Ps = S.dP.p;
    a.lmax = S.l_max;
    a.pseq = S.pseq.p;
    a.nseg = gmd::precond_nseg(S.Nr);
    a.wpart = S.wpart.p;
    return a;
}

// Thomas pivots are recomputed when dt, the E table or the CAP changed;
// factor_needed records the new key.
bool factor_needed(gm_solver &S)
{
    const double key[3] = {S.dt, S.E_key[4], S.E_key[3]};
    if (S.fac_valid && std::memcmp(key, S.fac_key, sizeof(key)) == 0 && S.fac_cap_version == S.cap_version) return false;
    std::memcpy(S.fac_key, key, sizeof(key));
    S.fac_cap_version = S.cap_version;
    S.fac_valid = true;
    return true;
}

void launch_factor(gm_solver &S)
{
    gmk::FactorArgs f{};
    f.N = S.Nr;
    f.K = S.K;
    f.lmax = S.l_max;
    f.dr = S.dr;
    f.eps = S.epsilon;
    f.mu = S.mu;
    f.off = -(1.0 / (S.dr * S.dr)) / (2 * S.mu);
    f.P = S.dP.p;
    f.Gdiag = S.Gdiag.p;
    f.E = S.E.p;
    f.cap = S.cap ? S.dcap.p : nullptr;
    f.inv = S.inv.p;
    f.inv32 = S.inv32.p;
    f.failed = S.dflags.p;
    f.pseq = S.pseq.p;
    launch(S, GM_PROF_FACTOR, [&] { return gmk::launch_factor(f, S.st); });
}

void factorize(gm_solver &S)
{
    if (factor_needed(S)) launch_factor(S);
}

gmk::SinkArgs sink_proj_args(gm_solver &S, const double2 *X, int dest, double scale, int by_nrm)
{
    gmk::SinkArgs a{};
    a.N = S.Nr;
    a.K = S.K;
    a.mode = gmk::SK_PROJ;
    a.dest = dest;
    a.scale = scale;
    a.scale_by_nrm = by_nrm;
    a.X = X;
    fill_sinks(S, a.sk);
    a.Kr = S.dK.p;
    a.P = S.dP.p;
    return a;
}

void sink_proj(gm_solver &S, const double2 *X, int dest, double scale, int by_nrm)
{
    const gmk::SinkArgs a = sink_proj_args(S, X, dest, scale, by_nrm);
    launch(S, GM_PROF_SMALL, [&] { return gmk::launch_sink(a, S.st); });
}

// The capacitance G = diag(1/((dt/4) g)) + V^H W and its inverse, every step
// (the stored columns are rotated to the step's angle).
gmk::SinkArgs woodbury_setup_args(gm_solver &S)
{
    gmk::SinkArgs a{};
    a.N = S.Nr;
    a.K = S.K;
    a.mode = gmk::SK_SETUP;
    a.lmax = S.l_max;
    fill_sinks(S, a.sk);
    a.Kr = S.dK.p;
    a.P = S.dP.p;
    return a;
}

// Woodbury columns W_a = P^-1 v_a (matrices.cpp:133-145), solved again only
// when P or an absorber support changed.
void woodbury_columns(gm_solver &S)
{
    const gmd::StepParams &P = *S.hP.p;
    gmk::PrecArgs pa = prec_args(S);
    // the exact non-zero node range of each column comes out of its solve
    S.hwr.p[0] = S.hwr.p[2] = S.Nr;
    S.hwr.p[1] = S.hwr.p[3] = 0;
    gmd::Krylov *dk = S.dK.p;
    if (S.w_dirty) {
        HIPCK(hipMemcpyAsync(&dk->wrange[0], S.hwr.p, 4 * sizeof(int), hipMemcpyHostToDevice, S.st));
        for (int s = 0; s < P.nsink; ++s) {
            Sink &sk = S.sink[P.sink_body[s]];
            if (sk.W.n != S.nvec()) sk.W.alloc(S.nvec());
        }
        // (fill_sinks again: the columns may just have been allocated)
        fill_sinks(S, pa.sk);
        pa.wood = 0;
        pa.ncol = P.nsink;
        pa.range = &dk->wrange[0];
        launch(S, GM_PROF_PRECOND_SETUP, [&] { return gmk::launch_precond(pa, gmk::PV_WCOL, S.st); });
        S.w_valid = true;
    }
}

// Woodbury columns and capacitance (matrices.cpp:133-145).
void woodbury_setup(gm_solver &S)
{
    woodbury_columns(S);
    const gmk::SinkArgs a = woodbury_setup_args(S);
    launch(S, GM_PROF_SMALL, [&] { return gmk::launch_sink(a, S.st); });
}

// y -= sum_a c_a W_a over the columns' support (matrices.cpp:151)
gmk::WcorrArgs wcorr_args(gm_solver &S, double2 *Y, void *Xt, int xt_f32)
{
    const gmd::StepParams &P = *S.hP.p;
    gmk::WcorrArgs w{};
    w.N = S.Nr;
    w.K = S.K;
    w.nsink = P.nsink;
    w.Kr = S.dK.p;
    w.wpart = S.wpart.p;
    w.W0 = S.sink[P.sink_body[0]].W.p;
    w.W1 = P.nsink > 1 ? S.sink[P.sink_body[1]].W.p : nullptr;
    w.Y = Y;
    w.Xt = Xt;
    w.xt_f32 = xt_f32;
    w.P = S.dP.p;
    w.lmax = S.l_max;
    return w;
}

void woodbury_correct(gm_solver &S, double2 *Y, void *Xt = nullptr, int xt_f32 = 0)
{
    const gmk::WcorrArgs w = wcorr_args(S, Y, Xt, xt_f32);
    launch(S, GM_PROF_SMALL, [&] { return gmk::launch_wcorr(w, S.st); });
}

// Arguments of the launches of one BiCGSTAB iteration (matrices.cpp:62-89):
// y = M^-1 p with p = r + beta (p - w v), v = A y, z = M^-1 s with s = r -
// alpha v, t = A z, then the x and r updates. Mixed precision: the Thomas
// sweeps in FP32 on FP32 pivots and the coupling of the two applies in FP32;
// rhs, residual check and diagnostics keep the FP64 operator. The
// preconditioner (and the Woodbury correction) also write the phased copy
// exp(i m theta) y that the next apply's coupling reads.
// The phased copy exp(i m theta) y that a Gaunt-form Krylov apply reads, or
// nullptr when the Krylov applies take the polar form (it phases x itself).
void *krylov_xt(gm_solver &S)
{
    const int mixed = S.flags.precision == GM_PRECISION_MIXED;
    const int ng = mixed && S.apply_ng[1] ? S.apply_ng[1] : (S.apply_ng[0] ? S.apply_ng[0] : 2);
    return ng == gmk::TAPPLY_NG ? nullptr : S.xt.p;
}

gmk::PrecArgs iter_prec_args(gm_solver &S, bool wood, int var)
{
    gmk::PrecArgs pa = prec_args(S);
    pa.wood = wood;
    const int mixed = S.flags.precision == GM_PRECISION_MIXED;
    if (mixed) pa.inv32 = S.inv32.p;
    pa.Xt = krylov_xt(S);
    pa.xt_f32 = mixed;
    pa.R = S.rr.p;
    pa.R0 = S.r0.p;
    pa.V = S.v.p;
    pa.Pv = var == gmk::PV_P ? S.p.p : S.s.p;
    pa.Pv2 = var == gmk::PV_P ? S.p2.p : nullptr;
    pa.Y = var == gmk::PV_P ? S.y.p : S.z.p;
    return pa;
}

gmk::ApplyArgs iter_apply_args(gm_solver &S, int mode)
{
    const int mixed = S.flags.precision == GM_PRECISION_MIXED;
    gmk::ApplyArgs a = mode == gmk::AM_V ? apply_args(S, S.y.p, S.v.p) : apply_args(S, S.z.p, S.t.p);
    a.R0 = mode == gmk::AM_V ? S.r0.p : S.s.p;
    a.mixed = mixed;
    a.xt_ready = 1;
    if (mixed && S.apply_ng[1]) a.ng = S.apply_ng[1];
    return a;
}

gmk::UpdateArgs update_args(gm_solver &S)
{
    gmk::UpdateArgs u{};
    u.n = S.nvec();
    u.Kr = S.dK.p;
    u.X = S.x.p;
    u.Yc = S.y.p;
    u.Zc = S.z.p;
    u.S = S.s.p;
    u.T = S.t.p;
    u.R0 = S.r0.p;
    u.R = S.rr.p;
    u.partial = S.partial.p;
    return u;
}

void enqueue_iteration(gm_solver &S, bool wood)
{
    const int mixed = S.flags.precision == GM_PRECISION_MIXED;
    const gmk::PrecArgs pp = iter_prec_args(S, wood, gmk::PV_P);
    launch(S, GM_PROF_PRECOND, [&] { return gmk::launch_precond(pp, gmk::PV_P, S.st); });
    if (wood) woodbury_correct(S, S.y.p, krylov_xt(S), mixed);
    const gmk::ApplyArgs av = iter_apply_args(S, gmk::AM_V);
    launch(S, GM_PROF_APPLY_KRYLOV, [&] { return gmk::launch_apply(av, gmk::AM_V, S.st); });
    const gmk::PrecArgs ps = iter_prec_args(S, wood, gmk::PV_S);
    launch(S, GM_PROF_PRECOND, [&] { return gmk::launch_precond(ps, gmk::PV_S, S.st); });
    if (wood) woodbury_correct(S, S.z.p, krylov_xt(S), mixed);
    const gmk::ApplyArgs at = iter_apply_args(S, gmk::AM_T);
    launch(S, GM_PROF_APPLY_KRYLOV, [&] { return gmk::launch_apply(at, gmk::AM_T, S.st); });
    const gmk::UpdateArgs u = update_args(S);
    launch(S, GM_PROF_UPDATE, [&] { return gmk::launch_update(u, S.st); });
}


constexpr int kMaxIt = 500;  // matrices.cpp:104

// Graph-captured parts of a step: the start of the solve (orbit profiles and
// pivots when the orbit moved, Woodbury columns, rhs, init, first
// iterations), a batch of iterations, and the tail (residual check, mask,
// diagnostics, readbacks). Per-step values (dt, orbit radii and weights,
// theta, phases, sink strengths, Krylov scalars) live in device memory, so a
// graph stays valid while the kernel arguments below do not change; it is
// captured once the same arguments were requested on two consecutive steps.
// An evolving orbit replays the same graphs until an absorber support crosses
// a radial node.
enum GraphKind { GK_BEGIN = 0, GK_BATCH = 1, GK_TAIL = 2 };

std::vector<uint64_t> graph_key(const gm_solver &S, int batch)
{
    std::vector<uint64_t> k;
    k.push_back(static_cast<uint64_t>(batch));
    k.push_back(S.step_wood);
    k.push_back(static_cast<uint64_t>(S.flags.precision));
    k.push_back(S.mask);
    k.push_back(static_cast<uint64_t>(S.apply_ng[0] | (S.apply_ng[1] << 8)));
    k.push_back(S.w_dirty);
    k.push_back(static_cast<uint64_t>(S.step_E) | (static_cast<uint64_t>(S.step_fac) << 1));
    gmd::SinkDev sk[2];
    fill_sinks(S, sk);
    for (const auto &d : sk) {
        k.push_back((static_cast<uint64_t>(static_cast<uint32_t>(d.lo)) << 32) | static_cast<uint32_t>(d.n));
        k.push_back(reinterpret_cast<uint64_t>(d.v));
        k.push_back(reinterpret_cast<uint64_t>(d.W));
    }
    const gmd::StepParams &P = *S.hP.p;
    k.push_back(static_cast<uint64_t>(P.nsink) | (static_cast<uint64_t>(P.sink_body[0]) << 8) |
                (static_cast<uint64_t>(P.sink_body[1]) << 16));
    return k;
}

bool graphs_enabled(const gm_solver &S)
{
    return S.flags.use_graphs && !S.prof_on;
}

// Enqueue `work` on the solver's stream, through a cached graph when enabled.
template <class F>
void run_captured(gm_solver &S, int kind, const std::vector<uint64_t> &key, F &&work)
{
    if (!graphs_enabled(S)) {
        work();
        return;
    }
    for (auto &e : S.graphs)
        if (e.kind == kind && e.key == key) {
            HIPCK(hipGraphLaunch(e.exec, S.st));
            return;
        }
    std::vector<uint64_t> &last = S.last_key[kind][S.flags.precision == GM_PRECISION_MIXED];
    const bool repeat = last == key;
    last = key;
    if (!repeat) {
        work();
        return;
    }
    hipGraph_t g = nullptr;
    HIPCK(hipStreamBeginCapture(S.st, hipStreamCaptureModeThreadLocal));
    try {
        work();
    } catch (...) {
        (void)hipStreamEndCapture(S.st, &g);
        if (g) (void)hipGraphDestroy(g);
        throw;
    }
    HIPCK(hipStreamEndCapture(S.st, &g));
    hipGraphExec_t ex = nullptr;
    const hipError_t e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    HIPCK(e);
    // a handful of shapes per solver: keep the newest
    if (S.graphs.size() >= 8) {
        (void)hipGraphExecDestroy(S.graphs.front().exec);
        S.graphs.erase(S.graphs.begin());
    }
    S.graphs.push_back({kind, key, ex});
    HIPCK(hipGraphLaunch(ex, S.st));
}

void drop_graphs(gm_solver &S)
{
    for (auto &e : S.graphs) (void)hipGraphExecDestroy(e.exec);
    S.graphs.clear();
    for (auto &kp : S.last_key)
        for (auto &k : kp) k.clear();
}

// A batch of BiCGSTAB iterations followed by the readback of the Krylov state.
void enqueue_batch(gm_solver &S, int batch)
{
    run_captured(S, GK_BATCH, graph_key(S, batch), [&] {
        for (int k = 0; k < batch; ++k) enqueue_iteration(S, S.step_wood);
        HIPCK(hipMemcpyAsync(S.hK.p, S.dK.p, sizeof(gmd::Krylov), hipMemcpyDeviceToHost, S.st));
    });
    S.step_enqueued += batch;
}

void solve_begin_work(gm_solver &S, int batch);

// solve_CN (matrices.cpp:94-215) with the block preconditioner, up to the
// first convergence check: pivots, Woodbury columns, rhs, BiCGSTAB init and a
// first batch of iterations (as many as the previous step needed).
void solve_begin(gm_solver &S)
{
    const gmd::StepParams &P = *S.hP.p;
    const bool wood = P.nsink > 0;
    S.step_wood = wood;
    S.step_enqueued = 0;
    S.step_E = S.E_pending;
    S.E_pending = false;
    S.step_fac = factor_needed(S);
    const int batch = std::max(1, S.prev_iters);
    run_captured(S, GK_BEGIN, graph_key(S, batch), [&] { solve_begin_work(S, batch); });
    S.step_enqueued = batch;
}

void rhs_and_init(gm_solver &S);

void solve_begin_work(gm_solver &S, int batch)
{
    if (S.step_E) launch_profiles(S);
    if (S.step_fac) launch_factor(S);
    const bool wood = S.step_wood;
    if (wood) woodbury_setup(S);
    rhs_and_init(S);
    for (int k = 0; k < batch; ++k) enqueue_iteration(S, wood);
    HIPCK(hipMemcpyAsync(S.hK.p, S.dK.p, sizeof(gmd::Krylov), hipMemcpyDeviceToHost, S.st));
}

// rhs = B psi - (dt/4) Gamma psi and the raw shadow residual rhs - A psi in
// one apply (b stays unnormalised in S.b; the acceptance divides).
gmk::ApplyArgs rhs_apply_args(gm_solver &S)
{
    gmk::ApplyArgs a = apply_args(S, S.psi.p, S.b.p);
    a.Y2 = S.r0.p;
    return a;
}

// Eigen's BiCGSTAB without sinks, the reference's own recurrence with them.
gmk::InitArgs init_args(gm_solver &S, bool wood)
{
    gmk::InitArgs ia{};
    ia.n = S.nvec();
    ia.Kr = S.dK.p;
    ia.Psi = S.psi.p;
    ia.X = S.x.p;
    ia.tol = 1e-10;
    ia.maxit = kMaxIt;
    ia.eigen_mode = wood ? 0 : 1;
    ia.rst_eps2 = S.rst_eps2;
    ia.rst_ref = S.rst_ref;
    return ia;
}

// Right-hand side, its norm and the BiCGSTAB start (b, x, r, r0 normalised).
void rhs_and_init(gm_solver &S)
{
    const bool wood = S.step_wood;
    if (wood) sink_proj(S, S.psi.p, gmk::SD_START, 1.0, 0);
    const gmk::ApplyArgs a = rhs_apply_args(S);
    launch(S, GM_PROF_APPLY_OTHER, [&] { return gmk::launch_apply(a, gmk::AM_RHS, S.st); });
    const gmk::InitArgs ia = init_args(S, wood);
    launch(S, GM_PROF_SMALL, [&] { return gmk::launch_init(ia, S.st); });
}

void solve_tail_work(gm_solver &S);
void ilut_host_solve(gm_solver &S);

// The restart branch of BiCGSTAB (matrices.cpp:67-74 and Eigen's): the
// iteration that found |rho| below the threshold stopped at its top; r = b -
// A x is formed with the FP64 operator and the rank-one sinks of x, then the
// restart apply's last block resets r0, rho, the counter (Eigen, first
// restart) and beta, and clears the stop so the next batch goes on.
void enqueue_restart(gm_solver &S)
{
    if (S.step_wood) sink_proj(S, S.x.p, gmk::SD_COEFA, 1.0, 0);
    gmk::ApplyArgs a = apply_args(S, S.x.p, S.rr.p);
    a.Y2 = S.r0.p;
    a.R0 = S.b.p;
    a.r0_raw = 1;
    launch(S, GM_PROF_APPLY_OTHER, [&] { return gmk::launch_apply(a, gmk::AM_RST, S.st); });
}

// After a convergence check (host copy of the Krylov state is current): more
// iterations (false) or the acceptance check, mask and diagnostics with their
// readback (true).
bool solve_advance(gm_solver &S)
{
    gmd::Krylov &K = *S.hK.p;
    if (!K.done || K.restart) {
        // Eigen resets its counter on the first restart, so a solve that
        // restarts may run up to twice the iteration limit
        if (S.step_enqueued > 3 * kMaxIt) throw std::runtime_error("BiCGSTAB did not report convergence");
        if (K.restart) enqueue_restart(S);
        enqueue_batch(S, 2);
        return false;
    }
    S.last_iterations = K.iter;
    S.last_restarts = K.restarts;
    S.prev_iters = std::max(1, S.last_iterations);
    run_captured(S, GK_TAIL, graph_key(S, 0), [&] { solve_tail_work(S); });
    return true;
}

// The acceptance apply: |A_full x - b| with the rank-one sinks of x (their
// coefficients from one projection of x), keeping V x for the diagnostics.
gmk::ApplyArgs res_apply_args(gm_solver &S)
{
    gmk::ApplyArgs a = apply_args(S, S.x.p, nullptr);
    a.R0 = S.b.p;
    a.r0_raw = 1;  // b = rhs / |rhs|
    a.Cout = S.tmp.p;  // V x, reused by the diagnostics of the accepted state
    return a;
}

gmk::AcceptArgs accept_args(gm_solver &S)
{
    gmk::AcceptArgs ac{};
    ac.n = S.nvec();
    ac.N = S.Nr;
    ac.Kr = S.dK.p;
    ac.X = S.x.p;
    ac.mask = S.mask ? S.dmask.p : nullptr;
    ac.Psi = S.psi.p;
    ac.status = S.dflags.p + 1;
    ac.resid = S.dresid.p;
    return ac;
}

// Diagnostics on the post-mask state (extraction.cpp), with Gamma psi. V psi
// = |rhs| mask V x: the coupling is node-local and the mask is the same for
// every mode, so only the stencil is recomputed.
gmk::ApplyArgs diag_apply_args(gm_solver &S)
{
    gmk::ApplyArgs a = apply_args(S, S.psi.p, nullptr);
    a.Cin = S.tmp.p;
    a.mask = S.mask ? S.dmask.p : nullptr;
    return a;
}

gmk::DiagArgs diag_final_args(gm_solver &S)
{
    gmk::DiagArgs da{};
    da.N = S.Nr;
    da.K = S.K;
    da.ix = S.idx_ex;
    da.ntiles = (S.Nr + gmd::DIAG_NODES - 1) / gmd::DIAG_NODES;  // the diagnostics apply's tiles
    da.dr = S.dr;
    da.eps = S.epsilon;
    da.mu = S.mu;
    da.Psi = S.psi.p;
    da.dpsi_ix = S.dpsi_ix.p;
    da.rowpart = S.rowpart.p;
    da.out = S.diagout.p;
    return da;
}

// The step's results back to the host: diagnostics, flags, residual, Krylov state.
void tail_readback(gm_solver &S, hipStream_t st)
{
    HIPCK(hipMemcpyAsync(S.hdiag.p, S.diagout.p, sizeof(double) * (8 * S.K), hipMemcpyDeviceToHost, st));
    HIPCK(hipMemcpyAsync(S.hflags.p, S.dflags.p, 2 * sizeof(int), hipMemcpyDeviceToHost, st));
    HIPCK(hipMemcpyAsync(S.hresid.p, S.dresid.p, sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCK(hipMemcpyAsync(S.hK.p, S.dK.p, sizeof(gmd::Krylov), hipMemcpyDeviceToHost, st));
}

void solve_tail_work(gm_solver &S)
{
    const gmd::StepParams &P = *S.hP.p;
    if (S.step_wood) sink_proj(S, S.x.p, gmk::SD_COEFA_QNEW, 1.0, 0);
    const gmk::ApplyArgs ar = res_apply_args(S);
    launch(S, GM_PROF_APPLY_OTHER, [&] { return gmk::launch_apply(ar, gmk::AM_RES, S.st); });
    const gmk::AcceptArgs ac = accept_args(S);
    launch(S, GM_PROF_SMALL, [&] { return gmk::launch_accept(ac, S.st); });
    if (P.nsink) sink_proj(S, S.psi.p, gmk::SD_POST, 1.0, 0);
    const gmk::ApplyArgs ad = diag_apply_args(S);
    launch(S, GM_PROF_DIAG, [&] { return gmk::launch_diag_apply(ad, S.st); });
    const gmk::DiagArgs da = diag_final_args(S);
    launch(S, GM_PROF_DIAG, [&] { return gmk::launch_diag_final(da, S.st); });
    tail_readback(S, S.st);
}

void alloc_device(gm_solver &S)
{
    const size_t n = S.nvec();
    for (auto *bf : {&S.psi, &S.b, &S.x, &S.rr, &S.r0, &S.p, &S.p2, &S.v, &S.y, &S.z, &S.s, &S.t, &S.inv, &S.tmp})
        if (bf->n != n) bf->alloc(n);
    const size_t nwp = static_cast<size_t>(4) * S.K * gmd::precond_nseg(S.Nr);
    if (S.wpart.n != nwp) S.wpart.alloc(nwp);
    const size_t nflags = static_cast<size_t>(S.K) * (gmd::precond_nseg(S.Nr) + 1);
    if (S.pseq.n != nflags) {
        S.pseq.alloc(nflags);
        HIPCK(hipMemsetAsync(S.pseq.p, 0, nflags * sizeof(int), S.st));
    }
    if (S.inv32.n != n) S.inv32.alloc(n);
    if (S.xt.n != n) S.xt.alloc(n);
    S.E.alloc(static_cast<size_t>(2) * (2 * S.l_max + 1) * S.Nr);
    S.E32.alloc(static_cast<size_t>(2) * (2 * S.l_max + 1) * S.Nr);
    S.lcut.alloc(S.ntiles);
    int per = 0;
    const int grid = gmk::apply_grid(S.Nr, S.nchunks, 1, &per);
    S.partial.alloc(static_cast<size_t>(3) * std::max({grid, gmk::tapply_grid(S.Nr), gmd::RED_BLOCKS}));
    S.sinkscr.alloc(1);
    S.rowpart.alloc(static_cast<size_t>(S.ntiles) * S.K * 6);
    S.dpsi_ix.alloc(S.K);
    S.diagout.alloc(static_cast<size_t>(8) * S.K);
    S.hdiag.alloc(static_cast<size_t>(8) * S.K);
    S.mchunks.alloc(S.mchunks_h.size());
    S.mblocks.alloc(S.mblocks_h.size());
    S.Gdiag.alloc(S.Gdiag_h.size());
    // tables are uploaded from pageable vectors: synchronous copies
    HIPCK(hipMemcpy(S.mchunks.p, S.mchunks_h.data(), S.mchunks_h.size() * sizeof(gmd::MChunk), hipMemcpyHostToDevice));
    HIPCK(hipMemcpy(S.mblocks.p, S.mblocks_h.data(), S.mblocks_h.size() * sizeof(gmd::MBlock), hipMemcpyHostToDevice));
    {
        // The coupling streams of each chunk: for every column (l1, m1), the
        // slot groups of four covering the slots where some row of the chunk
        // has a non-zero Gaunt coefficient (the groups stay inside the
        // block's nTp slots; slots outside that range hold zeros for this
        // column, so a group may extend over them). Consecutive groups of a
        // column go in pairs (groups 0 and 1, 2 and 3, ...) to the chunk's
        // pair list, an odd last group to its single list; both are sorted by
        // the l* of their first slot e. A tile with cut lc runs
        //   - the pairs with e + 8 <= lc, both groups chained (a prefix),
        //   - the pairs with e <= lc < e + 8 as their first group alone (the
        //     next entries of the same list: the second group lies above the
        //     cut),
        //   - the singles with e <= lc (a prefix),
        // so it runs the same MFMAs as with one step per group.
        const int Lm = S.l_max + 1, Lc = 2 * S.l_max + 1;
        const size_t CL = static_cast<size_t>(S.nchunks) * Lc;
        std::vector<gmd::MStep> msteps;  // g: source offset into GM_h until staged
        std::vector<int> nvals;          // fragment values of each step (64, or 128 paired)
        std::vector<char> padding;
        std::vector<int> sstart(S.nchunks, 0), pstart(S.nchunks, 0);
        std::vector<int> scount(CL, 0), pfull(CL, 0), pend(CL, 0);
        auto append = [&](const std::vector<gmd::MStep> &v, int vals) {
            msteps.insert(msteps.end(), v.begin(), v.end());
            nvals.insert(nvals.end(), v.size(), vals);
            padding.insert(padding.end(), v.size(), 0);
            // padding steps: the last step's fragment and column, the zero LDS rows
            const gmd::MStep last = v.empty() ? gmd::MStep{} : v.back();
            for (int k = 0; k + 1 < gmd::RING; ++k) {
                msteps.push_back(gmd::MStep{last.g, last.x, gmd::ZROW, 0});
                nvals.push_back(vals);
                padding.push_back(1);
            }
        };
        for (int c = 0; c < S.nchunks; ++c) {
            std::vector<gmd::MStep> v1, v2;
            for (int l1 = S.l_max; l1 >= 0; --l1) {
                const gmd::MBlock &B = S.mblocks_h[static_cast<size_t>(c) * Lm + l1];
                const int p = (S.mchunks_h[c].par + l1) & 1;
                const int nTp = (B.nT + 3) / 4 * 4;
                for (int m1i = 0; m1i <= l1; ++m1i) {
                    int tl = nTp, th = -1;
                    for (int t = 0; t < B.nT; ++t)
                        for (int r = 0; r < gmd::MR; ++r)
                            if (S.GM_h[static_cast<size_t>(B.goff) + (static_cast<size_t>(m1i) * nTp + t) * gmd::MR + r] != 0.0) {
                                tl = std::min(tl, t);
                                th = std::max(th, t);
                            }
                    if (th < 0) continue;
                    const int ng = (th - tl + 4) / 4;
                    const int t0 = std::min(tl, nTp - 4 * ng);
                    for (int k = 0; k < ng; k += 2) {
                        const int t = t0 + 4 * k;
                        const gmd::MStep st{B.goff + (m1i * nTp + t) * gmd::MR, l1 * (l1 + 1) / 2 + m1i,
                                            2 * (B.tlo + t) + p, 0};
                        (k + 1 < ng ? v2 : v1).push_back(st);
                    }
                }
            }
            const auto by_e = [](const gmd::MStep &x, const gmd::MStep &y) { return x.e < y.e; };
            std::stable_sort(v1.begin(), v1.end(), by_e);
            std::stable_sort(v2.begin(), v2.end(), by_e);
            for (int lc = 0; lc < Lc; ++lc) {
                const size_t ix = static_cast<size_t>(c) * Lc + lc;
                for (const gmd::MStep &e : v1) scount[ix] += e.e <= lc;
                for (const gmd::MStep &e : v2) {
                    pfull[ix] += e.e + 8 <= lc;
                    pend[ix] += e.e <= lc;
                }
            }
            pstart[c] = static_cast<int>(msteps.size());
            append(v2, 8 * gmd::MR);
            sstart[c] = static_cast<int>(msteps.size());
            append(v1, 4 * gmd::MR);
        }
        msteps.push_back(gmd::MStep{});
        nvals.push_back(4 * gmd::MR);
        padding.push_back(1);
        // Gaunt fragments in stream order: the slots x 16 rows of every step
        // stored contiguously (a tile streams prefixes and ranges of its
        // chunk's lists front to back)
        std::vector<double> gmc;
        gmc.reserve(msteps.size() * 4 * gmd::MR);
        for (size_t k = 0; k < msteps.size(); ++k) {
            gmd::MStep &e = msteps[k];
            if (padding[k]) {  // the previous step's fragment (the first entry is never padding)
                e.g = k > 0 ? msteps[k - 1].g : 0;
                continue;
            }
            const size_t src = static_cast<size_t>(e.g);
            e.g = static_cast<int>(gmc.size());
            for (int v = 0; v < nvals[k]; ++v) gmc.push_back(src + v < S.GM_h.size() ? S.GM_h[src + v] : 0.0);
        }
        S.gm_n = static_cast<unsigned>(gmc.size());
        S.GM.alloc(gmc.size());
        S.GM32.alloc(gmc.size());
        HIPCK(hipMemcpy(S.GM.p, gmc.data(), gmc.size() * sizeof(double), hipMemcpyHostToDevice));
        const std::vector<float> gm32(gmc.begin(), gmc.end());
        HIPCK(hipMemcpy(S.GM32.p, gm32.data(), gm32.size() * sizeof(float), hipMemcpyHostToDevice));
        if (S.theta_h) {
            const gmt::ThetaTables &T = *S.theta_h;
            S.TA.alloc(T.A.size());
            S.TC.alloc(T.C.size());
            S.TU.alloc(T.U.size());
            HIPCK(hipMemcpy(S.TA.p, T.A.data(), T.A.size() * sizeof(double), hipMemcpyHostToDevice));
            HIPCK(hipMemcpy(S.TC.p, T.C.data(), T.C.size() * sizeof(double), hipMemcpyHostToDevice));
            HIPCK(hipMemcpy(S.TU.p, T.U.data(), T.U.size() * sizeof(double), hipMemcpyHostToDevice));
            S.Ut.alloc(gmk::tapply_u_doubles(S.Nr, S.l_max));
        }
        S.msteps.alloc(msteps.size());
        S.mstep_start.alloc(sstart.size());
        S.pair_start.alloc(pstart.size());
        S.step_count.alloc(scount.size());
        S.pair_full.alloc(pfull.size());
        S.pair_end.alloc(pend.size());
        HIPCK(hipMemcpy(S.msteps.p, msteps.data(), msteps.size() * sizeof(gmd::MStep), hipMemcpyHostToDevice));
        HIPCK(hipMemcpy(S.mstep_start.p, sstart.data(), sstart.size() * sizeof(int), hipMemcpyHostToDevice));
        HIPCK(hipMemcpy(S.pair_start.p, pstart.data(), pstart.size() * sizeof(int), hipMemcpyHostToDevice));
        HIPCK(hipMemcpy(S.step_count.p, scount.data(), scount.size() * sizeof(int), hipMemcpyHostToDevice));
        HIPCK(hipMemcpy(S.pair_full.p, pfull.data(), pfull.size() * sizeof(int), hipMemcpyHostToDevice));
        HIPCK(hipMemcpy(S.pair_end.p, pend.data(), pend.size() * sizeof(int), hipMemcpyHostToDevice));
    }
    HIPCK(hipMemcpy(S.Gdiag.p, S.Gdiag_h.data(), S.Gdiag_h.size() * sizeof(double), hipMemcpyHostToDevice));
    HIPCK(hipMemsetAsync(S.dK.p, 0, sizeof(gmd::Krylov), S.st));
    HIPCK(hipMemsetAsync(S.dflags.p, 0, 2 * sizeof(int), S.st));
}

// ------------------------------------------------------------------ C ABI plumbing

template <class F>
int guarded(gm_solver *S, F &&f)
{
    try {
        f();
        return GM_OK;
    } catch (const std::invalid_argument &e) {
        if (S) S->err = e.what();
        else g_create_err = e.what();
        return GM_E_INVALID;
    } catch (const std::logic_error &e) {
        if (S) S->err = e.what();
        else g_create_err = e.what();
        return GM_E_LOGIC;
    } catch (const DeviceError &e) {
        if (S) S->err = e.what();
        else g_create_err = e.what();
        return GM_E_DEVICE;
    } catch (const std::exception &e) {
        if (S) S->err = e.what();
        else g_create_err = e.what();
        return GM_E_RUNTIME;
    }
}

void to_host(gm_solver &S, const double2 *d, double *h)
{
    HIPCK(hipMemcpyAsync(h, d, S.nvec() * sizeof(double2), hipMemcpyDeviceToHost, S.st));
    HIPCK(hipStreamSynchronize(S.st));
}

void to_dev(gm_solver &S, const double *h, double2 *d)
{
    HIPCK(hipMemcpyAsync(d, h, S.nvec() * sizeof(double2), hipMemcpyHostToDevice, S.st));
    HIPCK(hipStreamSynchronize(S.st));
}

// apply_sink (absorption.cpp:100-109) on device vectors: Y = factor sum_a g_a
// (v_a^H X) v_a over the active sinks' supports, added to Y when accumulate
// is set; the projections reduce through scr.
void sink_apply_dev(const gm_solver &S, const double2 *X, double2 *Y, double factor, bool accumulate,
                    gmk::SinkScratch *scr, hipStream_t st)
{
    gmk::SinkApplyArgs a{};
    a.N = S.Nr;
    a.K = S.K;
    if (S.sink_active)
        for (const Sink &sk : S.sink) {
            if (sk.strength == 0.0 || sk.n == 0) continue;
            a.sk[a.nsink].lo = sk.lo;
            a.sk[a.nsink].n = sk.n;
            a.sk[a.nsink].v = sk.dv.p;
            a.scale[a.nsink] = factor * sk.strength;
            a.nsink++;
        }
    a.X = X;
    a.Y = Y;
    a.zero_base = accumulate ? 0 : 1;
    a.scr = scr;
    HIPCK(gmk::launch_sink_apply(a, st));
}

void apply_H_dev(gm_solver &S, const double2 *X, double2 *Y)
{
    gmk::ApplyArgs a = apply_args(S, X, Y);
    HIPCK(gmk::launch_apply(a, gmk::AM_H, S.st));
}

void apply_H_host(gm_solver &S, const double *in, double *out)
{
    to_dev(S, in, S.tmp.p);
    apply_H_dev(S, S.tmp.p, S.t.p);
    to_host(S, S.t.p, out);
}

void require_initialized(const gm_solver &S)
{
    if (!S.initialized) throw std::logic_error("initialize() must be called first");
}

// ------------------------------------------------------------------ full-ILUT solve

// Host memory the ILUT solve may use: the factor of the node-major CN operator
// keeps about 3K entries per row (24 bytes each with its column), plus the
// extracted operator and ten work vectors. Larger systems are refused.
constexpr double kIlutHostBytes = 32e9;

double ilut_bytes(const gm_solver &S)
{
    const double n = static_cast<double>(S.nvec());
    return n * (3.0 * S.K + 3.0) * 24.0 + n * (S.K + 3.0) * 24.0 + 12.0 * 16.0 * n;
}

// H as node-major CSR (row i*K + j for mode j at node i), read off the device
// operator with 3K applies. The probe for (c, j) has ones on mode j at every
// node i = c (mod 3); H couples modes only within a node and nodes only within
// a mode and to their neighbours, so row (k, i') of the result is the entry of
// the single probed column (j, i) with |i - i'| <= 1.
gmh::CsrZ extract_H(gm_solver &S)
{
    const int K = S.K, N = S.Nr;
    const size_t n = S.nvec();
    std::vector<std::vector<std::pair<int, gmh::cplx>>> rows(n);
    std::vector<double> probe(2 * n), out(2 * n);
    for (int c = 0; c < 3; ++c)
        for (int j = 0; j < K; ++j) {
            std::fill(probe.begin(), probe.end(), 0.0);
            for (int i = c; i < N; i += 3) probe[2 * (static_cast<size_t>(j) * N + i)] = 1.0;
            to_dev(S, probe.data(), S.s.p);
            gmk::ApplyArgs a = apply_args(S, S.s.p, S.t.p);
            launch(S, GM_PROF_APPLY_OTHER, [&] { return gmk::launch_apply(a, gmk::AM_H, S.st); });
            to_host(S, S.t.p, out.data());
            for (int k = 0; k < K; ++k)
                for (int i2 = 0; i2 < N; ++i2) {
                    const size_t q = static_cast<size_t>(k) * N + i2;
                    const gmh::cplx v(out[2 * q], out[2 * q + 1]);
                    if (v == gmh::cplx(0.0)) continue;
                    const int d = ((c - i2) % 3 + 3) % 3;
                    const int i = i2 + (d == 2 ? -1 : d);
                    if (i < 0 || i >= N) throw std::logic_error("operator probe: entry outside the stencil");
                    rows[static_cast<size_t>(i2) * K + k].emplace_back(i * K + j, v);
                }
        }
    gmh::CsrZ H;
    H.n = static_cast<int>(n);
    H.ptr.assign(n + 1, 0);
    for (size_t r = 0; r < n; ++r) {
        auto &row = rows[r];
        std::sort(row.begin(), row.end(), [](const auto &x, const auto &y) { return x.first < y.first; });
        for (const auto &e : row) {
            H.col.push_back(e.first);
            H.val.push_back(e.second);
        }
        H.ptr[r + 1] = static_cast<int>(H.val.size());
        std::vector<std::pair<int, gmh::cplx>>().swap(row);
    }
    return H;
}

// ------------------------------------------------------------------ public operators

// The operators the reference keeps as public sparse members (H, H0, A, B, V1,
// V2, V, dVdt, Id, CAP; include/CrankNicolsonModes.hpp:197-206) at the current
// step, read off the device tables in the reference's stacking (row j Nr + i).
// H and the coupling V come from 3K probes of the apply (mode j at nodes c,
// c + 3, ...; H couples modes only within a node and a node only to its
// neighbours, V only within a node); V1 and V2 from the same probes with the
// radial profiles of one body (the primary's profile weight or the
// companion's set to zero); H0 and CAP from the host formulas
// (hamiltonian.cpp:68-86); dVdt, A and B from V and H (hamiltonian.cpp:
// 205-219, matrices.cpp:27-40). Explicit zeros of the reference's patterns
// are left out.
std::vector<std::vector<std::pair<int32_t, gmh::cplx>>> probe_rows(gm_solver &S, bool coupling)
{
    const int K = S.K, N = S.Nr;
    const size_t n = S.nvec();
    std::vector<std::vector<std::pair<int32_t, gmh::cplx>>> rows(n);
    std::vector<double> probe(2 * n), out(2 * n);
    for (int c = 0; c < 3; ++c)
        for (int j = 0; j < K; ++j) {
            std::fill(probe.begin(), probe.end(), 0.0);
            for (int i = c; i < N; i += 3) probe[2 * (static_cast<size_t>(j) * N + i)] = 1.0;
            to_dev(S, probe.data(), S.s.p);
            gmk::ApplyArgs a = apply_args(S, S.s.p, S.t.p);
            a.Cout = S.tmp.p;
            launch(S, GM_PROF_APPLY_OTHER, [&] { return gmk::launch_apply(a, gmk::AM_H, S.st); });
            to_host(S, coupling ? S.tmp.p : S.t.p, out.data());
            for (int k = 0; k < K; ++k)
                for (int i2 = 0; i2 < N; ++i2) {
                    const size_t q = static_cast<size_t>(k) * N + i2;
                    const gmh::cplx v(out[2 * q], out[2 * q + 1]);
                    if (v == gmh::cplx(0.0)) continue;
                    const int d = ((c - i2) % 3 + 3) % 3;
                    const int i = i2 + (d == 2 ? -1 : d);
                    if (coupling && d != 0) throw std::logic_error("operator probe: coupling across nodes");
                    if (i < 0 || i >= N) throw std::logic_error("operator probe: entry outside the stencil");
                    rows[q].emplace_back(static_cast<int32_t>(static_cast<size_t>(j) * N + i), v);
                }
        }
    return rows;
}

OpCsr to_csr(std::vector<std::vector<std::pair<int32_t, gmh::cplx>>> &rows)
{
    OpCsr o;
    o.ptr.assign(rows.size() + 1, 0);
    for (size_t r = 0; r < rows.size(); ++r) {
        auto &row = rows[r];
        std::sort(row.begin(), row.end(), [](const auto &x, const auto &y) { return x.first < y.first; });
        for (const auto &e : row) {
            o.col.push_back(e.first);
            o.val.push_back(e.second);
        }
        o.ptr[r + 1] = static_cast<int64_t>(o.col.size());
        std::vector<std::pair<int32_t, gmh::cplx>>().swap(row);
    }
    return o;
}

// The coupling with the radial profiles weighted (w1, w2) for the primary and
// the companion; the step's profiles are restored afterwards.
std::vector<std::vector<std::pair<int32_t, gmh::cplx>>> probe_coupling(gm_solver &S, double w1, double w2)
{
    gmd::StepParams &P = *S.hP.p;
    const gmd::StepParams keep = P;
    P.w1 = w1;
    P.prf = w2;
    HIPCK(hipMemcpyAsync(S.dP.p, S.hP.p, sizeof(gmd::StepParams), hipMemcpyHostToDevice, S.st));
    launch_profiles(S);
    std::vector<std::vector<std::pair<int32_t, gmh::cplx>>> rows;
    try {
        rows = probe_rows(S, true);
    } catch (...) {
        P = keep;
        HIPCK(hipMemcpyAsync(S.dP.p, S.hP.p, sizeof(gmd::StepParams), hipMemcpyHostToDevice, S.st));
        launch_profiles(S);
        throw;
    }
    P = keep;
    HIPCK(hipMemcpyAsync(S.dP.p, S.hP.p, sizeof(gmd::StepParams), hipMemcpyHostToDevice, S.st));
    launch_profiles(S);
    HIPCK(hipStreamSynchronize(S.st));
    return rows;
}

OpCsr materialise(gm_solver &S, int which)
{
    require_initialized(S);
    const int K = S.K, N = S.Nr;
    const size_t n = S.nvec();
    if (static_cast<double>(n) * (K + 3) > static_cast<double>(1 << 24))
        throw std::invalid_argument("operator materialisation is for small systems: rows x (n_modes + 3) <= 2^24");
    if (which < GM_OP_H || which > GM_OP_CAP) throw std::invalid_argument("unknown operator");
    if (S.E_pending) {
        launch_profiles(S);
        S.E_pending = false;
    }
    HIPCK(hipStreamSynchronize(S.st));
    std::vector<std::vector<std::pair<int32_t, gmh::cplx>>> rows(n);
    const double hdt = 0.5 * S.dt;
    switch (which) {
    case GM_OP_ID:
        for (size_t r = 0; r < n; ++r) rows[r].emplace_back(static_cast<int32_t>(r), 1.0);
        break;
    case GM_OP_CAP:
        if (S.cap)
            for (int j = 0; j < K; ++j)
                for (int i = 0; i < N; ++i) {
                    const size_t r = static_cast<size_t>(j) * N + i;
                    const gmh::cplx v(S.capv[i].x, S.capv[i].y);
                    if (v != gmh::cplx(0.0)) rows[r].emplace_back(static_cast<int32_t>(r), v);
                }
        break;
    case GM_OP_H0: {
        const double alpha = 1.0 / (S.dr * S.dr), off = -alpha / (2 * S.mu);
        for (int j = 0; j < K; ++j)
            for (int i = 1; i < N - 1; ++i) {
                const size_t r = static_cast<size_t>(j) * N + i;
                const double l = S.ml[j];
                rows[r].emplace_back(static_cast<int32_t>(r - 1), off);
                rows[r].emplace_back(static_cast<int32_t>(r), -(-2.0 * alpha - l * (l + 1) / (S.r[i] * S.r[i])) / (2 * S.mu));
                rows[r].emplace_back(static_cast<int32_t>(r + 1), off);
            }
        break;
    }
    case GM_OP_V1:
        rows = probe_coupling(S, 1.0, 0.0);
        break;
    case GM_OP_V2:
        rows = probe_coupling(S, 0.0, 1.0);
        break;
    case GM_OP_V:
    case GM_OP_DVDT:
        rows = probe_rows(S, true);
        if (which == GM_OP_DVDT) {
            const double sign = S.counter ? -1.0 : 1.0;
            for (size_t r = 0; r < n; ++r) {
                const int j = static_cast<int>(r / N);
                for (auto &e : rows[r]) {
                    const int k = e.first / N;
                    e.second *= gmh::cplx(0.0, sign * static_cast<double>(S.mm[k] - S.mm[j]) * S.Omega_V);
                }
            }
        }
        break;
    default: {  // H, A, B
        rows = probe_rows(S, false);
        if (which != GM_OP_H) {
            const gmh::cplx sc(0.0, which == GM_OP_A ? hdt : -hdt);
            for (size_t r = 0; r < n; ++r) {
                bool diag = false;
                for (auto &e : rows[r]) {
                    e.second *= sc;
                    if (static_cast<size_t>(e.first) == r) {
                        e.second += 1.0;
                        diag = true;
                    }
                }
                if (!diag) rows[r].emplace_back(static_cast<int32_t>(r), 1.0);
            }
        }
    }
    }
    return to_csr(rows);
}

// A = I + i (dt/2) H (update_CN_values, matrices.cpp:27-42), with every
// diagonal entry present.
gmh::CsrZ cn_operator(const gmh::CsrZ &H, double half_dt)
{
    gmh::CsrZ A;
    A.n = H.n;
    A.ptr.assign(static_cast<size_t>(H.n) + 1, 0);
    A.col.reserve(H.nnz() + H.n);
    A.val.reserve(H.nnz() + H.n);
    const gmh::cplx s(0.0, half_dt);
    for (int r = 0; r < H.n; ++r) {
        bool diag = false;
        for (int q = H.ptr[r]; q < H.ptr[r + 1]; ++q) {
            const int c = H.col[q];
            if (!diag && c > r) {
                A.col.push_back(r);
                A.val.push_back(1.0);
                diag = true;
            }
            A.col.push_back(c);
            A.val.push_back((c == r ? 1.0 : 0.0) + s * H.val[q]);
            diag = diag || c == r;
        }
        if (!diag) {
            A.col.push_back(r);
            A.val.push_back(1.0);
        }
        A.ptr[r + 1] = static_cast<int>(A.val.size());
    }
    return A;
}

// The full-ILUT solve of solve_CN (src/CrankNicolsonModes_matrices.cpp:173-185,
// 203-214) on the host: ILUT (fill factor 50, drop tolerance 1e-9) of the
// operator the device applies, under Eigen's BiCGSTAB without sinks or the
// reference's BiCGSTAB with the Woodbury sink correction, tolerance 1e-10, at
// most 5000 iterations, from the guess psi/|rhs|. Reads the raw rhs and |rhs|
// from the device (after rhs_and_init), solves for b = rhs/|rhs| and leaves
// x = solution/|rhs| in S.x, where the tail's acceptance check, mask and
// diagnostics take it over.
void ilut_host_solve(gm_solver &S)
{
    HIPCK(hipStreamSynchronize(S.st));
    if (ilut_bytes(S) > kIlutHostBytes)
        throw std::runtime_error("ILUT preconditioner: the factor of " + std::to_string(S.nvec()) +
                                 " unknowns would exceed the host memory budget of the ILUT solve");
    gmd::Krylov kr;
    HIPCK(hipMemcpy(&kr, S.dK.p, sizeof(kr), hipMemcpyDeviceToHost));
    S.last_iterations = 0;
    if (kr.zero_rhs) return;  // the acceptance writes psi = 0
    const gmd::StepParams &P = *S.hP.p;
    const int K = S.K, N = S.Nr;
    const size_t n = S.nvec();
    const gmh::CsrZ A = cn_operator(extract_H(S), P.half_dt);
    gmh::Ilut M;
    std::string why;
    if (!M.compute(A, 1e-9, 50, &why)) throw std::runtime_error("Crank-Nicolson ILUT factorization failed");
    std::vector<double> hb(2 * n), hpsi(2 * n);
    to_host(S, S.b.p, hb.data());
    to_host(S, S.psi.p, hpsi.data());
    auto node_major = [&](size_t q) { return (q % N) * K + q / N; };
    std::vector<gmh::cplx> b(n), x(n);
    for (size_t q = 0; q < n; ++q) {
        b[node_major(q)] = gmh::cplx(hb[2 * q] / kr.rhs_nrm, hb[2 * q + 1] / kr.rhs_nrm);
        x[node_major(q)] = gmh::cplx(hpsi[2 * q], hpsi[2 * q + 1]) / kr.rhs_nrm;
    }
    // active sinks: A_full = A + (dt/4) sum_a g_a v_a v_a^H (matrices.cpp:108-116)
    const int ns = P.nsink;
    std::vector<std::vector<std::pair<size_t, gmh::cplx>>> sv(ns);
    std::vector<double> beta(ns);
    for (int a = 0; a < ns; ++a) {
        const Sink &sk = sink_host(S, P.sink_body[a]);
        beta[a] = 0.5 * P.half_dt * P.g[a];
        for (int k = 0; k < K; ++k)
            for (int d = 0; d < sk.n; ++d) {
                const double2 v = sk.v[static_cast<size_t>(k) * sk.n + d];
                if (v.x != 0.0 || v.y != 0.0)
                    sv[a].emplace_back(static_cast<size_t>(sk.lo + d) * K + k, gmh::cplx(v.x, v.y));
            }
    }
    auto vdot = [&](int a, const gmh::cplx *y) {
        gmh::cplx s = 0.0;
        for (const auto &e : sv[a]) s += std::conj(e.second) * y[e.first];
        return s;
    };
    const gmh::LinOp apply_A = [&](const gmh::cplx *in, gmh::cplx *out) {
        A.multiply(in, out);
        for (int a = 0; a < ns; ++a) {
            const gmh::cplx c = beta[a] * vdot(a, in);
            for (const auto &e : sv[a]) out[e.first] += c * e.second;
        }
    };
    // Woodbury (matrices.cpp:133-159): W_a = M^-1 v_a, G = diag(1/beta) + V^H W
    std::vector<std::vector<gmh::cplx>> W(ns, std::vector<gmh::cplx>(n));
    gmh::cplx Ginv[2][2] = {{0.0, 0.0}, {0.0, 0.0}};
    if (ns > 0) {
        for (int a = 0; a < ns; ++a) {
            std::vector<gmh::cplx> va(n, 0.0);
            for (const auto &e : sv[a]) va[e.first] = e.second;
            M.solve(va.data(), W[a].data());
        }
        gmh::cplx G[2][2] = {{0.0, 0.0}, {0.0, 0.0}};
        for (int a = 0; a < ns; ++a) {
            G[a][a] += 1.0 / beta[a];
            for (int c = 0; c < ns; ++c) G[a][c] += vdot(a, W[c].data());
        }
        if (ns == 1) {
            Ginv[0][0] = 1.0 / G[0][0];
        } else {
            const gmh::cplx det = G[0][0] * G[1][1] - G[0][1] * G[1][0];
            Ginv[0][0] = G[1][1] / det;
            Ginv[0][1] = -G[0][1] / det;
            Ginv[1][0] = -G[1][0] / det;
            Ginv[1][1] = G[0][0] / det;
        }
    }
    const gmh::LinOp apply_M = [&](const gmh::cplx *in, gmh::cplx *out) {
        M.solve(in, out);
        if (ns == 0) return;
        gmh::cplx proj[2] = {0.0, 0.0}, corr[2] = {0.0, 0.0};
        for (int a = 0; a < ns; ++a) proj[a] = vdot(a, out);
        for (int a = 0; a < ns; ++a)
            for (int c = 0; c < ns; ++c) corr[a] += Ginv[a][c] * proj[c];
        for (size_t q = 0; q < n; ++q) {
            gmh::cplx t = 0.0;
            for (int a = 0; a < ns; ++a) t += W[a][q] * corr[a];
            out[q] -= t;
        }
    };
    int its = 0;
    gmh::bicgstab(apply_A, apply_M, b.data(), x.data(), n, 1e-10, 5000, ns == 0, &its);
    S.last_iterations = its;
    for (size_t q = 0; q < n; ++q) {
        const gmh::cplx v = x[node_major(q)];
        hb[2 * q] = v.real();
        hb[2 * q + 1] = v.imag();
    }
    to_dev(S, hb.data(), S.x.p);
}

}  // namespace

extern "C" {

const char *gm_version(void) { return "gm-mi355x-0.1"; }

int gm_device_count(void)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

const char *gm_last_error(const gm_solver *S) { return S ? S->err.c_str() : g_create_err.c_str(); }

int gm_create(const gm_params *p, gm_solver **out)
{
    *out = nullptr;
    return guarded(nullptr, [&] {
        if (p->Nr < 3 || p->l_max < 0 || p->L <= 0.0 || p->dt <= 0.0 || p->save_every <= 0)
            throw std::invalid_argument("Require Nr >= 3, l_max >= 0, and positive L, dt, save_every");
        if (p->l_max + 1 > gmd::MAXLM)
            throw std::invalid_argument("l_max above the largest supported value (" + std::to_string(gmd::MAXLM - 1) + ")");
        ensure_device();
        auto S = std::make_unique<gm_solver>();
        S->Nr = p->Nr;
        S->save_every = p->save_every;
        S->l_max = p->l_max;
        S->epsilon = p->epsilon;
        S->mu = p->mu;
        S->q = p->q;
        S->R = p->R;
        S->a = p->a;
        S->L = p->L;
        S->dr = p->L / (p->Nr - 1);
        S->dt = p->dt;
        S->counter = p->counter != 0;
        S->M_c = p->M_c;
        S->ramp_periods = p->ramp_periods;
        S->initial_period = 2.0 * M_PI / std::sqrt((1 + p->q) / std::pow(p->R, 3));
        S->absorption = p->absorption != 0;
        S->epsilon_smear = p->epsilon_smear;
        S->Omega = 0.0;
        S->Omega_0 = std::sqrt((1.0 + p->q) / std::pow(p->R, 3));
        S->P = S->initial_period;
        S->R1 = p->q / (1 + p->q) * p->R;
        S->R2 = 1 / (1 + p->q) * p->R;
        S->M1 = 1.0;
        S->M2 = p->q;
        S->flags = gm_flags{1, 1, 50.0, 1, 2.0, GM_PRECISION_FP64, 1, 0};
        S->avg = std::make_unique<gmh::RollingPhaseAverage>(2.0);
        S->r.resize(S->Nr);
        for (int i = 0; i < S->Nr; ++i) S->r[i] = S->epsilon + i * S->dr;
        for (int l = 0; l <= S->l_max; ++l)
            for (int m = -l; m <= l; ++m)
                if (std::abs(l - m) % 2 == 0) {
                    S->ml.push_back(l);
                    S->mm.push_back(m);
                }
        S->K = static_cast<int>(S->ml.size());
        HIPCK(hipStreamCreateWithFlags(&S->st, hipStreamNonBlocking));
        S->dP.alloc(1);
        S->dK.alloc(1);
        S->dflags.alloc(2);
        S->dresid.alloc(1);
        S->hP.alloc(1);
        S->hK.alloc(1);
        S->hflags.alloc(2);
        S->hresid.alloc(1);
        S->hwr.alloc(4);
        std::memset(S->hP.p, 0, sizeof(gmd::StepParams));
        host_tables(*S);
        alloc_device(*S);
        // set-up of the allocation-free entry point: this GPU's static scratch
        HIPCK(gmk::sink_scratch_address(&S->static_scr));
        HIPCK(hipMemsetAsync(S->psi.p, 0, S->nvec() * sizeof(double2), S->st));
        HIPCK(hipStreamSynchronize(S->st));
        *out = S.release();
    });
}

void gm_destroy(gm_solver *S)
{
    if (!S) return;
    if (S->st) {
        (void)hipStreamSynchronize(S->st);
        drop_graphs(*S);
        (void)hipStreamDestroy(S->st);
    }
    for (auto &p : S->pending) S->ev_pool.insert(S->ev_pool.end(), {p.a, p.b});
    for (hipEvent_t e : S->ev_pool) (void)hipEventDestroy(e);
    delete S;
}

int gm_apply_config(const gm_solver *S, int32_t precision, int32_t *tile_nodes, double *ms)
{
    if (!S || (precision != GM_PRECISION_FP64 && precision != GM_PRECISION_MIXED) || !tile_nodes) return GM_E_INVALID;
    const int m = precision == GM_PRECISION_MIXED;
    *tile_nodes = S->apply_ng[m] == gmk::TAPPLY_NG ? 8 : 16 * S->apply_ng[m];
    if (ms)
        for (int c = 0; c < 4; ++c) ms[c] = S->apply_ms[m][c];
    return GM_OK;
}

int gm_set_coupling_tolerance(gm_solver *S, double tol)
{
    return guarded(S, [&] {
        if (!(tol >= 0.0) || tol >= 1e-6) throw std::invalid_argument("coupling tolerance must lie in [0, 1e-6)");
        S->coupling_tol = tol;
        S->E_valid = false;
        // the cut changes the work per tile: tune again (a fixed form stays)
        S->apply_ng[0] = S->apply_ng[1] =
            S->flags.apply_tile ? (S->flags.apply_tile == 8 ? gmk::TAPPLY_NG : S->flags.apply_tile / 16) : 0;
        if (S->initialized) {
            build_perturbation(*S, S->t_i);
            HIPCK(hipStreamSynchronize(S->st));
        }
    });
}

int gm_set_restart_thresholds(gm_solver *S, double eigen_eps2, double sink_factor)
{
    return guarded(S, [&] {
        if (!(eigen_eps2 > 0.0) || !(sink_factor > 0.0))
            throw std::invalid_argument("restart thresholds must be positive");
        S->rst_eps2 = eigen_eps2;
        S->rst_ref = sink_factor;
        drop_graphs(*S);
    });
}

int gm_restart_count(const gm_solver *S, int32_t *n)
{
    return guarded(const_cast<gm_solver *>(S), [&] {
        if (!n) throw std::invalid_argument("restart count pointer is NULL");
        *n = S->last_restarts;
    });
}

int gm_coupling_work(gm_solver *S, double *terms, double *pairs)
{
    return guarded(S, [&] {
        require_initialized(*S);
        std::vector<int> lc(S->ntiles);
        HIPCK(hipMemcpyAsync(lc.data(), S->lcut.p, sizeof(int) * S->ntiles, hipMemcpyDeviceToHost, S->st));
        HIPCK(hipStreamSynchronize(S->st));
        // per tile: Gaunt terms (|m*| <= l* <= lcut) and mode pairs with a kept term
        double t = 0.0, pr = 0.0;
        std::vector<double> tcount(2 * S->l_max + 1, 0.0), pcount(2 * S->l_max + 1, 0.0);
        for (int j = 0; j < S->K; ++j)
            for (int k = 0; k < S->K; ++k) {
                const int l = S->ml[j], l1 = S->ml[k], ms = std::abs(S->mm[k] - S->mm[j]);
                int lo = std::max(std::abs(l - l1), ms);
                if ((lo + l + l1) % 2) ++lo;
                if (lo > l + l1) continue;
                pcount[lo] += 1.0;
                for (int ls = lo; ls <= l + l1; ls += 2) tcount[ls] += 1.0;
            }
        for (int c = 1; c <= 2 * S->l_max; ++c) {
            tcount[c] += tcount[c - 1];
            pcount[c] += pcount[c - 1];
        }
        for (int tile = 0; tile < S->ntiles; ++tile) {
            const int nodes = std::min(gmd::TILE, S->Nr - tile * gmd::TILE);
            t += nodes * tcount[lc[tile]];
            pr += nodes * pcount[lc[tile]];
        }
        *terms = t;
        *pairs = pr;
    });
}

int gm_multipole_cut(gm_solver *S, int32_t *lcut, int32_t n)
{
    return guarded(S, [&] {
        require_initialized(*S);
        if (!lcut || n < S->ntiles) throw std::invalid_argument("lcut needs one entry per 64-node tile");
        std::vector<int> lc(S->ntiles);
        HIPCK(hipMemcpyAsync(lc.data(), S->lcut.p, sizeof(int) * S->ntiles, hipMemcpyDeviceToHost, S->st));
        HIPCK(hipStreamSynchronize(S->st));
        for (int t = 0; t < S->ntiles; ++t) lcut[t] = lc[t];
    });
}

int gm_profile_enable(gm_solver *S, int32_t on)
{
    return guarded(S, [&] {
        HIPCK(hipStreamSynchronize(S->st));
        prof_collect(*S);
        S->prof_on = on != 0;
        for (int c = 0; c < GM_PROF_NCAT; ++c) {
            S->prof_ms[c] = 0.0;
            S->prof_n[c] = 0;
        }
    });
}

int gm_profile_read(const gm_solver *S, double *ms, int64_t *launches, int32_t ncat)
{
    for (int c = 0; c < ncat && c < GM_PROF_NCAT; ++c) {
        ms[c] = S->prof_ms[c];
        launches[c] = S->prof_n[c];
    }
    return GM_OK;
}

int gm_n_modes(const gm_solver *S) { return S->K; }

int gm_modes(const gm_solver *S, int32_t *lm)
{
    for (int j = 0; j < S->K; ++j) {
        lm[2 * j] = S->ml[j];
        lm[2 * j + 1] = S->mm[j];
    }
    return GM_OK;
}

int gm_grid(const gm_solver *S, double *r)
{
    std::memcpy(r, S->r.data(), sizeof(double) * S->Nr);
    return GM_OK;
}

int gm_set_flags(gm_solver *S, const gm_flags *f)
{
    return guarded(S, [&] {
        if (!(f->power_average_orbits > 0.0)) throw std::invalid_argument("RollingPhaseAverage: n_orbits must be positive");
        if (f->precision != GM_PRECISION_FP64 && f->precision != GM_PRECISION_MIXED)
            throw std::invalid_argument("precision must be GM_PRECISION_FP64 or GM_PRECISION_MIXED");
        if (f->apply_tile != 0 && f->apply_tile != 8 && f->apply_tile != 16 && f->apply_tile != 32 &&
            f->apply_tile != 64)
            throw std::invalid_argument("apply_tile must be 0 (autotune), 8 (polar form), 16, 32 or 64");
        if (f->apply_tile == 8 && gmk::tapply_class(S->l_max, 0) == 0)
            throw std::invalid_argument("apply_tile 8 (the polar-quadrature form) needs l_max <= 20");
        if (f->power_average_orbits != S->flags.power_average_orbits)
            S->avg = std::make_unique<gmh::RollingPhaseAverage>(f->power_average_orbits);
        if (f->ramp_secondary_potential != S->flags.ramp_secondary_potential) S->E_valid = false;
        if (f->apply_tile != S->flags.apply_tile) S->apply_ng[0] = S->apply_ng[1] = 0;
        // a fixed form applies at once (to the solver's own applies as well)
        if (f->apply_tile) S->apply_ng[0] = S->apply_ng[1] = f->apply_tile == 8 ? gmk::TAPPLY_NG : f->apply_tile / 16;
        S->flags = *f;
        S->retry_mixed = false;
    });
}

int gm_get_flags(const gm_solver *S, gm_flags *f)
{
    *f = S->flags;
    return GM_OK;
}

int gm_configure_isolated_black_hole(gm_solver *S, double x)
{
    return guarded(S, [&] {
        if (!std::isfinite(x) || std::abs(x) >= S->L)
            throw std::invalid_argument("The isolated BH must lie inside the radial domain");
        if (S->initialized) throw std::logic_error("Configure the isolated BH before initialize()");
        S->isolated = true;
        S->flags.evolve_orbit = 0;
        S->M1 = 1.0;
        S->M2 = S->q = 0.0;
        S->R = S->R1 = std::abs(x);
        S->R2 = 0.0;
        S->theta = x < 0.0 ? M_PI : 0.0;
        S->phase = 0.0;
        S->Omega = S->Omega_0 = 0.0;
        S->ramp_periods = 0.0;
    });
}

int gm_set_initial_state_full(gm_solver *S, const double *psi, int64_t n)
{
    return guarded(S, [&] {
        if (n != static_cast<int64_t>(S->nvec()))
            throw std::invalid_argument("Initial state size does not match the radial grid and modes");
        const double scale = std::sqrt(S->M_c / S->mu);
        std::vector<double> tmp(psi, psi + 2 * S->nvec());
        for (double &v : tmp) v *= scale;
        to_dev(*S, tmp.data(), S->psi.p);
    });
}

int gm_set_state(gm_solver *S, const double *psi)
{
    return guarded(S, [&] { to_dev(*S, psi, S->psi.p); });
}

int gm_get_state(const gm_solver *S, double *psi)
{
    gm_solver *M = const_cast<gm_solver *>(S);
    return guarded(M, [&] { to_host(*M, M->psi.p, psi); });
}

int gm_initialize(gm_solver *S)
{
    return guarded(S, [&] {
        if (S->absorption) {
            S->prof[0] = Profile{};
            S->prof[1] = Profile{};
            S->mode00 = -1;
            for (int j = 0; j < S->K; ++j)
                if (S->ml[j] == 0 && S->mm[j] == 0) S->mode00 = j;
            absorber_tables(*S);
            for (Sink &sk : S->sink) {
                sk.v.clear();
                sk.n = 0;
                sk.strength = 0.0;
                sk.norm2 = 0.0;
                sk.v_host = false;
            }
        }
        S->t_i = 0.0;
        build_perturbation(*S, 0.0);
        S->initialized = true;
        HIPCK(hipStreamSynchronize(S->st));
    });
}

int gm_build_damping_mask(gm_solver *S, double f, double W, int32_t n)
{
    return guarded(S, [&] {
        S->damping.assign(S->Nr, 1.0);
        S->mask = true;
        const int idx = static_cast<int>(std::floor(S->L * f / S->dr));
        const double r_mask = f * S->L, L_mask = (1 - f) * S->L;
        for (int i = 0; i < S->Nr; ++i)
            if (i >= idx) S->damping[i] = std::exp(-W * std::pow((S->r[i] - r_mask) / L_mask, n));
        S->dmask.alloc(S->Nr);
        HIPCK(hipMemcpyAsync(S->dmask.p, S->damping.data(), sizeof(double) * S->Nr, hipMemcpyHostToDevice, S->st));
        HIPCK(hipStreamSynchronize(S->st));
    });
}

int gm_build_cap(gm_solver *S, double f, double W, int32_t n)
{
    return guarded(S, [&] {
        S->capv.assign(S->Nr, make_double2(0.0, 0.0));
        S->cap = true;
        const int idx = static_cast<int>(std::floor(S->L * f / S->dr));
        const double r_cap = f * S->L, L_cap = (1 - f) * S->L;
        for (int i = 0; i < S->Nr; ++i)
            if (i >= idx) S->capv[i] = make_double2(0.0, -W * std::pow((S->r[i] - r_cap) / (S->r[i] - L_cap), n));
        S->dcap.alloc(S->Nr);
        HIPCK(hipMemcpyAsync(S->dcap.p, S->capv.data(), sizeof(double2) * S->Nr, hipMemcpyHostToDevice, S->st));
        S->cap_version++;
        HIPCK(hipStreamSynchronize(S->st));
    });
}

int gm_begin_run(gm_solver *S, double r_ex, int32_t n_steps)
{
    return guarded(S, [&] {
        require_initialized(*S);
        S->r_ex = r_ex;
        S->idx_ex = static_cast<int>(std::floor(r_ex / S->dr));
        if (n_steps < 1 || S->idx_ex < 1 || S->idx_ex >= S->Nr - 1)
            throw std::invalid_argument("Require at least one time sample and an interior extraction radius");
        S->t_i = 0.0;
        build_perturbation(*S, S->t_i);
        HIPCK(hipStreamSynchronize(S->st));
    });
}

}  // extern "C"

namespace {

// Autotuner of the operator-apply tile: 16, 32 or 64 radial nodes per
// workgroup tile (register pressure and occupancy against reuse of the Gaunt
// fragments and load balance). Each candidate is timed with HIP events on the
// solver's stream over a few applies of A to psi (into the Krylov vector t,
// which every iteration overwrites); the fastest is kept for this shape in a
// process-wide table, so later solvers of the same shape skip the timing.
struct TuneKey {
    int N, lmax, mixed;
    double tol;
    bool operator<(const TuneKey &o) const
    {
        return std::tie(N, lmax, mixed, tol) < std::tie(o.N, o.lmax, o.mixed, o.tol);
    }
};
std::mutex tune_mu;
std::map<TuneKey, std::pair<int, std::array<double, 4>>> tune_table;

void tune_apply(gm_solver &S, int mixed)
{
    const bool polar_ok = gmk::tapply_class(S.l_max, mixed) != 0;
    if (S.flags.apply_tile) {
        S.apply_ng[mixed] = S.flags.apply_tile == 8 ? gmk::TAPPLY_NG : S.flags.apply_tile / 16;
        return;
    }
    if (S.apply_ng[mixed]) return;
    const TuneKey key{S.Nr, S.l_max, mixed, S.coupling_tol};
    {
        std::lock_guard<std::mutex> lk(tune_mu);
        auto it = tune_table.find(key);
        if (it != tune_table.end()) {
            S.apply_ng[mixed] = it->second.first;
            for (int c = 0; c < 4; ++c) S.apply_ms[mixed][c] = it->second.second[c];
            return;
        }
    }
    hipEvent_t e0, e1;
    HIPCK(hipEventCreate(&e0));
    HIPCK(hipEventCreate(&e1));
    const int cand[4] = {1, 2, 4, gmk::TAPPLY_NG};
    const int ncand = polar_ok ? 4 : 3;
    std::array<double, 4> ms{};
    int best = 2;
    double best_ms = 1e300;
    try {
        // the residual form: full coupling plus a fused reduction, like the
        // Krylov applies (its result in the Krylov state is overwritten by the
        // step's own residual check); two interleaved rounds, best of each
        for (int c = 0; c < ncand; ++c) ms[c] = 1e300;
        for (int round = 0; round < 2; ++round)
            for (int c = 0; c < ncand; ++c) {
                gmk::ApplyArgs a = apply_args(S, S.psi.p, nullptr);
                a.R0 = S.psi.p;
                a.mixed = mixed;
                a.ng = cand[c];
                HIPCK(gmk::launch_apply(a, gmk::AM_RES, S.st));  // warm
                HIPCK(hipEventRecord(e0, S.st));
                for (int r = 0; r < 5; ++r) HIPCK(gmk::launch_apply(a, gmk::AM_RES, S.st));
                HIPCK(hipEventRecord(e1, S.st));
                HIPCK(hipEventSynchronize(e1));
                float t = 0.0f;
                HIPCK(hipEventElapsedTime(&t, e0, e1));
                ms[c] = std::min(ms[c], t / 5.0);
            }
        for (int c = 0; c < ncand; ++c)
            if (ms[c] < best_ms) {
                best_ms = ms[c];
                best = cand[c];
            }
        if (!polar_ok) ms[3] = 0.0;
    } catch (...) {
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        throw;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    S.apply_ng[mixed] = best;
    for (int c = 0; c < 4; ++c) S.apply_ms[mixed][c] = ms[c];
    std::lock_guard<std::mutex> lk(tune_mu);
    tune_table[key] = {best, ms};
}

// step (src/CrankNicolsonModes.cpp:145-234), phase 1: advance t, theta and the
// phase, rebuild the perturbation and enqueue the solve up to its first
// convergence check.
// A mixed step's FP64 retry that ended in an error leaves the precision to restore.
void restore_precision(gm_solver &S)
{
    if (S.retry_mixed) {
        S.flags.precision = GM_PRECISION_MIXED;
        S.retry_mixed = false;
    }
}

void step_begin(gm_solver &S)
{
    require_initialized(S);
    restore_precision(S);
    S.step_start = S.t_i;
    S.t_i += S.dt;
    S.theta += (S.counter ? -1.0 : 1.0) * S.Omega * S.dt;
    S.phase += S.Omega * S.dt;
    // with the block preconditioner the profiles of a moved orbit are
    // rebuilt inside the captured start of the solve
    build_perturbation(S, S.t_i, S.flags.use_block_preconditioner != 0);
    S.step_ilut = false;
    S.step_fp64 = false;
    if (!S.flags.use_block_preconditioner) {
        // the "ilut" preconditioner key: right-hand side on the device, solve
        // on the host, then the tail as after a converged BiCGSTAB
        S.step_wood = S.hP.p->nsink > 0;
        S.step_enqueued = 0;
        S.step_ilut = true;
        rhs_and_init(S);
        ilut_host_solve(S);
        gmd::Krylov &K = *S.hK.p;
        K.done = 1;
        K.restart = 0;
        K.iter = S.last_iterations;
        return;
    }
    tune_apply(S, 0);
    if (S.flags.precision == GM_PRECISION_MIXED) tune_apply(S, 1);
    solve_begin(S);
}

// Phase 3a, once the stream has drained after the tail's readback: the
// residual check. A mixed-precision step that failed it is enqueued again
// with the FP64 operator (from the step's start state: psi is untouched until
// the step is accepted) and false is returned: the caller polls solve_advance
// again and calls step_accept once more, so ensemble and server loops keep
// advancing their other solvers meanwhile. A second failure, or a failure in
// FP64, goes to the host ILUT (matrices.cpp:173,203). True once the step's
// result is accepted; repeated calls then return true at once.
bool step_accept(gm_solver &S)
{
    for (;;) {
        const bool bad_factor = S.hflags.p[0] != 0;
        if (bad_factor) {
            S.fac_valid = false;
            S.hflags.p[0] = 0;
            int zero = 0;
            HIPCK(hipMemcpy(S.dflags.p, &zero, sizeof(int), hipMemcpyHostToDevice));
        }
        if (S.hflags.p[1] && !bad_factor) {
            restore_precision(S);
            return true;
        }
        if (S.step_ilut) {
            restore_precision(S);
            throw std::runtime_error("Crank-Nicolson solve failed the relative residual check");
        }
        if (S.flags.precision == GM_PRECISION_MIXED && !S.step_fp64) {
            // The FP32 coupling kept the solve above the FP64 residual
            // threshold (its error grows with dt |V|): the same radial solve
            // again with the FP64 operator, before the host ILUT.
            S.step_fp64 = true;
            S.retry_mixed = true;
            S.flags.precision = GM_PRECISION_FP64;
            try {
                tune_apply(S, 0);
                solve_begin(S);
            } catch (...) {
                S.flags.precision = GM_PRECISION_MIXED;
                S.retry_mixed = false;
                throw;
            }
            return false;
        }
        // "Radial preconditioner did not converge; retrying with full ILUT."
        // (matrices.cpp:173,203): psi is still the step's start state
        S.step_ilut = true;
        ilut_host_solve(S);
        run_captured(S, GK_TAIL, graph_key(S, 0), [&] { solve_tail_work(S); });
        HIPCK(hipStreamSynchronize(S.st));
        prof_collect(S);
    }
}

// Phase 3, once the stream has drained after the diagnostics readback: the
// residual check (a failed mixed step is solved again here, synchronously),
// absorbed mass, orbit update and diagnostics.
void step_finish(gm_solver &S, gm_diag *d)
{
    const double step_start = S.step_start;
    const gmd::StepParams &P = *S.hP.p;
    {
        while (!step_accept(S)) {
            for (;;) {
                HIPCK(hipStreamSynchronize(S.st));
                prof_collect(S);
                if (solve_advance(S)) break;
            }
            HIPCK(hipStreamSynchronize(S.st));
            prof_collect(S);
        }
        S.last_residual = S.hresid.p[0];
        const gmd::Krylov &K = *S.hK.p;
        // absorbed mass: mu dr dt sum_a g_a |v_a^H psi_mid|^2 (CN balance)
        double absorbed = 0.0, Mdot[2] = {0.0, 0.0};
        for (int s = 0; s < P.nsink; ++s) {
            const double mr = 0.5 * (K.qold[s].x + K.qnew[s].x), mi = 0.5 * (K.qold[s].y + K.qnew[s].y);
            absorbed += S.mu * S.dr * S.dt * P.g[s] * (mr * mr + mi * mi);
            // update_time_dependent_increment uses the post-mask midpoint (evolution.cpp:20-30)
            const double pr = 0.5 * (K.qpost[s].x + K.qold[s].x), pi = 0.5 * (K.qpost[s].y + K.qold[s].y);
            Mdot[P.sink_body[s]] = S.mu * S.dr * P.g[s] * (pr * pr + pi * pi);
        }
        const double Omega_diag = S.Omega_V;
        if (!S.isolated && S.flags.evolve_orbit && S.t_i > S.ramp_end_time()) {
            // update_time_dependent_increment (evolution.cpp:17-72)
            const double orbit_dt = S.t_i - std::max(step_start, S.ramp_end_time());
            const double Mdot1 = S.absorption ? Mdot[0] : 0.0, Mdot2 = S.absorption ? Mdot[1] : 0.0;
            double ramp = 1.0;
            if (S.t_i < S.ramp_end_time()) ramp = 0.0;
            else if (S.t_i < S.ramp_end_time() + S.initial_period) {
                const double xx = (S.t_i - S.ramp_end_time()) / S.initial_period;
                ramp = 3.0 * xx * xx - 2.0 * xx * xx * xx;
            }
            const double dR_gw = -(64.0 / 5.0) * S.M1 * S.M2 * (S.M1 + S.M2) / std::pow(S.R, 3);
            const double dR_cloud = -2.0 * S.R * S.R * S.p_bin / (S.M1 * S.M2);
            const double dRdt = (dR_gw + dR_cloud) * ramp;
            S.R += dRdt * orbit_dt;
            S.M1 += Mdot1 * orbit_dt;
            S.M2 += Mdot2 * orbit_dt;
            S.q = S.M2 / S.M1;
            const double Mt = S.M1 + S.M2;
            S.R1 = S.M2 / Mt * S.R;
            S.R2 = S.M1 / Mt * S.R;
            S.Omega = std::sqrt(Mt / std::pow(S.R, 3));
            S.P = 2.0 * M_PI / S.Omega;
            if (S.ramp_factor == 1.0) S.dt = S.P / S.flags.steps_per_period;
        }
        const double *h = S.hdiag.p;
        const int Kn = S.K;
        // global sums: per-mode shares added in mode order
        double gsum[4] = {0.0, 0.0, 0.0, 0.0};
        for (int q = 0; q < 4; ++q)
            for (int j = 0; j < Kn; ++j) gsum[q] += h[(4 + q) * Kn + j];
        const double sign = S.counter ? -1.0 : 1.0;
        const double p_bin_last = 2.0 * sign * Omega_diag * gsum[1] * S.dr;
        S.p_bin = S.isolated ? 0.0 : S.avg->update(S.t_i, S.phase, p_bin_last);
        if (d) {
            d->energy = gsum[0] * S.dr;
            d->p_bin = S.p_bin;
            d->p_bin_inst = p_bin_last;
            d->dV_bin_dt = gsum[2] * S.dr;
            d->cloud_mass = S.mu * S.dr * gsum[3];
            d->absorbed = absorbed;
            if (d->flux) std::memcpy(d->flux, h, sizeof(double) * Kn);
            if (d->energy_flux) std::memcpy(d->energy_flux, h + Kn, sizeof(double) * Kn);
            if (d->norms) std::memcpy(d->norms, h + 2 * Kn, sizeof(double) * Kn);
            if (d->energies) std::memcpy(d->energies, h + 3 * Kn, sizeof(double) * Kn);
        }
    }
}

}  // namespace

extern "C" {

int gm_step(gm_solver *S, int32_t step_id, gm_diag *d)
{
    (void)step_id;
    return guarded(S, [&] {
        step_begin(*S);
        for (;;) {
            HIPCK(hipStreamSynchronize(S->st));
            prof_collect(*S);
            if (solve_advance(*S)) break;
        }
        HIPCK(hipStreamSynchronize(S->st));
        prof_collect(*S);
        step_finish(*S, d);
    });
}

int gm_ensemble_step(gm_solver *const *members, int32_t n, int32_t step_id, gm_diag *diags)
{
    (void)step_id;
    if (n < 0 || (n > 0 && !members)) return GM_E_INVALID;
    for (int i = 0; i < n; ++i)
        for (int k = 0; k < i; ++k)
            if (!members[i] || members[i] == members[k]) return GM_E_INVALID;
    // every member runs on its own stream: enqueue all, then walk the members
    // through their convergence checks; a member that fails keeps its error
    // (gm_last_error) and drops out of the step
    std::vector<int> st(n, GM_OK), phase(n, 0);
    for (int i = 0; i < n; ++i) {
        st[i] = guarded(members[i], [&] { step_begin(*members[i]); });
        phase[i] = st[i] == GM_OK ? 1 : -1;
    }
    // members are advanced as their streams go idle (non-blocking queries), so
    // a member whose batch finished early is refilled at once instead of
    // waiting behind the members before it; a mixed member that fails the
    // FP64 residual check goes back to phase 1 with its FP64 retry enqueued
    for (bool busy = true; busy;) {
        busy = false;
        bool progressed = false;
        for (int i = 0; i < n; ++i) {
            if (phase[i] != 1 && phase[i] != 2) continue;
            gm_solver *S = members[i];
            const hipError_t q = hipStreamQuery(S->st);
            if (q == hipErrorNotReady) {
                busy = true;
                continue;
            }
            st[i] = guarded(S, [&] {
                HIPCK(q);
                prof_collect(*S);
                if (phase[i] == 1) {
                    if (solve_advance(*S)) phase[i] = 2;
                } else {
                    phase[i] = step_accept(*S) ? 3 : 1;
                }
            });
            progressed = true;
            if (st[i] != GM_OK) phase[i] = -1;
            busy = busy || phase[i] == 1 || phase[i] == 2;
        }
        if (busy && !progressed) std::this_thread::yield();
    }
    int first = GM_OK;
    for (int i = 0; i < n; ++i) {
        if (phase[i] == 3) {
            gm_solver *S = members[i];
            st[i] = guarded(S, [&] {
                HIPCK(hipStreamSynchronize(S->st));
                prof_collect(*S);
                step_finish(*S, diags ? diags + i : nullptr);
            });
        }
        if (first == GM_OK && st[i] != GM_OK) first = st[i];
    }
    return first;
}

int gm_get_scalars(const gm_solver *S, gm_scalars *o)
{
    o->t = S->t_i;
    o->theta = S->theta;
    o->phase = S->phase;
    o->R = S->R;
    o->M1 = S->M1;
    o->M2 = S->M2;
    o->q = S->q;
    o->R1 = S->R1;
    o->R2 = S->R2;
    o->Omega = S->Omega;
    o->P = S->P;
    o->dt = S->dt;
    o->p_bin = S->p_bin;
    o->Omega_0 = S->Omega_0;
    o->last_solver_residual = S->last_residual;
    o->last_solver_iterations = S->last_iterations;
    o->sink_active = S->sink_active;
    return GM_OK;
}

// checkpoint.bin (src/CrankNicolsonModes_restart.cpp:139-233): the magic
// "GMCKPT01", int32 {step_id, Nr, n_modes}, the 13 doubles t, theta, phase, R,
// M1, M2, q, R1, R2, Omega, P, dt, p_bin, uint64 psi size, psi as complex128,
// then the rolling phase average block; written to <path>.tmp and renamed, so
// a crash mid-write keeps the previous checkpoint.
constexpr char kCheckpointMagic[8] = {'G', 'M', 'C', 'K', 'P', 'T', '0', '1'};

int gm_save_checkpoint(gm_solver *S, const char *path, int32_t step_id)
{
    return guarded(S, [&] {
        require_initialized(*S);
        if (!path) throw std::invalid_argument("checkpoint path is NULL");
        std::vector<double> psi(2 * S->nvec());
        to_host(*S, S->psi.p, psi.data());
        const std::string final_path(path), tmp = final_path + ".tmp";
        {
            std::ofstream out(tmp, std::ios::binary | std::ios::trunc);
            if (!out) throw std::runtime_error("Could not write checkpoint " + tmp);
            out.write(kCheckpointMagic, sizeof(kCheckpointMagic));
            const int32_t header[3] = {step_id, S->Nr, S->K};
            out.write(reinterpret_cast<const char *>(header), sizeof(header));
            const double v[13] = {S->t_i, S->theta, S->phase, S->R, S->M1, S->M2, S->q,
                                  S->R1, S->R2, S->Omega, S->P, S->dt, S->p_bin};
            out.write(reinterpret_cast<const char *>(v), sizeof(v));
            const uint64_t size = S->nvec();
            out.write(reinterpret_cast<const char *>(&size), sizeof(size));
            out.write(reinterpret_cast<const char *>(psi.data()), static_cast<std::streamsize>(psi.size() * 8));
            S->avg->save(out);
            if (!out) throw std::runtime_error("Could not write checkpoint " + tmp);
        }
        if (std::rename(tmp.c_str(), final_path.c_str()) != 0)
            throw std::runtime_error("Could not replace checkpoint " + final_path);
    });
}

int gm_load_checkpoint(gm_solver *S, const char *path, int32_t step_id, int32_t *loaded)
{
    return guarded(S, [&] {
        require_initialized(*S);
        if (!path || !loaded) throw std::invalid_argument("checkpoint path or result is NULL");
        *loaded = 0;
        std::ifstream in(path, std::ios::binary);
        if (!in) return;
        char magic[8];
        int32_t header[3] = {-1, -1, -1};
        in.read(magic, sizeof(magic));
        in.read(reinterpret_cast<char *>(header), sizeof(header));
        if (!in || std::memcmp(magic, kCheckpointMagic, 8) != 0 || header[0] != step_id || header[1] != S->Nr ||
            header[2] != S->K)
            return;
        double v[13];
        in.read(reinterpret_cast<char *>(v), sizeof(v));
        uint64_t size = 0;
        in.read(reinterpret_cast<char *>(&size), sizeof(size));
        if (!in || size != S->nvec()) return;
        std::vector<double> psi(2 * size);
        in.read(reinterpret_cast<char *>(psi.data()), static_cast<std::streamsize>(psi.size() * 8));
        if (!in) return;
        auto avg = std::make_unique<gmh::RollingPhaseAverage>(S->flags.power_average_orbits);
        try {
            avg->load(in);
        } catch (const std::exception &) {
            return;
        }
        S->t_i = v[0];
        S->theta = v[1];
        S->phase = v[2];
        S->R = v[3];
        S->M1 = v[4];
        S->M2 = v[5];
        S->q = v[6];
        S->R1 = v[7];
        S->R2 = v[8];
        S->Omega = v[9];
        S->P = v[10];
        S->dt = v[11];
        S->p_bin = v[12];
        to_dev(*S, psi.data(), S->psi.p);
        S->avg = std::move(avg);
        *loaded = 1;
    });
}

int gm_ilut_bicgstab(int32_t n, const int32_t *ptr, const int32_t *col, const double *val, const double *b,
                     double *x, double droptol, int32_t fill_factor, double tol, int32_t max_iterations,
                     int32_t eigen_mode, int32_t *iterations)
{
    if (iterations) *iterations = 0;
    if (n < 1 || !ptr || !col || !val || !b || !x || !iterations || fill_factor < 1 || !(droptol >= 0.0) ||
        !(tol > 0.0) || max_iterations < 0) {
        g_create_err = "gm_ilut_bicgstab: invalid argument";
        return GM_E_INVALID;
    }
    try {
        gmh::CsrZ A;
        A.n = n;
        A.ptr.assign(ptr, ptr + n + 1);
        if (A.ptr[0] != 0) throw std::invalid_argument("gm_ilut_bicgstab: ptr[0] must be 0");
        for (int i = 0; i < n; ++i)
            if (A.ptr[i + 1] < A.ptr[i]) throw std::invalid_argument("gm_ilut_bicgstab: ptr must not decrease");
        const size_t nnz = static_cast<size_t>(A.ptr[n]);
        A.col.assign(col, col + nnz);
        for (int i = 0; i < n; ++i)
            for (int q = A.ptr[i]; q < A.ptr[i + 1]; ++q)
                if (A.col[q] < 0 || A.col[q] >= n || (q > A.ptr[i] && A.col[q] <= A.col[q - 1]))
                    throw std::invalid_argument("gm_ilut_bicgstab: columns must be in range and increasing per row");
        A.val.resize(nnz);
        for (size_t q = 0; q < nnz; ++q) A.val[q] = gmh::cplx(val[2 * q], val[2 * q + 1]);
        gmh::Ilut M;
        std::string why;
        if (!M.compute(A, droptol, fill_factor, &why)) {
            g_create_err = "Crank-Nicolson ILUT factorization failed: " + why;
            return GM_E_RUNTIME;
        }
        std::vector<gmh::cplx> bb(static_cast<size_t>(n)), xx(static_cast<size_t>(n));
        for (int i = 0; i < n; ++i) {
            bb[i] = gmh::cplx(b[2 * i], b[2 * i + 1]);
            xx[i] = gmh::cplx(x[2 * i], x[2 * i + 1]);
        }
        int its = 0;
        const bool ok = gmh::bicgstab([&](const gmh::cplx *in, gmh::cplx *out) { A.multiply(in, out); },
                                      [&](const gmh::cplx *in, gmh::cplx *out) { M.solve(in, out); }, bb.data(),
                                      xx.data(), static_cast<size_t>(n), tol, max_iterations, eigen_mode != 0, &its);
        *iterations = its;
        for (int i = 0; i < n; ++i) {
            x[2 * i] = xx[i].real();
            x[2 * i + 1] = xx[i].imag();
        }
        if (!ok) {
            g_create_err = "gm_ilut_bicgstab: BiCGSTAB stopped above the tolerance";
            return GM_E_RUNTIME;
        }
        return GM_OK;
    } catch (const std::invalid_argument &e) {
        g_create_err = e.what();
        return GM_E_INVALID;
    } catch (const std::exception &e) {
        g_create_err = e.what();
        return GM_E_RUNTIME;
    }
}

// reconstruct_state_from_diagnostics (src/CrankNicolsonModes_restart.cpp:238-285).
int gm_restore_from_diagnostics(gm_solver *S, int32_t n_rows, const double *rows)
{
    return guarded(S, [&] {
        require_initialized(*S);
        if (n_rows < 0 || (n_rows > 0 && !rows)) throw std::invalid_argument("diagnostic rows are missing");
        // before step 1 Omega is what the perturbation at t = 0 set
        double Omega_previous = S->isolated ? 0.0 : (S->ramp_end_time() > 0.0 ? 0.0 : S->Omega_0);
        double t_previous = 0.0;
        S->phase = 0.0;
        S->avg->reset();
        for (int k = 0; k < n_rows; ++k) {
            const double *row = rows + 8 * static_cast<size_t>(k);
            S->phase += Omega_previous * (row[0] - t_previous);
            if (!S->isolated) S->avg->update(row[0], S->phase, row[7]);
            Omega_previous = row[3];
            t_previous = row[0];
        }
        if (n_rows == 0) return;
        const double *last = rows + 8 * static_cast<size_t>(n_rows - 1);
        S->t_i = last[0];
        S->R = last[1];
        S->theta = last[2];
        S->M1 = last[4];
        S->q = last[5];
        S->M2 = S->q * S->M1;
        S->R1 = S->M2 / (S->M1 + S->M2) * S->R;
        S->R2 = S->M1 / (S->M1 + S->M2) * S->R;
        S->Omega = last[3];
        S->p_bin = last[6];
        // past the ramp an evolving orbit sets P and dt = P / steps_per_period
        if (!S->isolated && S->flags.evolve_orbit && S->t_i > S->ramp_end_time()) {
            S->P = 2.0 * M_PI / S->Omega;
            S->dt = S->P / S->flags.steps_per_period;
        }
    });
}

int gm_set_scalars(gm_solver *S, const gm_scalars *o)
{
    S->t_i = o->t;
    S->theta = o->theta;
    S->phase = o->phase;
    S->R = o->R;
    S->M1 = o->M1;
    S->M2 = o->M2;
    S->q = o->q;
    S->R1 = o->R1;
    S->R2 = o->R2;
    S->Omega = o->Omega;
    S->P = o->P;
    S->dt = o->dt;
    S->p_bin = o->p_bin;
    return GM_OK;
}

int gm_apply_H(gm_solver *S, const double *in, double *out)
{
    return guarded(S, [&] {
        require_initialized(*S);
        apply_H_host(*S, in, out);
    });
}

int gm_apply_sink(gm_solver *S, const double *in, double *out)
{
    return guarded(S, [&] {
        to_dev(*S, in, S->tmp.p);
        sink_apply_dev(*S, S->tmp.p, S->t.p, 1.0, false, S->sinkscr.p, S->st);
        to_host(*S, S->t.p, out);
    });
}

int gm_apply_sink_device(gm_solver *S, const double *d_in, double *d_out, void *stream)
{
    return guarded(S, [&] {
        if (!d_in || !d_out) throw std::invalid_argument("apply_sink_device: NULL device buffer");
        if (S->K > gmk::SINK_MAX_MODES)
            throw std::invalid_argument("apply_sink_device: " + std::to_string(S->K) +
                                        " modes exceed the static scratch's " +
                                        std::to_string(gmk::SINK_MAX_MODES));
        if (!S->static_scr) throw DeviceError("apply_sink_device: static scratch not set up");
        hipStream_t st = stream ? static_cast<hipStream_t>(stream) : S->st;
        sink_apply_dev(*S, reinterpret_cast<const double2 *>(d_in), reinterpret_cast<double2 *>(d_out), 1.0, false,
                       S->static_scr, st);
        if (!stream) HIPCK(hipStreamSynchronize(st));
    });
}

int gm_schrodinger_rhs(gm_solver *S, const double *in, double *out)
{
    return guarded(S, [&] {
        require_initialized(*S);
        // -i H psi - Gamma psi / 2 (absorption.cpp:111-117)
        to_dev(*S, in, S->tmp.p);
        apply_H_dev(*S, S->tmp.p, S->t.p);
        HIPCK(gmk::launch_cscale(S->t.p, S->nvec(), make_double2(0.0, -1.0), S->st));
        sink_apply_dev(*S, S->tmp.p, S->t.p, -0.5, true, S->sinkscr.p, S->st);
        to_host(*S, S->t.p, out);
    });
}

int gm_operator_csr(gm_solver *S, int32_t which, int64_t *nnz, int64_t *row_ptr, int32_t *col, double *val)
{
    return guarded(S, [&] {
        if (!nnz) throw std::invalid_argument("nnz must not be null");
        if (!row_ptr) {
            S->op_cache = materialise(*S, which);
            S->op_which = which;
            *nnz = static_cast<int64_t>(S->op_cache.col.size());
            return;
        }
        if (S->op_which != which) {
            S->op_cache = materialise(*S, which);
            S->op_which = which;
        }
        const OpCsr &o = S->op_cache;
        const int64_t m = static_cast<int64_t>(o.col.size());
        if (*nnz < m || (m > 0 && (!col || !val))) throw std::invalid_argument("col and val must hold nnz entries");
        std::memcpy(row_ptr, o.ptr.data(), o.ptr.size() * sizeof(int64_t));
        if (m > 0) {
            std::memcpy(col, o.col.data(), m * sizeof(int32_t));
            std::memcpy(val, o.val.data(), m * sizeof(gmh::cplx));
        }
        *nnz = m;
        S->op_cache = OpCsr{};
        S->op_which = -1;
    });
}

double gm_sink_strength(const gm_solver *S, int32_t body)
{
    const Sink &sk = S->sink[body == 1 ? 0 : 1];
    if (!S->sink_active || sk.strength == 0.0) return 0.0;
    return sk.strength * sink_host(*S, body == 1 ? 0 : 1).norm2;
}

}  // extern "C"

// serving loop (gm_server_*): same translation unit, it drives the step phases
#include "gm_server.inl"
// batched ragged ensemble (gm_batch_*): same translation unit, it reuses the
// step's argument builders
#include "gm_batch.inl"
