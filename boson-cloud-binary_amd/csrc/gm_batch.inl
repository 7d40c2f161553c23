// gm_batch.inl - the ensemble as batched ragged launches (part of the
// gm_solver.cpp translation unit: it reuses the step's argument builders).
//
// A batch steps many independent solvers together. Every launch of the CN
// step (sink projections, the rhs apply, the BiCGSTAB start, the
// preconditioner sweeps, the Woodbury corrections, the Krylov applies, the
// updates, the acceptance apply, the mask and the diagnostics) is one grid
// over all members: member m owns a contiguous range of blocks sized as its
// own launch would be, reads its arguments from a device table, and reduces
// over its own blocks only, so every member's arithmetic is bitwise that of
// its solo step. Members whose launch differs in template (apply tile,
// precision) fall into separate grids; a member that sits a launch out (no
// sinks) owns no blocks. Converged members' iteration launches return at once
// (their Krylov state says done) while the others iterate. The begin of the
// step with its first iterations, a batch of iterations, and the tail are
// captured as HIP graphs on the batch's stream and replayed while the
// members' arguments repeat.
#include <numeric>

namespace {

struct BatchSeg {
    gmk::BatchLaunch L{};
    DevBuf<unsigned char> args;
    DevBuf<int> off;
    int pool_device = -1;  // BK_APPLY: the argument entries held in the constant pool
    ~BatchSeg()
    {
        if (L.pool >= 0) gmk::apply_pool_release(pool_device, L.pool, L.M);
    }
};

// One ragged launch of `kernel` over the members in `ms`, with per-member
// arguments and block counts (0: the member sits out).
template <class A>
void add_seg(std::vector<std::unique_ptr<BatchSeg>> &out, int kernel, const std::vector<A> &args,
             const std::vector<int> &blocks, int ng = 0, int mixed = 0, int var = 0, int f32 = 0)
{
    const int M = static_cast<int>(args.size());
    std::vector<int> off(M + 1, 0);
    for (int m = 0; m < M; ++m) off[m + 1] = off[m] + blocks[m];
    if (off[M] == 0) return;
    auto seg = std::make_unique<BatchSeg>();
    seg->args.alloc(sizeof(A) * M);
    seg->off.alloc(M + 1);
    HIPCK(hipMemcpy(seg->args.p, args.data(), sizeof(A) * M, hipMemcpyHostToDevice));
    HIPCK(hipMemcpy(seg->off.p, off.data(), sizeof(int) * (M + 1), hipMemcpyHostToDevice));
    seg->L.kernel = kernel;
    seg->L.args = seg->args.p;
    seg->L.off = seg->off.p;
    seg->L.M = M;
    seg->L.total = off[M];
    seg->L.ng = ng;
    seg->L.mixed = mixed;
    seg->L.var = var;
    seg->L.f32 = f32;
    seg->L.pool = -1;
    if constexpr (std::is_same<A, gmk::ApplyArgs>::value) {
        // the apply reads its arguments from the constant pool while it has room
        int first = -1, dev = -1;
        HIPCK(hipGetDevice(&dev));
        if (kernel == gmk::BK_APPLY && ng != gmk::TAPPLY_NG && gmk::apply_pool_acquire(M, &first) == hipSuccess) {
            seg->L.pool = first;
            seg->pool_device = dev;
            HIPCK(gmk::apply_pool_write(first, args.data(), M));
        }
    }
    out.push_back(std::move(seg));
}

}  // namespace

// One lane of a batch: a subset of the members stepped as ragged grids on
// the lane's own stream.
struct BatchLane {
    std::vector<gm_solver *> mem;
    hipStream_t st = nullptr;
    // the batched step: begin (rhs and BiCGSTAB start), one iteration, tail
    std::vector<std::unique_ptr<BatchSeg>> begin, iter, tail;
    std::vector<uint64_t> plan_key;  // members' launch arguments the plan was built from
    DevBuf<gmd::Krylov *> kr;
    DevBuf<int> kstate;
    PinBuf<int> hkstate;
    // captured graphs: begin with its first iterations, a batch of iterations, the tail
    struct Graph {
        int kind = 0, batch = 0;
        hipGraphExec_t exec = nullptr;
    };
    std::vector<Graph> graphs;
    std::vector<uint64_t> last_key[3];
    // progress of the step in flight: 1 iterating, 2 tail enqueued, 3 done
    int phase = 0, enqueued = 0;
    bool capture_ok = false;
    ~BatchLane()
    {
        for (auto &g : graphs) (void)hipGraphExecDestroy(g.exec);
        if (st) (void)hipStreamDestroy(st);
    }
};

// Lanes overlap on the device: while one lane's grid of a memory-bound launch
// runs, another lane's matrix-core apply can fill the chip.
struct gm_batch {
    std::vector<gm_solver *> mem;
    std::vector<std::unique_ptr<BatchLane>> lanes;
    std::string err;
};

namespace {

// Everything the plan's launch arguments depend on, per member.
std::vector<uint64_t> batch_key(BatchLane &B)
{
    std::vector<uint64_t> k;
    for (gm_solver *S : B.mem) {
        const std::vector<uint64_t> g = graph_key(*S, 0);
        k.push_back(g.size());
        k.insert(k.end(), g.begin(), g.end());
        k.push_back(static_cast<uint64_t>(S->idx_ex));
    }
    return k;
}

int apply_blocks(gm_solver &S, gmk::ApplyArgs &a, int mode)
{
    a.mode = mode;
    if (a.ng == gmk::TAPPLY_NG) return gmk::tapply_grid(S.Nr);
    int ngroups = 0;
    const int grid = gmk::apply_grid(S.Nr, S.nchunks, a.ng, &ngroups);
    a.ngroups = ngroups;
    return grid;
}

// Ragged apply launches, one per (tile, precision) of the Gaunt form and one
// per l_max class of the polar form (FP64 for both precisions) present among
// the members.
void add_apply_segs(std::vector<std::unique_ptr<BatchSeg>> &out, std::vector<gmk::ApplyArgs> &args,
                    std::vector<int> &blocks, const std::vector<int> &run)
{
    const int M = static_cast<int>(args.size());
    for (int mixed = 0; mixed < 2; ++mixed)
        for (int ng : {1, 2, 4}) {
            std::vector<int> b(M, 0);
            bool any = false;
            for (int m = 0; m < M; ++m)
                if (run[m] && args[m].ng == ng && args[m].mixed == mixed) {
                    b[m] = blocks[m];
                    any = true;
                }
            if (any) add_seg(out, gmk::BK_APPLY, args, b, ng, mixed);
        }
    for (int cls = 4; cls <= 20; cls += 4) {
        std::vector<int> b(M, 0);
        bool any = false;
        for (int m = 0; m < M; ++m)
            if (run[m] && args[m].ng == gmk::TAPPLY_NG && gmk::tapply_class(args[m].g.lmax, 0) == cls) {
                b[m] = blocks[m];
                any = true;
            }
        if (any) add_seg(out, gmk::BK_APPLY, args, b, gmk::TAPPLY_NG, 0, cls);
    }
}

// The phase pass exp(i m theta) X that an apply without a ready copy needs.
void add_phase_seg(std::vector<std::unique_ptr<BatchSeg>> &out, const std::vector<gmk::ApplyArgs> &args,
                   const std::vector<int> &run)
{
    const int M = static_cast<int>(args.size());
    std::vector<gmk::PhaseArgs> pa(M);
    std::vector<int> b(M, 0);
    for (int m = 0; m < M; ++m) {
        pa[m] = gmk::phase_args(args[m]);
        if (run[m] && args[m].ng != gmk::TAPPLY_NG) b[m] = pa[m].nbx * args[m].g.K;
    }
    add_seg(out, gmk::BK_PHASE, pa, b);
}

void build_plan(BatchLane &B)
{
    const int M = static_cast<int>(B.mem.size());
    B.begin.clear();
    B.iter.clear();
    B.tail.clear();
    std::vector<int> all(M, 1), wood(M, 0), sinks(M, 0);
    for (int m = 0; m < M; ++m) {
        wood[m] = B.mem[m]->step_wood ? 1 : 0;
        sinks[m] = B.mem[m]->hP.p->nsink > 0 ? 1 : 0;
    }
    auto sink_seg = [&](std::vector<std::unique_ptr<BatchSeg>> &out, const std::vector<int> &run, auto &&make) {
        std::vector<gmk::SinkArgs> a(M);
        std::vector<int> b(M, 0);
        for (int m = 0; m < M; ++m) {
            a[m] = make(*B.mem[m]);
            b[m] = run[m];
        }
        add_seg(out, gmk::BK_SINK, a, b);
    };
    // ---- begin: capacitance, v^H psi, rhs with the shadow residual, BiCGSTAB start
    sink_seg(B.begin, wood, [](gm_solver &S) { return woodbury_setup_args(S); });
    sink_seg(B.begin, wood, [](gm_solver &S) { return sink_proj_args(S, S.psi.p, gmk::SD_START, 1.0, 0); });
    {
        std::vector<gmk::ApplyArgs> a(M);
        std::vector<int> b(M);
        for (int m = 0; m < M; ++m) {
            a[m] = rhs_apply_args(*B.mem[m]);
            b[m] = apply_blocks(*B.mem[m], a[m], gmk::AM_RHS);
        }
        add_phase_seg(B.begin, a, all);
        add_apply_segs(B.begin, a, b, all);
    }
    {
        std::vector<gmk::InitArgs> a(M);
        std::vector<int> b(M);
        for (int m = 0; m < M; ++m) {
            a[m] = init_args(*B.mem[m], B.mem[m]->step_wood);
            b[m] = gmk::stream_blocks(a[m].n);
        }
        add_seg(B.begin, gmk::BK_INIT, a, b);
    }
    // ---- one BiCGSTAB iteration
    auto prec_seg = [&](int var) {
        std::vector<gmk::PrecArgs> a(M);
        for (int m = 0; m < M; ++m) a[m] = iter_prec_args(*B.mem[m], B.mem[m]->step_wood, var);
        for (int f32 = 0; f32 < 2; ++f32) {
            std::vector<int> b(M, 0);
            bool any = false;
            for (int m = 0; m < M; ++m)
                if ((a[m].inv32 != nullptr) == (f32 != 0)) {
                    b[m] = gmk::precond_blocks(a[m].K, a[m].N);
                    any = true;
                }
            if (any) add_seg(B.iter, gmk::BK_PRECOND, a, b, 0, 0, var, f32);
        }
    };
    auto wcorr_seg = [&](bool z) {
        std::vector<gmk::WcorrArgs> a(M);
        std::vector<int> b(M, 0);
        for (int m = 0; m < M; ++m) {
            gm_solver &S = *B.mem[m];
            if (!wood[m]) continue;
            a[m] = wcorr_args(S, z ? S.z.p : S.y.p, krylov_xt(S), S.flags.precision == GM_PRECISION_MIXED);
            b[m] = 1;
        }
        // the coefficients (one block per member), then the correction
        add_seg(B.iter, gmk::BK_WCOEF, a, b);
        for (int m = 0; m < M; ++m)
            if (wood[m]) b[m] = gmk::stream_blocks(static_cast<size_t>(B.mem[m]->Nr) * B.mem[m]->K);
        add_seg(B.iter, gmk::BK_WCORR, a, b);
    };
    auto krylov_apply_seg = [&](int mode) {
        std::vector<gmk::ApplyArgs> a(M);
        std::vector<int> b(M);
        for (int m = 0; m < M; ++m) {
            a[m] = iter_apply_args(*B.mem[m], mode);
            b[m] = apply_blocks(*B.mem[m], a[m], mode);
        }
        add_apply_segs(B.iter, a, b, all);
    };
    prec_seg(gmk::PV_P);
    wcorr_seg(false);
    krylov_apply_seg(gmk::AM_V);
    prec_seg(gmk::PV_S);
    wcorr_seg(true);
    krylov_apply_seg(gmk::AM_T);
    {
        std::vector<gmk::UpdateArgs> a(M);
        std::vector<int> b(M);
        for (int m = 0; m < M; ++m) {
            a[m] = update_args(*B.mem[m]);
            b[m] = gmk::stream_blocks(a[m].n);
        }
        add_seg(B.iter, gmk::BK_UPDATE, a, b);
    }
    // ---- tail: acceptance, mask, diagnostics
    sink_seg(B.tail, wood, [](gm_solver &S) { return sink_proj_args(S, S.x.p, gmk::SD_COEFA_QNEW, 1.0, 0); });
    {
        std::vector<gmk::ApplyArgs> a(M);
        std::vector<int> b(M);
        for (int m = 0; m < M; ++m) {
            a[m] = res_apply_args(*B.mem[m]);
            b[m] = apply_blocks(*B.mem[m], a[m], gmk::AM_RES);
        }
        add_phase_seg(B.tail, a, all);
        add_apply_segs(B.tail, a, b, all);
    }
    {
        std::vector<gmk::AcceptArgs> a(M);
        std::vector<int> b(M);
        for (int m = 0; m < M; ++m) {
            a[m] = accept_args(*B.mem[m]);
            b[m] = gmk::stream_blocks(a[m].n);
        }
        add_seg(B.tail, gmk::BK_ACCEPT, a, b);
    }
    sink_seg(B.tail, sinks, [](gm_solver &S) { return sink_proj_args(S, S.psi.p, gmk::SD_POST, 1.0, 0); });
    {
        std::vector<gmk::ApplyArgs> a(M);
        std::vector<int> b(M);
        for (int m = 0; m < M; ++m) {
            a[m] = diag_apply_args(*B.mem[m]);
            b[m] = gmk::diag_apply_blocks(B.mem[m]->Nr, B.mem[m]->K);
        }
        add_seg(B.tail, gmk::BK_DIAG_APPLY, a, b);
    }
    {
        std::vector<gmk::DiagArgs> a(M);
        std::vector<int> b(M);
        for (int m = 0; m < M; ++m) {
            a[m] = diag_final_args(*B.mem[m]);
            b[m] = B.mem[m]->K;
        }
        add_seg(B.tail, gmk::BK_DIAG_FINAL, a, b);
    }
}

void run_segs(BatchLane &B, const std::vector<std::unique_ptr<BatchSeg>> &segs)
{
    for (const auto &s : segs) HIPCK(gmk::launch_batch(s->L, B.st));
}

void batch_gather(BatchLane &B)
{
    const int M = static_cast<int>(B.mem.size());
    HIPCK(gmk::launch_krylov_gather(B.kr.p, B.kstate.p, M, B.st));
    HIPCK(hipMemcpyAsync(B.hkstate.p, B.kstate.p, sizeof(int) * 3 * M, hipMemcpyDeviceToHost, B.st));
}

enum { BG_BEGIN = 0, BG_BATCH = 1, BG_TAIL = 2 };

// Enqueue `work` on the batch stream, through a captured graph once the same
// plan and batch size were requested twice in a row.
template <class F>
void batch_captured(BatchLane &B, int kind, int batch, bool capture_ok, F &&work)
{
    std::vector<uint64_t> key = B.plan_key;
    key.push_back(static_cast<uint64_t>(batch));
    if (!capture_ok) {
        work();
        return;
    }
    for (auto &g : B.graphs)
        if (g.kind == kind && g.batch == batch) {
            HIPCK(hipGraphLaunch(g.exec, B.st));
            return;
        }
    const bool repeat = B.last_key[kind] == key;
    B.last_key[kind] = key;
    if (!repeat) {
        work();
        return;
    }
    hipGraph_t gr = nullptr;
    HIPCK(hipStreamBeginCapture(B.st, hipStreamCaptureModeThreadLocal));
    try {
        work();
    } catch (...) {
        (void)hipStreamEndCapture(B.st, &gr);
        if (gr) (void)hipGraphDestroy(gr);
        throw;
    }
    HIPCK(hipStreamEndCapture(B.st, &gr));
    hipGraphExec_t ex = nullptr;
    const hipError_t e = hipGraphInstantiate(&ex, gr, nullptr, nullptr, 0);
    (void)hipGraphDestroy(gr);
    HIPCK(e);
    B.graphs.push_back({kind, batch, ex});
    HIPCK(hipGraphLaunch(ex, B.st));
}

void batch_drop_graphs(BatchLane &B)
{
    for (auto &g : B.graphs) (void)hipGraphExecDestroy(g.exec);
    B.graphs.clear();
    for (auto &k : B.last_key) k.clear();
}

// step's host part before the solve (src/CrankNicolsonModes.cpp:145-168):
// time, angle, phase, the time-dependent operator, pivots and, when they
// changed, the Woodbury columns; enqueued on the batch stream.
void batch_member_begin(gm_solver &S)
{
    require_initialized(S);
    restore_precision(S);
    S.step_start = S.t_i;
    S.t_i += S.dt;
    S.theta += (S.counter ? -1.0 : 1.0) * S.Omega * S.dt;
    S.phase += S.Omega * S.dt;
    build_perturbation(S, S.t_i);
    S.step_ilut = false;
    S.step_fp64 = false;
    tune_apply(S, 0);
    if (S.flags.precision == GM_PRECISION_MIXED) tune_apply(S, 1);
    S.step_wood = S.hP.p->nsink > 0;
    S.step_enqueued = 0;
    factorize(S);
    if (S.step_wood) woodbury_columns(S);
}


// Enqueue the begin of a lane's step (after its members' host parts).
void lane_begin(BatchLane &L)
{
    bool dirty = false, graphs_ok = true;
    for (gm_solver *S : L.mem) {
        dirty = dirty || S->w_dirty;
        graphs_ok = graphs_ok && graphs_enabled(*S);
    }
    const std::vector<uint64_t> key = batch_key(L);
    if (key != L.plan_key) {
        HIPCK(hipStreamSynchronize(L.st));
        batch_drop_graphs(L);
        build_plan(L);
        L.plan_key = key;
    }
    L.capture_ok = graphs_ok && !dirty;
    int batch = 1;
    for (gm_solver *S : L.mem) batch = std::max(batch, S->prev_iters);
    batch_captured(L, BG_BEGIN, batch, L.capture_ok, [&] {
        run_segs(L, L.begin);
        for (int k = 0; k < batch; ++k) run_segs(L, L.iter);
        batch_gather(L);
    });
    L.enqueued = batch;
    L.phase = 1;
}

// After the lane's stream went idle: restarts, more iterations, or the tail.
void lane_advance(BatchLane &L)
{
    if (L.phase == 2) {
        L.phase = 3;
        return;
    }
    bool all_done = true;
    for (size_t m = 0; m < L.mem.size(); ++m) {
        const int *k = L.hkstate.p + 3 * m;
        if (k[1]) {
            enqueue_restart(*L.mem[m]);
            all_done = false;
        } else if (!k[0]) {
            all_done = false;
        }
    }
    if (all_done) {
        batch_captured(L, BG_TAIL, 0, L.capture_ok, [&] {
            run_segs(L, L.tail);
            for (gm_solver *S : L.mem) tail_readback(*S, L.st);
        });
        L.phase = 2;
        return;
    }
    if (L.enqueued > 3 * kMaxIt) throw std::runtime_error("BiCGSTAB did not report convergence");
    batch_captured(L, BG_BATCH, 2, L.capture_ok, [&] {
        for (int k = 0; k < 2; ++k) run_segs(L, L.iter);
        batch_gather(L);
    });
    L.enqueued += 2;
}

}  // namespace

extern "C" {

int gm_batch_create_lanes(gm_solver *const *members, int32_t n, int32_t lanes, gm_batch **out)
{
    if (!out) return GM_E_INVALID;
    *out = nullptr;
    if (n < 1 || !members || lanes < 1) return GM_E_INVALID;
    for (int i = 0; i < n; ++i)
        for (int k = 0; k < i; ++k)
            if (!members[i] || members[i] == members[k]) return GM_E_INVALID;
    try {
        auto B = std::make_unique<gm_batch>();
        B->mem.assign(members, members + n);
        lanes = std::min(lanes, n);
        // members by decreasing work, each to the lane with the least work so far
        std::vector<int> order(n);
        std::iota(order.begin(), order.end(), 0);
        auto work = [&](int i) { return static_cast<double>(members[i]->K) * members[i]->Nr; };
        std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return work(a) > work(b); });
        std::vector<std::vector<int>> part(lanes);
        std::vector<double> load(lanes, 0.0);
        for (int i : order) {
            const int l = static_cast<int>(std::min_element(load.begin(), load.end()) - load.begin());
            part[l].push_back(i);
            load[l] += work(i);
        }
        for (auto &idx : part) {
            std::sort(idx.begin(), idx.end());
            auto L = std::make_unique<BatchLane>();
            for (int i : idx) L->mem.push_back(members[i]);
            HIPCK(hipStreamCreateWithFlags(&L->st, hipStreamNonBlocking));
            const int M = static_cast<int>(idx.size());
            L->kr.alloc(M);
            std::vector<gmd::Krylov *> kp(M);
            for (int m = 0; m < M; ++m) kp[m] = L->mem[m]->dK.p;
            HIPCK(hipMemcpy(L->kr.p, kp.data(), sizeof(gmd::Krylov *) * M, hipMemcpyHostToDevice));
            L->kstate.alloc(3 * static_cast<size_t>(M));
            L->hkstate.alloc(3 * static_cast<size_t>(M));
            B->lanes.push_back(std::move(L));
        }
        *out = B.release();
        return GM_OK;
    } catch (const std::exception &e) {
        g_create_err = e.what();
        return GM_E_DEVICE;
    }
}

int gm_batch_create(gm_solver *const *members, int32_t n, gm_batch **out)
{
    return gm_batch_create_lanes(members, n, 1, out);
}

void gm_batch_destroy(gm_batch *B) { delete B; }

const char *gm_batch_last_error(const gm_batch *B) { return B ? B->err.c_str() : ""; }

int gm_batch_step(gm_batch *B, int32_t step_id, gm_diag *diags)
{
    if (!B) return GM_E_INVALID;
    const int M = static_cast<int>(B->mem.size());
    bool ilut = false;
    for (gm_solver *S : B->mem) ilut = ilut || !S->flags.use_block_preconditioner;
    int first = GM_OK;
    if (ilut) {
        // a member on the host ILUT path: the whole step per member
        for (int m = 0; m < M; ++m) {
            gm_solver *S = B->mem[m];
            const int st = gm_ensemble_step(&S, 1, step_id, diags ? diags + m : nullptr);
            if (first == GM_OK && st != GM_OK) first = st;
        }
        return first;
    }
    std::vector<hipStream_t> own(M);
    for (int m = 0; m < M; ++m) own[m] = B->mem[m]->st;
    auto restore = [&] {
        for (int m = 0; m < M; ++m) B->mem[m]->st = own[m];
    };
    try {
        // every member's work of the step goes to its lane's stream
        for (auto &L : B->lanes)
            for (gm_solver *S : L->mem) {
                S->st = L->st;
                batch_member_begin(*S);
            }
        for (auto &L : B->lanes) lane_begin(*L);
        // lanes advance as their streams go idle
        for (bool busy = true; busy;) {
            busy = false;
            bool progressed = false;
            for (auto &L : B->lanes) {
                if (L->phase == 3) continue;
                const hipError_t q = hipStreamQuery(L->st);
                if (q == hipErrorNotReady) {
                    busy = true;
                    continue;
                }
                HIPCK(q);
                lane_advance(*L);
                progressed = true;
                busy = busy || L->phase != 3;
            }
            if (busy && !progressed) std::this_thread::yield();
        }
        for (auto &L : B->lanes)
            for (size_t i = 0; i < L->mem.size(); ++i) {
                gm_solver &S = *L->mem[i];
                const gmd::Krylov &K = *S.hK.p;
                S.last_iterations = K.iter;
                S.last_restarts = K.restarts;
                S.prev_iters = std::max(1, S.last_iterations);
            }
        for (int m = 0; m < M; ++m) {
            gm_solver &S = *B->mem[m];
            const int st = guarded(&S, [&] { step_finish(S, diags ? diags + m : nullptr); });
            if (first == GM_OK && st != GM_OK) first = st;
        }
        restore();
    } catch (const std::exception &e) {
        restore();
        for (auto &L : B->lanes) (void)hipStreamSynchronize(L->st);
        B->err = e.what();
        return GM_E_DEVICE;
    }
    if (first != GM_OK) {
        for (gm_solver *S : B->mem)
            if (!S->err.empty()) B->err = S->err;
    }
    return first;
}

}  // extern "C"
