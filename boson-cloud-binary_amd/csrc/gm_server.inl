// gm_server.inl - serving loop over device-resident solvers (part of the
// gm_solver.cpp translation unit: it drives the step phases directly).
//
// Requests (constructor arguments, flags, initial state, step count) queue up
// in arrival order. Each call of gm_server_run admits queued requests while
// their HBM footprint fits the budget and a slot is free, creates their
// solvers (state resident on the device until the request completes), and
// then walks every active solver through its step phases as soon as its own
// stream is idle (a non-blocking query): steps of different sizes from
// different requests overlap on the device, and a request admitted while
// others run joins at once. Every completed step's latency (begin of the
// step to the end of its host-side finish) is recorded.
#include <chrono>
#include <deque>
#include <thread>

namespace {

using Clock = std::chrono::steady_clock;

struct ServeRequest {
    int64_t ticket = 0;
    gm_params params{};
    gm_flags flags{};
    std::vector<double> psi0;
    double mask_f = 0, mask_W = 0;
    int mask_n = 0;
    double r_ex = 0;
    int n_steps = 0;
    double footprint = 0;
    int probe_step = 0;
    double *probe_psi = nullptr;
    Clock::time_point submitted;
};

struct ServeActive {
    ServeRequest req;
    gm_solver *S = nullptr;
    int phase = 0;  // 0: between steps, 1: iterating, 2: tail enqueued
    int steps_done = 0;
    Clock::time_point t0;
};

struct ServeResult {
    int state = 0;  // GM_SERVE_QUEUED / _RUNNING / _DONE / _FAILED
    int steps_done = 0;
    std::vector<double> psi;
    gm_scalars sc{};
    std::string error;
};

// Device bytes of one solver: 14 state-sized work vectors, the phased copy,
// the FP32 pivots, up to two full-length Woodbury columns,
// the profiles (FP64 + FP32), the Gaunt tables, the U sums of the
// polar-quadrature apply, the per-tile reduction rows and small buffers.
double footprint_bytes(const gm_params &p)
{
    const double K = (p.l_max + 1.0) * (p.l_max + 2.0) / 2.0, N = p.Nr;
    const double Ls = 2.0 * p.l_max + 1.0;
    const double vectors = (14.0 + 1.0 + 0.5 + 2.0) * K * N * 16.0;
    const double profiles = 2.0 * Ls * N * (8.0 + 4.0);
    const double tables = 12.0 * K * K * (p.l_max + 1.0);
    const double usums = 8.0 * static_cast<double>(gmk::tapply_u_doubles(p.Nr, p.l_max));
    const double rows = (N / gmd::TILE + 1.0) * K * 6.0 * 8.0;
    return vectors + profiles + tables + usums + rows + (4 << 20);
}

}  // namespace

struct gm_server {
    double budget = 0, in_use = 0;
    int max_active = 0;
    int64_t next_ticket = 1;
    std::deque<ServeRequest> queue;
    std::vector<ServeActive> active;
    std::map<int64_t, ServeResult> results;
    std::vector<double> lat_ms;
    std::vector<double> req_ms;  // submission to completion of finished requests
    double work = 0;  // mode-node-steps completed
    double busy_s = 0;  // wall time inside gm_server_run
    std::string err;
};

namespace {

void serve_fail(gm_server &V, ServeActive &a, const char *what)
{
    ServeResult &r = V.results[a.req.ticket];
    r.state = GM_SERVE_FAILED;
    r.steps_done = a.steps_done;
    r.error = std::string(what) + ": " + (a.S ? a.S->err : std::string("solver creation failed"));
}

// Create and prepare the solver of an admitted request.
bool serve_admit(gm_server &V, ServeRequest &&q)
{
    ServeActive a;
    a.req = std::move(q);
    gm_solver *S = nullptr;
    int st = gm_create(&a.req.params, &S);
    if (st != GM_OK) {
        V.results[a.req.ticket] = ServeResult{GM_SERVE_FAILED, 0, {}, {}, "gm_create: " + g_create_err};
        return false;
    }
    a.S = S;
    const ServeRequest &r = a.req;
    if (st == GM_OK) st = gm_set_flags(S, &r.flags);
    if (st == GM_OK) st = gm_set_initial_state_full(S, r.psi0.data(), static_cast<int64_t>(r.psi0.size() / 2));
    if (st == GM_OK) st = gm_initialize(S);
    if (st == GM_OK && r.mask_n > 0) st = gm_build_damping_mask(S, r.mask_f, r.mask_W, r.mask_n);
    if (st == GM_OK) st = gm_begin_run(S, r.r_ex, r.n_steps + 1);
    if (st != GM_OK) {
        serve_fail(V, a, "setup");
        gm_destroy(S);
        return false;
    }
    V.in_use += r.footprint;
    V.results[r.ticket].state = GM_SERVE_RUNNING;
    V.active.push_back(std::move(a));
    return true;
}

void serve_retire(gm_server &V, size_t i, bool ok)
{
    ServeActive &a = V.active[i];
    ServeResult &r = V.results[a.req.ticket];
    if (ok) {
        r.state = GM_SERVE_DONE;
        r.steps_done = a.steps_done;
        r.psi.resize(static_cast<size_t>(2) * a.S->Nr * a.S->K);
        if (gm_get_state(a.S, r.psi.data()) != GM_OK || gm_get_scalars(a.S, &r.sc) != GM_OK) {
            r.state = GM_SERVE_FAILED;
            r.error = a.S->err;
        } else {
            V.req_ms.push_back(std::chrono::duration<double, std::milli>(Clock::now() - a.req.submitted).count());
        }
    }
    gm_destroy(a.S);
    V.in_use -= a.req.footprint;
    V.active.erase(V.active.begin() + static_cast<std::ptrdiff_t>(i));
}

}  // namespace

extern "C" {

int gm_footprint(const gm_params *params, double *bytes)
{
    if (!params || !bytes || params->Nr < 3 || params->l_max < 0) return GM_E_INVALID;
    *bytes = footprint_bytes(*params);
    return GM_OK;
}

int gm_server_create(double hbm_budget_bytes, int32_t max_active, gm_server **out)
{
    if (!out || !(hbm_budget_bytes > 0) || max_active < 1) return GM_E_INVALID;
    *out = new gm_server;
    (*out)->budget = hbm_budget_bytes;
    (*out)->max_active = max_active;
    return GM_OK;
}

void gm_server_destroy(gm_server *V)
{
    if (!V) return;
    for (auto &a : V->active) gm_destroy(a.S);
    delete V;
}

int gm_server_submit(gm_server *V, const gm_request *req, int64_t *ticket)
{
    if (!V || !req || !ticket || !req->psi0 || req->n_steps < 1) return GM_E_INVALID;
    if (req->params.Nr < 3 || req->params.l_max < 0) return GM_E_INVALID;
    ServeRequest q;
    q.params = req->params;
    q.flags = req->flags;
    q.footprint = footprint_bytes(q.params);
    if (q.footprint > V->budget) return GM_E_INVALID;  // could never be admitted
    const size_t K = static_cast<size_t>(req->params.l_max + 1) * (req->params.l_max + 2) / 2;
    q.psi0.assign(req->psi0, req->psi0 + 2 * K * static_cast<size_t>(req->params.Nr));
    q.mask_f = req->mask_L_factor;
    q.mask_W = req->mask_W;
    q.mask_n = req->mask_n;
    q.r_ex = req->r_ex;
    q.n_steps = req->n_steps;
    q.probe_step = req->probe_step;
    q.probe_psi = req->probe_psi;
    q.submitted = Clock::now();
    q.ticket = V->next_ticket++;
    *ticket = q.ticket;
    V->results[q.ticket].state = GM_SERVE_QUEUED;
    V->queue.push_back(std::move(q));
    return GM_OK;
}

int gm_server_run(gm_server *V, double seconds, int32_t *completed)
{
    if (!V) return GM_E_INVALID;
    const Clock::time_point start = Clock::now();
    int done = 0;
    for (;;) {
        bool progressed = false;
        // admission in arrival order, by HBM footprint and free slots
        while (!V->queue.empty() && static_cast<int>(V->active.size()) < V->max_active &&
               V->in_use + V->queue.front().footprint <= V->budget) {
            ServeRequest q = std::move(V->queue.front());
            V->queue.pop_front();
            if (!serve_admit(*V, std::move(q))) ++done;
            progressed = true;
        }
        for (size_t i = 0; i < V->active.size();) {
            ServeActive &a = V->active[i];
            gm_solver *S = a.S;
            int st = GM_OK;
            bool retire = false, ok = true;
            if (a.phase == 0) {
                if (a.steps_done >= a.req.n_steps) {
                    retire = true;
                } else {
                    a.t0 = Clock::now();
                    st = guarded(S, [&] { step_begin(*S); });
                    a.phase = 1;
                    progressed = true;
                }
            } else {
                const hipError_t qe = hipStreamQuery(S->st);
                if (qe == hipSuccess) {
                    if (a.phase == 1) {
                        st = guarded(S, [&] {
                            prof_collect(*S);
                            if (solve_advance(*S)) a.phase = 2;
                        });
                    } else {
                        bool retry = false;
                        st = guarded(S, [&] {
                            prof_collect(*S);
                            // a mixed step that failed the FP64 check is enqueued
                            // again in FP64 and polled like any other solve
                            if (!step_accept(*S)) {
                                retry = true;
                                return;
                            }
                            step_finish(*S, nullptr);
                        });
                        if (st == GM_OK && retry) {
                            a.phase = 1;
                            progressed = true;
                            ++i;
                            continue;
                        }
                        const double ms =
                            std::chrono::duration<double, std::milli>(Clock::now() - a.t0).count();
                        if (st == GM_OK) {
                            V->lat_ms.push_back(ms);
                            V->work += static_cast<double>(S->K) * S->Nr;
                            ++a.steps_done;
                            V->results[a.req.ticket].steps_done = a.steps_done;
                            // spot-check probe: the state after the requested step
                            if (a.req.probe_psi && a.steps_done == a.req.probe_step)
                                st = gm_get_state(S, a.req.probe_psi);
                        }
                        a.phase = 0;
                    }
                    progressed = true;
                } else if (qe != hipErrorNotReady) {
                    S->err = std::string("HIP: ") + hipGetErrorString(qe);
                    st = GM_E_DEVICE;
                }
            }
            if (st != GM_OK) {
                serve_fail(*V, a, "step");
                retire = true;
                ok = false;
            }
            if (retire) {
                serve_retire(*V, i, ok);
                ++done;
                continue;
            }
            ++i;
        }
        const double el = std::chrono::duration<double>(Clock::now() - start).count();
        if (V->queue.empty() && V->active.empty()) break;
        if (el >= seconds) break;
        if (!progressed) std::this_thread::yield();
    }
    V->busy_s += std::chrono::duration<double>(Clock::now() - start).count();
    if (completed) *completed = done;
    return GM_OK;
}

int gm_server_status(const gm_server *V, int64_t ticket, int32_t *state, int32_t *steps_done)
{
    if (!V) return GM_E_INVALID;
    auto it = V->results.find(ticket);
    if (it == V->results.end()) return GM_E_INVALID;
    if (state) *state = it->second.state;
    if (steps_done) *steps_done = it->second.steps_done;
    return GM_OK;
}

int gm_server_take(gm_server *V, int64_t ticket, double *psi, gm_scalars *sc, char *error, int32_t error_len)
{
    if (!V) return GM_E_INVALID;
    auto it = V->results.find(ticket);
    if (it == V->results.end() || (it->second.state != GM_SERVE_DONE && it->second.state != GM_SERVE_FAILED))
        return GM_E_LOGIC;
    const ServeResult &r = it->second;
    if (error && error_len > 0) {
        std::strncpy(error, r.error.c_str(), static_cast<size_t>(error_len) - 1);
        error[error_len - 1] = 0;
    }
    const int state = r.state;
    if (state == GM_SERVE_DONE) {
        if (psi) std::memcpy(psi, r.psi.data(), sizeof(double) * r.psi.size());
        if (sc) *sc = r.sc;
    }
    V->results.erase(it);
    return state == GM_SERVE_DONE ? GM_OK : GM_E_RUNTIME;
}

int gm_server_stats(const gm_server *V, gm_serve_stats *out)
{
    if (!V || !out) return GM_E_INVALID;
    *out = gm_serve_stats{};
    out->steps = static_cast<int64_t>(V->lat_ms.size());
    out->busy_seconds = V->busy_s;
    out->mode_node_steps = V->work;
    out->queued = static_cast<int32_t>(V->queue.size());
    out->active = static_cast<int32_t>(V->active.size());
    out->hbm_in_use = V->in_use;
    out->completed = static_cast<int64_t>(V->req_ms.size());
    // nearest-rank percentiles
    auto pct = [](std::vector<double> v, double &p50, double &p99, double &mx) {
        if (v.empty()) return;
        std::sort(v.begin(), v.end());
        auto at = [&](double f) {
            const size_t k = static_cast<size_t>(std::ceil(f * v.size())) - 1;
            return v[std::min(k, v.size() - 1)];
        };
        p50 = at(0.50);
        p99 = at(0.99);
        mx = v.back();
    };
    pct(V->lat_ms, out->p50_ms, out->p99_ms, out->max_ms);
    pct(V->req_ms, out->req_p50_ms, out->req_p99_ms, out->req_max_ms);
    return GM_OK;
}

}  // extern "C"
