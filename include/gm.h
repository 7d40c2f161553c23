/*
 * gm.h - C ABI of the MI355X Crank-Nicolson step of the gravitational-molecule
 * solver.
 *
 * Every entry point mirrors one member of the reference class
 * CrankNicolsonModes (include/CrankNicolsonModes.hpp:60-455 of
 * rdellamonica/boson-cloud-binary) and is cited beside its declaration. The
 * handle owns the device state (psi, the Krylov work vectors, the Gaunt and
 * radial tables, the preconditioner factors) on one GPU; the caller only ever
 * passes plain host pointers and sizes.
 *
 * Complex arrays are interleaved (re, im) doubles, mode-major: element
 * j*Nr + i is mode j at radial node i (include/CrankNicolsonModes.hpp:10-13).
 *
 * Status codes mirror the reference's exception types:
 *   GM_OK, GM_E_INVALID (std::invalid_argument), GM_E_LOGIC (std::logic_error),
 *   GM_E_RUNTIME (std::runtime_error), GM_E_DEVICE (HIP failure or a missing
 *   device). gm_last_error() returns the message of the last failure.
 *
 * Threading: one host thread drives a handle; handles are independent.
 */
#ifndef GM_H
#define GM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    GM_OK = 0,
    GM_E_INVALID = 1,
    GM_E_LOGIC = 2,
    GM_E_RUNTIME = 3,
    GM_E_DEVICE = 4
};

enum {
    GM_PRECISION_FP64 = 0,  /* plain FP64 drop-in */
    GM_PRECISION_MIXED = 1  /* FP32 preconditioner sweeps (and FP32 coupling in the Gaunt
                               form; the polar form is FP64) inside FP64 BiCGSTAB; a step
                               that fails the FP64 residual check is solved again in FP64 */
};

typedef struct gm_solver gm_solver;

/* Constructor arguments, CrankNicolsonModes::CrankNicolsonModes
 * (include/CrankNicolsonModes.hpp:84, src/CrankNicolsonModes.cpp:11-50). */
typedef struct gm_params {
    int32_t Nr;
    double L;
    double dt;
    double epsilon;
    int32_t l_max;
    double mu;
    double M_c;
    double q;
    double a;
    double R;
    int32_t counter;
    int32_t absorption;
    int32_t save_every;
    double ramp_periods;
    double epsilon_smear;
} gm_params;

/* Public flags of the class (include/CrankNicolsonModes.hpp:137-156) plus the
 * rolling-average window (set_power_average_orbits, :153) and the arithmetic
 * mode of this build. */
typedef struct gm_flags {
    int32_t evolve_orbit;             /* default 1 */
    int32_t ramp_secondary_potential; /* default 1 */
    double steps_per_period;          /* default 50 */
    int32_t use_block_preconditioner; /* default 1 */
    double power_average_orbits;      /* default 2 */
    int32_t precision;                /* GM_PRECISION_FP64 (default) or _MIXED */
    int32_t use_graphs;               /* replay captured parts of the step (default 1);
                                         profiling (gm_profile_enable) runs without graphs */
    int32_t apply_tile;               /* operator-apply form: 0 = autotune (default); 16, 32, 64 = Gaunt-form
                                         tiles of that many nodes; 8 = the polar-quadrature form
                                         (8-node tiles; l_max <= 20) */
} gm_flags;

/* Orbit state: the 13 doubles of the checkpoint (restart.cpp:153-163) in the
 * same order, followed by solver statistics (include/CrankNicolsonModes.hpp:157-158). */
typedef struct gm_scalars {
    double t, theta, phase, R, M1, M2, q, R1, R2, Omega, P, dt, p_bin;
    double Omega_0;
    double last_solver_residual;
    int32_t last_solver_iterations;
    int32_t sink_active;
} gm_scalars;

/* Diagnostics appended by one step (src/CrankNicolsonModes.cpp:197-226,
 * src/CrankNicolsonModes_extraction.cpp:8-129). The per-mode arrays are
 * caller-owned, of length n_modes, and may be NULL. */
typedef struct gm_diag {
    double energy;       /* get_energy */
    double p_bin;        /* rolling average written to p_bin.txt */
    double p_bin_inst;   /* get_binary_power */
    double dV_bin_dt;    /* get_dVexpectation_dt */
    double cloud_mass;   /* mu * dr * |psi|^2 after the step */
    double absorbed;     /* mu * dr * dt * Re<mid|Gamma|mid>, CN sink balance */
    double *flux;        /* get_flux */
    double *energy_flux; /* get_energy_flux */
    double *norms;       /* get_norms */
    double *energies;    /* get_energy_per_mode */
} gm_diag;

/* Library version string and whether a GPU is usable. */
const char *gm_version(void);
int gm_device_count(void);

/* CrankNicolsonModes::CrankNicolsonModes (include/CrankNicolsonModes.hpp:84). */
int gm_create(const gm_params *params, gm_solver **out);
/* ~CrankNicolsonModes. */
void gm_destroy(gm_solver *solver);
/* Message of the last failed call on this handle (or of gm_create when solver is NULL). */
const char *gm_last_error(const gm_solver *solver);

/* n_modes and the (l, m) list (include/CrankNicolsonModes.hpp:91,174);
 * lm receives 2*n_modes int32. */
int gm_n_modes(const gm_solver *solver);
int gm_modes(const gm_solver *solver, int32_t *lm);
/* r (include/CrankNicolsonModes.hpp:161); receives Nr doubles. */
int gm_grid(const gm_solver *solver, double *r);

/* Public flags (include/CrankNicolsonModes.hpp:137-156) and set_power_average_orbits (:153). */
int gm_set_flags(gm_solver *solver, const gm_flags *flags);
int gm_get_flags(const gm_solver *solver, gm_flags *flags);

/* configure_isolated_black_hole (include/CrankNicolsonModes.hpp:193). */
int gm_configure_isolated_black_hole(gm_solver *solver, double x);
/* set_initial_state_full (include/CrankNicolsonModes.hpp:171): n = Nr*n_modes
 * complex values; the state is scaled by sqrt(M_c/mu). */
int gm_set_initial_state_full(gm_solver *solver, const double *psi, int64_t n);
/* initialize (include/CrankNicolsonModes.hpp:183). */
int gm_initialize(gm_solver *solver);
/* build_damping_mask / build_cap (include/CrankNicolsonModes.hpp:230,238). */
int gm_build_damping_mask(gm_solver *solver, double L_factor, double W, int32_t n);
int gm_build_cap(gm_solver *solver, double L_factor, double W, int32_t n);

/* The compute part of run() (src/CrankNicolsonModes.cpp:236-259) without its
 * file output: set_extraction_radius, t = 0, and the refresh of H(0). */
int gm_begin_run(gm_solver *solver, double r_ex, int32_t n_steps);
/* step (src/CrankNicolsonModes.cpp:145-234) without the text output; the
 * diagnostics that step() writes are returned in diag (may be NULL). */
int gm_step(gm_solver *solver, int32_t step_id, gm_diag *diag);

/* Ensemble step (no counterpart in the reference, whose drivers run one
 * binary per process): one step() of each of n independent solvers, which may
 * differ in every parameter (q, R, mu, N, l_max, precision). Each solver runs
 * on its own HIP stream; their kernels overlap on the device. diags may be
 * NULL or hold n entries (filled as by gm_step). Returns GM_OK or the status of
 * the first member that failed; a failed member keeps its message in
 * gm_last_error and the others complete their step. A solver may appear once. */
int gm_ensemble_step(gm_solver *const *members, int32_t n, int32_t step_id, gm_diag *diags);

/* Batched ragged ensemble (no counterpart in the reference): n independent
 * solvers (any q, R, mu, N, l_max, precision) stepped together, every launch
 * of the CN step being one grid over all members, each member owning the
 * blocks its own launch would use, so each member's result is bitwise that of
 * its own gm_step (at a fixed apply tile). The step is captured in HIP graphs
 * and replayed while the members' launch arguments repeat. The members stay
 * usable on their own between batch steps; a solver may be in one batch. A
 * member that fails keeps its message in gm_last_error; gm_batch_step
 * returns the first failing status. */
typedef struct gm_batch gm_batch;
int gm_batch_create(gm_solver *const *members, int32_t n, gm_batch **out);
/* The same with the members split into `lanes` subsets of about equal work,
 * each with its own stream, ragged grids and graphs: the lanes' launches
 * overlap on the device (gm_batch_create is one lane). */
int gm_batch_create_lanes(gm_solver *const *members, int32_t n, int32_t lanes, gm_batch **out);
void gm_batch_destroy(gm_batch *batch);
int gm_batch_step(gm_batch *batch, int32_t step_id, gm_diag *diags);
const char *gm_batch_last_error(const gm_batch *batch);

/* psi (include/CrankNicolsonModes.hpp:162): read, or overwrite without scaling. */
int gm_get_state(const gm_solver *solver, double *psi);
int gm_set_state(gm_solver *solver, const double *psi);
/* Orbit scalars: read, or restore for resume (restart.cpp:179-233). */
int gm_get_scalars(const gm_solver *solver, gm_scalars *out);
int gm_set_scalars(gm_solver *solver, const gm_scalars *in);

/* save_checkpoint / load_checkpoint (src/CrankNicolsonModes_restart.cpp:139-233):
 * the reference's checkpoint.bin byte layout at `path` (written to path.tmp,
 * then renamed). gm_load_checkpoint sets *loaded to 0 and changes nothing when
 * the file is missing, truncated, or its magic, step_id, Nr or n_modes differ
 * (load_checkpoint returning false); otherwise restores psi, the 13 orbit
 * scalars and the rolling power average, and sets *loaded to 1. */
int gm_save_checkpoint(gm_solver *solver, const char *path, int32_t step_id);
int gm_load_checkpoint(gm_solver *solver, const char *path, int32_t step_id, int32_t *loaded);
/* reconstruct_state_from_diagnostics (src/CrankNicolsonModes_restart.cpp:238-285),
 * for a resume without a matching checkpoint: rebuilds the orbital phase and
 * the rolling power average from the first n_rows diagnostic rows, and the
 * orbit scalars from the last of them. rows holds n_rows x 8 doubles, row k
 * being what step k+1 wrote to t.txt, R_orb.txt, theta_orb.txt,
 * Omega_orb.txt, M_1.txt, q.txt, p_bin.txt and p_bin_inst.txt, in that order.
 * n_rows = 0 resets the phase and the average only. psi is not touched. */
int gm_restore_from_diagnostics(gm_solver *solver, int32_t n_rows, const double *rows);

/* Operators on host buffers of Nr*n_modes complex values, with the current H:
 * H*psi (include/CrankNicolsonModes.hpp:197), apply_sink (:210),
 * schrodinger_rhs (:213) and sink_strength (:216). */
int gm_apply_H(gm_solver *solver, const double *in, double *out);
int gm_apply_sink(gm_solver *solver, const double *in, double *out);
int gm_schrodinger_rhs(gm_solver *solver, const double *in, double *out);
double gm_sink_strength(const gm_solver *solver, int32_t body);

/* The public sparse operators of the reference class (H, H0, A, B, V1, V2, V,
 * dVdt, Id, CAP; include/CrankNicolsonModes.hpp:197-206) at the current step,
 * materialised on the host in the reference's stacking (row j*Nr + i) as CSR
 * with columns sorted and complex values as (re, im) pairs; explicit zeros of
 * the reference's sparsity patterns are left out. For small systems only
 * (Nr*n_modes*(n_modes+3) <= 2^24). Call with row_ptr NULL for the entry
 * count in *nnz, then with row_ptr (Nr*n_modes+1), col and val (2 *nnz
 * doubles) to fill them; the operator is built once for the pair of calls. */
enum {
    GM_OP_H = 0, GM_OP_H0 = 1, GM_OP_A = 2, GM_OP_B = 3, GM_OP_V1 = 4, GM_OP_V2 = 5,
    GM_OP_V = 6, GM_OP_DVDT = 7, GM_OP_ID = 8, GM_OP_CAP = 9
};
int gm_operator_csr(gm_solver *solver, int32_t which, int64_t *nnz, int64_t *row_ptr, int32_t *col, double *val);

/* apply_sink without any allocation (no counterpart in the reference; for
 * callers that cannot allocate device memory: code under HIP graph capture,
 * serving loops that run with allocation forbidden). d_in and d_out are
 * device buffers of Nr*n_modes complex values on the solver's GPU; the result
 * equals gm_apply_sink's. The projections reduce through a static scratch
 * array of about 26 KB per GPU, set up by gm_create and held until the
 * process ends; solvers with more than 820 modes (l_max above 39) are refused.
 * stream NULL runs on the solver's stream and waits for the result; otherwise
 * the work is only enqueued on that hipStream_t. Every call on one GPU shares
 * that GPU's scratch: calls that overlap on one GPU race. */
int gm_apply_sink_device(gm_solver *solver, const double *d_in, double *d_out, void *stream);

/* The reference's ilut_solver (include/CrankNicolsonModes.hpp:264, configured at
 * src/CrankNicolsonModes_matrices.cpp:21-24) on a caller's system, host only:
 * ILUT(fill_factor, droptol) of the n x n complex CSR matrix (ptr[n+1],
 * col[nnz] sorted per row, val[2*nnz] interleaved re/im) under right-
 * preconditioned BiCGSTAB from the guess in x (2n doubles, overwritten with the
 * solution) for b (2n doubles), to |A x - b| <= tol |b| in at most
 * max_iterations. eigen_mode 1 is Eigen's recurrence, 0 the reference's
 * (src/CrankNicolsonModes_matrices.cpp:47-91). The step uses the same code
 * with fill 50, droptol 1e-9, tol 1e-10, 5000 iterations, on the operator in
 * node-major order. GM_E_RUNTIME when the factorization fails (an empty row)
 * or BiCGSTAB stops above the tolerance; *iterations is set in both cases. */
int gm_ilut_bicgstab(int32_t n, const int32_t *ptr, const int32_t *col, const double *val, const double *b,
                     double *x, double droptol, int32_t fill_factor, double tol, int32_t max_iterations,
                     int32_t eigen_mode, int32_t *iterations);

/* Restart thresholds of BiCGSTAB, as fractions of |r0|^2 (a test hook with no
 * counterpart in the reference, whose thresholds are fixed): without sinks
 * Eigen's eps^2 = 4.93e-32, with sinks the reference's 1e-32
 * (src/CrankNicolsonModes_matrices.cpp:56,67-74). Raising them forces the
 * restart branch. gm_restart_count gives the restarts of the last solve. */
int gm_set_restart_thresholds(gm_solver *solver, double eigen_eps2, double sink_factor);
int gm_restart_count(const gm_solver *solver, int32_t *n);

/* Multipole truncation of the coupling (no counterpart in the reference, which
 * evaluates every l*): at each tile of 64 radial nodes, multipoles whose radial
 * profile is below tol times the tile's leading profile are not evaluated.
 * Default 2^-64 (their sum lies below the rounding of the kept terms); 0 keeps
 * every term. Allowed range [0, 1e-6). */
int gm_set_coupling_tolerance(gm_solver *solver, double tol);
/* Work of one coupling apply under the current truncation, summed over radial
 * nodes: Gaunt terms evaluated (terms) and mode pairs coupled (pairs). */
int gm_coupling_work(gm_solver *solver, double *terms, double *pairs);
/* The truncation itself (no counterpart in the reference): the largest
 * multipole l* evaluated at each tile of 64 radial nodes, tile t covering
 * nodes [64 t, 64 t + 64); n must be at least ceil(Nr / 64). */
int gm_multipole_cut(gm_solver *solver, int32_t *lcut, int32_t n);

/* Operator-apply form chosen by the autotuner (no counterpart in the
 * reference): on a solver's first step the Gaunt-form tiles of 16, 32 and 64
 * radial nodes and the polar-quadrature form (8-node tiles, where its l_max
 * class exists) are timed on the solver's stream and the fastest kept, per
 * (N, l_max, precision, coupling tolerance) for the whole process. tile_nodes
 * receives the choice (16, 32, 64, or 8 for the polar form; 0 before the first
 * step); ms, if not NULL, the 4 measured times in ms (16, 32, 64, polar; 0 for
 * a form not timed). */
int gm_apply_config(const gm_solver *solver, int32_t precision, int32_t *tile_nodes, double *ms);

/* Measurement hook (no counterpart in the reference; for benchmarks): device
 * time per kernel class, recorded with HIP events on the solver's stream.
 * gm_profile_enable resets the counters; gm_profile_read fills ncat entries. */
enum {
    GM_PROF_APPLY_KRYLOV = 0, /* operator apply inside BiCGSTAB (v = A y, t = A z) */
    GM_PROF_APPLY_OTHER = 1,  /* rhs and acceptance applies (same coupling work) */
    GM_PROF_PRECOND = 2,      /* Thomas solves inside BiCGSTAB */
    GM_PROF_PRECOND_SETUP = 3,/* Woodbury columns P^-1 v */
    GM_PROF_UPDATE = 4,       /* x, r update and loop control */
    GM_PROF_FACTOR = 5,       /* Thomas pivots and radial profiles */
    GM_PROF_SMALL = 6,        /* sink projections, init, accept, Woodbury correction */
    GM_PROF_DIAG = 7,         /* diagnostics stencil and per-mode sums */
    GM_PROF_NCAT = 8
};
int gm_profile_enable(gm_solver *solver, int32_t on);
int gm_profile_read(const gm_solver *solver, double *ms, int64_t *launches, int32_t ncat);

/* ---------------------------------------------------------------------------
 * Serving (no counterpart in the reference, whose drivers run one binary per
 * process): a request loop over device-resident solvers.
 *
 * Requests queue in arrival order. gm_server_run admits them while their HBM
 * footprint (gm_footprint) fits the budget and fewer than max_active run,
 * creates their solvers (state resident until the request completes), and
 * advances every active solver through its step as soon as its own stream is
 * idle: steps of different sizes overlap on the device and requests that
 * arrive while others run join at once. Results are bitwise those of the same
 * request run alone. Per-step latency (start of a step to the end of its
 * host-side finish) is recorded for every step, and every finished request's
 * latency from its submission (queueing included) to its completion.
 * ------------------------------------------------------------------------- */
typedef struct gm_server gm_server;

enum { GM_SERVE_QUEUED = 0, GM_SERVE_RUNNING = 1, GM_SERVE_DONE = 2, GM_SERVE_FAILED = 3 };

typedef struct gm_request {
    gm_params params;
    gm_flags flags;
    const double *psi0;    /* initial state, Nr*n_modes complex (copied at submit) */
    double mask_L_factor;  /* damping mask (build_damping_mask), used when mask_n > 0 */
    double mask_W;
    int32_t mask_n;
    double r_ex;           /* extraction radius */
    int32_t n_steps;       /* steps to run */
    int32_t probe_step;    /* > 0: after this step, psi is copied into probe_psi (for spot checks) */
    double *probe_psi;     /* caller buffer of Nr*n_modes complex values, kept valid until the request ends */
} gm_request;

typedef struct gm_serve_stats {
    int64_t steps;           /* steps completed */
    double mode_node_steps;  /* sum of K*N over completed steps */
    double busy_seconds;     /* wall time spent inside gm_server_run */
    double p50_ms, p99_ms, max_ms;  /* per-step latency */
    int32_t queued, active;
    double hbm_in_use;       /* footprint of the active solvers, bytes */
    int64_t completed;       /* requests finished */
    double req_p50_ms, req_p99_ms, req_max_ms;  /* request latency: submission to completion */
} gm_serve_stats;

/* Estimated device bytes of one solver with these constructor arguments. */
int gm_footprint(const gm_params *params, double *bytes);
int gm_server_create(double hbm_budget_bytes, int32_t max_active, gm_server **out);
void gm_server_destroy(gm_server *server);
/* Queue a request; GM_E_INVALID if its footprint exceeds the whole budget. */
int gm_server_submit(gm_server *server, const gm_request *req, int64_t *ticket);
/* Run the loop for up to `seconds` of wall time, or until nothing is queued or
 * active; completed (may be NULL) receives the requests finished in this call. */
int gm_server_run(gm_server *server, double seconds, int32_t *completed);
int gm_server_status(const gm_server *server, int64_t ticket, int32_t *state, int32_t *steps_done);
/* Final state and scalars of a finished request (psi: Nr*n_modes complex; any
 * pointer may be NULL), then forget it. GM_E_RUNTIME with the message in
 * error for a failed request; GM_E_LOGIC while it is queued or running. */
int gm_server_take(gm_server *server, int64_t ticket, double *psi, gm_scalars *sc, char *error, int32_t error_len);
int gm_server_stats(const gm_server *server, gm_serve_stats *out);

#ifdef __cplusplus
}
#endif

#endif /* GM_H */
