/*
 * gm_oracle.c - FP64 CPU oracle of the Crank-Nicolson step.
 *
 * TEST INFRASTRUCTURE ONLY. This file is the correctness checker for the HIP
 * implementation: only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg load it, and never as the thing measured or shipped.
 *
 * It restates, in plain C99 + OpenMP, the algorithm of the reference solver
 * rdellamonica/boson-cloud-binary (Eigen/GSL are absent here, so the Eigen
 * BiCGSTAB recurrence and the GSL special functions are restated too). Every
 * function cites the reference lines it follows. The reference stores seven
 * complex CSR matrices; this oracle is matrix-free: per step it stores the
 * real coupling W_jk(r_i) (the reference's V without its orbital phase) and
 * applies every operator row by row in the reference's column order.
 *
 * Parity pins (see tests/): 3j and Y_lm(pi/2) against exact sympy values;
 * operators against a dense restatement of tests/parallel_correctness.cpp;
 * one CN step against a dense LU solve; Thomas solve against the synthetic
 * tridiagonal known answer; physics invariants of tests/physics.cpp.
 * BiCGSTAB iteration counts: parity unpinned except where the mathematics
 * fixes them (one mode without absorbers, where the radial preconditioner is
 * the exact inverse and the loop stops after one iteration); no reference test
 * checks them. The recurrence is restated from Eigen 3.4's BiCGSTAB.h as the
 * reference uses it (src/CrankNicolsonModes_matrices.cpp:44-91,188-196).
 * The Gaussian-quadrature MODESBIN generator and the ILUT solve are pinned in
 * tests/test_initial.py and tests/test_ilut.py.
 */
#include <complex.h>
#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "gm.h"

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

typedef double complex zc;

/* ------------------------------------------------------------------------ */
/* Special functions (GSL restated).                                         */
/* ------------------------------------------------------------------------ */

typedef __float128 q128;

static q128 q_sqrt(q128 v)
{
    if (v <= 0) return 0;
    /* bring v into the double range by exact powers of 4 (the Racah
     * prefactors pass 1e308 from l ~ 25 on), then Newton from the double root */
    q128 scale = 1;
    while (v > 1e300Q) {
        v *= 0x1p-200Q;
        scale *= 0x1p100Q;
    }
    while (v < 1e-300Q) {
        v *= 0x1p200Q;
        scale *= 0x1p-100Q;
    }
    q128 x = (q128)sqrt((double)v);
    for (int it = 0; it < 3; ++it) x = 0.5Q * (x + v / x);
    return x * scale;
}

static q128 q_fact(int n)
{
    static q128 table[400];
    static int ready = 0;
    if (!ready) {
        table[0] = 1;
        for (int i = 1; i < 400; ++i) table[i] = table[i - 1] * (q128)i;
        ready = 1;
    }
    return table[n];
}

/* Wigner 3j (j1 j2 j3; m1 m2 m3) for integer arguments by the Racah formula
 * in binary128, the quantity gsl_sf_coupling_3j returns
 * (src/CrankNicolsonModes_hamiltonian.cpp:85-86 passes doubled arguments). */
double om_wigner3j(int j1, int j2, int j3, int m1, int m2, int m3)
{
    if (m1 + m2 + m3 != 0) return 0.0;
    if (j1 < 0 || j2 < 0 || j3 < 0) return 0.0;
    if (abs(m1) > j1 || abs(m2) > j2 || abs(m3) > j3) return 0.0;
    if (j3 < abs(j1 - j2) || j3 > j1 + j2) return 0.0;
    int tmin = 0;
    if (j2 - j3 - m1 > tmin) tmin = j2 - j3 - m1;
    if (j1 - j3 + m2 > tmin) tmin = j1 - j3 + m2;
    int tmax = j1 + j2 - j3;
    if (j1 - m1 < tmax) tmax = j1 - m1;
    if (j2 + m2 < tmax) tmax = j2 + m2;
    q128 sum = 0;
    for (int t = tmin; t <= tmax; ++t) {
        q128 den = q_fact(t) * q_fact(j3 - j2 + t + m1) * q_fact(j3 - j1 + t - m2) *
                   q_fact(j1 + j2 - j3 - t) * q_fact(j1 - t - m1) * q_fact(j2 - t + m2);
        q128 term = 1 / den;
        sum += (t & 1) ? -term : term;
    }
    q128 delta = q_fact(j1 + j2 - j3) * q_fact(j1 - j2 + j3) * q_fact(-j1 + j2 + j3) /
                 q_fact(j1 + j2 + j3 + 1);
    q128 pre = delta * q_fact(j1 + m1) * q_fact(j1 - m1) * q_fact(j2 + m2) * q_fact(j2 - m2) *
               q_fact(j3 + m3) * q_fact(j3 - m3);
    q128 val = q_sqrt(pre) * sum;
    int ph = j1 - j2 - m3;
    if (ph & 1) val = -val;
    return (double)val;
}

/* gsl_sf_legendre_sphPlm(l, m, 0) (Condon-Shortley phase included), m >= 0:
 * sqrt((2l+1)/(4pi) (l-m)!/(l+m)!) P_l^m(0), with
 * P_l^m(0) = (-1)^((l+m)/2) (l+m-1)!!/(l-m)!! for even l+m, else 0. */
double om_sphPlm0(int l, int m)
{
    if (m < 0 || m > l) return 0.0;
    if ((l + m) & 1) return 0.0;
    q128 dfn = 1; /* (l+m-1)!! */
    for (int k = l + m - 1; k > 1; k -= 2) dfn *= k;
    q128 dfd = 1; /* (l-m)!! */
    for (int k = l - m; k > 1; k -= 2) dfd *= k;
    q128 p = dfn / dfd;
    if (((l + m) / 2) & 1) p = -p;
    q128 norm = (q128)(2 * l + 1) / (4 * (q128)M_PI) * q_fact(l - m) / q_fact(l + m);
    return (double)(q_sqrt(norm) * p);
}

/* Gauss-Legendre nodes and weights on [0, 1], ascending, as
 * gsl_integration_glfixed_point(0, 1, k, ...) returns them
 * (src/CrankNicolsonModes_absorption.cpp:27-34). Newton on P_n in binary128. */
void om_gauss_legendre01(int n, double *x, double *w)
{
    for (int i = 0; i < n; ++i) {
        /* i-th root from the top of (-1, 1) */
        q128 z = (q128)cos(M_PI * (i + 0.75) / (n + 0.5));
        q128 dp = 0;
        for (int it = 0; it < 100; ++it) {
            q128 p0 = 1, p1 = z;
            for (int k = 2; k <= n; ++k) {
                q128 p2 = ((2 * k - 1) * z * p1 - (k - 1) * p0) / k;
                p0 = p1;
                p1 = p2;
            }
            if (n == 1) { p1 = z; p0 = 1; }
            dp = n * (z * p1 - p0) / (z * z - 1);
            q128 dz = p1 / dp;
            z -= dz;
            double adz = (double)dz;
            if (fabs(adz) < 1e-33) break;
        }
        {
            q128 p0 = 1, p1 = z;
            for (int k = 2; k <= n; ++k) {
                q128 p2 = ((2 * k - 1) * z * p1 - (k - 1) * p0) / k;
                p0 = p1;
                p1 = p2;
            }
            dp = n * (z * p1 - p0) / (z * z - 1);
        }
        q128 wt = 2 / ((1 - z * z) * dp * dp);
        /* z descends with i: store ascending on [0,1] */
        /* z descends with i; (1 + z)/2 maps [-1, 1] onto [0, 1] as GSL's
         * B + A*x with A = B = 1/2. */
        x[n - 1 - i] = (double)((1 + z) / 2);
        w[n - 1 - i] = (double)(wt / 2);
    }
}

/* ------------------------------------------------------------------------ */
/* Rolling phase average (src/RollingPhase.cpp:14-129).                      */
/* ------------------------------------------------------------------------ */

typedef struct { double time0, time1, phase0, phase1, value0, value1, integral; } om_segment;

typedef struct {
    double window_phase;
    om_segment *seg;
    size_t head, count, cap;
    double prev_time, prev_phase, prev_value;
    int has_previous, ready;
    double initial_phase, integral, duration, average;
} om_rolling;

static void rolling_init(om_rolling *a, double n_orbits)
{
    free(a->seg);
    memset(a, 0, sizeof(*a));
    a->window_phase = 2.0 * M_PI * n_orbits;
}

static int rolling_update(om_rolling *a, double time, double phase, double value, double *out)
{
    if (!a->has_previous) {
        a->prev_time = time; a->prev_phase = phase; a->prev_value = value;
        a->initial_phase = phase;
        a->has_previous = 1;
        a->average = value;
        *out = a->average;
        return 0;
    }
    if (time <= a->prev_time) return -1;
    if (phase <= a->prev_phase) return -2;
    om_segment s = {a->prev_time, time, a->prev_phase, phase, a->prev_value, value, 0.0};
    s.integral = 0.5 * (s.value0 + s.value1) * (s.time1 - s.time0);
    if (a->head + a->count == a->cap) {
        if (a->head > 0) {
            memmove(a->seg, a->seg + a->head, a->count * sizeof(om_segment));
            a->head = 0;
        }
        if (a->count == a->cap) {
            a->cap = a->cap ? 2 * a->cap : 64;
            a->seg = (om_segment *)realloc(a->seg, a->cap * sizeof(om_segment));
        }
    }
    a->seg[a->head + a->count++] = s;
    a->integral += s.integral;
    a->duration += s.time1 - s.time0;
    if (phase - a->initial_phase >= a->window_phase) a->ready = 1;
    const double phase_cut = phase - a->window_phase;
    while (a->count > 0 && a->seg[a->head].phase1 <= phase_cut) {
        om_segment *f = &a->seg[a->head];
        a->integral -= f->integral;
        a->duration -= f->time1 - f->time0;
        a->head++;
        a->count--;
    }
    if (a->count > 0) {
        om_segment *f = &a->seg[a->head];
        if (f->phase0 < phase_cut && phase_cut < f->phase1) {
            const double fraction = (phase_cut - f->phase0) / (f->phase1 - f->phase0);
            const double time_cut = f->time0 + fraction * (f->time1 - f->time0);
            const double value_cut = f->value0 + fraction * (f->value1 - f->value0);
            const double removed = 0.5 * (f->value0 + value_cut) * (time_cut - f->time0);
            a->integral -= removed;
            a->duration -= time_cut - f->time0;
            f->time0 = time_cut;
            f->phase0 = phase_cut;
            f->value0 = value_cut;
            f->integral = 0.5 * (f->value0 + f->value1) * (f->time1 - f->time0);
        }
    }
    if (a->duration < 0.0) a->duration = 0.0;
    a->average = a->duration > 0.0 ? a->integral / a->duration : value;
    a->prev_time = time; a->prev_phase = phase; a->prev_value = value;
    *out = a->average;
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Solver state (include/CrankNicolsonModes.hpp:60-455).                     */
/* ------------------------------------------------------------------------ */

typedef struct {
    int row, col, m_star, t0, nt;
} om_pair;

typedef struct {
    double radius, width;
    double *mult; /* (2 l_max + 1) * Nr */
} om_profile;

typedef struct {
    zc *v;            /* Nr * K */
    double strength;
} om_sink;

struct gm_solver {
    /* constructor */
    int Nr, save_every, l_max, K;
    double epsilon, mu, q, R, theta, phase, Omega, Omega_0, P, R1, R2, M1, M2, a, L, dr, dt;
    int counter;
    double M_c;
    int isolated;
    double ramp_periods, initial_period;
    int absorption;
    double epsilon_smear;
    gm_flags flags;
    int initialized;
    int *ml, *mm;
    double *r;
    zc *psi, *psi_last;
    /* masks */
    int mask, cap;
    double *damping; /* Nr (same for all modes) */
    zc *capv;        /* Nr diagonal */
    /* coupling cache (src/CrankNicolsonModes_hamiltonian.cpp:68-104) */
    int npairs;
    om_pair *pairs;
    int *term_l;
    double *term_ang;
    /* per-step potential */
    double *f1, *f2; /* (2 l_max + 1) * Nr */
    double *W;       /* Nr * K * K, node-major */
    int32_t *lcut;   /* Nr: largest l* kept at each node, or NULL for every term (a test probe) */
    zc *phases;      /* 4 l_max + 1 */
    double ramp_factor, prf, Omega_V;
    /* absorption */
    om_profile prof1, prof2;
    om_sink sink1, sink2;
    int sink_active, mode00;
    double Ah1, Ah2;
    /* run */
    double r_ex;
    int idx_ex;
    double t_i, p_bin;
    om_rolling avg;
    double last_residual;
    int last_iterations;
    double rst_eps2, rst_ref; /* BiCGSTAB restart thresholds relative to |r0|^2 */
    int last_restarts;
    char err[512];
};

static int fail(gm_solver *s, int code, const char *msg)
{
    if (s) snprintf(s->err, sizeof(s->err), "%s", msg);
    return code;
}

static char g_create_err[512];

static double ramp_end_time(const gm_solver *s) { return s->ramp_periods * s->initial_period; }

static size_t nvec(const gm_solver *s) { return (size_t)s->Nr * (size_t)s->K; }

const char *om_version(void) { return "gm-oracle-1"; }

const char *om_last_error(const gm_solver *s) { return s ? s->err : g_create_err; }

/* build_grid / build_modes (src/CrankNicolsonModes.cpp:100-128). */
int om_create(const gm_params *p, gm_solver **out)
{
    *out = NULL;
    if (p->Nr < 3 || p->l_max < 0 || p->L <= 0.0 || p->dt <= 0.0 || p->save_every <= 0) {
        snprintf(g_create_err, sizeof(g_create_err), "%s",
                 "Require Nr >= 3, l_max >= 0, and positive L, dt, save_every");
        return GM_E_INVALID;
    }
    gm_solver *s = (gm_solver *)calloc(1, sizeof(gm_solver));
    s->Nr = p->Nr;
    s->save_every = p->save_every;
    s->l_max = p->l_max;
    s->epsilon = p->epsilon;
    s->mu = p->mu;
    s->q = p->q;
    s->R = p->R;
    s->a = p->a;
    s->L = p->L;
    s->dr = p->L / (p->Nr - 1);
    s->dt = p->dt;
    s->counter = p->counter;
    s->M_c = p->M_c;
    s->ramp_periods = p->ramp_periods;
    s->initial_period = 2.0 * M_PI / sqrt((1 + p->q) / pow(p->R, 3));
    s->absorption = p->absorption;
    s->epsilon_smear = p->epsilon_smear;
    s->Omega = 0.0;
    s->Omega_0 = sqrt((1.0 + p->q) / pow(p->R, 3));
    s->P = s->initial_period;
    s->R1 = p->q / (1 + p->q) * p->R;
    s->R2 = 1 / (1 + p->q) * p->R;
    s->theta = 0.0;
    s->phase = 0.0;
    s->M1 = 1.0;
    s->M2 = p->q;
    s->flags.evolve_orbit = 1;
    s->flags.ramp_secondary_potential = 1;
    s->flags.steps_per_period = 50.0;
    s->flags.use_block_preconditioner = 1;
    s->flags.power_average_orbits = 2.0;
    s->flags.precision = GM_PRECISION_FP64;
    s->flags.use_graphs = 1;
    s->flags.apply_tile = 0;
    s->rst_eps2 = DBL_EPSILON * DBL_EPSILON;
    s->rst_ref = 1e-32;
    rolling_init(&s->avg, 2.0);
    s->r = (double *)malloc(sizeof(double) * s->Nr);
    for (int i = 0; i < s->Nr; i++) s->r[i] = s->epsilon + i * s->dr;
    int K = 0;
    for (int l = 0; l <= s->l_max; ++l)
        for (int m = -l; m <= l; ++m)
            if (abs(l - m) % 2 == 0) K++;
    s->K = K;
    s->ml = (int *)malloc(sizeof(int) * K);
    s->mm = (int *)malloc(sizeof(int) * K);
    K = 0;
    for (int l = 0; l <= s->l_max; ++l)
        for (int m = -l; m <= l; ++m)
            if (abs(l - m) % 2 == 0) { s->ml[K] = l; s->mm[K] = m; K++; }
    s->psi = (zc *)calloc(nvec(s), sizeof(zc));
    s->psi_last = (zc *)calloc(nvec(s), sizeof(zc));
    s->mode00 = -1;
    *out = s;
    return GM_OK;
}

static void free_profile(om_profile *p) { free(p->mult); p->mult = NULL; p->radius = p->width = -1.0; }

void om_destroy(gm_solver *s)
{
    if (!s) return;
    free(s->ml); free(s->mm); free(s->r); free(s->psi); free(s->psi_last);
    free(s->damping); free(s->capv); free(s->pairs); free(s->term_l); free(s->term_ang);
    free(s->f1); free(s->f2); free(s->W); free(s->lcut); free(s->phases);
    free_profile(&s->prof1); free_profile(&s->prof2);
    free(s->sink1.v); free(s->sink2.v);
    free(s->avg.seg);
    free(s);
}

int om_n_modes(const gm_solver *s) { return s->K; }

int om_modes(const gm_solver *s, int32_t *lm)
{
    for (int j = 0; j < s->K; ++j) { lm[2 * j] = s->ml[j]; lm[2 * j + 1] = s->mm[j]; }
    return GM_OK;
}

int om_grid(const gm_solver *s, double *r)
{
    memcpy(r, s->r, sizeof(double) * s->Nr);
    return GM_OK;
}

int om_set_flags(gm_solver *s, const gm_flags *f)
{
    if (!(f->power_average_orbits > 0.0))
        return fail(s, GM_E_INVALID, "RollingPhaseAverage: n_orbits must be positive");
    if (f->power_average_orbits != s->flags.power_average_orbits)
        rolling_init(&s->avg, f->power_average_orbits);
    s->flags = *f;
    return GM_OK;
}

int om_get_flags(const gm_solver *s, gm_flags *f) { *f = s->flags; return GM_OK; }

/* configure_isolated_black_hole (src/CrankNicolsonModes.cpp:58-74). */
int om_configure_isolated_black_hole(gm_solver *s, double x)
{
    if (!isfinite(x) || fabs(x) >= s->L)
        return fail(s, GM_E_INVALID, "The isolated BH must lie inside the radial domain");
    if (s->initialized)
        return fail(s, GM_E_LOGIC, "Configure the isolated BH before initialize()");
    s->isolated = 1;
    s->flags.evolve_orbit = 0;
    s->M1 = 1.0;
    s->M2 = s->q = 0.0;
    s->R = s->R1 = fabs(x);
    s->R2 = 0.0;
    s->theta = x < 0.0 ? M_PI : 0.0;
    s->phase = 0.0;
    s->Omega = s->Omega_0 = 0.0;
    s->ramp_periods = 0.0;
    return GM_OK;
}

/* set_initial_state_full (src/CrankNicolsonModes.cpp:130-143). */
int om_set_initial_state_full(gm_solver *s, const double *psi, int64_t n)
{
    if (n != (int64_t)nvec(s))
        return fail(s, GM_E_INVALID, "Initial state size does not match the radial grid and modes");
    const double scale = sqrt(s->M_c / s->mu);
    for (size_t k = 0; k < nvec(s); ++k) s->psi[k] = (psi[2 * k] + I * psi[2 * k + 1]) * scale;
    memset(s->psi_last, 0, sizeof(zc) * nvec(s));
    return GM_OK;
}

int om_set_state(gm_solver *s, const double *psi)
{
    for (size_t k = 0; k < nvec(s); ++k) s->psi[k] = psi[2 * k] + I * psi[2 * k + 1];
    return GM_OK;
}

int om_get_state(const gm_solver *s, double *psi)
{
    for (size_t k = 0; k < nvec(s); ++k) { psi[2 * k] = creal(s->psi[k]); psi[2 * k + 1] = cimag(s->psi[k]); }
    return GM_OK;
}

/* build_coupling_cache (src/CrankNicolsonModes_hamiltonian.cpp:68-104). */
static void build_coupling_cache(gm_solver *s)
{
    const int K = s->K;
    int cap_terms = 1024, nterms = 0;
    free(s->pairs); free(s->term_l); free(s->term_ang);
    s->pairs = (om_pair *)malloc(sizeof(om_pair) * (size_t)K * K);
    s->term_l = (int *)malloc(sizeof(int) * cap_terms);
    s->term_ang = (double *)malloc(sizeof(double) * cap_terms);
    s->npairs = 0;
    for (int j = 0; j < K; ++j)
        for (int k = 0; k < K; ++k) {
            const int l = s->ml[j], m = s->mm[j];
            const int l1 = s->ml[k], m1 = s->mm[k];
            const int m_star = m1 - m;
            om_pair pr = {j, k, m_star, nterms, 0};
            for (int ls = abs(l1 - l); ls <= l1 + l; ++ls) {
                if (abs(m_star) > ls || (l + l1 + ls) % 2 != 0) continue;
                const double w1 = om_wigner3j(ls, l1, l, 0, 0, 0);
                const double w2 = om_wigner3j(ls, l1, l, -m_star, m1, -m);
                double y = om_sphPlm0(ls, abs(m_star));
                if (m_star < 0 && abs(m_star) % 2 != 0) y = -y;
                const double pref = (m1 % 2 == 0 ? 1.0 : -1.0) *
                                    sqrt((2.0 * l + 1) * (2.0 * l1 + 1) * (2.0 * ls + 1) / (4.0 * M_PI));
                const double angular = pref * w1 * w2 * y;
                if (angular != 0.0) {
                    if (nterms == cap_terms) {
                        cap_terms *= 2;
                        s->term_l = (int *)realloc(s->term_l, sizeof(int) * cap_terms);
                        s->term_ang = (double *)realloc(s->term_ang, sizeof(double) * cap_terms);
                    }
                    s->term_l[nterms] = ls;
                    s->term_ang[nterms] = angular;
                    nterms++;
                    pr.nt++;
                }
            }
            if (pr.nt > 0) s->pairs[s->npairs++] = pr;
        }
}

/* F1, F2 and the radial profile loop (src/CrankNicolsonModes_hamiltonian.cpp:12-42,179-185). */
static double F1(const gm_solver *s, double r, int ls)
{
    if (r >= s->R1) return 1 / (1 + s->q) * pow(s->R1, ls) / pow(r, ls + 1);
    return 1 / (1 + s->q) * pow(r, ls) / pow(s->R1, ls + 1);
}

static double F2(const gm_solver *s, double r, int ls)
{
    if (r >= s->R2) return s->q / (1 + s->q) * pow(s->R2, ls) / pow(r, ls + 1);
    return s->q / (1 + s->q) * pow(r, ls) / pow(s->R2, ls + 1);
}

/* build_static_perturbation (src/CrankNicolsonModes_hamiltonian.cpp:176-222),
 * storing the real coupling W = v1 + prf*v2 without the phase exp(i m* theta). */
static void build_static_perturbation(gm_solver *s)
{
    const int N = s->Nr, K = s->K, Ls = 2 * s->l_max + 1;
    if (!s->f1) {
        s->f1 = (double *)calloc((size_t)Ls * N, sizeof(double));
        s->f2 = (double *)calloc((size_t)Ls * N, sizeof(double));
        s->W = (double *)calloc((size_t)N * K * K, sizeof(double));
        s->phases = (zc *)calloc(4 * s->l_max + 1, sizeof(zc));
    }
#pragma omp parallel for schedule(static)
    for (int l = 0; l < Ls; ++l)
        for (int i = 1; i < N - 1; ++i) {
            s->f1[(size_t)l * N + i] = F1(s, s->r[i], l);
            s->f2[(size_t)l * N + i] = F2(s, s->r[i], l);
        }
    for (int m = -2 * s->l_max; m <= 2 * s->l_max; ++m)
        s->phases[m + 2 * s->l_max] = cexp(I * (double)m * s->theta);
    const double prf = s->prf;
    const double Mt = s->M1 + s->M2;
#pragma omp parallel for schedule(static)
    for (int i = 1; i < N - 1; ++i) {
        double *Wi = s->W + (size_t)i * K * K;
        for (int p = 0; p < s->npairs; ++p) {
            const om_pair *pr = &s->pairs[p];
            const double parity = pr->m_star % 2 == 0 ? 1.0 : -1.0;
            double v1 = 0.0, v2 = 0.0;
            for (int t = pr->t0; t < pr->t0 + pr->nt; ++t) {
                const int ls = s->term_l[t];
                if (s->lcut && ls > s->lcut[i]) continue;
                const double common = -s->mu * Mt * 4.0 * M_PI / (2.0 * ls + 1.0) * s->term_ang[t];
                v1 += common * s->f1[(size_t)ls * N + i];
                v2 += common * (parity * s->f2[(size_t)ls * N + i]);
            }
            Wi[(size_t)pr->row * K + pr->col] = v1 + v2 * prf;
        }
    }
    s->Omega_V = s->Omega;
}

/* update_absorption_profile (src/CrankNicolsonModes_absorption.cpp:12-81). */
static void update_absorption_profile(gm_solver *s, om_profile *prof, double radius, double width)
{
    const int N = s->Nr, degree = 2 * s->l_max;
    const size_t size = (size_t)(degree + 1) * N;
    if (prof->mult && prof->radius == radius && prof->width == width) return;
    prof->radius = radius;
    prof->width = width;
    free(prof->mult);
    prof->mult = (double *)calloc(size, sizeof(double));
    if (width < s->dr || width <= 0.0 || radius < 0.0) return;
    const int nq = 128 > degree + 16 ? 128 : degree + 16;
    double *nodes = (double *)malloc(sizeof(double) * nq), *weights = (double *)malloc(sizeof(double) * nq);
    om_gauss_legendre01(nq, nodes, weights);
    for (int i = 1; i < N - 1; ++i) {
        const double delta = s->r[i] - radius;
        if (fabs(delta) >= width) continue;
        if (radius == 0.0) {
            const double x = s->r[i] / width;
            prof->mult[i] = 4.0 * M_PI * exp(-1.0 / (1.0 - x * x));
            continue;
        }
        const double product = 2.0 * s->r[i] * radius;
        double cap = (width * width - delta * delta) / product;
        if (cap > 2.0) cap = 2.0;
        for (int k = 0; k < nq; ++k) {
            const double z = cap * nodes[k];
            const double u = 1.0 - z;
            const double distance2 = delta * delta + product * z;
            const double weight = 2.0 * M_PI * cap * weights[k] * exp(-1.0 / (1.0 - distance2 / (width * width)));
            double previous = 1.0, current = u;
            prof->mult[i] += weight;
            if (degree > 0) prof->mult[N + i] += weight * current;
            for (int l = 2; l <= degree; ++l) {
                const double next = ((2 * l - 1) * u * current - (l - 1) * previous) / l;
                prof->mult[(size_t)l * N + i] += weight * next;
                previous = current;
                current = next;
            }
        }
    }
    free(nodes);
    free(weights);
    double norm = 0.0;
    for (int i = 1; i < N - 1; ++i) norm += s->dr * s->r[i] * s->r[i] * prof->mult[i];
    if (norm > 0.0)
        for (size_t k = 0; k < size; ++k) prof->mult[k] /= norm;
}

/* build_swave_projector (src/CrankNicolsonModes_absorption.cpp:127-151). */
static void build_swave_projector(gm_solver *s, const om_profile *prof, int secondary, om_sink *sink)
{
    const int N = s->Nr;
    if (!sink->v) sink->v = (zc *)calloc(nvec(s), sizeof(zc));
    memset(sink->v, 0, sizeof(zc) * nvec(s));
    if (s->mode00 < 0) return;
    const double sqrt4pi = sqrt(4.0 * M_PI);
    for (int p = 0; p < s->npairs; ++p) {
        const om_pair *pr = &s->pairs[p];
        if (pr->col != s->mode00) continue;
        const zc phase = cexp(I * (double)pr->m_star * s->theta);
        const double parity = secondary && pr->m_star % 2 != 0 ? -1.0 : 1.0;
        for (int i = 1; i < N - 1; ++i) {
            zc kappa = 0.0;
            for (int t = pr->t0; t < pr->t0 + pr->nt; ++t)
                kappa += prof->mult[(size_t)s->term_l[t] * N + i] * s->term_ang[t] * phase;
            sink->v[(size_t)pr->row * N + i] = s->r[i] * sqrt4pi * parity * kappa;
        }
    }
}

/* build_spherical_absorption (src/CrankNicolsonModes_absorption.cpp:153-182). */
static void build_spherical_absorption(gm_solver *s)
{
    update_absorption_profile(s, &s->prof1, s->R1, s->epsilon_smear * s->M1);
    update_absorption_profile(s, &s->prof2, s->R2, s->epsilon_smear * s->M2);
    double ramp = 1.0;
    if (!s->isolated && s->t_i < ramp_end_time(s)) ramp = 0.0;
    else if (!s->isolated && s->t_i < ramp_end_time(s) + s->initial_period) {
        const double x = (s->t_i - ramp_end_time(s)) / s->initial_period;
        ramp = 3.0 * x * x - 2.0 * x * x * x;
    }
    s->Ah1 = 16.0 * M_PI * s->M1 * s->M1 * ramp;
    s->Ah2 = 16.0 * M_PI * s->M2 * s->M2 * ramp;
    build_swave_projector(s, &s->prof1, 0, &s->sink1);
    s->sink1.strength = s->Ah1 * s->dr;
    build_swave_projector(s, &s->prof2, 1, &s->sink2);
    s->sink2.strength = s->Ah2 * s->dr;
}

/* build_time_dependent_perturbation (src/CrankNicolsonModes_hamiltonian.cpp:245-268). */
static void build_time_dependent_perturbation(gm_solver *s, double time)
{
    s->ramp_factor = 1.0;
    if (s->isolated) {
        s->Omega = 0.0;
    } else if (time < ramp_end_time(s)) {
        const double x = time / ramp_end_time(s);
        s->ramp_factor = 3.0 * x * x - 2.0 * x * x * x;
        s->Omega = s->Omega_0 * s->ramp_factor;
    } else {
        s->Omega = sqrt((s->M1 + s->M2) / pow(s->R, 3));
    }
    s->prf = s->flags.ramp_secondary_potential ? s->ramp_factor : 1.0;
    build_static_perturbation(s);
    if (s->absorption) build_spherical_absorption(s);
    s->sink_active = s->absorption && time >= ramp_end_time(s);
}

/* initialize (src/CrankNicolsonModes.cpp:76-98). */
int om_initialize(gm_solver *s)
{
    build_coupling_cache(s);
    if (s->absorption) {
        free_profile(&s->prof1);
        free_profile(&s->prof2);
        s->mode00 = -1;
        for (int j = 0; j < s->K; ++j)
            if (s->ml[j] == 0 && s->mm[j] == 0) s->mode00 = j;
        free(s->sink1.v); free(s->sink2.v);
        s->sink1.v = s->sink2.v = NULL;
        s->sink1.strength = s->sink2.strength = 0.0;
    }
    s->t_i = 0.0;
    build_time_dependent_perturbation(s, 0.0);
    s->initialized = 1;
    return GM_OK;
}

/* build_damping_mask (src/CrankNicolsonModes_outgoing.cpp:10-35). */
int om_build_damping_mask(gm_solver *s, double f, double W, int32_t n)
{
    free(s->damping);
    s->damping = (double *)malloc(sizeof(double) * s->Nr);
    s->mask = 1;
    const int idx_mask = (int)floor(s->L * f / s->dr);
    const double r_mask = f * s->L, L_mask = (1 - f) * s->L;
    for (int i = 0; i < s->Nr; i++)
        s->damping[i] = i >= idx_mask ? exp(-W * pow((s->r[i] - r_mask) / L_mask, n)) : 1.0;
    return GM_OK;
}

/* build_cap (src/CrankNicolsonModes_outgoing.cpp:37-62), including the
 * reference's (r - L_cap) denominator. */
int om_build_cap(gm_solver *s, double f, double W, int32_t n)
{
    free(s->capv);
    s->capv = (zc *)calloc(s->Nr, sizeof(zc));
    s->cap = 1;
    const int idx_cap = (int)floor(s->L * f / s->dr);
    const double r_cap = f * s->L, L_cap = (1 - f) * s->L;
    for (int i = 0; i < s->Nr; i++)
        if (i >= idx_cap) s->capv[i] = -I * W * pow((s->r[i] - r_cap) / (s->r[i] - L_cap), n);
    return GM_OK;
}

/* ------------------------------------------------------------------------ */
/* Operators, matrix-free, in the reference's CSR column order.              */
/* ------------------------------------------------------------------------ */

enum { OP_A, OP_B, OP_H, OP_V, OP_DVDT };

/* Value of H at (row j*N+i, col k*N+i') following assemble_hamiltonian
 * (src/CrankNicolsonModes_hamiltonian.cpp:224-243); the CN matrices follow
 * update_CN_values (src/CrankNicolsonModes_matrices.cpp:27-40). Rows are
 * summed in increasing column order like Eigen's row-major product. */
static void apply_op(const gm_solver *s, int op, const zc *x, zc *y)
{
    const int N = s->Nr, K = s->K;
    const double alpha = 1.0 / (s->dr * s->dr);
    const double off = -alpha / (2 * s->mu);
    const zc scale = I * (s->dt * 0.5);
    const double phase_sign = s->counter ? -1.0 : 1.0;
    const int L2 = 2 * s->l_max;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < N; ++i) {
        const double *Wi = s->W + (size_t)i * K * K;
        for (int j = 0; j < K; ++j) {
            const size_t row = (size_t)j * N + i;
            zc acc = 0.0;
            if (i == 0 || i == N - 1) {
                if (op == OP_V || op == OP_DVDT) { y[row] = 0.0; continue; }
                zc h = 0.0;
                if (s->cap) h += s->capv[i];
                zc c = op == OP_A ? 1.0 + scale * h : op == OP_B ? 1.0 - scale * h : h;
                y[row] = c * x[row];
                continue;
            }
            const double h0d = -(-2.0 * alpha - s->ml[j] * (s->ml[j] + 1) / (s->r[i] * s->r[i])) / (2 * s->mu);
            for (int k = 0; k < K; ++k) {
                const int ms = s->mm[k] - s->mm[j];
                const zc v = Wi[(size_t)j * K + k] * s->phases[ms + L2];
                if (k != j) {
                    zc c;
                    switch (op) {
                    case OP_A: c = scale * v; break;
                    case OP_B: c = -(scale * v); break;
                    case OP_DVDT: c = v * phase_sign * I * (double)ms * s->Omega_V; break;
                    default: c = v;
                    }
                    acc += c * x[(size_t)k * N + i];
                    continue;
                }
                if (op == OP_V) { acc += v * x[row]; continue; }
                if (op == OP_DVDT) { acc += 0.0 * x[row]; continue; }
                zc hd = h0d;
                hd += v;
                if (s->cap) hd += s->capv[i];
                zc cm, cd, cp;
                if (op == OP_A) { cm = scale * off; cd = 1.0 + scale * hd; cp = scale * off; }
                else if (op == OP_B) { cm = -(scale * off); cd = 1.0 - scale * hd; cp = -(scale * off); }
                else { cm = off; cd = hd; cp = off; }
                acc += cm * x[row - 1];
                acc += cd * x[row];
                acc += cp * x[row + 1];
            }
            y[row] = acc;
        }
    }
}

/* Eigen-style reductions. Partial sums over fixed chunks of OM_CHUNK
 * elements, added in chunk order, so results do not depend on the thread
 * count or on the order in which threads finish. */
#define OM_CHUNK 4096

static zc vdot(const zc *a, const zc *b, size_t n) /* sum conj(a) b */
{
    const size_t nc = (n + OM_CHUNK - 1) / OM_CHUNK;
    double *part = (double *)calloc(2 * (nc ? nc : 1), sizeof(double));
#pragma omp parallel for schedule(static)
    for (size_t c = 0; c < nc; ++c) {
        double re = 0, im = 0;
        const size_t e = (c + 1) * OM_CHUNK < n ? (c + 1) * OM_CHUNK : n;
        for (size_t k = c * OM_CHUNK; k < e; ++k) {
            zc p = conj(a[k]) * b[k];
            re += creal(p);
            im += cimag(p);
        }
        part[2 * c] = re;
        part[2 * c + 1] = im;
    }
    double re = 0, im = 0;
    for (size_t c = 0; c < nc; ++c) { re += part[2 * c]; im += part[2 * c + 1]; }
    free(part);
    return re + I * im;
}

static double vnorm2(const zc *a, size_t n)
{
    return creal(vdot(a, a, n));
}

/* apply_sink (src/CrankNicolsonModes_absorption.cpp:100-109). */
static void apply_sink_internal(const gm_solver *s, const zc *x, zc *y)
{
    const size_t n = nvec(s);
    memset(y, 0, sizeof(zc) * n);
    if (!s->sink_active) return;
    const om_sink *sk[2] = {&s->sink1, &s->sink2};
    for (int a = 0; a < 2; ++a)
        if (sk[a]->strength != 0.0 && sk[a]->v) {
            const zc c = sk[a]->strength * vdot(sk[a]->v, x, n);
            for (size_t k = 0; k < n; ++k) y[k] += c * sk[a]->v[k];
        }
}

/* schrodinger_rhs (src/CrankNicolsonModes_absorption.cpp:111-117). */
static void schrodinger_rhs_internal(const gm_solver *s, const zc *x, zc *y)
{
    const size_t n = nvec(s);
    apply_op(s, OP_H, x, y);
    for (size_t k = 0; k < n; ++k) y[k] = -I * y[k];
    if (s->sink_active) {
        zc *t = (zc *)malloc(sizeof(zc) * n);
        apply_sink_internal(s, x, t);
        for (size_t k = 0; k < n; ++k) y[k] -= 0.5 * t[k];
        free(t);
    }
}

/* RadialBlockPreconditioner::factorize / _solve_impl
 * (include/RadialBlockPreconditioner.hpp:43-97). */
typedef struct { zc *inv, *lower, *upper; int N, K; } om_precond;

static int precond_factor(const gm_solver *s, om_precond *pc)
{
    const int N = s->Nr, K = s->K;
    const double alpha = 1.0 / (s->dr * s->dr);
    const double off = -alpha / (2 * s->mu);
    const zc scale = I * (s->dt * 0.5);
    pc->N = N;
    pc->K = K;
    pc->inv = (zc *)malloc(sizeof(zc) * nvec(s));
    pc->lower = (zc *)malloc(sizeof(zc) * nvec(s));
    pc->upper = (zc *)malloc(sizeof(zc) * nvec(s));
    int failed = 0;
#pragma omp parallel for schedule(static) reduction(|:failed)
    for (int j = 0; j < K; ++j) {
        const size_t st = (size_t)j * N;
        for (int i = 0; i < N; ++i) {
            const size_t row = st + i;
            const int interior = i > 0 && i < N - 1;
            /* A(row,row-1) exists for interior rows (incl. i=1 -> boundary col). */
            pc->lower[row] = (i > 0 && interior) ? scale * off : 0.0;
            pc->upper[row] = (i + 1 < N && interior) ? scale * off : 0.0;
            zc h = 0.0;
            if (interior) {
                h = -(-2.0 * alpha - s->ml[j] * (s->ml[j] + 1) / (s->r[i] * s->r[i])) / (2 * s->mu);
                h += s->W[(size_t)i * K * K + (size_t)j * K + j] * s->phases[2 * s->l_max];
            }
            if (s->cap) h += s->capv[i];
            zc pivot = 1.0 + scale * h;
            if (i > 0) pivot -= pc->lower[row] * pc->inv[row - 1] * pc->upper[row - 1];
            if (cabs(pivot) == 0.0 || !isfinite(cabs(pivot))) {
                failed = 1;
                pc->inv[row] = 0.0;
            } else
                pc->inv[row] = 1.0 / pivot;
        }
    }
    return failed;
}

static void precond_free(om_precond *pc) { free(pc->inv); free(pc->lower); free(pc->upper); }

static void precond_solve(const om_precond *pc, const zc *b, zc *x)
{
    const int N = pc->N, K = pc->K;
#pragma omp parallel for schedule(static)
    for (int j = 0; j < K; ++j) {
        const size_t st = (size_t)j * N, en = st + N;
        x[st] = b[st] * pc->inv[st];
        for (size_t row = st + 1; row < en; ++row)
            x[row] = (b[row] - pc->lower[row] * x[row - 1]) * pc->inv[row];
        for (long row = (long)en - 2; row >= (long)st; --row)
            x[row] -= pc->upper[row] * pc->inv[row] * x[row + 1];
    }
}

/* Active sinks and the CN operator with sinks (src/CrankNicolsonModes_matrices.cpp:105-116). */
typedef struct {
    const gm_solver *s;
    const om_sink *sinks[2];
    int k;
    zc *tmp;
} om_opctx;

static void apply_A_full(const om_opctx *c, const zc *x, zc *y)
{
    const size_t n = nvec(c->s);
    apply_op(c->s, OP_A, x, y);
    for (int a = 0; a < c->k; ++a) {
        const zc coef = 0.25 * c->s->dt * c->sinks[a]->strength * vdot(c->sinks[a]->v, x, n);
        for (size_t q = 0; q < n; ++q) y[q] += coef * c->sinks[a]->v[q];
    }
}

typedef struct {
    const om_precond *pc;
    const om_sink *sinks[2];
    int k;
    zc *Wcol[2];
    zc Ginv[2][2];
} om_mctx;

static void apply_M_full(const om_mctx *m, size_t n, const zc *x, zc *y)
{
    precond_solve(m->pc, x, y);
    if (m->k == 0) return;
    zc proj[2] = {0, 0}, corr[2] = {0, 0};
    for (int a = 0; a < m->k; ++a) proj[a] = vdot(m->sinks[a]->v, y, n);
    for (int a = 0; a < m->k; ++a)
        for (int b = 0; b < m->k; ++b) corr[a] += m->Ginv[a][b] * proj[b];
    /* y -= W * corr, Eigen evaluates W * (G_inverse * projections) as a product. */
    for (size_t q = 0; q < n; ++q) {
        zc t = 0.0;
        for (int a = 0; a < m->k; ++a) t += m->Wcol[a][q] * corr[a];
        y[q] -= t;
    }
}

/* Right-preconditioned BiCGSTAB. eigen_mode = 1 restates Eigen 3.4's
 * internal::bicgstab (restart threshold eps^2 |r0|^2, counter reset on the
 * first restart); eigen_mode = 0 restates the reference's own recurrence for
 * the sink path (src/CrankNicolsonModes_matrices.cpp:47-91). */
static void bicgstab(const om_opctx *A, const om_mctx *M, const zc *b, zc *x, double tol,
                     int max_iterations, int *iterations, int eigen_mode, int *restarts_out)
{
    const size_t n = nvec(A->s);
    const double b_norm2 = vnorm2(b, n);
    zc *r = (zc *)malloc(sizeof(zc) * n), *r0 = (zc *)malloc(sizeof(zc) * n);
    zc *v = (zc *)calloc(n, sizeof(zc)), *p = (zc *)calloc(n, sizeof(zc));
    zc *y = (zc *)malloc(sizeof(zc) * n), *z = (zc *)malloc(sizeof(zc) * n);
    zc *sv = (zc *)malloc(sizeof(zc) * n), *t = (zc *)malloc(sizeof(zc) * n);
    *iterations = 0;
    *restarts_out = 0;
    if (b_norm2 == 0.0) {
        memset(x, 0, sizeof(zc) * n);
        goto done;
    }
    {
        const double threshold = tol * tol * b_norm2;
        apply_A_full(A, x, r);
        for (size_t q = 0; q < n; ++q) r[q] = b[q] - r[q];
        memcpy(r0, r, sizeof(zc) * n);
        double r0_norm2 = vnorm2(r0, n);
        /* restart thresholds: 1e-32 |r0|^2 (matrices.cpp:56) and Eigen's
         * eps^2 |r0|^2, both settable for tests (om_set_restart_thresholds) */
        const double fixed_restart = r0_norm2 * A->s->rst_ref;
        const double eps2 = A->s->rst_eps2;
        zc rho = 1.0, alpha = 1.0, w = 1.0;
        int i = 0, restarts = 0;
        while (vnorm2(r, n) > threshold && i < max_iterations) {
            const zc rho_old = rho;
            rho = vdot(r0, r, n);
            const double restart_threshold = eigen_mode ? eps2 * r0_norm2 : fixed_restart;
            if (cabs(rho) < restart_threshold) {
                apply_A_full(A, x, r);
                for (size_t q = 0; q < n; ++q) r[q] = b[q] - r[q];
                memcpy(r0, r, sizeof(zc) * n);
                rho = r0_norm2 = vnorm2(r, n);
                if (restarts++ == 0 && eigen_mode) i = 0;
            }
            const zc beta = (rho / rho_old) * (alpha / w);
#pragma omp parallel for schedule(static)
            for (size_t q = 0; q < n; ++q) p[q] = r[q] + beta * (p[q] - w * v[q]);
            apply_M_full(M, n, p, y);
            apply_A_full(A, y, v);
            alpha = rho / vdot(r0, v, n);
#pragma omp parallel for schedule(static)
            for (size_t q = 0; q < n; ++q) sv[q] = r[q] - alpha * v[q];
            apply_M_full(M, n, sv, z);
            apply_A_full(A, z, t);
            const double t_norm2 = vnorm2(t, n);
            w = t_norm2 > 0.0 ? vdot(t, sv, n) / t_norm2 : 0.0;
#pragma omp parallel for schedule(static)
            for (size_t q = 0; q < n; ++q) {
                x[q] += alpha * y[q] + w * z[q];
                r[q] = sv[q] - w * t[q];
            }
            ++i;
        }
        *iterations = i;
        *restarts_out = restarts;
    }
done:
    free(r); free(r0); free(v); free(p); free(y); free(z); free(sv); free(t);
}

static zc matrix_inverse_2x2(zc G[2][2], zc out[2][2])
{
    const zc det = G[0][0] * G[1][1] - G[0][1] * G[1][0];
    const zc inv = 1.0 / det;
    out[0][0] = G[1][1] * inv;
    out[0][1] = -G[0][1] * inv;
    out[1][0] = -G[1][0] * inv;
    out[1][1] = G[0][0] * inv;
    return det;
}

/* solve_CN (src/CrankNicolsonModes_matrices.cpp:94-215). The full-ILUT
 * fallback is not restated; a failed radial solve is reported. */
static int solve_CN(gm_solver *s, const zc *rhs)
{
    const size_t n = nvec(s);
    const double rhs_norm = sqrt(vnorm2(rhs, n));
    if (rhs_norm == 0.0) {
        memset(s->psi, 0, sizeof(zc) * n);
        s->last_residual = 0.0;
        s->last_iterations = 0;
        return GM_OK;
    }
    om_opctx A = {s, {NULL, NULL}, 0, NULL};
    if (s->sink_active) {
        const om_sink *sk[2] = {&s->sink1, &s->sink2};
        for (int a = 0; a < 2; ++a)
            if (sk[a]->strength != 0.0 && sk[a]->v && vnorm2(sk[a]->v, n) > 0.0) A.sinks[A.k++] = sk[a];
    }
    if (!s->flags.use_block_preconditioner)
        return fail(s, GM_E_RUNTIME, "ILUT preconditioner is not available in this build");
    zc *b = (zc *)malloc(sizeof(zc) * n), *x = (zc *)malloc(sizeof(zc) * n);
    zc *res = (zc *)malloc(sizeof(zc) * n);
    for (size_t q = 0; q < n; ++q) { b[q] = rhs[q] / rhs_norm; x[q] = s->psi[q] / rhs_norm; }
    om_precond pc;
    const int failed = precond_factor(s, &pc);
    om_mctx M = {&pc, {NULL, NULL}, 0, {NULL, NULL}, {{0, 0}, {0, 0}}};
    int status = GM_OK;
    if (failed) {
        status = fail(s, GM_E_RUNTIME, "Radial preconditioner did not converge; ILUT fallback is not available");
        goto out;
    }
    if (A.k > 0) {
        M.k = A.k;
        zc G[2][2] = {{0, 0}, {0, 0}};
        for (int a = 0; a < A.k; ++a) {
            M.sinks[a] = A.sinks[a];
            M.Wcol[a] = (zc *)malloc(sizeof(zc) * n);
            precond_solve(&pc, A.sinks[a]->v, M.Wcol[a]);
        }
        for (int a = 0; a < A.k; ++a) {
            G[a][a] += 1.0 / (0.25 * s->dt * A.sinks[a]->strength);
            for (int c = 0; c < A.k; ++c) G[a][c] += vdot(A.sinks[a]->v, M.Wcol[c], n);
        }
        if (A.k == 1) M.Ginv[0][0] = 1.0 / G[0][0];
        else matrix_inverse_2x2(G, M.Ginv);
        bicgstab(&A, &M, b, x, 1e-10, 500, &s->last_iterations, 0, &s->last_restarts);
    } else {
        bicgstab(&A, &M, b, x, 1e-10, 500, &s->last_iterations, 1, &s->last_restarts);
    }
    for (size_t q = 0; q < n; ++q) x[q] *= rhs_norm;
    {
        int finite = 1;
        for (size_t q = 0; q < n; ++q)
            if (!isfinite(creal(x[q])) || !isfinite(cimag(x[q]))) { finite = 0; break; }
        if (!finite) s->last_residual = INFINITY;
        else {
            apply_A_full(&A, x, res);
            for (size_t q = 0; q < n; ++q) res[q] -= rhs[q];
            s->last_residual = sqrt(vnorm2(res, n)) / rhs_norm;
        }
    }
    if (s->last_residual <= 1e-9) memcpy(s->psi, x, sizeof(zc) * n);
    else status = fail(s, GM_E_RUNTIME, "Crank-Nicolson solve failed the relative residual check");
out:
    for (int a = 0; a < M.k; ++a) free(M.Wcol[a]);
    precond_free(&pc);
    free(b); free(x); free(res);
    return status;
}

static double absorption_rate(const gm_solver *s, const om_sink *sk, const zc *state)
{
    if (sk->strength == 0.0 || !sk->v) return 0.0;
    const zc d = vdot(sk->v, state, nvec(s));
    return s->mu * s->dr * sk->strength * (creal(d) * creal(d) + cimag(d) * cimag(d));
}

/* update_time_dependent_increment (src/CrankNicolsonModes_evolution.cpp:17-72). */
static void update_time_dependent_increment(gm_solver *s, double orbit_dt)
{
    const size_t n = nvec(s);
    zc *mid = (zc *)malloc(sizeof(zc) * n);
    for (size_t q = 0; q < n; ++q) { mid[q] = s->psi[q] + s->psi_last[q]; mid[q] *= 0.5; }
    double Mdot1 = 0.0, Mdot2 = 0.0;
    if (s->absorption) {
        Mdot1 = absorption_rate(s, &s->sink1, mid);
        Mdot2 = absorption_rate(s, &s->sink2, mid);
    }
    free(mid);
    double ramp = 1.0;
    if (s->t_i < ramp_end_time(s)) ramp = 0.0;
    else if (s->t_i < ramp_end_time(s) + s->initial_period) {
        const double x = (s->t_i - ramp_end_time(s)) / s->initial_period;
        ramp = 3.0 * x * x - 2.0 * x * x * x;
    }
    const double dR_dt_gw = -(64.0 / 5.0) * s->M1 * s->M2 * (s->M1 + s->M2) / pow(s->R, 3);
    const double dR_dt_cloud = -2.0 * s->R * s->R * s->p_bin / (s->M1 * s->M2);
    const double dR_dt = (dR_dt_gw + dR_dt_cloud) * ramp;
    s->R += dR_dt * orbit_dt;
    s->M1 += Mdot1 * orbit_dt;
    s->M2 += Mdot2 * orbit_dt;
    s->q = s->M2 / s->M1;
    const double Mt = s->M1 + s->M2;
    s->R1 = s->M2 / Mt * s->R;
    s->R2 = s->M1 / Mt * s->R;
    s->Omega = sqrt(Mt / pow(s->R, 3));
    s->P = 2.0 * M_PI / s->Omega;
    if (s->ramp_factor == 1.0) s->dt = s->P / s->flags.steps_per_period;
}

/* set_extraction_radius and the compute part of run() (src/CrankNicolsonModes.cpp:236-270). */
int om_begin_run(gm_solver *s, double r_ex, int32_t n_steps)
{
    s->r_ex = r_ex;
    s->idx_ex = (int)floor(r_ex / s->dr);
    if (n_steps < 1 || s->idx_ex < 1 || s->idx_ex >= s->Nr - 1)
        return fail(s, GM_E_INVALID, "Require at least one time sample and an interior extraction radius");
    s->t_i = 0.0;
    build_time_dependent_perturbation(s, s->t_i);
    return GM_OK;
}

/* step (src/CrankNicolsonModes.cpp:145-234) and the diagnostics of
 * src/CrankNicolsonModes_extraction.cpp:8-129. */
int om_step(gm_solver *s, int32_t step_id, gm_diag *d)
{
    (void)step_id;
    const int N = s->Nr, K = s->K;
    const size_t n = nvec(s);
    memcpy(s->psi_last, s->psi, sizeof(zc) * n);
    const double step_start = s->t_i;
    s->t_i += s->dt;
    s->theta += (s->counter ? -1.0 : 1.0) * s->Omega * s->dt;
    s->phase += s->Omega * s->dt;
    build_time_dependent_perturbation(s, s->t_i);
    zc *rhs = (zc *)malloc(sizeof(zc) * n);
    zc *tmp = (zc *)malloc(sizeof(zc) * n);
    apply_op(s, OP_B, s->psi, rhs);
    if (s->sink_active) {
        apply_sink_internal(s, s->psi, tmp);
        for (size_t q = 0; q < n; ++q) rhs[q] -= (0.25 * s->dt) * tmp[q];
    }
    int st = solve_CN(s, rhs);
    if (st != GM_OK) { free(rhs); free(tmp); return st; }
    double absorbed = 0.0;
    if (s->sink_active) {
        for (size_t q = 0; q < n; ++q) rhs[q] = 0.5 * (s->psi_last[q] + s->psi[q]);
        apply_sink_internal(s, rhs, tmp);
        absorbed = s->mu * s->dr * s->dt * creal(vdot(rhs, tmp, n));
    }
    if (s->mask)
        for (int j = 0; j < K; ++j)
            for (int i = 0; i < N; ++i) s->psi[(size_t)j * N + i] *= s->damping[i];
    zc *dpsi = (zc *)malloc(sizeof(zc) * n);
    schrodinger_rhs_internal(s, s->psi, dpsi);
    if (!s->isolated && s->flags.evolve_orbit && s->t_i > ramp_end_time(s)) {
        const double start = step_start > ramp_end_time(s) ? step_start : ramp_end_time(s);
        update_time_dependent_increment(s, s->t_i - start);
    }
    const int ix = s->idx_ex;
    if (d) {
        for (int j = 0; j < K; ++j) {
            const zc *u = s->psi + (size_t)j * N;
            const zc dU = (u[ix + 1] - u[ix - 1]) / (2.0 * s->dr) - u[ix] / s->r[ix];
            if (d->flux) d->flux[j] = cimag(conj(u[ix]) * dU);
            if (d->energy_flux) d->energy_flux[j] = creal(conj(dpsi[(size_t)j * N + ix]) * dU) / s->mu;
        }
    }
    /* energy and energies per mode: H psi */
    apply_op(s, OP_H, s->psi, tmp);
    zc e_all = 0.0;
    for (int j = 0; j < K; ++j)
        for (int i = 0; i <= ix; ++i) e_all += conj(s->psi[(size_t)j * N + i]) * tmp[(size_t)j * N + i];
    if (d && d->energies)
        for (int j = 0; j < K; ++j) {
            zc e = 0.0;
            for (int i = 0; i < ix; ++i) e += conj(s->psi[(size_t)j * N + i]) * tmp[(size_t)j * N + i];
            d->energies[j] = creal(e) * s->dr;
        }
    if (d && d->norms)
        for (int j = 0; j < K; ++j) {
            double acc = 0.0;
            for (int i = 0; i < ix; ++i) {
                const zc u = s->psi[(size_t)j * N + i];
                acc += creal(u) * creal(u) + cimag(u) * cimag(u);
            }
            d->norms[j] = acc * s->dr;
        }
    /* binary power: dVdt expectation */
    apply_op(s, OP_DVDT, s->psi, tmp);
    zc pb = 0.0;
    for (int j = 0; j < K; ++j)
        for (int i = 0; i <= ix; ++i) pb += conj(s->psi[(size_t)j * N + i]) * tmp[(size_t)j * N + i];
    const double p_bin_last = creal(pb) * s->dr;
    double avg = 0.0;
    if (!s->isolated) {
        if (rolling_update(&s->avg, s->t_i, s->phase, p_bin_last, &avg) != 0) {
            free(rhs); free(tmp); free(dpsi);
            return fail(s, GM_E_RUNTIME, "RollingPhaseAverage: time and phase must increase");
        }
    }
    s->p_bin = s->isolated ? 0.0 : avg;
    apply_op(s, OP_V, s->psi, tmp);
    zc dv = 0.0;
    for (int j = 0; j < K; ++j)
        for (int i = 0; i <= ix; ++i) dv += conj(dpsi[(size_t)j * N + i]) * tmp[(size_t)j * N + i];
    if (d) {
        d->energy = creal(e_all) * s->dr;
        d->p_bin = s->p_bin;
        d->p_bin_inst = p_bin_last;
        d->dV_bin_dt = creal(dv) * s->dr;
        d->cloud_mass = s->mu * s->dr * vnorm2(s->psi, n);
        d->absorbed = absorbed;
    }
    free(rhs); free(tmp); free(dpsi);
    return GM_OK;
}

int om_get_scalars(const gm_solver *s, gm_scalars *o)
{
    o->t = s->t_i; o->theta = s->theta; o->phase = s->phase; o->R = s->R; o->M1 = s->M1; o->M2 = s->M2;
    o->q = s->q; o->R1 = s->R1; o->R2 = s->R2; o->Omega = s->Omega; o->P = s->P; o->dt = s->dt;
    o->p_bin = s->p_bin; o->Omega_0 = s->Omega_0;
    o->last_solver_residual = s->last_residual;
    o->last_solver_iterations = s->last_iterations;
    o->sink_active = s->sink_active;
    return GM_OK;
}

int om_set_scalars(gm_solver *s, const gm_scalars *o)
{
    s->t_i = o->t; s->theta = o->theta; s->phase = o->phase; s->R = o->R; s->M1 = o->M1; s->M2 = o->M2;
    s->q = o->q; s->R1 = o->R1; s->R2 = o->R2; s->Omega = o->Omega; s->P = o->P; s->dt = o->dt;
    s->p_bin = o->p_bin;
    return GM_OK;
}

static void unpack(const double *in, zc *out, size_t n)
{
    for (size_t k = 0; k < n; ++k) out[k] = in[2 * k] + I * in[2 * k + 1];
}

static void pack(const zc *in, double *out, size_t n)
{
    for (size_t k = 0; k < n; ++k) { out[2 * k] = creal(in[k]); out[2 * k + 1] = cimag(in[k]); }
}

int om_apply_H(gm_solver *s, const double *in, double *out)
{
    const size_t n = nvec(s);
    zc *x = (zc *)malloc(sizeof(zc) * n), *y = (zc *)malloc(sizeof(zc) * n);
    unpack(in, x, n);
    apply_op(s, OP_H, x, y);
    pack(y, out, n);
    free(x); free(y);
    return GM_OK;
}

/* Test probe, no counterpart in the reference: evaluate only the multipoles
 * l* <= lcut[i / tile] at node i (a truncated coupling given from outside, to
 * check a truncating implementation term for term); lcut = NULL restores every
 * term. The coupling is rebuilt at once and at every later rebuild. */
int om_set_multipole_cut(gm_solver *s, const int32_t *lcut, int32_t tile)
{
    free(s->lcut);
    s->lcut = NULL;
    if (lcut) {
        if (tile <= 0) return GM_E_INVALID;
        s->lcut = (int32_t *)malloc(sizeof(int32_t) * (size_t)s->Nr);
        for (int i = 0; i < s->Nr; ++i) s->lcut[i] = lcut[i / tile];
    }
    if (s->f1) build_static_perturbation(s);
    return GM_OK;
}

/* Extra operator probes used by the tests: V, dVdt, A, B. */
int om_apply_op(gm_solver *s, int32_t op, const double *in, double *out)
{
    const size_t n = nvec(s);
    zc *x = (zc *)malloc(sizeof(zc) * n), *y = (zc *)malloc(sizeof(zc) * n);
    unpack(in, x, n);
    apply_op(s, op, x, y);
    pack(y, out, n);
    free(x); free(y);
    return GM_OK;
}

int om_apply_sink(gm_solver *s, const double *in, double *out)
{
    const size_t n = nvec(s);
    zc *x = (zc *)malloc(sizeof(zc) * n), *y = (zc *)malloc(sizeof(zc) * n);
    unpack(in, x, n);
    apply_sink_internal(s, x, y);
    pack(y, out, n);
    free(x); free(y);
    return GM_OK;
}

int om_schrodinger_rhs(gm_solver *s, const double *in, double *out)
{
    const size_t n = nvec(s);
    zc *x = (zc *)malloc(sizeof(zc) * n), *y = (zc *)malloc(sizeof(zc) * n);
    unpack(in, x, n);
    schrodinger_rhs_internal(s, x, y);
    pack(y, out, n);
    free(x); free(y);
    return GM_OK;
}

/* sink_strength (src/CrankNicolsonModes_absorption.cpp:119-125). */
double om_sink_strength(const gm_solver *s, int32_t body)
{
    const om_sink *sk = body == 1 ? &s->sink1 : &s->sink2;
    if (!s->sink_active || sk->strength == 0.0 || !sk->v) return 0.0;
    return sk->strength * vnorm2(sk->v, nvec(s));
}

/* Test probes: the sink projector v_a and the Gaunt cache. */
int om_sink_vector(const gm_solver *s, int32_t body, double *out)
{
    const om_sink *sk = body == 1 ? &s->sink1 : &s->sink2;
    const size_t n = nvec(s);
    if (!sk->v) { memset(out, 0, sizeof(double) * 2 * n); return GM_OK; }
    pack(sk->v, out, n);
    return GM_OK;
}

int om_coupling_terms(const gm_solver *s, int32_t *rows, int32_t *cols, int32_t *ls, double *ang, int32_t cap)
{
    int c = 0;
    for (int p = 0; p < s->npairs; ++p)
        for (int t = s->pairs[p].t0; t < s->pairs[p].t0 + s->pairs[p].nt; ++t) {
            if (c < cap && rows) {
                rows[c] = s->pairs[p].row;
                cols[c] = s->pairs[p].col;
                ls[c] = s->term_l[t];
                ang[c] = s->term_ang[t];
            }
            c++;
        }
    return c;
}

/* Thomas solve probe on an explicit tridiagonal (tests/parallel_correctness.cpp:16-47). */
int om_thomas_probe(int32_t N, int32_t K, const double *diag, const double *lower, const double *upper,
                    const double *rhs, double *x)
{
    const size_t n = (size_t)N * K;
    om_precond pc;
    pc.N = N;
    pc.K = K;
    pc.inv = (zc *)malloc(sizeof(zc) * n);
    pc.lower = (zc *)malloc(sizeof(zc) * n);
    pc.upper = (zc *)malloc(sizeof(zc) * n);
    int failed = 0;
    for (int j = 0; j < K; ++j)
        for (int i = 0; i < N; ++i) {
            const size_t row = (size_t)j * N + i;
            pc.lower[row] = i > 0 ? lower[2 * row] + I * lower[2 * row + 1] : 0.0;
            pc.upper[row] = i + 1 < N ? upper[2 * row] + I * upper[2 * row + 1] : 0.0;
            zc pivot = diag[2 * row] + I * diag[2 * row + 1];
            if (i > 0) pivot -= pc.lower[row] * pc.inv[row - 1] * pc.upper[row - 1];
            if (cabs(pivot) == 0.0 || !isfinite(cabs(pivot))) { failed = 1; pc.inv[row] = 0.0; }
            else pc.inv[row] = 1.0 / pivot;
        }
    zc *b = (zc *)malloc(sizeof(zc) * n), *y = (zc *)malloc(sizeof(zc) * n);
    unpack(rhs, b, n);
    precond_solve(&pc, b, y);
    pack(y, x, n);
    free(b); free(y);
    precond_free(&pc);
    return failed ? GM_E_RUNTIME : GM_OK;
}

/* ------------------------------------------------------------------------ */
/* Checkpoints and resume (src/CrankNicolsonModes_restart.cpp:139-285).      */
/* ------------------------------------------------------------------------ */

static const char om_magic[8] = {'G', 'M', 'C', 'K', 'P', 'T', '0', '1'};

/* RollingPhaseAverage::save / load (src/RollingPhase.cpp:169-209): window
 * phase, previous (time, phase, value), has_previous and ready as one byte
 * each, initial phase, integral, duration, average, then the uint64 segment
 * count and the segments as 7 doubles each. */
static int rolling_save(const om_rolling *a, FILE *f)
{
    const double prev[3] = {a->prev_time, a->prev_phase, a->prev_value};
    const unsigned char flags[2] = {(unsigned char)(a->has_previous != 0), (unsigned char)(a->ready != 0)};
    const double tail[4] = {a->initial_phase, a->integral, a->duration, a->average};
    const uint64_t n = a->count;
    int ok = fwrite(&a->window_phase, 8, 1, f) == 1 && fwrite(prev, 8, 3, f) == 3 && fwrite(flags, 1, 2, f) == 2 &&
             fwrite(tail, 8, 4, f) == 4 && fwrite(&n, 8, 1, f) == 1;
    if (ok && n) ok = fwrite(a->seg + a->head, sizeof(om_segment), n, f) == n;
    return ok;
}

static int rolling_load(om_rolling *a, FILE *f)
{
    double prev[3], tail[4];
    unsigned char flags[2];
    uint64_t n = 0;
    if (fread(&a->window_phase, 8, 1, f) != 1 || fread(prev, 8, 3, f) != 3 || fread(flags, 1, 2, f) != 2 ||
        fread(tail, 8, 4, f) != 4 || fread(&n, 8, 1, f) != 1)
        return 0;
    a->prev_time = prev[0]; a->prev_phase = prev[1]; a->prev_value = prev[2];
    a->has_previous = flags[0]; a->ready = flags[1];
    a->initial_phase = tail[0]; a->integral = tail[1]; a->duration = tail[2]; a->average = tail[3];
    free(a->seg);
    a->seg = NULL;
    a->head = a->count = a->cap = 0;
    if (n > ((uint64_t)1 << 32)) return 0;
    if (n) {
        a->seg = (om_segment *)malloc(sizeof(om_segment) * n);
        a->cap = a->count = n;
        if (fread(a->seg, sizeof(om_segment), n, f) != n) return 0;
    }
    return 1;
}

int om_save_checkpoint(gm_solver *s, const char *path, int32_t step_id)
{
    if (!s->initialized) return fail(s, GM_E_LOGIC, "initialize() has not been called");
    if (!path) return fail(s, GM_E_INVALID, "checkpoint path is NULL");
    char tmp[4096];
    snprintf(tmp, sizeof(tmp), "%s.tmp", path);
    FILE *f = fopen(tmp, "wb");
    if (!f) return fail(s, GM_E_RUNTIME, "Could not write checkpoint");
    const int32_t header[3] = {step_id, s->Nr, s->K};
    const double v[13] = {s->t_i, s->theta, s->phase, s->R, s->M1, s->M2, s->q,
                          s->R1, s->R2, s->Omega, s->P, s->dt, s->p_bin};
    const uint64_t n = nvec(s);
    int ok = fwrite(om_magic, 1, 8, f) == 8 && fwrite(header, 4, 3, f) == 3 && fwrite(v, 8, 13, f) == 13 &&
             fwrite(&n, 8, 1, f) == 1 && fwrite(s->psi, sizeof(zc), n, f) == n && rolling_save(&s->avg, f);
    ok = (fclose(f) == 0) && ok;
    if (!ok) return fail(s, GM_E_RUNTIME, "Could not write checkpoint");
    if (rename(tmp, path) != 0) return fail(s, GM_E_RUNTIME, "Could not replace checkpoint");
    return GM_OK;
}

int om_load_checkpoint(gm_solver *s, const char *path, int32_t step_id, int32_t *loaded)
{
    if (!s->initialized) return fail(s, GM_E_LOGIC, "initialize() has not been called");
    if (!path || !loaded) return fail(s, GM_E_INVALID, "checkpoint path or result is NULL");
    *loaded = 0;
    FILE *f = fopen(path, "rb");
    if (!f) return GM_OK;
    char magic[8];
    int32_t header[3] = {-1, -1, -1};
    double v[13];
    uint64_t n = 0;
    int ok = fread(magic, 1, 8, f) == 8 && fread(header, 4, 3, f) == 3 && memcmp(magic, om_magic, 8) == 0 &&
             header[0] == step_id && header[1] == s->Nr && header[2] == s->K && fread(v, 8, 13, f) == 13 &&
             fread(&n, 8, 1, f) == 1 && n == nvec(s);
    zc *psi = ok ? (zc *)malloc(sizeof(zc) * n) : NULL;
    ok = ok && fread(psi, sizeof(zc), n, f) == n;
    om_rolling avg;
    memset(&avg, 0, sizeof(avg));
    ok = ok && rolling_load(&avg, f);
    fclose(f);
    if (!ok) {
        free(psi);
        free(avg.seg);
        return GM_OK;
    }
    s->t_i = v[0]; s->theta = v[1]; s->phase = v[2]; s->R = v[3]; s->M1 = v[4]; s->M2 = v[5]; s->q = v[6];
    s->R1 = v[7]; s->R2 = v[8]; s->Omega = v[9]; s->P = v[10]; s->dt = v[11]; s->p_bin = v[12];
    memcpy(s->psi, psi, sizeof(zc) * n);
    free(psi);
    free(s->avg.seg);
    s->avg = avg;
    *loaded = 1;
    return GM_OK;
}

/* reconstruct_state_from_diagnostics (src/CrankNicolsonModes_restart.cpp:238-285).
 * rows holds n_rows rows of 8 doubles: t, R, theta, Omega, M1, q, p_bin,
 * p_bin_inst (row k is the one step k+1 wrote). */
int om_restore_from_diagnostics(gm_solver *s, int32_t n_rows, const double *rows)
{
    if (!s->initialized) return fail(s, GM_E_LOGIC, "initialize() has not been called");
    if (n_rows < 0 || (n_rows > 0 && !rows)) return fail(s, GM_E_INVALID, "diagnostic rows are missing");
    double Omega_previous = s->isolated ? 0.0 : (ramp_end_time(s) > 0.0 ? 0.0 : s->Omega_0);
    double t_previous = 0.0, out = 0.0;
    const double window = s->avg.window_phase;
    s->phase = 0.0;
    rolling_init(&s->avg, 1.0);
    s->avg.window_phase = window;
    for (int k = 0; k < n_rows; ++k) {
        const double *row = rows + 8 * (size_t)k;
        s->phase += Omega_previous * (row[0] - t_previous);
        if (!s->isolated && rolling_update(&s->avg, row[0], s->phase, row[7], &out) != 0)
            return fail(s, GM_E_RUNTIME, "RollingPhaseAverage: time and phase must increase");
        Omega_previous = row[3];
        t_previous = row[0];
    }
    if (n_rows == 0) return GM_OK;
    const double *last = rows + 8 * (size_t)(n_rows - 1);
    s->t_i = last[0];
    s->R = last[1];
    s->theta = last[2];
    s->M1 = last[4];
    s->q = last[5];
    s->M2 = s->q * s->M1;
    s->R1 = s->M2 / (s->M1 + s->M2) * s->R;
    s->R2 = s->M1 / (s->M1 + s->M2) * s->R;
    s->Omega = last[3];
    s->p_bin = last[6];
    if (!s->isolated && s->flags.evolve_orbit && s->t_i > ramp_end_time(s)) {
        s->P = 2.0 * M_PI / s->Omega;
        s->dt = s->P / s->flags.steps_per_period;
    }
    return GM_OK;
}

/* BiCGSTAB restart thresholds, as fractions of |r0|^2 (Eigen's eps^2 when
 * there are no sinks, the reference's 1e-32 with them), so that tests can
 * force the restart branch (src/CrankNicolsonModes_matrices.cpp:67-74); and
 * the restarts of the last solve. */
int om_set_restart_thresholds(gm_solver *s, double eigen_eps2, double sink_factor)
{
    if (!(eigen_eps2 > 0.0) || !(sink_factor > 0.0)) return fail(s, GM_E_INVALID, "restart thresholds must be positive");
    s->rst_eps2 = eigen_eps2;
    s->rst_ref = sink_factor;
    return GM_OK;
}

int om_restart_count(const gm_solver *s, int32_t *n)
{
    *n = s->last_restarts;
    return GM_OK;
}

/* Rolling average probe (tests/rolling_average.cpp). A window of n_orbits <= 0
 * is refused (NULL), as the constructor's std::invalid_argument
 * (src/RollingPhase.cpp:14-22). */
void *om_rolling_new(double n_orbits)
{
    if (!(n_orbits > 0.0)) return NULL;
    om_rolling *a = (om_rolling *)calloc(1, sizeof(om_rolling));
    rolling_init(a, n_orbits);
    return a;
}

/* ready() and value() (src/RollingPhase.cpp:129-137). */
int om_rolling_ready(void *h) { return ((om_rolling *)h)->ready; }

double om_rolling_value(void *h) { return ((om_rolling *)h)->average; }

/* reset() (src/RollingPhase.cpp:139-150): samples dropped, window kept. */
void om_rolling_reset(void *h)
{
    om_rolling *a = (om_rolling *)h;
    const double window = a->window_phase;
    rolling_init(a, 1.0);
    a->window_phase = window;
}

/* save() / load() (src/RollingPhase.cpp:167-209) through a file: 0 on
 * success, -1 when the file cannot be written or read whole (load's
 * "truncated checkpoint"). */
int om_rolling_save(void *h, const char *path)
{
    FILE *f = fopen(path, "wb");
    if (!f) return -1;
    const int ok = rolling_save((om_rolling *)h, f);
    return (fclose(f) == 0 && ok) ? 0 : -1;
}

int om_rolling_load(void *h, const char *path)
{
    FILE *f = fopen(path, "rb");
    if (!f) return -1;
    om_rolling tmp;
    memset(&tmp, 0, sizeof(tmp));
    const int ok = rolling_load(&tmp, f);
    fclose(f);
    if (!ok) {
        free(tmp.seg);
        return -1;
    }
    om_rolling *a = (om_rolling *)h;
    free(a->seg);
    *a = tmp;
    return 0;
}

int om_rolling_update(void *h, double time, double phase, double value, double *out)
{
    return rolling_update((om_rolling *)h, time, phase, value, out);
}

void om_rolling_free(void *h)
{
    om_rolling *a = (om_rolling *)h;
    free(a->seg);
    free(a);
}
