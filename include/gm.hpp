// gm.hpp - C++17 class over the C ABI of gm.h with the member surface of the
// reference class CrankNicolsonModes (include/CrankNicolsonModes.hpp:60-455).
//
// Header-only; link with libgm.so. The state is exchanged as
// std::vector<std::complex<double>> in the mode-major layout of the reference
// (element j*Nr + i is mode j at node i). Failures rethrow the reference's
// exception types with the library's message:
//   GM_E_INVALID -> std::invalid_argument   (src/CrankNicolsonModes.cpp:18-19)
//   GM_E_LOGIC   -> std::logic_error        (src/CrankNicolsonModes.cpp:62-63)
//   GM_E_RUNTIME -> std::runtime_error      (src/CrankNicolsonModes_matrices.cpp:183,213)
//   GM_E_DEVICE  -> std::runtime_error
#pragma once

#include <complex>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "gm.h"

namespace gm {

using cvec = std::vector<std::complex<double>>;

inline void check(int status, const gm_solver *h)
{
    if (status == GM_OK) return;
    const std::string msg = gm_last_error(h);
    switch (status) {
    case GM_E_INVALID: throw std::invalid_argument(msg);
    case GM_E_LOGIC: throw std::logic_error(msg);
    default: throw std::runtime_error(msg);
    }
}

// Per-step diagnostics (src/CrankNicolsonModes.cpp:197-226).
struct StepDiagnostics {
    double energy = 0, p_bin = 0, p_bin_inst = 0, dV_bin_dt = 0, cloud_mass = 0, absorbed = 0;
    std::vector<double> flux, energy_flux, norms, energies;
};

class CrankNicolsonModes {
public:
    // include/CrankNicolsonModes.hpp:84 (config_file is the caller's business:
    // this class writes no files)
    CrankNicolsonModes(int Nx, double L, double dt, double epsilon, int l_max, double mu, double M_c, double q,
                       double a, double R, bool counter, bool absorption, int save_every, double ramp_periods,
                       double epsilon_smear)
    {
        gm_params p{};
        p.Nr = Nx;
        p.L = L;
        p.dt = dt;
        p.epsilon = epsilon;
        p.l_max = l_max;
        p.mu = mu;
        p.M_c = M_c;
        p.q = q;
        p.a = a;
        p.R = R;
        p.counter = counter;
        p.absorption = absorption;
        p.save_every = save_every;
        p.ramp_periods = ramp_periods;
        p.epsilon_smear = epsilon_smear;
        check(gm_create(&p, &h_), nullptr);
        n_modes = gm_n_modes(h_);
        Nr = Nx;
        lm_.resize(2 * static_cast<size_t>(n_modes));
        check(gm_modes(h_, lm_.data()), h_);
        r.resize(Nx);
        check(gm_grid(h_, r.data()), h_);
    }
    ~CrankNicolsonModes() { gm_destroy(h_); }
    CrankNicolsonModes(const CrankNicolsonModes &) = delete;
    CrankNicolsonModes &operator=(const CrankNicolsonModes &) = delete;
    CrankNicolsonModes(CrankNicolsonModes &&o) noexcept : h_(std::exchange(o.h_, nullptr)) {}

    int Nr = 0, n_modes = 0;
    std::vector<double> r;

    // modes (include/CrankNicolsonModes.hpp:174): (l, m) of mode j
    std::pair<int, int> mode(int j) const { return {lm_[2 * j], lm_[2 * j + 1]}; }

    // public flags (include/CrankNicolsonModes.hpp:137-156)
    gm_flags flags() const
    {
        gm_flags f{};
        check(gm_get_flags(h_, &f), h_);
        return f;
    }
    void set_flags(const gm_flags &f) { check(gm_set_flags(h_, &f), h_); }
    void set_power_average_orbits(double orbits)  // :153
    {
        gm_flags f = flags();
        f.power_average_orbits = orbits;
        set_flags(f);
    }

    void configure_isolated_black_hole(double x) { check(gm_configure_isolated_black_hole(h_, x), h_); }  // :193
    void set_initial_state_full(const cvec &psi)                                                          // :171
    {
        check(gm_set_initial_state_full(h_, reinterpret_cast<const double *>(psi.data()),
                                        static_cast<int64_t>(psi.size())),
              h_);
    }
    void initialize() { check(gm_initialize(h_), h_); }  // :183
    void build_damping_mask(double L_factor, double W, int n) { check(gm_build_damping_mask(h_, L_factor, W, n), h_); }
    void build_cap(double L_factor, double W, int n) { check(gm_build_cap(h_, L_factor, W, n), h_); }

    // run() without its files: extraction radius, t = 0, H(0)
    void begin_run(double r_ex, int n_steps) { check(gm_begin_run(h_, r_ex, n_steps), h_); }

    // step (src/CrankNicolsonModes.cpp:145-234); diagnostics returned instead of written
    StepDiagnostics step(int step_id)
    {
        StepDiagnostics d;
        d.flux.resize(n_modes);
        d.energy_flux.resize(n_modes);
        d.norms.resize(n_modes);
        d.energies.resize(n_modes);
        gm_diag g{};
        g.flux = d.flux.data();
        g.energy_flux = d.energy_flux.data();
        g.norms = d.norms.data();
        g.energies = d.energies.data();
        check(gm_step(h_, step_id, &g), h_);
        d.energy = g.energy;
        d.p_bin = g.p_bin;
        d.p_bin_inst = g.p_bin_inst;
        d.dV_bin_dt = g.dV_bin_dt;
        d.cloud_mass = g.cloud_mass;
        d.absorbed = g.absorbed;
        return d;
    }

    cvec psi() const  // :162
    {
        cvec out(static_cast<size_t>(Nr) * n_modes);
        check(gm_get_state(h_, reinterpret_cast<double *>(out.data())), h_);
        return out;
    }
    void set_psi(const cvec &v) { check(gm_set_state(h_, reinterpret_cast<const double *>(v.data())), h_); }

    gm_scalars scalars() const
    {
        gm_scalars s{};
        check(gm_get_scalars(h_, &s), h_);
        return s;
    }
    void set_scalars(const gm_scalars &s) { check(gm_set_scalars(h_, &s), h_); }

    cvec apply_H(const cvec &x) const { return op(gm_apply_H, x); }                  // :197
    cvec apply_sink(const cvec &x) const { return op(gm_apply_sink, x); }            // :210
    cvec schrodinger_rhs(const cvec &x) const { return op(gm_schrodinger_rhs, x); }  // :213
    double sink_strength(int body) const { return gm_sink_strength(h_, body); }      // :216

    gm_solver *handle() const { return h_; }

private:
    template <class F>
    cvec op(F f, const cvec &x) const
    {
        if (x.size() != static_cast<size_t>(Nr) * n_modes) throw std::invalid_argument("Vector size mismatch");
        cvec y(x.size());
        check(f(h_, reinterpret_cast<const double *>(x.data()), reinterpret_cast<double *>(y.data())), h_);
        return y;
    }
    gm_solver *h_ = nullptr;
    std::vector<int32_t> lm_;
};

}  // namespace gm
