// gm_cnm.hpp - the reference class CrankNicolsonModes
// (include/CrankNicolsonModes.hpp:60-455 of rdellamonica/boson-cloud-binary)
// over the C ABI of gm.h: a drop-in for the reference drivers
// (src/main.cpp:55-106, src/isolated_bh_main.cpp:108-126), header-only,
// linked with libgm.so.
//
// The public members keep the reference's names and meaning: the constructor
// arguments, the orbit scalars (q, R, theta, phase, Omega, Omega_0, P, R1, R2,
// M1, M2, dt, ...), the flags, t, r, psi, modes, mask, cap,
// last_solver_residual and last_solver_iterations. They are host mirrors of
// the device solver: a step pushes scalars and flags the caller changed (and
// psi, when the caller wrote it) before it runs, and pulls them back after
// it. run() and resume() write and read the reference's run folder
// (Simulation_%Y_%m_%d_%H_%M_%S with config.toml, r.txt, modes.txt, psi/ and
// psi_dt/ snapshots, the 14 diagnostic files and checkpoint.bin) in the
// reference's formats (src/CrankNicolsonModes_utils.cpp:10-100,
// src/CrankNicolsonModes_restart.cpp:22-399, src/Utils.cpp:11-27).
//
// Vectors are std::vector where the reference has Eigen vectors: psi is a
// std::vector<std::complex<double>> in the reference's mode-major layout
// (element j*Nr + i is mode j at node i), t and r are std::vector<double>.
// Within run() and resume() the state stays on the device between snapshots;
// a step() called by the caller syncs psi after it (compat_sync, default on).
//
// Members this build adds: `precision` (GM_PRECISION_FP64 or _MIXED) and
// handle(). The sparse operator members of the reference (H, A, B, V, ...) are
// not materialised: the operator is matrix-free on the device.
#pragma once

#include <algorithm>
#include <chrono>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstring>
#include <ctime>
#include <filesystem>
#include <fstream>
#include <iostream>
#include <map>
#include <regex>
#include <sstream>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "gm.h"

namespace gm {

struct mode {
    int l, m;
};

namespace cnm_detail {

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

// diagnostic_files() (src/CrankNicolsonModes_restart.cpp:109-127), in its order
inline const std::vector<std::string> &diagnostic_files()
{
    static const std::vector<std::string> f = {"flux.txt", "energy_flux.txt", "energy.txt", "energy_modes.txt",
                                               "norms.txt", "t.txt", "p_bin.txt", "p_bin_inst.txt",
                                               "dV_bin_dt.txt", "R_orb.txt", "Omega_orb.txt", "M_1.txt",
                                               "q.txt", "theta_orb.txt"};
    return f;
}

// write_quantity_line / write_quantity_line_modes (utils.cpp:73-88)
inline std::string line(double v)
{
    char b[64];
    std::snprintf(b, sizeof b, "%.12e\n", v);
    return b;
}
inline std::string line_modes(const std::vector<double> &v)
{
    std::string s;
    char b[64];
    for (size_t k = 0; k < v.size(); ++k) {
        std::snprintf(b, sizeof b, k ? " %.12e" : "%.12e", v[k]);
        s += b;
    }
    return s + "\n";
}

// byte offsets just past each newline (line_ends, restart.cpp:38-57)
inline std::vector<size_t> line_ends(const std::string &path)
{
    std::ifstream f(path, std::ios::binary);
    if (!f) throw std::runtime_error("Cannot open " + path);
    std::vector<size_t> ends;
    char buf[1 << 16];
    size_t off = 0;
    while (f) {
        f.read(buf, sizeof buf);
        const std::streamsize n = f.gcount();
        for (std::streamsize k = 0; k < n; ++k)
            if (buf[k] == '\n') ends.push_back(off + k + 1);
        off += static_cast<size_t>(n);
    }
    return ends;
}

// the leading number of each of the first n_rows lines (read_column, restart.cpp:60-75)
inline std::vector<double> read_column(const std::string &path, int n_rows)
{
    std::ifstream f(path);
    if (!f) throw std::runtime_error("Cannot open " + path);
    std::vector<double> out;
    std::string ln;
    while (static_cast<int>(out.size()) < n_rows && std::getline(f, ln)) {
        std::istringstream is(ln);
        double v = 0.0;
        if (!(is >> v)) v = 0.0;
        out.push_back(v);
    }
    if (static_cast<int>(out.size()) < n_rows) throw std::runtime_error("Too few rows in " + path);
    return out;
}

// load_snapshot (restart.cpp:78-105): exactly `size` lines of two numbers
inline bool load_snapshot(const std::string &path, size_t size, std::vector<std::complex<double>> &out)
{
    std::ifstream f(path);
    if (!f) return false;
    std::vector<std::complex<double>> v;
    v.reserve(size);
    std::string ln;
    while (std::getline(f, ln)) {
        if (ln.empty()) continue;
        std::istringstream is(ln);
        double re, im;
        if (!(is >> re >> im)) return false;
        v.emplace_back(re, im);
        if (v.size() > size) return false;
    }
    if (v.size() != size) return false;
    out.swap(v);
    return true;
}

}  // namespace cnm_detail

// Stand-in for the reference's SparseComplex (Eigen::SparseMatrix<dcomplex,
// RowMajor>): CSR with sorted columns, explicit zeros of the reference's
// pattern left out.
struct SparseComplex {
    int64_t n = 0;
    std::vector<int64_t> outer;
    std::vector<int32_t> inner;
    std::vector<std::complex<double>> values;
    int64_t rows() const { return n; }
    int64_t cols() const { return n; }
    int64_t nonZeros() const { return static_cast<int64_t>(values.size()); }
    std::complex<double> coeff(int64_t row, int64_t col) const
    {
        if (row < 0 || row >= n) return {0.0, 0.0};
        const auto b = inner.begin() + outer[row], e = inner.begin() + outer[row + 1];
        const auto it = std::lower_bound(b, e, static_cast<int32_t>(col));
        return it != e && *it == col ? values[it - inner.begin()] : std::complex<double>(0.0, 0.0);
    }
    std::vector<std::complex<double>> operator*(const std::vector<std::complex<double>> &x) const
    {
        std::vector<std::complex<double>> y(n);
        for (int64_t r = 0; r < n; ++r)
            for (int64_t q = outer[r]; q < outer[r + 1]; ++q) y[r] += values[q] * x[inner[q]];
        return y;
    }
};

class CrankNicolsonModes {
public:
    using cvec = std::vector<std::complex<double>>;

    // include/CrankNicolsonModes.hpp:84
    CrankNicolsonModes(int Nx, double L_, double dt_, double epsilon_, int l_max_, double mu_, double M_c_,
                       double q_, double a_, double R_, bool counter_, bool absorption_, int save_every_,
                       double ramp_periods_, double epsilon_smear_, std::string config_file)
        : config_file_(std::move(config_file))
    {
        gm_params p{};
        p.Nr = Nx;
        p.L = L_;
        p.dt = dt_;
        p.epsilon = epsilon_;
        p.l_max = l_max_;
        p.mu = mu_;
        p.M_c = M_c_;
        p.q = q_;
        p.a = a_;
        p.R = R_;
        p.counter = counter_;
        p.absorption = absorption_;
        p.save_every = save_every_;
        p.ramp_periods = ramp_periods_;
        p.epsilon_smear = epsilon_smear_;
        cnm_detail::check(gm_create(&p, &h_), nullptr);
        Nr = Nx;
        save_every = save_every_;
        l_max = l_max_;
        epsilon = epsilon_;
        mu = mu_;
        a = a_;
        L = L_;
        dr = L_ / (Nx - 1);
        counter = counter_;
        M_c = M_c_;
        n_modes = gm_n_modes(h_);
        std::vector<int32_t> lm(2 * static_cast<size_t>(n_modes));
        cnm_detail::check(gm_modes(h_, lm.data()), h_);
        for (int j = 0; j < n_modes; ++j) modes.push_back({lm[2 * j], lm[2 * j + 1]});
        r.resize(Nx);
        cnm_detail::check(gm_grid(h_, r.data()), h_);
        gm_flags f{};
        cnm_detail::check(gm_get_flags(h_, &f), h_);
        pulled_flags_ = f;
        flags_to_members(f);
        pull_scalars();
        psi.assign(static_cast<size_t>(Nr) * n_modes, {0.0, 0.0});
        pulled_psi_ = psi;
    }
    ~CrankNicolsonModes()
    {
        close_diagnostics();
        gm_destroy(h_);
    }
    CrankNicolsonModes(const CrankNicolsonModes &) = delete;
    CrankNicolsonModes &operator=(const CrankNicolsonModes &) = delete;

    // ---- public members (include/CrankNicolsonModes.hpp:88-238)
    int Nr = 0, save_every = 0, l_max = 0, n_modes = 0;
    double epsilon = 0, mu = 0, q = 0, R = 0, theta = 0, phase = 0, Omega = 0, Omega_0 = 0, P = 0;
    double R1 = 0, R2 = 0, M1 = 0, M2 = 0, a = 0, L = 0, dr = 0, dt = 0;
    bool counter = false;
    double M_c = 0;
    bool evolve_orbit = true;
    bool ramp_secondary_potential = true;
    double steps_per_period = 50.0;
    bool use_block_preconditioner = true;
    double last_solver_residual = 0.0;
    int last_solver_iterations = 0;
    std::vector<double> t;
    std::vector<double> r;
    cvec psi;
    bool mask = false;
    bool cap = false;
    std::vector<mode> modes;
    // this build: arithmetic of the step, and the psi sync after a caller's step()
    int precision = GM_PRECISION_FP64;
    bool compat_sync = true;
    // operators (include/CrankNicolsonModes.hpp:197-206): the device keeps them
    // as tables; refresh_operators() materialises them here (small systems)
    SparseComplex H, H0, A, B, V1, V2, V, dVdt, Id, CAP;

    void refresh_operators()
    {
        SparseComplex *dst[] = {&H, &H0, &A, &B, &V1, &V2, &V, &dVdt, &Id, &CAP};
        const int32_t which[] = {GM_OP_H, GM_OP_H0, GM_OP_A, GM_OP_B, GM_OP_V1,
                                 GM_OP_V2, GM_OP_V, GM_OP_DVDT, GM_OP_ID, GM_OP_CAP};
        for (int k = 0; k < 10; ++k) {
            SparseComplex &m = *dst[k];
            int64_t nnz = 0;
            cnm_detail::check(gm_operator_csr(h_, which[k], &nnz, nullptr, nullptr, nullptr), h_);
            m.n = static_cast<int64_t>(Nr) * n_modes;
            m.outer.assign(m.n + 1, 0);
            m.inner.assign(nnz, 0);
            m.values.assign(nnz, {0.0, 0.0});
            cnm_detail::check(gm_operator_csr(h_, which[k], &nnz, m.outer.data(), m.inner.data(),
                                              reinterpret_cast<double *>(m.values.data())),
                              h_);
        }
    }

    void set_power_average_orbits(double n_orbits)
    {
        push_flags(true, n_orbits);
    }
    void set_initial_state_full(cvec psi_0)
    {
        cnm_detail::check(gm_set_initial_state_full(h_, reinterpret_cast<const double *>(psi_0.data()),
                                                    static_cast<int64_t>(psi_0.size())),
                          h_);
        pull_psi();
    }
    void initialize()
    {
        push();
        cnm_detail::check(gm_initialize(h_), h_);
        pull();
    }
    void configure_isolated_black_hole(double x)
    {
        push();
        cnm_detail::check(gm_configure_isolated_black_hole(h_, x), h_);
        isolated_ = true;
        pull();
    }
    void build_damping_mask(double mask_L_factor, double mask_W, int mask_n)
    {
        cnm_detail::check(gm_build_damping_mask(h_, mask_L_factor, mask_W, mask_n), h_);
        mask = true;
    }
    void build_cap(double cap_L_factor, double cap_W, int cap_n)
    {
        cnm_detail::check(gm_build_cap(h_, cap_L_factor, cap_W, cap_n), h_);
        cap = true;
    }

    cvec apply_sink(const cvec &state) const { return op(gm_apply_sink, state); }
    cvec schrodinger_rhs(const cvec &state) const { return op(gm_schrodinger_rhs, state); }
    double sink_strength(int body) const { return gm_sink_strength(h_, body); }
    bool absorption_sink_active() const
    {
        gm_scalars s{};
        cnm_detail::check(gm_get_scalars(h_, &s), h_);
        return s.sink_active != 0;
    }

    // step (src/CrankNicolsonModes.cpp:145-234): one CN step, its diagnostic
    // rows (when a run opened the files) and, every save_every steps, the
    // snapshot and checkpoint
    void step(int step_id) { step_impl(step_id, compat_sync); }

    // run (src/CrankNicolsonModes.cpp:236-259)
    void run(int n_steps, double r_ex_)
    {
        std::cout << "Running simulation..." << std::endl;
        push();
        cnm_detail::check(gm_begin_run(h_, r_ex_, n_steps), h_);
        pull();
        t.assign(std::max(n_steps, 1), 0.0);
        t[0] = 0.0;
        folder_ = create_output_folder();
        std::cout << "Output folder: " << folder_ << std::endl;
        open_diagnostics(std::ios::out);
        save_run_info();
        save_snapshot(0, psi, "psi");
        save_snapshot(0, schrodinger_rhs(psi), "psi_dt");
        save_checkpoint(0);
        evolve(1, n_steps);
    }

    // resume (src/CrankNicolsonModes_restart.cpp:287-399): from the latest
    // complete snapshot that has all its diagnostic rows, exactly from
    // checkpoint.bin when it belongs to that snapshot, else from the snapshot
    // with the orbit rebuilt from the diagnostics; rows written after it are
    // dropped and the run goes on to n_steps - 1
    void resume(const std::string &folder, int n_steps, double r_ex_)
    {
        namespace fs = std::filesystem;
        std::cout << "Resuming simulation in " << folder << std::endl;
        push();
        cnm_detail::check(gm_begin_run(h_, r_ex_, n_steps), h_);
        pull();
        t.assign(std::max(n_steps, 1), 0.0);
        folder_ = folder;
        check_modes(folder);
        std::map<std::string, std::vector<size_t>> ends;
        size_t complete = static_cast<size_t>(-1);
        for (const auto &name : cnm_detail::diagnostic_files()) {
            ends[name] = cnm_detail::line_ends((fs::path(folder) / name).string());
            complete = std::min(complete, ends[name].size());
        }
        std::vector<int> steps;
        const std::regex snap("snapshot_(\\d+)\\.txt");
        const fs::path psi_dir = fs::path(folder) / "psi";
        std::error_code ec;
        fs::directory_iterator it(psi_dir, ec);
        if (ec) throw std::runtime_error("Cannot read " + psi_dir.string());
        for (const auto &e : it) {
            std::smatch m;
            const std::string nm = e.path().filename().string();
            if (std::regex_match(nm, m, snap)) steps.push_back(std::stoi(m[1]));
        }
        std::sort(steps.rbegin(), steps.rend());
        int start = -1;
        cvec state;
        for (int cand : steps) {
            if (static_cast<size_t>(cand) > complete) continue;
            const std::string path = snapshot_path("psi", cand);
            if (cnm_detail::load_snapshot(path, static_cast<size_t>(Nr) * n_modes, state)) {
                start = cand;
                break;
            }
            std::cout << "Skipping incomplete snapshot " << path << std::endl;
        }
        if (start < 0) throw std::runtime_error("No usable snapshot found in " + psi_dir.string());
        std::cout << "Restarting from snapshot " << start << std::endl;
        cnm_detail::check(gm_set_state(h_, reinterpret_cast<const double *>(state.data())), h_);
        int32_t loaded = 0;
        cnm_detail::check(gm_load_checkpoint(h_, (fs::path(folder) / "checkpoint.bin").c_str(), start, &loaded),
                          h_);
        if (loaded) {
            std::cout << "Exact state restored from checkpoint.bin." << std::endl;
        } else {
            std::cout << "No checkpoint for this snapshot: rebuilding the orbit state from the diagnostics."
                      << std::endl;
            static const char *cols[8] = {"t.txt", "R_orb.txt", "theta_orb.txt", "Omega_orb.txt",
                                          "M_1.txt", "q.txt", "p_bin.txt", "p_bin_inst.txt"};
            std::vector<double> rows(8 * static_cast<size_t>(start));
            for (int c = 0; c < 8; ++c) {
                const auto col = cnm_detail::read_column((fs::path(folder) / cols[c]).string(), start);
                for (int k = 0; k < start; ++k) rows[8 * static_cast<size_t>(k) + c] = col[k];
            }
            cnm_detail::check(gm_restore_from_diagnostics(h_, start, start ? rows.data() : nullptr), h_);
        }
        for (const auto &name : cnm_detail::diagnostic_files()) {
            const size_t keep = start == 0 ? 0 : ends[name][start - 1];
            fs::resize_file(fs::path(folder) / name, keep);
        }
        pull();
        if (start < static_cast<int>(t.size())) t[start] = t_i_;
        open_diagnostics(std::ios::app);
        if (start + 1 >= n_steps) {
            std::cout << "The run had already reached its last step (" << n_steps - 1 << ")." << std::endl;
            close_diagnostics();
            return;
        }
        evolve(start + 1, n_steps);
    }

    gm_solver *handle() const { return h_; }
    const std::string &output_folder() const { return folder_; }

private:
    void flags_to_members(const gm_flags &f)
    {
        evolve_orbit = f.evolve_orbit != 0;
        ramp_secondary_potential = f.ramp_secondary_potential != 0;
        steps_per_period = f.steps_per_period;
        use_block_preconditioner = f.use_block_preconditioner != 0;
        precision = f.precision;
    }
    // flags the caller changed, and the rolling-average window when given
    void push_flags(bool set_window = false, double window = 0.0)
    {
        gm_flags f = pulled_flags_;
        f.evolve_orbit = evolve_orbit;
        f.ramp_secondary_potential = ramp_secondary_potential;
        f.steps_per_period = steps_per_period;
        f.use_block_preconditioner = use_block_preconditioner;
        f.precision = precision;
        if (set_window) f.power_average_orbits = window;
        if (std::memcmp(&f, &pulled_flags_, sizeof f) != 0) {
            cnm_detail::check(gm_set_flags(h_, &f), h_);
            pulled_flags_ = f;
        }
    }
    void push()
    {
        push_flags();
        gm_scalars s = pulled_;
        s.theta = theta;
        s.phase = phase;
        s.R = R;
        s.M1 = M1;
        s.M2 = M2;
        s.q = q;
        s.R1 = R1;
        s.R2 = R2;
        s.Omega = Omega;
        s.P = P;
        s.dt = dt;
        s.Omega_0 = Omega_0;
        if (std::memcmp(&s, &pulled_, sizeof s) != 0) {
            cnm_detail::check(gm_set_scalars(h_, &s), h_);
            pulled_ = s;
        }
        if (psi.size() == pulled_psi_.size() &&
            std::memcmp(psi.data(), pulled_psi_.data(), psi.size() * sizeof(psi[0])) != 0) {
            cnm_detail::check(gm_set_state(h_, reinterpret_cast<const double *>(psi.data())), h_);
            pulled_psi_ = psi;
        }
    }
    void pull_scalars()
    {
        gm_scalars s{};
        cnm_detail::check(gm_get_scalars(h_, &s), h_);
        pulled_ = s;
        t_i_ = s.t;
        theta = s.theta;
        phase = s.phase;
        R = s.R;
        M1 = s.M1;
        M2 = s.M2;
        q = s.q;
        R1 = s.R1;
        R2 = s.R2;
        Omega = s.Omega;
        P = s.P;
        dt = s.dt;
        Omega_0 = s.Omega_0;
        last_solver_residual = s.last_solver_residual;
        last_solver_iterations = s.last_solver_iterations;
    }
    void pull_psi()
    {
        psi.resize(static_cast<size_t>(Nr) * n_modes);
        cnm_detail::check(gm_get_state(h_, reinterpret_cast<double *>(psi.data())), h_);
        pulled_psi_ = psi;
    }
    void pull()
    {
        pull_scalars();
        pull_psi();
    }

    void step_impl(int step_id, bool sync_psi)
    {
        push();
        std::vector<double> flux(n_modes), eflux(n_modes), norms(n_modes), energies(n_modes);
        gm_diag d{};
        d.flux = flux.data();
        d.energy_flux = eflux.data();
        d.norms = norms.data();
        d.energies = energies.data();
        cnm_detail::check(gm_step(h_, step_id, &d), h_);
        pull_scalars();
        if (step_id >= 0) {
            if (static_cast<size_t>(step_id) >= t.size()) t.resize(step_id + 1, 0.0);
            t[step_id] = t_i_;
        }
        const bool snapshot = step_id % save_every == 0;
        if (sync_psi || snapshot) pull_psi();
        if (!files_.empty()) {
            using cnm_detail::line;
            using cnm_detail::line_modes;
            files_["R_orb.txt"] << line(R);
            files_["theta_orb.txt"] << line(theta);
            files_["Omega_orb.txt"] << line(Omega);
            files_["M_1.txt"] << line(M1);
            files_["q.txt"] << line(q);
            files_["t.txt"] << line(t_i_);
            files_["energy_flux.txt"] << line_modes(eflux);
            files_["flux.txt"] << line_modes(flux);
            files_["energy.txt"] << line(d.energy);
            files_["p_bin.txt"] << line(d.p_bin);
            files_["p_bin_inst.txt"] << line(d.p_bin_inst);
            files_["dV_bin_dt.txt"] << line(d.dV_bin_dt);
            files_["norms.txt"] << line_modes(norms);
            files_["energy_modes.txt"] << line_modes(energies);
        }
        if (snapshot && !folder_.empty()) {
            save_snapshot(step_id, psi, "psi");
            flush_diagnostics();
            save_checkpoint(step_id);
        }
    }

    // evolve (src/CrankNicolsonModes.cpp:272-289)
    void evolve(int first_step, int n_steps)
    {
        for (int n = first_step; n < n_steps; ++n) {
            if (!isolated_ && evolve_orbit && R <= 6.0) break;
            step_impl(n, false);
        }
        pull_psi();
        std::cout << std::endl;
        flush_diagnostics();
        std::cout << "Simulation completed!" << std::endl;
    }

    // create_output_folder (src/Utils.cpp:11-27)
    static std::string create_output_folder()
    {
        const std::time_t now = std::chrono::system_clock::to_time_t(std::chrono::system_clock::now());
        std::tm tm{};
        localtime_r(&now, &tm);
        char name[64];
        std::strftime(name, sizeof name, "Simulation_%Y_%m_%d_%H_%M_%S", &tm);
        std::filesystem::create_directories(std::filesystem::path(name) / "psi");
        std::filesystem::create_directories(std::filesystem::path(name) / "psi_dt");
        return name;
    }
    void open_diagnostics(std::ios::openmode mode)
    {
        close_diagnostics();
        for (const auto &name : cnm_detail::diagnostic_files()) {
            const std::string path = (std::filesystem::path(folder_) / name).string();
            files_[name].open(path, mode | std::ios::out);
            if (!files_[name]) throw std::runtime_error("Cannot open " + path);
        }
    }
    void flush_diagnostics()
    {
        for (auto &kv : files_) kv.second.flush();
    }
    void close_diagnostics() { files_.clear(); }
    // save_run_info (utils.cpp:10-45)
    void save_run_info()
    {
        namespace fs = std::filesystem;
        std::error_code ec;
        if (!config_file_.empty() && fs::exists(config_file_))
            fs::copy_file(config_file_, fs::path(folder_) / "config.toml", fs::copy_options::overwrite_existing, ec);
        std::ofstream fr(fs::path(folder_) / "r.txt");
        for (double x : r) fr << x << "\n";
        std::ofstream fm(fs::path(folder_) / "modes.txt");
        for (const auto &md : modes) fm << md.l << "," << md.m << "\n";
    }
    std::string snapshot_path(const std::string &name, int step_id) const
    {
        char b[64];
        std::snprintf(b, sizeof b, "snapshot_%06d.txt", step_id);
        return (std::filesystem::path(folder_) / name / b).string();
    }
    // save_snapshot (utils.cpp:52-71): 're im' per element, stream default format
    void save_snapshot(int step_id, const cvec &v, const std::string &name) const
    {
        std::ofstream f(snapshot_path(name, step_id));
        for (const auto &c : v) f << c.real() << " " << c.imag() << "\n";
    }
    void save_checkpoint(int step_id)
    {
        cnm_detail::check(
            gm_save_checkpoint(h_, (std::filesystem::path(folder_) / "checkpoint.bin").c_str(), step_id), h_);
    }
    void check_modes(const std::string &folder) const
    {
        std::ifstream f(std::filesystem::path(folder) / "modes.txt");
        if (!f) return;
        std::string ln;
        size_t count = 0;
        while (std::getline(f, ln)) {
            if (ln.empty()) continue;
            int l = -1, m = 0;
            char comma = 0;
            std::istringstream is(ln);
            const bool ok = static_cast<bool>(is >> l >> comma >> m) && comma == ',';
            if (!ok || count >= modes.size() || modes[count].l != l || modes[count].m != m)
                throw std::runtime_error("modes.txt does not match the modes built from config.toml");
            ++count;
        }
        if (count != modes.size()) throw std::runtime_error("modes.txt does not match the modes built from config.toml");
    }

    template <class F>
    cvec op(F f, const cvec &x) const
    {
        if (x.size() != static_cast<size_t>(Nr) * n_modes) throw std::invalid_argument("Vector size mismatch");
        cvec y(x.size());
        cnm_detail::check(f(h_, reinterpret_cast<const double *>(x.data()), reinterpret_cast<double *>(y.data())), h_);
        return y;
    }

    gm_solver *h_ = nullptr;
    std::string config_file_;
    std::string folder_;
    std::map<std::string, std::ofstream> files_;
    gm_scalars pulled_{};
    gm_flags pulled_flags_{};
    cvec pulled_psi_;
    double t_i_ = 0.0;
    bool isolated_ = false;
};

}  // namespace gm
