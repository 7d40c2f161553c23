// The operator checks of the reference's physics test (tests/physics.cpp:80-102)
// written against the drop-in class of include/gm_cnm.hpp: after a one-step
// run and three more steps of a small binary with absorbers, the conservative
// Hamiltonian H is Hermitian on the interior nodes (the Dirichlet rows only
// pin the boundary values), dV/dt is Hermitian, and the sink is Hermitian and
// positive semi-definite. Also: the CN operators are A = 1 + i dt H / 2 and
// B = 1 - i dt H / 2, and H = H0 + V. Prints one line per check; exit status 0
// when all hold. Run it in a scratch folder: run() writes its output folder.
#include <cmath>
#include <complex>
#include <cstdio>
#include <vector>

#include "gm_cnm.hpp"

using gm::CrankNicolsonModes;
using gm::SparseComplex;
using cplx = std::complex<double>;
using cvec = std::vector<cplx>;

static int failures = 0;

static void require(bool ok, const char *what)
{
    std::printf("%s: %s\n", ok ? "ok" : "FAILED", what);
    if (!ok) ++failures;
}

// A smooth state with every mode populated and zero boundary values.
static cvec generic_state(const CrankNicolsonModes &cn)
{
    cvec psi(static_cast<size_t>(cn.Nr) * cn.n_modes);
    for (int j = 0; j < cn.n_modes; ++j)
        for (int i = 1; i < cn.Nr - 1; ++i) {
            const double x = static_cast<double>(i) / (cn.Nr - 1);
            psi[static_cast<size_t>(j) * cn.Nr + i] =
                std::polar(std::sin(M_PI * x) * std::exp(-3.0 * x) / (1.0 + j), 0.37 * j + 1.3 * x);
        }
    return psi;
}

// max |M - M^H| over the entries of M restricted to rows and columns in `keep`
static double hermitian_defect(const SparseComplex &m, const std::vector<char> &keep)
{
    double d = 0.0;
    for (int64_t r = 0; r < m.n; ++r) {
        if (!keep[r]) continue;
        for (int64_t q = m.outer[r]; q < m.outer[r + 1]; ++q) {
            const int64_t c = m.inner[q];
            if (!keep[c]) continue;
            d = std::max(d, std::abs(m.values[q] - std::conj(m.coeff(c, r))));
        }
    }
    return d;
}

static double max_abs(const SparseComplex &m)
{
    double v = 0.0;
    for (const cplx &x : m.values) v = std::max(v, std::abs(x));
    return v;
}

static cplx dot(const cvec &a, const cvec &b)
{
    cplx s = 0.0;
    for (size_t k = 0; k < a.size(); ++k) s += std::conj(a[k]) * b[k];
    return s;
}

int main()
{
    // l_max 4 with absorbers on a small grid; the sinks act after the ramp-free start
    CrankNicolsonModes cn(121, 60.0, 0.25, 0.0, 4, 0.2, 0.01, 0.5, 0.0, 12.0, false, true, 1000, 0.0, 2.0, "");
    cn.set_initial_state_full(generic_state(cn));
    cn.initialize();
    cn.run(1, 40.0);
    cn.t.resize(4);
    for (int step = 1; step <= 3; ++step) cn.step(step);
    cn.refresh_operators();

    const int64_t n = static_cast<int64_t>(cn.Nr) * cn.n_modes;
    std::vector<char> interior(n, 1);
    for (int j = 0; j < cn.n_modes; ++j) {
        interior[static_cast<size_t>(j) * cn.Nr] = 0;
        interior[static_cast<size_t>(j) * cn.Nr + cn.Nr - 1] = 0;
    }
    const std::vector<char> all(n, 1);
    require(hermitian_defect(cn.H, interior) < 1e-12 * max_abs(cn.H), "conservative Hamiltonian is Hermitian");
    require(cn.dVdt.nonZeros() > 0 && hermitian_defect(cn.dVdt, all) < 1e-12 * (1.0 + max_abs(cn.dVdt)),
            "dV/dt is Hermitian");

    // the sink: <a, G b> = conj(<b, G a>) and <a, G a> >= 0
    const cvec a = generic_state(cn);
    cvec b(a.size());
    for (size_t k = 0; k < a.size(); ++k) b[k] = std::conj(a[a.size() - 1 - k]) * cplx(0.3, -0.8);
    const cplx ab = dot(a, cn.apply_sink(b)), ba = dot(b, cn.apply_sink(a));
    require(std::abs(ab - std::conj(ba)) < 1e-12 * (1.0 + std::abs(ab)), "sink is Hermitian");
    require(dot(a, cn.apply_sink(a)).real() >= -1e-14, "sink is positive semi-definite");

    // A = 1 + i dt H / 2, B = 1 - i dt H / 2, H = H0 + V (no CAP here), over
    // every entry of any of them
    double dA = 0.0, dB = 0.0, dH = 0.0;
    for (const SparseComplex *m : {&cn.H, &cn.H0, &cn.V, &cn.A, &cn.B})
        for (int64_t r = 0; r < n; ++r)
            for (int64_t q = m->outer[r]; q < m->outer[r + 1]; ++q) {
                const int64_t c = m->inner[q];
                const cplx h = cn.H.coeff(r, c), id = r == c ? 1.0 : 0.0;
                dA = std::max(dA, std::abs(cn.A.coeff(r, c) - (id + cplx(0.0, 0.5 * cn.dt) * h)));
                dB = std::max(dB, std::abs(cn.B.coeff(r, c) - (id - cplx(0.0, 0.5 * cn.dt) * h)));
                dH = std::max(dH, std::abs(h - cn.H0.coeff(r, c) - cn.V.coeff(r, c)));
            }
    const double hs = max_abs(cn.H);
    require(dA < 1e-14 * hs + 1e-15 && dB < 1e-14 * hs + 1e-15, "A and B are the CN operators of H");
    require(dH < 1e-13 * hs, "H = H0 + V");
    require(cn.Id.nonZeros() == n && cn.CAP.nonZeros() == 0, "identity and an empty CAP");
    std::printf("%s\n", failures ? "FAILED" : "all checks passed");
    return failures ? 1 : 0;
}
