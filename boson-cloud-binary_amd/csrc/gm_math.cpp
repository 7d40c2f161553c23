// gm_math.cpp - see gm_math.hpp.
#include "gm_math.hpp"

#include <cmath>
#include <istream>
#include <map>
#include <mutex>
#include <ostream>
#include <stdexcept>

namespace gmh {

namespace {

using f128 = __float128;

f128 sqrt128(f128 v)
{
    if (v <= 0) return 0;
    // into the double range by exact powers of 4 (3j prefactors pass 1e308
    // from degree ~25 on), then Newton from the double root
    f128 scale = 1;
    const f128 big = static_cast<f128>(1e300), small = static_cast<f128>(1e-300);
    const f128 down = static_cast<f128>(std::ldexp(1.0, -200)), up = static_cast<f128>(std::ldexp(1.0, 200));
    while (v > big) {
        v *= down;
        scale *= static_cast<f128>(std::ldexp(1.0, 100));
    }
    while (v < small) {
        v *= up;
        scale *= static_cast<f128>(std::ldexp(1.0, -100));
    }
    f128 x = static_cast<f128>(std::sqrt(static_cast<double>(v)));
    for (int k = 0; k < 3; ++k) x = (x + v / x) / 2;
    return x * scale;
}

const std::vector<f128> &factorials()
{
    static const std::vector<f128> table = [] {
        std::vector<f128> t(512);
        t[0] = 1;
        for (std::size_t n = 1; n < t.size(); ++n) t[n] = t[n - 1] * static_cast<f128>(n);
        return t;
    }();
    return table;
}

}  // namespace

double wigner3j(int j1, int j2, int j3, int m1, int m2, int m3)
{
    if (m1 + m2 + m3 != 0 || j1 < 0 || j2 < 0 || j3 < 0) return 0.0;
    if (std::abs(m1) > j1 || std::abs(m2) > j2 || std::abs(m3) > j3) return 0.0;
    if (j3 < std::abs(j1 - j2) || j3 > j1 + j2) return 0.0;
    const auto &f = factorials();
    // Racah: sum over k of (-1)^k / [k! (j3-j2+k+m1)! (j3-j1+k-m2)! (j1+j2-j3-k)! (j1-k-m1)! (j2-k+m2)!]
    const int kmin = std::max({0, j2 - j3 - m1, j1 - j3 + m2});
    const int kmax = std::min({j1 + j2 - j3, j1 - m1, j2 + m2});
    f128 sum = 0;
    for (int k = kmin; k <= kmax; ++k) {
        const f128 d = f[k] * f[j3 - j2 + k + m1] * f[j3 - j1 + k - m2] * f[j1 + j2 - j3 - k] *
                       f[j1 - k - m1] * f[j2 - k + m2];
        sum += (k % 2 ? -1 : 1) / d;
    }
    const f128 tri = f[j1 + j2 - j3] * f[j1 - j2 + j3] * f[j2 + j3 - j1] / f[j1 + j2 + j3 + 1];
    const f128 mfac = f[j1 + m1] * f[j1 - m1] * f[j2 + m2] * f[j2 - m2] * f[j3 + m3] * f[j3 - m3];
    f128 value = sqrt128(tri * mfac) * sum;
    if ((j1 - j2 - m3) % 2 != 0) value = -value;
    return static_cast<double>(value);
}

double sph_plm0(int l, int m)
{
    if (m < 0 || m > l || (l + m) % 2 != 0) return 0.0;
    const auto &f = factorials();
    f128 num = 1, den = 1;
    for (int k = l + m - 1; k > 1; k -= 2) num *= k;
    for (int k = l - m; k > 1; k -= 2) den *= k;
    f128 p = num / den;
    if (((l + m) / 2) % 2 != 0) p = -p;
    const f128 pi = static_cast<f128>(3.14159265358979323846264338327950288Q);
    const f128 norm = static_cast<f128>(2 * l + 1) / (4 * pi) * f[l - m] / f[l + m];
    return static_cast<double>(sqrt128(norm) * p);
}

void gauss_legendre01(int n, std::vector<double> &x, std::vector<double> &w)
{
    // the rule is computed once per order and process (its binary128 Newton
    // iterations cost milliseconds, and an evolving orbit asks for it at every
    // step when the absorber profiles move)
    static std::mutex mu;
    static std::map<int, std::pair<std::vector<double>, std::vector<double>>> cache;
    {
        std::lock_guard<std::mutex> lk(mu);
        const auto it = cache.find(n);
        if (it != cache.end()) {
            x = it->second.first;
            w = it->second.second;
            return;
        }
    }
    x.assign(n, 0.0);
    w.assign(n, 0.0);
    auto legendre = [n](f128 z, f128 &pn, f128 &dpn) {
        f128 p0 = 1, p1 = z;
        for (int k = 2; k <= n; ++k) {
            const f128 p2 = ((2 * k - 1) * z * p1 - (k - 1) * p0) / k;
            p0 = p1;
            p1 = p2;
        }
        pn = p1;
        dpn = n * (z * p1 - p0) / (z * z - 1);
    };
    for (int k = 0; k < n; ++k) {
        // k-th largest root of P_n
        f128 z = static_cast<f128>(std::cos(M_PI * (k + 0.75) / (n + 0.5)));
        f128 pn, dpn;
        for (int it = 0; it < 60; ++it) {
            legendre(z, pn, dpn);
            const f128 dz = pn / dpn;
            z -= dz;
            if (std::fabs(static_cast<double>(dz)) < 1e-32) break;
        }
        legendre(z, pn, dpn);
        x[n - 1 - k] = static_cast<double>((1 + z) / 2);
        w[n - 1 - k] = static_cast<double>(1 / ((1 - z * z) * dpn * dpn));
    }
    std::lock_guard<std::mutex> lk(mu);
    cache.emplace(n, std::make_pair(x, w));
}

// ------------------------------------------------------------------ rolling average

RollingPhaseAverage::RollingPhaseAverage(double n_orbits) : window_phase_(2.0 * M_PI * n_orbits)
{
    if (n_orbits <= 0.0) throw std::invalid_argument("RollingPhaseAverage: n_orbits must be positive");
}

double RollingPhaseAverage::update(double time, double phase, double value)
{
    if (!has_previous_) {
        previous_ = {time, phase, value};
        initial_phase_ = phase;
        has_previous_ = true;
        average_ = value;
        return average_;
    }
    if (time <= previous_.time) throw std::runtime_error("RollingPhaseAverage: time must increase");
    if (phase <= previous_.phase) throw std::runtime_error("RollingPhaseAverage: phase must increase");
    Segment seg{previous_.time, time, previous_.phase, phase, previous_.value, value};
    seg.integral = 0.5 * (seg.value0 + seg.value1) * (seg.time1 - seg.time0);
    segments_.push_back(seg);
    integral_ += seg.integral;
    duration_ += seg.time1 - seg.time0;
    if (phase - initial_phase_ >= window_phase_) ready_ = true;
    const double cut = phase - window_phase_;
    while (!segments_.empty() && segments_.front().phase1 <= cut) {
        integral_ -= segments_.front().integral;
        duration_ -= segments_.front().time1 - segments_.front().time0;
        segments_.pop_front();
    }
    if (!segments_.empty()) {
        Segment &f = segments_.front();
        if (f.phase0 < cut && cut < f.phase1) {
            const double frac = (cut - f.phase0) / (f.phase1 - f.phase0);
            const double tc = f.time0 + frac * (f.time1 - f.time0);
            const double vc = f.value0 + frac * (f.value1 - f.value0);
            integral_ -= 0.5 * (f.value0 + vc) * (tc - f.time0);
            duration_ -= tc - f.time0;
            f.time0 = tc;
            f.phase0 = cut;
            f.value0 = vc;
            f.integral = 0.5 * (f.value0 + f.value1) * (f.time1 - f.time0);
        }
    }
    duration_ = std::max(duration_, 0.0);
    average_ = duration_ > 0.0 ? integral_ / duration_ : value;
    previous_ = {time, phase, value};
    return average_;
}

void RollingPhaseAverage::reset()
{
    segments_.clear();
    has_previous_ = ready_ = false;
    integral_ = duration_ = average_ = initial_phase_ = 0.0;
}

namespace {
template <class T>
void put(std::ostream &o, const T &v) { o.write(reinterpret_cast<const char *>(&v), sizeof(T)); }
template <class T>
void get(std::istream &i, T &v) { i.read(reinterpret_cast<char *>(&v), sizeof(T)); }
}  // namespace

void RollingPhaseAverage::save(std::ostream &out) const
{
    put(out, window_phase_);
    put(out, previous_);
    put(out, has_previous_);
    put(out, ready_);
    put(out, initial_phase_);
    put(out, integral_);
    put(out, duration_);
    put(out, average_);
    const std::uint64_t n = segments_.size();
    put(out, n);
    for (const Segment &s : segments_) put(out, s);
}

void RollingPhaseAverage::load(std::istream &in)
{
    get(in, window_phase_);
    get(in, previous_);
    get(in, has_previous_);
    get(in, ready_);
    get(in, initial_phase_);
    get(in, integral_);
    get(in, duration_);
    get(in, average_);
    std::uint64_t n = 0;
    get(in, n);
    segments_.clear();
    for (std::uint64_t k = 0; k < n && in; ++k) {
        Segment s;
        get(in, s);
        segments_.push_back(s);
    }
    if (!in) throw std::runtime_error("RollingPhaseAverage: truncated checkpoint");
}

}  // namespace gmh
