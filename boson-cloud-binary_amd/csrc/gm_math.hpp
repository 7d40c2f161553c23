// gm_math.hpp - host special functions and the rolling phase average.
//
// GSL is not part of this build, so the three functions the reference takes
// from it are written here from their definitions:
//   wigner3j          gsl_sf_coupling_3j (src/CrankNicolsonModes_hamiltonian.cpp:85-86),
//                     Racah's closed sum evaluated in binary128;
//   sph_plm0          gsl_sf_legendre_sphPlm(l, m, 0) (hamiltonian.cpp:87), from
//                     P_l^m(0) = (-1)^((l+m)/2) (l+m-1)!!/(l-m)!! (Condon-Shortley phase);
//   gauss_legendre01  gsl_integration_glfixed_point on [0, 1] (absorption.cpp:27-34),
//                     Newton iteration on P_n in binary128, nodes ascending.
#pragma once
#include <cstdint>
#include <deque>
#include <iosfwd>
#include <vector>

namespace gmh {

double wigner3j(int j1, int j2, int j3, int m1, int m2, int m3);
double sph_plm0(int l, int m);
void gauss_legendre01(int n, std::vector<double> &x, std::vector<double> &w);

// RollingPhaseAverage (include/RollingPhase.hpp:22-78, src/RollingPhase.cpp:14-209):
// trapezoidal time average over the last n_orbits of orbital phase. The
// save/load byte layout is the reference's checkpoint block (restart.cpp:163).
class RollingPhaseAverage {
public:
    explicit RollingPhaseAverage(double n_orbits);
    double update(double time, double phase, double value);
    bool ready() const { return ready_; }
    double value() const { return average_; }
    void reset();
    void save(std::ostream &out) const;
    void load(std::istream &in);

private:
    struct Sample {
        double time, phase, value;
    };
    struct Segment {
        double time0, time1, phase0, phase1, value0, value1;
        double integral = 0.0;
    };
    double window_phase_;
    std::deque<Segment> segments_;
    Sample previous_{};
    bool has_previous_ = false;
    bool ready_ = false;
    double initial_phase_ = 0.0;
    double integral_ = 0.0;
    double duration_ = 0.0;
    double average_ = 0.0;
};

}  // namespace gmh
