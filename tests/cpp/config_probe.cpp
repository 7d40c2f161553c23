// Prints the Config that include/gm_config.hpp reads from a file as one JSON
// object, or the exception it throws ("error": type, "what": message).
#include <cmath>
#include <cstdio>
#include <stdexcept>
#include <string>

#include "gm_config.hpp"

static void num(const char *k, double v, bool last = false)
{
    std::printf("\"%s\": %.17g%s", k, v, last ? "" : ", ");
}

int main(int argc, char **argv)
{
    if (argc != 2) return 2;
    try {
        const gm::Config c = gm::load_config(argv[1]);
        std::printf("{");
        num("N", c.N); num("L", c.L); num("dt", c.dt); num("num_periods", c.num_periods);
        num("num_steps", c.num_steps); num("save_every", c.save_every); num("l_max", c.l_max);
        num("r_ex", c.r_ex); num("ramp_periods", c.ramp_periods);
        num("ramp_secondary_potential", c.ramp_secondary_potential); num("steps_per_period", c.steps_per_period);
        num("power_average_orbits", c.power_average_orbits); num("epsilon", c.epsilon);
        num("epsilon_smear", c.epsilon_smear); num("evolve_orbit", c.evolve_orbit); num("absorption", c.absorption);
        num("q", c.q); num("R", c.R); num("counter", c.counter); num("mu", c.mu); num("a", c.a); num("M_c", c.M_c);
        num("n", c.n); num("l", c.l); num("m", c.m); num("threads", c.threads);
        num("mask", c.mask); num("mask_L_factor", c.mask_L_factor); num("mask_W", c.mask_W); num("mask_n", c.mask_n);
        num("cap", c.cap); num("cap_L_factor", c.cap_L_factor); num("cap_W", c.cap_W); num("cap_n", c.cap_n);
        std::printf("\"preconditioner\": \"%s\"}\n", c.preconditioner.c_str());
    } catch (const std::invalid_argument &e) {
        std::printf("{\"error\": \"invalid_argument\", \"what\": \"%s\"}\n", e.what());
    } catch (const std::exception &e) {
        std::printf("{\"error\": \"exception\", \"what\": \"%s\"}\n", e.what());
    }
    return 0;
}
