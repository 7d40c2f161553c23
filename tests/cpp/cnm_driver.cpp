// Test driver of include/gm.hpp: the C++ class a reference caller would use.
//   cnm_driver errors           exception mapping without a device
//   cnm_driver step <in> <out>  one step of a small binary from a raw complex128 state
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>

#include "gm.hpp"

static int errors()
{
    // Nx < 3 is rejected before any device work (src/CrankNicolsonModes.cpp:18-19)
    try {
        gm::CrankNicolsonModes bad(2, 10.0, 0.1, 0.0, 2, 0.2, 0.1, 0.5, 0.0, 20.0, false, false, 10, 0.0, 4.0);
        std::puts("no exception");
        return 1;
    } catch (const std::invalid_argument &e) {
        std::printf("invalid_argument: %s\n", e.what());
    }
    // dt <= 0 too
    try {
        gm::CrankNicolsonModes bad(65, 64.0, -1.0, 0.0, 2, 0.2, 0.1, 0.5, 0.0, 20.0, false, false, 10, 0.0, 4.0);
        return 1;
    } catch (const std::invalid_argument &e) {
        std::printf("invalid_argument: %s\n", e.what());
    }
    return 0;
}

static int step(const char *in, const char *out)
{
    gm::CrankNicolsonModes s(129, 128.0, 1.0, 0.0, 4, 0.2, 0.05, 0.5, 0.0, 20.0, false, true, 100, 0.0, 4.0);
    gm::cvec psi(static_cast<size_t>(s.Nr) * s.n_modes);
    std::ifstream f(in, std::ios::binary);
    f.read(reinterpret_cast<char *>(psi.data()), static_cast<std::streamsize>(psi.size() * sizeof(psi[0])));
    if (!f) {
        std::cerr << "short input\n";
        return 2;
    }
    gm_flags fl = s.flags();
    fl.evolve_orbit = 0;
    s.set_flags(fl);
    s.set_initial_state_full(psi);
    s.initialize();
    s.begin_run(100.0, 2);
    const gm::StepDiagnostics d = s.step(1);
    const gm::cvec after = s.psi();
    std::ofstream o(out, std::ios::binary);
    o.write(reinterpret_cast<const char *>(after.data()), static_cast<std::streamsize>(after.size() * sizeof(after[0])));
    std::printf("{\"energy\": %.17g, \"iterations\": %d, \"n_modes\": %d}\n", d.energy,
                s.scalars().last_solver_iterations, s.n_modes);
    return 0;
}

int main(int argc, char **argv)
{
    try {
        if (argc >= 2 && !std::strcmp(argv[1], "errors")) return errors();
        if (argc >= 4 && !std::strcmp(argv[1], "step")) return step(argv[2], argv[3]);
    } catch (const std::exception &e) {
        std::cerr << "exception: " << e.what() << "\n";
        return 3;
    }
    std::cerr << "usage: cnm_driver errors | step <in> <out>\n";
    return 2;
}
