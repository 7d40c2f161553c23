// The reference's GravitationalMolecule driver (src/main.cpp:20-106 of
// rdellamonica/boson-cloud-binary), statement for statement, over the
// MI355X solver's drop-in headers: include/gm_config.hpp (load_config),
// include/gm_cnm.hpp (CrankNicolsonModes) and include/gm_initial_data.hpp
// (load_initial_modes_binary). std::vector stands in for Eigen and there is
// no OpenMP thread setting (the step runs on the GPU).
//
//   gm_main config.toml initial_modes.bin   # new run
//   gm_main Simulation_YYYY_MM_DD_HH_MM_SS   # resume a run
#include <cmath>
#include <filesystem>
#include <iostream>
#include <string>

#include "gm_cnm.hpp"
#include "gm_config.hpp"
#include "gm_initial_data.hpp"

int main(int argc, char **argv)
{
    const bool resume = argc == 2 && std::filesystem::is_directory(argv[1]);
    if (argc != 3 && !resume) {
        std::cerr << "Usage: " << argv[0] << " config.toml initial_modes.bin\n"
                  << "       " << argv[0] << " simulation_folder   (resume)\n";
        return 1;
    }
    try {
        std::string config_file = "config.toml";
        std::string resume_folder;
        if (resume) {
            resume_folder = std::filesystem::path(argv[1]).lexically_normal().string();
            while (resume_folder.size() > 1 && resume_folder.back() == '/') resume_folder.pop_back();
            config_file = resume_folder + "/config.toml";
        } else {
            config_file = argv[1];
        }
        gm::Config cfg = gm::load_config(config_file);
        std::cout << "Loaded configuration from " << config_file << "\n";

        gm::CrankNicolsonModes solver(cfg.N, cfg.L, cfg.dt, cfg.epsilon, cfg.l_max, cfg.mu, cfg.M_c, cfg.q, cfg.a,
                                      cfg.R, cfg.counter, cfg.absorption, cfg.save_every, cfg.ramp_periods,
                                      cfg.epsilon_smear, config_file);
        if (!resume) solver.set_initial_state_full(gm::load_initial_modes_binary(argv[2], solver.modes));

        solver.evolve_orbit = cfg.evolve_orbit;
        solver.ramp_secondary_potential = cfg.ramp_secondary_potential;
        solver.steps_per_period = cfg.steps_per_period;
        solver.set_power_average_orbits(cfg.power_average_orbits);
        solver.use_block_preconditioner = cfg.preconditioner == "radial";
        solver.initialize();
        if (cfg.mask) solver.build_damping_mask(cfg.mask_L_factor, cfg.mask_W, cfg.mask_n);
        if (cfg.cap) solver.build_cap(cfg.cap_L_factor, cfg.cap_W, cfg.cap_n);

        // num_steps takes precedence; otherwise num_periods initial periods
        const int num_steps =
            cfg.num_steps ? cfg.num_steps : static_cast<int>(std::floor(cfg.num_periods * solver.P / solver.dt));
        if (resume)
            solver.resume(resume_folder, num_steps, cfg.r_ex);
        else
            solver.run(num_steps, cfg.r_ex);
        std::cout << "folder " << solver.output_folder() << "\n";
    } catch (const std::exception &e) {
        std::cerr << "GravitationalMolecule: " << e.what() << "\n";
        return 1;
    }
    return 0;
}
