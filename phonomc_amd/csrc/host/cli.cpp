// Command-line driver with the reference CLI's contract (src/main.cpp:61-141):
// pmc [input.toml] -> indexed result folder with summary.txt, grid_centers.csv
// and convergence.txt; "Fatal error: ..." and exit status 1 on failure.
#include "solver.hpp"

#include <algorithm>
#include <chrono>
#include <iomanip>
#include <exception>
#include <filesystem>
#include <fstream>
#include <iostream>

int main(int argc, char** argv) {
    try {
        namespace fs = std::filesystem;
        std::cout << "PhonoMC deviational Monte Carlo step on MI355X\n";
        const std::string input = argc > 1 ? argv[1] : "input.toml";
        pmc::SimConfig config = pmc::parse_toml_file(input);
        if (config.heat_source_enabled && config.compute_kappa) {
            std::cout << "Config note: heat_source.enabled=true, so thermal conductivity estimation is disabled "
                         "for this run.\n";
            config.compute_kappa = false;
        }
        if (!config.output_folder.empty()) {
            fs::path r(config.output_folder);
            if (!r.is_absolute()) config.output_folder = (fs::path(config.input_directory) / r).lexically_normal().string();
        }
        config.output_folder = pmc::create_indexed_output_folder(config.output_folder);
        std::cout << "Results folder: " << config.output_folder << '\n';
        const auto start = std::chrono::steady_clock::now();
        pmc::MonteCarloSolver solver(pmc::build_problem(config), config.output_folder, &std::cout);
        const int interval = std::max(1, config.iterations / 100);
        while (solver.current_timestep() < config.iterations && !solver.converged()) {
            const int n = std::min(interval - solver.current_timestep() % interval,
                                   config.iterations - solver.current_timestep());
            solver.run_steps(n);
            const auto t = solver.cell_temperatures();
            std::cout << "--- Progress: " << std::fixed << std::setprecision(1)
                      << 100.0 * solver.current_timestep() / config.iterations << "% ---\n";
            if (config.compute_kappa) {
                std::cout << "Current Conductivity (Int): " << solver.kappa_int() << " W/mK\n";
                std::cout << "Current Conductivity (Eff): " << solver.kappa_eff() << " W/mK\n";
            }
            std::cout.unsetf(std::ios::floatfield);
        }
        const auto end = std::chrono::steady_clock::now();
        const auto sec = std::chrono::duration_cast<std::chrono::seconds>(end - start).count();
        const auto days = sec / 86400, hours = (sec % 86400) / 3600, minutes = (sec % 3600) / 60, seconds = sec % 60;
        std::cout << "Total time: " << days << " days " << hours << " h " << minutes << " min " << seconds << " s\n";
        std::ofstream summary(fs::path(config.output_folder) / "summary.txt", std::ios::app);
        if (!summary) throw std::runtime_error("Cannot append runtime summary.txt.");
        summary << "\n[runtime]\n";
        summary << "total_seconds = " << sec << '\n';
        summary << "total_days = " << days << '\n';
        summary << "total_hours = " << hours << '\n';
        summary << "total_minutes = " << minutes << '\n';
        summary << "total_seconds_remainder = " << seconds << '\n';
        summary << "total_human_readable = " << days << " days " << hours << " h " << minutes << " min " << seconds
                << " s\n";
        solver.append_profile_summary(summary);
        return 0;
    } catch (const std::exception& e) {
        std::cerr << "Fatal error: " << e.what() << '\n';
        return 1;
    }
}
