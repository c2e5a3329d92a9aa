// Sectioned-TOML input for the Monte Carlo step, with the reference's key set,
// units (nm in, Angstrom inside) and rejection messages
// (src/config/ConfigParsing.cpp:330-971, include/SimulationConfig.h:42-146).
#pragma once
#include <array>
#include <cstdint>
#include <optional>
#include <string>
#include <vector>

namespace pmc {

enum class CollisionModel { Rta, Callaway };
enum class HeatProfile { Uniform, Gaussian };

struct SimConfig {
    std::string input_directory;
    std::string model = "box";
    std::vector<double> sizes{200.0, 200.0, 200.0}; // Angstrom (include/SimulationConfig.h:46)
    bool merge_coplanar_facets = true;
    double particle_count = 1e4;
    double time_step = 1.0; // ps
    int iterations = 10000;
    int convergence_write_interval = 10;
    std::uint64_t random_seed = 12345;
    bool compute_kappa = false;
    bool write_cell_heat_flux = false;
    int transport_axis = -1;
    bool profile_timers = false;
    bool progress_temperature_summary_only = false;
    double temperature_lookup_dt = 0.1;
    double background_temperature = 300.0;
    bool lifetime_local = true;
    double lifetime_temperature = 300.0;
    bool initial_linear = false;
    double initial_uniform_temperature = 300.0;
    std::array<int, 3> grid{1, 1, 1};
    CollisionModel collision = CollisionModel::Rta;
    bool callaway_nonlinear = false;
    bool conservative_boundaries = false;

    // Boundary selectors: relative boxes [xmin,ymin,zmin,xmax,ymax,zmax].
    std::vector<std::array<double, 6>> boundary_boxes;
    std::vector<char> boundary_conditions; // 'T', 'P', 'R'
    std::vector<double> boundary_values;   // raw TOML values (K or nm)
    std::vector<std::array<double, 6>> periodic_pairs;

    std::string material_folder;
    std::string output_folder;

    // Prescribed heat source (Angstrom inside; include/SimulationConfig.h:103-139).
    bool heat_source_enabled = false;
    HeatProfile heat_profile = HeatProfile::Uniform;
    std::vector<double> heat_min, heat_max, heat_center, heat_sigma;
    std::vector<double> heat_centers;         // flattened [x,y,z] rows
    std::vector<double> heat_half_width;
    std::vector<double> heat_power_densities; // per Gaussian centre, W/m^3
    double heat_power_density = 0.0;          // W/m^3
    std::optional<double> heat_total_power;   // W
    double heat_frequency_min = -1.0, heat_frequency_max = -1.0; // THz, <0: no cutoff
    std::vector<int> heat_branches;
    int heat_spectrum = 0;                    // 0 thermal, 1 weighted, 2 gaussian
    std::vector<double> heat_branch_weights;  // weighted spectrum, per branch
    std::optional<double> heat_frequency_center, heat_frequency_sigma; // gaussian spectrum, THz
    bool heat_time_square = false;
    double heat_time_start = 0.0, heat_time_end = -1.0; // ps
    double heat_period = 0.0, heat_on_duration = -1.0, heat_duty_cycle = 0.5, heat_amplitude = 1.0;

    // Keys parsed for compatibility whose features this build rejects at solver set-up.
    bool has_gradient_drive = false;
    bool has_layered_materials = false;
    // [materials] (src/config/ConfigParsing.cpp:600-656): layered materials along one axis.
    std::vector<std::string> material_folders;
    int interface_axis = 0;
    std::vector<double> interface_positions; // Angstrom
    std::string interface_model = "dmm";
    double interface_bin_thz = 0.1;
    double interface_amm_fraction = 0.0;
    double interface_parallel_bin = 0.05; // 1/Angstrom
    std::vector<double> mass_densities;   // kg/m^3, one per layer (AMM/MMM)
    bool has_convergence_stop = false;
    bool has_mode_excitation = false;        // mode_excitation.enabled
    double excite_fmin = -1.0, excite_fmax = -1.0, excite_multiplier = 1.0; // THz (<0: no cut), factor
};

SimConfig parse_toml_file(const std::string& path);
SimConfig parse_toml_text(const std::string& text, const std::string& input_directory);
void validate_config(const SimConfig& c);
std::string create_indexed_output_folder(const std::string& base);

} // namespace pmc
