// Phonon mode tables for the deviational Monte Carlo step.
//
// RawMaterial mirrors the reference's MaterialData (include/material/MaterialData.h:19-33).
// ModeTables holds the quantities the reference's PhononMaterial derives once at
// construction (src/PhononMaterial.cpp:31-160, src/material/Thermodynamics.cpp:71-224),
// laid out by active-mode index for upload to the device.
#pragma once
#include <array>
#include <cstdint>
#include <string>
#include <vector>

namespace pmc {

constexpr double kHbarEvPs = 6.582119569e-4;     // include/PhononMaterial.h:100
constexpr double kBoltzmannEvK = 8.617333262145e-5; // include/PhononMaterial.h:101

struct RawMaterial {
    double volume_a3 = 0;
    std::array<std::array<double, 3>, 3> reciprocal{}; // angular, 1/Angstrom, row-major
    int qpoints = 0, branches = 0;
    std::array<int, 3> mesh{};
    std::vector<double> frequency_thz;               // nq*nb
    std::vector<std::array<double, 3>> qpoint_fractions; // nq
    std::vector<std::array<double, 3>> group_velocity;   // nq*nb, Angstrom/ps
    std::vector<double> temperatures;                // nT
    std::vector<double> gamma;                       // nT*nq*nb, THz linewidth
    std::vector<double> gamma_normal, gamma_umklapp; // optional, like gamma
    std::vector<double> gamma_isotope;               // optional, nq*nb
    unsigned bz_refinement = 1;                      // copies per mesh point after first-BZ completion
};

// Parameters of the seeded synthetic Si-like table.
struct SyntheticSpec {
    int mesh = 16;
    double cell_volume_a3 = 5.431 * 5.431 * 5.431 / 4.0; // Si primitive-cell volume
    double ta_velocity = 58.0;                           // Angstrom/ps (5.8 km/s)
    double la_velocity = 84.0;                           // Angstrom/ps (8.4 km/s)
    double optical_thz[3] = {15.0, 14.0, 13.0};
    double optical_drop_thz = 0.5;  // zone-centre to zone-corner frequency drop
    double rate_a = 1.3e-45;        // s^3     (isotope-like w^4 term)
    double rate_b = 1.7e-19;        // s/K     (anharmonic w^2 T exp(-C/T) term)
    double rate_c = 137.0;          // K
    double t_max = 1000.0, t_step = 10.0;
    bool callaway_split = false;    // gamma = anharmonic; N/U = 0.6/0.4; isotope separate
};

RawMaterial make_synthetic_material(const SyntheticSpec& spec);
void validate_raw_material(const RawMaterial& raw);

struct ModeTables {
    int qpoints = 0, branches = 0, n_flat = 0, n_active = 0;
    double unit_cell_volume = 0;
    std::vector<int> active_flat;            // active -> flat (q*nb+b)
    std::vector<int> flat_to_active;         // flat -> active or -1
    std::vector<double> omega_flat;          // rad/ps, all modes
    std::vector<std::array<double, 3>> velocity_flat;
    std::vector<double> knorm_flat, window_flat;
    std::vector<std::array<double, 3>> wavevector_flat;
    std::vector<double> temps;               // gamma temperature samples
    std::vector<double> gamma_total;         // nT*n_flat: gamma + resistive extras
    std::vector<double> gamma_table;         // nT*n_flat: gamma as supplied
    std::vector<double> gamma_extra;         // nT*n_flat: isotope (and other resistive) linewidths
    std::vector<double> gamma_normal, gamma_umklapp; // optional
    bool include_stationary = false;         // active = omega > 0 instead of |v| > 1e-12
    std::vector<double> lut_t, lut_e;        // temperature lookup table
    double zero_point_density = 0;           // eV/A^3
    double energy_normalization = 1;         // max(1, Nq*Vuc)

    double omega(int active) const { return omega_flat[active_flat[active]]; }
    double bose(double temperature, int flat) const;
    double heat_capacity(double temperature, int flat) const;
    double gamma_at(double temperature, int flat) const;
    double relaxation_weight(double temperature, int flat, double dt) const;
    double crystal_energy_density(double temperature) const;
    double energy_density_from_temperature(double temperature) const;
    double temperature_from_energy_density(double e) const;
};

// include_stationary: the active bank holds every positive-frequency mode, as the
// reference's MaterialFactory requests for conservative transport
// (src/material/MaterialFactory.cpp:51-55, src/PhononMaterial.cpp:140-160).
ModeTables build_mode_tables(const RawMaterial& raw, double temperature_lookup_dt, bool include_stationary = false);

// Callaway quadrature data by active index at a fixed reference temperature:
// PhononMaterial::callaway_data (src/material/PhononMaterialCollisionData.cpp:32-165).
struct CallawayTable {
    std::vector<double> omega, capacity, normal_rate, resistive_rate; // rad/ps, eV/K, 1/ps, 1/ps
    std::vector<std::array<double, 3>> wavevector;                    // 1/Angstrom
};
CallawayTable callaway_data(const ModeTables& t, const RawMaterial& raw, double temperature);

// Mesh and first-Brillouin-zone checks of complete_callaway_bz
// (src/material/BzQuadrature.cpp:11-110) for materials with N/U linewidths.
// Returns the equal-weight refinement the reference would apply (1 when every
// q point has a unique shortest representative).
unsigned callaway_bz_refinement(const RawMaterial& raw);
// Equal-weight first-Brillouin-zone completion of an N/U table
// (complete_callaway_bz, src/material/BzQuadrature.cpp:11-161): every q point
// becomes bz_refinement points, its equally short representatives repeated
// evenly, with every per-mode array expanded alike.
RawMaterial complete_callaway_bz(const RawMaterial& raw);

// Loads a material named by io.material_folder: "synthetic:si16[:callaway]" or a
// directory holding material.pmcmat (the raw binary layout written by save_raw_material).
RawMaterial load_material(const std::string& spec, const std::string& input_directory);
void save_raw_material(const RawMaterial& raw, const std::string& path);

} // namespace pmc
