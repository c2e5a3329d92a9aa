// Host facade with the surface of the reference's MonteCarloSolver
// (include/MonteCarloSolver.h:42-103): builds every immutable table once on the
// host, hands them to the device context through the C ABI (pmc_capi.h) and
// advances whole time steps there.  Observables and result files follow the
// reference cadence and formats (src/solver/Diagnostics.cpp:200-293,
// src/solver/ResultWriter.cpp:25-119).
#pragma once
#include "amm.hpp"
#include "boundary_tables.hpp"
#include "heat_source.hpp"
#include "config.hpp"
#include "domain.hpp"
#include "material.hpp"
#include "pmc_capi.h"

#include <array>
#include <memory>
#include <ostream>
#include <string>
#include <vector>

namespace pmc {

// Everything the device needs, built on the host from one configuration.
struct Problem {
    SimConfig config;
    Domain domain;
    RawMaterial raw;
    ModeTables modes;
    std::vector<RoughTable> rough;
    std::vector<ReservoirTable> reservoirs;
    std::vector<int> facet_to_rough, facet_to_reservoir;
    double particle_volume = 0;   // V / N
    double particle_density = 0;  // N / V
    double energy_weight = 0;
    double background_energy_density = 0;
    std::vector<double> neq_background; // per active mode
    std::vector<double> omega_active, velocity_active; // per active mode (3 per mode for velocity)
    HeatSource heat;
    pmc_heat_source heat_view{};
    // Linearized Callaway transport (scattering.model = 'callaway').
    bool linearized = false;
    CallawayTable callaway;
    std::vector<double> callaway_wavevector; // 3 per active mode
    double heat_capacity_density = 0;
    pmc_callaway_table callaway_view{};
    // Conservative transport (strict-elastic walls, fixed incoming flux,
    // stationary modes in the bank): linearized Callaway, nonlinear Callaway or
    // simulation.conservative_boundaries.
    bool conservative = false;
    // Nonlinear Callaway (scattering.callaway_formulation = 'nonlinear').
    bool nonlinear = false;
    struct NonlinearRates {
        std::vector<double> temps, gamma_normal, gamma_umklapp, gamma_extra, wavevector; // by (T sample, active)
    } nl_rates;
    pmc_callaway_rates nl_view{};
    // Layered materials behind pmc_layers (materials.folders with more than one entry).
    struct LayerArrays {
        int count = 1;
        std::vector<double> planes, background_energy_density, spatial_volume;
        std::vector<int32_t> mode_begin, cell_material, lut_begin, bank_begin, bank_mode;
        std::vector<double> bank_omega, bank_prefix, lut_t, lut_e, gamma_total;
        std::vector<int32_t> active_flat;
        // AMM rows (pmc_layers.amm_*), global active indices; audits per interface
        double amm_fraction = 0;
        std::vector<int32_t> amm_block, amm_reflected, amm_edge_begin, amm_edge_mode;
        std::vector<double> amm_edge_prob;
        std::vector<AmmAudit> amm_audit;
        // Reflection table of facet f for material k at k * n_facets + f (pmc_layers.facet_rough)
        std::vector<int32_t> facet_rough;
        // Linearized transport: heat-capacity density C_bg of every material (pmc_layers.heat_capacity_density)
        std::vector<double> heat_capacity_density;
    } layers;
    pmc_layers layers_view{};
    std::vector<const ModeTables*> layer_modes;            // per material (layered problems)
    std::vector<const RawMaterial*> layer_raw;
    std::vector<std::shared_ptr<const void>> layer_materials; // keeps them alive
    // Mesh model arrays behind pmc_mesh (geometry kind 1).
    struct MeshArrays {
        std::vector<double> vertices, face_normal, face_k, face_bounds;
        std::vector<int32_t> faces, face_facet;
        Bvh face_bvh, simplex_bvh;
        std::vector<double> simplex_vertices, simplex_inverse, simplex_bounds, simplex_volume;
        DistanceField field;
    } mesh;
    pmc_mesh mesh_view{};
    // C-ABI views (pointers into the vectors above)
    std::vector<pmc_facet> facets;
    std::vector<pmc_rough_table> rough_views;
    std::vector<pmc_reservoir> reservoir_views;
    pmc_mode_table mode_view{};
    pmc_geometry geometry_view{};
    pmc_setup setup{};
    pmc_init init{};
};

struct ProblemOptions {
    int mode_sampling = 0;       // 0 alias, 1 inverse CDF
    double capacity_factor = 1.25;
};

std::unique_ptr<Problem> build_problem(const SimConfig& config, const ProblemOptions& options = {});

class MonteCarloSolver {
public:
    MonteCarloSolver(std::unique_ptr<Problem> problem, std::string output_folder, std::ostream* log = nullptr);
    ~MonteCarloSolver();
    MonteCarloSolver(const MonteCarloSolver&) = delete;
    MonteCarloSolver& operator=(const MonteCarloSolver&) = delete;

    void run_timestep();
    // Advances n steps, recording observables at the reference cadence.
    void run_steps(int n);
    int current_timestep() const { return step_; }
    bool converged() const { return false; }
    long long particle_count() const;
    void append_profile_summary(std::ostream& out) const;

    const Problem& problem() const { return *problem_; }
    pmc_ctx* context() const { return ctx_; }
    std::vector<double> cell_temperatures() const;
    double kappa_eff() const { return kappa_eff_; }
    double kappa_int() const { return kappa_int_; }
    double average_heat_flux() const { return mean_flux_; }
    pmc_ledger ledger() const;
    double step_seconds() const { return step_seconds_; }

private:
    void throw_if_failed(int rc) const;
    void record_observables();
    void write_convergence_header();
    void append_convergence_row();
    double source_prescribed_ev(const pmc_ledger& l) const;
    void write_source_ledger(const pmc_ledger& l) const;
    void append_source_summary(std::ostream& out, const pmc_ledger& l) const;

    std::unique_ptr<Problem> problem_;
    std::string folder_;
    std::ostream* log_;
    pmc_ctx* ctx_ = nullptr;
    int step_ = 0;
    double kappa_eff_ = 0, kappa_int_ = 0, mean_flux_ = 0;
    std::vector<double> cell_T_;
    std::vector<std::array<double, 3>> flux_;
    double step_seconds_ = 0;
};

} // namespace pmc
