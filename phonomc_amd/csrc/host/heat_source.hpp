// Prescribed lattice heat source: per-cell power from the TOML profile and the
// integrated on-time of the time profile (src/solver/PrescribedHeatSource.cpp:22-230,
// src/solver/HeatSourceDriving.cpp:14-46).
#pragma once
#include "config.hpp"
#include "domain.hpp"
#include "material.hpp"

#include <vector>

namespace pmc {

struct HeatSource {
    bool enabled = false;
    std::vector<double> rates;     // eV/ps per cell
    std::vector<int> mode_eligible; // per active mode: the spectrum's frequency/branch filter (and weight > 0)
    // Spectrum (PhononSourceOperator, src/solver/PhononSourceOperator.cpp:20-118):
    // 0 thermal, 1 weighted (mode_weight = the branch weight over the largest),
    // 2 gaussian (mode_weight = -((f - f_c) / sigma)^2 / 2, f = omega / 2 pi in THz).
    int spectrum = 0;
    std::vector<double> mode_weight;
    double frequency_center = 0.0, frequency_sigma = 1.0; // THz, gaussian
    int selected_cells = 0;
    bool square = false;
    double start = 0.0, end = -1.0, period = 0.0, on_duration = -1.0, duty_cycle = 0.5, amplitude = 1.0;

    // Integrated time factor (ps) of the drive over [t0, t1).
    double integrated_time_factor(double t0, double t1) const;
    double base_power_w() const;
};

HeatSource build_heat_source(const SimConfig& c, const Domain& d, const ModeTables& m);
// Layered materials: every material's per-mode arrays, concatenated in the order
// of the global active bank, each material checked as PhononSourceOperator::configure
// checks it (src/solver/PhononSourceOperator.cpp:41-101).
HeatSource build_heat_source(const SimConfig& c, const Domain& d, const std::vector<const ModeTables*>& materials);

} // namespace pmc
