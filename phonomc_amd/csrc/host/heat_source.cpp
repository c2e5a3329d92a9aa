#include "heat_source.hpp"

#include <algorithm>
#include <cmath>
#include <limits>
#include <numeric>
#include <stdexcept>

namespace pmc {
namespace {
constexpr double kEvPerWattPs = 6.241509074e6;
constexpr double kTwoPi = 6.283185307179586476925286766559;
} // namespace

HeatSource build_heat_source(const SimConfig& c, const Domain& d, const std::vector<const ModeTables*>& materials) {
    HeatSource hs;
    if (!c.heat_source_enabled) return hs;
    const int ncell = d.cell_count();
    // Cell weights (PrescribedHeatSource::configure, src/solver/PrescribedHeatSource.cpp:22-190).
    const double density = c.heat_total_power ? 1.0 : c.heat_power_density;
    std::vector<double> w(ncell, 0.0);
    int selected = 0;
    if (c.heat_profile == HeatProfile::Uniform) {
        Vec3 lo{c.heat_min[0], c.heat_min[1], c.heat_min[2]}, hi{c.heat_max[0], c.heat_max[1], c.heat_max[2]};
        for (int k = 0; k < 3; ++k)
            if (lo[k] > hi[k]) std::swap(lo[k], hi[k]);
        for (int i = 0; i < ncell; ++i) {
            const Vec3& p = d.cell_center[i];
            if (p[0] >= lo[0] && p[0] <= hi[0] && p[1] >= lo[1] && p[1] <= hi[1] && p[2] >= lo[2] && p[2] <= hi[2]) {
                w[i] = 1.0;
                ++selected;
            }
        }
        if (selected == 0) throw std::runtime_error("uniform heat source region does not include any grid center.");
    } else {
        std::vector<Vec3> centers;
        if (!c.heat_centers.empty()) {
            for (size_t j = 0; j + 2 < c.heat_centers.size(); j += 3)
                centers.push_back({c.heat_centers[j], c.heat_centers[j + 1], c.heat_centers[j + 2]});
        } else {
            centers.push_back({c.heat_center[0], c.heat_center[1], c.heat_center[2]});
        }
        std::vector<double> peak(centers.size(), 1.0);
        if (!c.heat_power_densities.empty())
            for (size_t j = 0; j < centers.size(); ++j) peak[j] = c.heat_power_densities[j] / density;
        Vec3 sigma{1.0, 1.0, 1.0}, lo{}, hi{};
        bool gauss_axis[3];
        double half[3] = {std::numeric_limits<double>::infinity(), std::numeric_limits<double>::infinity(),
                          std::numeric_limits<double>::infinity()};
        const bool clip = c.heat_min.size() == 3 && c.heat_max.size() == 3;
        for (int k = 0; k < 3; ++k) {
            if (clip) {
                lo[k] = std::min(c.heat_min[k], c.heat_max[k]);
                hi[k] = std::max(c.heat_min[k], c.heat_max[k]);
            }
            if (c.heat_half_width.size() == 3 && std::isfinite(c.heat_half_width[k]) && c.heat_half_width[k] > 0.0)
                half[k] = c.heat_half_width[k];
            gauss_axis[k] = std::isfinite(c.heat_sigma[k]) && c.heat_sigma[k] > 0.0;
            sigma[k] = gauss_axis[k] ? std::max(1e-12, c.heat_sigma[k]) : 1.0; // sigma <= 0: uniform axis
        }
        for (int i = 0; i < ncell; ++i) {
            const Vec3& p = d.cell_center[i];
            if (clip && (p[0] < lo[0] || p[0] > hi[0] || p[1] < lo[1] || p[1] > hi[1] || p[2] < lo[2] || p[2] > hi[2]))
                continue;
            double sum = 0.0;
            for (size_t s = 0; s < centers.size(); ++s) {
                double r2 = 0.0;
                bool outside = false;
                for (int k = 0; k < 3 && !outside; ++k) {
                    if (gauss_axis[k]) {
                        const double u = std::abs(p[k] - centers[s][k]) / sigma[k];
                        if (u > 2.0) outside = true; // the profile is cut at two widths
                        r2 += u * u;
                    } else if (std::isfinite(half[k]) && std::abs(p[k] - centers[s][k]) > half[k]) {
                        outside = true;
                    }
                }
                if (!outside) sum += peak[s] * std::exp(-0.5 * r2);
            }
            if (sum > 0.0) {
                w[i] = sum;
                ++selected;
            }
        }
        if (selected == 0) throw std::runtime_error("gaussian heat source cannot be applied because grid is empty.");
    }
    double weighted_volume = 0.0;
    for (int i = 0; i < ncell; ++i) weighted_volume += w[i] * d.cell_volume[i];
    if (!(weighted_volume > 0.0) || !std::isfinite(weighted_volume))
        throw std::runtime_error("Heat source has no finite positive sampled volume; refine the grid or source region.");
    hs.rates.resize(ncell);
    for (int i = 0; i < ncell; ++i) {
        // W/m^3 * A^3 -> eV/ps: 1e-30 m^3/A^3 * 6.241509074e6 eV/(W ps).
        hs.rates[i] = c.heat_total_power ? *c.heat_total_power * kEvPerWattPs * (w[i] * d.cell_volume[i] / weighted_volume)
                                         : density * 6.241509074e-24 * w[i] * d.cell_volume[i];
        if (!std::isfinite(hs.rates[i]) || hs.rates[i] < 0.0) throw std::runtime_error("Non-finite heat-source cell power.");
    }
    const double total = std::accumulate(hs.rates.begin(), hs.rates.end(), 0.0);
    if (!std::isfinite(total) || !(total > 0.0))
        throw std::runtime_error("Total heat-source rate must be finite and representably positive.");
    // Modes the thermal spectrum deposits into (PhononSourceOperator::mode_weight,
    // src/solver/PhononSourceOperator.cpp:105-118): frequency window and branch filter.
    // PhononSourceOperator::configure (:20-103): branch indices and weights per
    // material, weights over the largest; the Gaussian's log weights over the
    // positive-frequency bank.
    std::vector<double> branch_weight = c.heat_branch_weights;
    if (!branch_weight.empty()) {
        const double scale = *std::max_element(branch_weight.begin(), branch_weight.end());
        for (double& w : branch_weight) w /= scale;
    }
    hs.spectrum = c.heat_spectrum;
    if (c.heat_spectrum == 2) {
        hs.frequency_center = *c.heat_frequency_center;
        hs.frequency_sigma = *c.heat_frequency_sigma;
    }
    hs.mode_eligible.clear();
    hs.mode_weight.clear();
    // One material after the other, in the order of the global active bank.
    for (const ModeTables* mp : materials) {
        const ModeTables& m = *mp;
        for (int b : c.heat_branches)
            if (b >= m.branches) throw std::invalid_argument("heat_source.branches index exceeds a material's branch count.");
        if (c.heat_spectrum == 1 && branch_weight.size() != static_cast<size_t>(m.branches))
            throw std::invalid_argument("heat_source.branch_weights must match each material's branch count.");
        const size_t base = hs.mode_eligible.size();
        hs.mode_eligible.resize(base + m.n_active, 1);
        hs.mode_weight.resize(base + m.n_active, 0.0);
        for (int a = 0; a < m.n_active; ++a) {
            const int f = m.active_flat[a];
            const double thz = m.omega_flat[f] / kTwoPi;
            bool ok = std::isfinite(thz) && !(c.heat_frequency_min >= 0.0 && thz < c.heat_frequency_min) &&
                      !(c.heat_frequency_max >= 0.0 && thz > c.heat_frequency_max);
            if (ok && !c.heat_branches.empty())
                ok = std::find(c.heat_branches.begin(), c.heat_branches.end(), f % m.branches) != c.heat_branches.end();
            double weight = ok ? 1.0 : 0.0;
            if (ok && c.heat_spectrum == 1) weight = branch_weight[static_cast<size_t>(f % m.branches)];
            hs.mode_eligible[base + a] = weight > 0.0 ? 1 : 0;
            hs.mode_weight[base + a] = c.heat_spectrum == 1 ? weight : 0.0;
        }
        if (c.heat_spectrum == 2) {
            size_t positive = 0;
            for (double w : m.omega_flat)
                if (w > 0) ++positive;
            if (positive != static_cast<size_t>(m.n_active))
                throw std::invalid_argument("Gaussian source requires all positive-frequency modes, including stationary modes.");
            bool any = false;
            for (int a = 0; a < m.n_active; ++a) {
                if (!hs.mode_eligible[base + a]) continue;
                const long double f = m.omega_flat[m.active_flat[a]] / 6.283185307179586476925286766559L;
                const long double z = (f - hs.frequency_center) / hs.frequency_sigma;
                hs.mode_weight[base + a] = static_cast<double>(-0.5L * z * z);
                any = true;
            }
            if (!any)
                throw std::invalid_argument("Gaussian source branch/frequency filters select no positive-frequency modes.");
        }
    }
    hs.enabled = true;
    hs.selected_cells = selected;
    hs.square = c.heat_time_square;
    hs.start = c.heat_time_start;
    hs.end = c.heat_time_end;
    hs.period = c.heat_period;
    hs.on_duration = c.heat_on_duration;
    hs.duty_cycle = c.heat_duty_cycle;
    hs.amplitude = c.heat_amplitude;
    return hs;
}

double HeatSource::integrated_time_factor(double t0, double t1) const {
    // MonteCarloSolver::local_heat_source_integrated_time_factor (src/solver/HeatSourceDriving.cpp:14-46).
    const double b = std::max(t0, start);
    const double e = end >= 0.0 ? std::min(t1, end) : t1;
    if (e <= b) return 0.0;
    if (!square) return amplitude * (e - b);
    const double on = on_duration < 0.0 ? duty_cycle * period : on_duration;
    if (on <= 0.0) return 0.0;
    auto on_time = [&](double rel) {
        if (!(rel > 0.0)) return 0.0;
        const double cycles = std::floor(rel / period);
        return cycles * on + std::min(rel - cycles * period, on);
    };
    return amplitude * std::max(0.0, on_time(e - start) - on_time(b - start));
}

double HeatSource::base_power_w() const { return std::accumulate(rates.begin(), rates.end(), 0.0) / kEvPerWattPs; }

} // namespace pmc

namespace pmc {
HeatSource build_heat_source(const SimConfig& c, const Domain& d, const ModeTables& m) {
    return build_heat_source(c, d, std::vector<const ModeTables*>{&m});
}
} // namespace pmc
