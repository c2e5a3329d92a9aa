#include "solver.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <filesystem>
#include <fstream>
#include <iomanip>
#include <limits>
#include <map>
#include <mutex>
#include <numeric>
#include <stdexcept>
#include <string>

namespace pmc {

namespace {
// Material tables depend only on the material, the lookup spacing and the
// stationary-mode policy; the temperature lookup table alone takes about half a
// second.  Problems built in one process (a batch of instances) share them.
struct MaterialEntry {
    RawMaterial raw;       // as loaded (the material view)
    RawMaterial completed; // N/U tables: equal-weight first-BZ completion, which the mode tables use
    ModeTables modes;
};

std::shared_ptr<const MaterialEntry> cached_material(const std::string& spec, const std::string& input_directory,
                                                     double lookup_dt, bool include_stationary) {
    namespace fs = std::filesystem;
    std::string key = spec + '\n' + std::to_string(lookup_dt) + '\n' + (include_stationary ? "1" : "0");
    if (spec.rfind("synthetic:", 0) != 0) { // a file: its path and modification time
        fs::path path(spec);
        if (path.is_relative()) path = fs::path(input_directory) / path;
        std::error_code ec;
        const auto stamp = fs::last_write_time(fs::is_directory(path, ec) ? path / "material.pmcmat" : path, ec);
        key += '\n' + fs::absolute(path, ec).string() + '\n' +
               std::to_string(ec ? 0 : static_cast<long long>(stamp.time_since_epoch().count()));
    }
    static std::mutex lock;
    static std::map<std::string, std::shared_ptr<const MaterialEntry>> cache;
    {
        std::lock_guard<std::mutex> g(lock);
        auto it = cache.find(key);
        if (it != cache.end()) return it->second;
    }
    auto entry = std::make_shared<MaterialEntry>();
    entry->raw = load_material(spec, input_directory);
    // PhononMaterial completes an N/U table at construction (src/PhononMaterial.cpp:41-43).
    entry->completed = entry->raw.gamma_normal.empty() ? entry->raw : complete_callaway_bz(entry->raw);
    entry->modes = build_mode_tables(entry->completed, lookup_dt, include_stationary);
    std::lock_guard<std::mutex> g(lock);
    if (cache.size() >= 16) cache.clear(); // bounded: a process rarely uses more materials
    return cache.emplace(key, std::move(entry)).first->second;
}

// Layered materials (src/solver/LayeredMaterialLayout.cpp:11-115,
// src/MonteCarloSolver.cpp:81-100, src/solver/InterfaceModeSampler.cpp:11-44):
// one global active-mode bank over the materials in layer order, cell
// materials, the common energy weight, per-material background energies and
// lookup tables, and the DMM banks.
void build_layers(Problem& p) {
    const SimConfig& c = p.config;
    const Domain& d = p.domain;
    auto& L = p.layers;
    L.count = static_cast<int>(c.material_folders.size());
    const int axis = c.interface_axis;
    if (!d.is_box) throw std::runtime_error("Layered materials on a mesh model are not supported by this build.");
    if (c.interface_positions.size() != static_cast<size_t>(L.count - 1))
        throw std::invalid_argument("Invalid layered material count, axis or interface count.");
    L.planes = c.interface_positions;
    for (size_t i = 0; i < L.planes.size(); ++i)
        if (!std::isfinite(L.planes[i]) || (i && L.planes[i] <= L.planes[i - 1]))
            throw std::invalid_argument("Material interface positions must be finite and strictly increasing.");
    auto material_at = [&](const Vec3& x) {
        return static_cast<int>(std::upper_bound(L.planes.begin(), L.planes.end(), x[axis]) - L.planes.begin());
    };
    // LayeredMaterialLayout::validated_cell_materials (:62-113).
    const double lo = d.bmin[axis], hi = d.bmax[axis];
    const double width = (hi - lo) / static_cast<double>(c.grid[axis]);
    const double tolerance = std::max(1e-10, 1e-10 * std::abs(hi - lo));
    for (double plane : L.planes) {
        if (!std::isfinite(plane) || !(plane > lo) || !(plane < hi))
            throw std::runtime_error("Each material interface must lie strictly inside the loaded geometry bounds.");
        const double cc = (plane - lo) / width;
        if (std::abs(cc - std::round(cc)) > tolerance / width)
            throw std::runtime_error("Each materials.interface_positions value must coincide with a grid-cell boundary.");
    }
    L.cell_material.resize(d.cell_count());
    std::vector<bool> occupied(L.count, false);
    for (int cell = 0; cell < d.cell_count(); ++cell) occupied[L.cell_material[cell] = material_at(d.cell_center[cell])] = true;
    if (std::find(occupied.begin(), occupied.end(), false) != occupied.end())
        throw std::runtime_error("Every material layer must contain at least one active geometry cell.");
    for (const auto& f : d.facets) {
        if (f.condition == 'P' && std::abs(f.normal[axis]) > 0.99)
            throw std::runtime_error("Periodic boundaries normal to materials.interface_axis are not supported, because "
                                     "they would join different end materials.");
        if (f.condition == 'T' && std::abs(f.normal[axis]) < 0.99)
            throw std::runtime_error("In the layered-material model, thermal-reservoir facets must be normal to "
                                     "materials.interface_axis so each reservoir touches one material.");
    }
    // Materials: equal temperature grids, one global bank.
    std::vector<std::shared_ptr<const MaterialEntry>> mats;
    for (const auto& f : c.material_folders)
        mats.push_back(cached_material(f, c.input_directory, c.temperature_lookup_dt,
                                       p.conservative || (c.heat_source_enabled && c.heat_spectrum == 2)));
    for (const auto& e : mats)
        if (e->modes.temps != mats[0]->modes.temps)
            throw std::runtime_error("Layered materials need equal lifetime temperature grids in this build.");
    std::vector<double> density(L.count);
    for (int k = 0; k < L.count; ++k) {
        const auto& t = mats[k]->modes;
        density[k] = t.n_active / t.energy_normalization;
        if (!(density[k] > 0.0) || !std::isfinite(density[k])) throw std::runtime_error("Invalid material phase-space density.");
    }
    double states = 0.0;
    for (int cell = 0; cell < d.cell_count(); ++cell) states += d.cell_volume[cell] * density[L.cell_material[cell]];
    const double common = states / std::llround(c.particle_count);
    p.energy_weight = common;
    L.spatial_volume.resize(L.count);
    for (int k = 0; k < L.count; ++k) L.spatial_volume[k] = common / density[k];
    p.particle_volume = L.spatial_volume[0];
    p.particle_density = 1.0 / L.spatial_volume[0];
    const auto& temps = mats[0]->modes.temps;
    const int nt = static_cast<int>(temps.size());
    p.omega_active.clear();
    p.neq_background.clear();
    p.velocity_active.clear();
    L.mode_begin.assign(1, 0);
    L.lut_begin.assign(1, 0);
    std::vector<std::vector<double>> gamma_cols;
    for (int k = 0; k < L.count; ++k) {
        const auto& t = mats[k]->modes;
        for (int a = 0; a < t.n_active; ++a) {
            const int f = t.active_flat[a];
            p.omega_active.push_back(t.omega_flat[f]);
            p.neq_background.push_back(t.bose(c.background_temperature, f));
            for (int q = 0; q < 3; ++q) p.velocity_active.push_back(t.velocity_flat[f][q]);
        }
        L.mode_begin.push_back(L.mode_begin.back() + t.n_active);
        L.background_energy_density.push_back(t.energy_density_from_temperature(c.background_temperature));
        L.lut_t.insert(L.lut_t.end(), t.lut_t.begin(), t.lut_t.end());
        L.lut_e.insert(L.lut_e.end(), t.lut_e.begin(), t.lut_e.end());
        L.lut_begin.push_back(static_cast<int32_t>(L.lut_t.size()));
    }
    const int total = L.mode_begin.back();
    L.active_flat.resize(total);
    std::iota(L.active_flat.begin(), L.active_flat.end(), 0);
    L.gamma_total.assign(static_cast<size_t>(nt) * total, 0.0);
    for (int k = 0; k < L.count; ++k) {
        const auto& t = mats[k]->modes;
        for (int a = 0; a < t.n_active; ++a)
            for (int r = 0; r < nt; ++r)
                L.gamma_total[static_cast<size_t>(r) * total + L.mode_begin[k] + a] =
                    t.gamma_total[static_cast<size_t>(r) * t.n_flat + t.active_flat[a]];
    }
    p.background_energy_density = L.background_energy_density[0];
    // Linearized Callaway (CarrierCollisionOperator::configure per material,
    // src/solver/CarrierCollisionOperator.cpp:13-36; BackgroundCache per material,
    // include/solver/BackgroundCache.h:13-30): each material's callaway_data at T_bg,
    // concatenated over the global bank, its exact crystal energy density as E_bg
    // and its heat-capacity density C_bg.
    L.heat_capacity_density.clear();
    if (p.linearized) {
        const double tb = c.background_temperature;
        p.callaway = CallawayTable{};
        for (int k = 0; k < L.count; ++k) {
            const auto& t = mats[k]->modes;
            double hc = 0.0;
            for (int q = 0; q < t.qpoints; ++q)
                for (int b = 0; b < t.branches; ++b) hc += t.heat_capacity(tb, q * t.branches + b) / t.energy_normalization;
            if (!(hc > 0.0) || !std::isfinite(hc))
                throw std::invalid_argument("Linearized temperature requires positive heat capacity.");
            L.heat_capacity_density.push_back(hc);
            L.background_energy_density[k] = t.crystal_energy_density(tb);
            const CallawayTable ct = callaway_data(t, mats[k]->completed, tb);
            auto append = [](auto& dst, const auto& src) { dst.insert(dst.end(), src.begin(), src.end()); };
            append(p.callaway.omega, ct.omega);
            append(p.callaway.capacity, ct.capacity);
            append(p.callaway.normal_rate, ct.normal_rate);
            append(p.callaway.resistive_rate, ct.resistive_rate);
            append(p.callaway.wavevector, ct.wavevector);
        }
        p.heat_capacity_density = L.heat_capacity_density[0];
        p.background_energy_density = L.background_energy_density[0];
        const size_t total = p.callaway.omega.size();
        p.callaway_wavevector.resize(3 * total);
        for (size_t a = 0; a < total; ++a)
            for (int k = 0; k < 3; ++k) p.callaway_wavevector[3 * a + k] = p.callaway.wavevector[a][k];
        auto& cv = p.callaway_view;
        cv.capacity = p.callaway.capacity.data();
        cv.wavevector = p.callaway_wavevector.data();
        cv.normal_rate = p.callaway.normal_rate.data();
        cv.resistive_rate = p.callaway.resistive_rate.data();
        cv.heat_capacity_density = p.heat_capacity_density;
        cv.background_energy_density = p.background_energy_density;
    }
    // Nonlinear Callaway (NonlinearCallawayCollision::configure per material,
    // src/solver/NonlinearCallawayCollision.cpp:287-301): every material's N/U split
    // validated through callaway_data at T_bg, the rate rows and Cartesian q of
    // each material on the common temperature grid, over the global bank.
    if (p.nonlinear) {
        const int total = L.mode_begin.back();
        auto& r = p.nl_rates;
        r.temps = temps;
        r.gamma_normal.assign(static_cast<size_t>(nt) * total, 0.0);
        r.gamma_umklapp.assign(static_cast<size_t>(nt) * total, 0.0);
        r.gamma_extra.assign(static_cast<size_t>(nt) * total, 0.0);
        r.wavevector.assign(3 * static_cast<size_t>(total), 0.0);
        for (int k = 0; k < L.count; ++k) {
            const auto& t = mats[k]->modes;
            const CallawayTable checked = callaway_data(t, mats[k]->completed, c.background_temperature);
            const size_t nf = static_cast<size_t>(t.n_flat);
            for (int a = 0; a < t.n_active; ++a) {
                const size_t g = static_cast<size_t>(L.mode_begin[k] + a);
                for (int row = 0; row < nt; ++row) {
                    const size_t f = static_cast<size_t>(row) * nf + static_cast<size_t>(t.active_flat[a]);
                    r.gamma_normal[static_cast<size_t>(row) * total + g] = t.gamma_normal[f];
                    r.gamma_umklapp[static_cast<size_t>(row) * total + g] = t.gamma_umklapp[f];
                    r.gamma_extra[static_cast<size_t>(row) * total + g] = t.gamma_extra[f];
                }
                for (int q = 0; q < 3; ++q) r.wavevector[3 * g + q] = checked.wavevector[a][q];
            }
        }
        auto& v = p.nl_view;
        v.n_temps = nt;
        v.temps = r.temps.data();
        v.gamma_normal = r.gamma_normal.data();
        v.gamma_umklapp = r.gamma_umklapp.data();
        v.gamma_extra = r.gamma_extra.data();
        v.wavevector = r.wavevector.data();
    }
    // DMM banks: per material and direction along the axis, modes sorted by
    // frequency with the cumulative |v_axis| / normalization (InterfaceModeSampler::configure).
    L.bank_begin.assign(1, 0);
    for (int k = 0; k < L.count; ++k) {
        const auto& t = mats[k]->modes;
        for (int dir = 0; dir < 2; ++dir) {
            std::vector<std::pair<double, int>> ordered;
            for (int a = 0; a < t.n_active; ++a) {
                const double comp = t.velocity_flat[t.active_flat[a]][axis];
                if ((dir == 0 && comp < 0.0) || (dir == 1 && comp > 0.0)) ordered.push_back({t.omega_flat[t.active_flat[a]], a});
            }
            std::sort(ordered.begin(), ordered.end());
            double acc = 0.0;
            L.bank_prefix.push_back(0.0);
            for (const auto& o : ordered) {
                L.bank_omega.push_back(o.first);
                L.bank_mode.push_back(L.mode_begin[k] + o.second);
                acc += std::abs(t.velocity_flat[t.active_flat[o.second]][axis]) / t.energy_normalization;
                L.bank_prefix.push_back(acc);
            }
            L.bank_begin.push_back(static_cast<int32_t>(L.bank_mode.size()));
        }
    }
    p.layer_modes.clear();
    p.layer_raw.clear();
    for (const auto& e : mats) {
        p.layer_modes.push_back(&e->modes);
        p.layer_raw.push_back(&e->raw);
    }
    p.layer_materials.assign(mats.begin(), mats.end());
    // Acoustic/mixed mismatch (InterfaceInteraction::configure,
    // src/solver/InterfaceInteraction.cpp:11-44): the AMM rows, in global
    // active indices, blocked by (interface, side).
    L.amm_fraction = c.interface_model == "amm" ? 1.0 : c.interface_model == "mmm" ? c.interface_amm_fraction : 0.0;
    L.amm_block.assign(1, 0);
    L.amm_reflected.clear();
    L.amm_edge_begin.clear();
    L.amm_edge_mode.clear();
    L.amm_edge_prob.clear();
    L.amm_audit.clear();
    if (L.amm_fraction > 0) {
        const auto tables = build_amm_tables(p.layer_modes, p.layer_raw, axis, c.mass_densities, c.interface_bin_thz,
                                             c.interface_parallel_bin);
        for (size_t f = 0; f < tables.size(); ++f) {
            const AmmInterface& t = tables[f];
            for (int side = 0; side < 2; ++side) {
                const int src = static_cast<int>(f) + side, dst = static_cast<int>(f) + 1 - side;
                const int n = L.mode_begin[src + 1] - L.mode_begin[src];
                const int e0 = static_cast<int>(L.amm_edge_mode.size());
                for (int i = 0; i < n; ++i) {
                    const int r = t.reflected[side][i];
                    L.amm_reflected.push_back(r < 0 ? -1 : L.mode_begin[src] + r);
                    L.amm_edge_begin.push_back(e0 + t.edge_begin[side][i]);
                }
                L.amm_reflected.push_back(-1);
                L.amm_edge_begin.push_back(e0 + t.edge_begin[side][n]);
                for (size_t e = 0; e < t.edge_mode[side].size(); ++e) {
                    L.amm_edge_mode.push_back(L.mode_begin[dst] + t.edge_mode[side][e]);
                    L.amm_edge_prob.push_back(t.edge_prob[side][e]);
                }
                L.amm_block.push_back(static_cast<int32_t>(L.amm_reflected.size()));
            }
            L.amm_audit.push_back(t.audit);
        }
    }
}

// Reflection tables of a layered box: RoughBoundarySampler builds one table per
// (material, facet) from that material's own modes (src/solver/RoughBoundarySampler.cpp:77-368,
// table_index :68-74).  Each material's tables are moved into the global bank:
// mode indices shifted by the material's first active mode, per-mode rows
// widened to every active mode (rows of other materials' modes have no
// partner or window; no carrier of theirs meets this table).
void layered_rough_tables(Problem& p, const std::vector<bool>& enabled, double strict_temperature) {
    auto& L = p.layers;
    const Domain& d = p.domain;
    const int nf = static_cast<int>(d.facets.size());
    const int total = L.mode_begin.back();
    p.rough.clear();
    L.facet_rough.assign(static_cast<size_t>(L.count) * nf, -1);
    for (int k = 0; k < L.count; ++k) {
        std::vector<int> f2r;
        auto tables = build_rough_tables(d, *p.layer_modes[k], enabled, strict_temperature, f2r);
        const int b = L.mode_begin[k], n = L.mode_begin[k + 1] - b;
        for (int f = 0; f < nf; ++f)
            if (f2r[f] >= 0) L.facet_rough[static_cast<size_t>(k) * nf + f] = static_cast<int32_t>(p.rough.size()) + f2r[f];
        for (auto& t : tables) {
            RoughTable g;
            g.facet = t.facet;
            g.strict_elastic = t.strict_elastic;
            g.mode_begin = b;
            g.mode_count = n;
            g.specularity.assign(total, 0.0);
            g.spec_match.assign(total, -1);
            g.diffuse_begin.assign(total, -1);
            g.diffuse_end.assign(total, -1);
            for (int a = 0; a < n; ++a) {
                g.specularity[b + a] = t.specularity[a];
                g.spec_match[b + a] = t.spec_match[a] < 0 ? -1 : b + t.spec_match[a];
                g.diffuse_begin[b + a] = t.diffuse_begin[a];
                g.diffuse_end[b + a] = t.diffuse_end[a];
            }
            for (int& x : t.outgoing_active) x += b;
            for (int& x : t.outgoing_sorted) x += b;
            for (int& x : t.roulette_active) x += b;
            g.outgoing_active = std::move(t.outgoing_active);
            g.outgoing_sorted = std::move(t.outgoing_sorted);
            g.residual_prefix = std::move(t.residual_prefix);
            g.roulette_active = std::move(t.roulette_active);
            g.roulette_cdf = std::move(t.roulette_cdf);
            p.rough.push_back(std::move(g));
        }
    }
    // the facet's own entry names material 0's table
    for (int f = 0; f < nf; ++f) p.facet_to_rough[f] = L.facet_rough[f];
}

// Reservoirs of a layered box: each reservoir facet touches one material
// (validated), so its table is the material's, with that material's carrier
// density, placed in the global bank.
void layered_reservoirs(Problem& p) {
    const auto& L = p.layers;
    const SimConfig& c = p.config;
    const Domain& d = p.domain;
    const int total = L.mode_begin.back();
    p.facet_to_reservoir.assign(d.facets.size(), -1);
    p.reservoirs.clear();
    for (int k = 0; k < L.count; ++k) {
        std::vector<int> f2r;
        auto tables = build_reservoir_tables(d, *p.layer_modes[k], 1.0 / L.spatial_volume[k], c.time_step, f2r);
        for (auto& r : tables) {
            const Vec3& centre = d.facets[r.facet].centroid;
            const int mat = static_cast<int>(std::upper_bound(L.planes.begin(), L.planes.end(),
                                                              centre[c.interface_axis]) - L.planes.begin());
            if (mat != k) continue;
            const int b = L.mode_begin[k], n = L.mode_begin[k + 1] - b;
            std::vector<double> cdf(total, 0.0), w(total, 0.0), occ(total, 0.0);
            for (int j = 0; j < total; ++j) cdf[j] = j < b ? 0.0 : (j >= b + n ? 1.0 : r.cdf[j - b]);
            for (int a = 0; a < n; ++a) {
                w[b + a] = r.cdf[a] - (a ? r.cdf[a - 1] : 0.0);
                occ[b + a] = r.occupation[a];
            }
            r.cdf = std::move(cdf);
            r.occupation = std::move(occ);
            build_alias(w, r.alias_prob, r.alias_index);
            p.facet_to_reservoir[r.facet] = static_cast<int>(p.reservoirs.size());
            p.reservoirs.push_back(std::move(r));
        }
    }
}
} // namespace

std::unique_ptr<Problem> build_problem(const SimConfig& c, const ProblemOptions& opt) {
    // Transport policies (include/SimulationConfig.h:148-156): linearized Callaway
    // is linearized, conservative transport; nonlinear Callaway and
    // simulation.conservative_boundaries are conservative transport with Bose
    // occupations.
    const bool nonlinear = c.collision == CollisionModel::Callaway && c.callaway_nonlinear;
    const bool linearized = c.collision == CollisionModel::Callaway && !nonlinear;
    const bool conservative = linearized || nonlinear || c.conservative_boundaries;
    if (c.collision == CollisionModel::Callaway &&
        (!std::isfinite(c.background_temperature) || c.background_temperature <= 0))
        throw std::runtime_error("Callaway requires a positive fixed background_temperature.");
    if (c.has_gradient_drive) throw std::runtime_error("driving.temperature_gradient is not supported by this build.");
    const bool layered = c.material_folders.size() > 1;
    if (layered) {
        if (!std::isfinite(c.interface_bin_thz) || c.interface_bin_thz <= 0)
            throw std::runtime_error("Interface frequency bin must be finite and positive.");
    }
    if (c.has_convergence_stop) throw std::runtime_error("convergence.enabled requires gradient driving.");
    if (c.material_folder.empty() && c.material_folders.empty()) throw std::runtime_error("io.material_folder is required.");

    auto p = std::make_unique<Problem>();
    p->config = c;
    // Stationary modes join the active bank under conservative transport and
    // for the Gaussian source spectrum (src/material/MaterialFactory.cpp:51-55).
    const auto material = cached_material(c.material_folders.empty() ? c.material_folder : c.material_folders.front(),
                                          c.input_directory, c.temperature_lookup_dt,
                                          conservative || (c.heat_source_enabled && c.heat_spectrum == 2));
    p->raw = material->raw;
    p->modes = material->modes;
    p->linearized = linearized;
    p->conservative = conservative;
    p->nonlinear = nonlinear;
    p->domain = build_domain(c);
    const auto& m = p->modes;
    const auto& d = p->domain;
    const long long n = std::llround(c.particle_count);
    // src/MonteCarloSolver.cpp:72-77.
    p->particle_density = static_cast<double>(n) / d.volume;
    p->particle_volume = d.volume / static_cast<double>(n);
    // include/solver/BackgroundCache.h:18-21.
    p->energy_weight = static_cast<double>(m.n_active) * p->particle_volume / m.energy_normalization;
    // BackgroundCache (include/solver/BackgroundCache.h:13-30): the exact crystal
    // energy density and the heat-capacity density over all modes when linearized.
    p->background_energy_density = linearized ? m.crystal_energy_density(c.background_temperature)
                                              : m.energy_density_from_temperature(c.background_temperature);
    if (linearized) {
        double hc = 0.0;
        for (int q = 0; q < m.qpoints; ++q)
            for (int b = 0; b < m.branches; ++b)
                hc += m.heat_capacity(c.background_temperature, q * m.branches + b) / m.energy_normalization;
        if (!(hc > 0.0) || !std::isfinite(hc))
            throw std::invalid_argument("Linearized temperature requires positive heat capacity.");
        p->heat_capacity_density = hc;
        p->callaway = callaway_data(m, material->completed, c.background_temperature);
        p->callaway_wavevector.resize(3 * static_cast<size_t>(m.n_active));
        for (int a = 0; a < m.n_active; ++a)
            for (int k = 0; k < 3; ++k) p->callaway_wavevector[3 * a + k] = p->callaway.wavevector[a][k];
        auto& cv = p->callaway_view;
        cv.capacity = p->callaway.capacity.data();
        cv.wavevector = p->callaway_wavevector.data();
        cv.normal_rate = p->callaway.normal_rate.data();
        cv.resistive_rate = p->callaway.resistive_rate.data();
        cv.heat_capacity_density = p->heat_capacity_density;
        cv.background_energy_density = p->background_energy_density;
    }
    if (nonlinear) {
        // NonlinearCallawayCollision::configure validates callaway_data(T_bg)
        // (src/solver/NonlinearCallawayCollision.cpp:287-301); apply then reads
        // mode_callaway_rates at the rate temperature and the Cartesian q
        // (src/material/PhononMaterialCollisionData.cpp:167-196).
        const CallawayTable checked = callaway_data(m, material->completed, c.background_temperature);
        auto& r = p->nl_rates;
        const size_t na = static_cast<size_t>(m.n_active), nt = m.temps.size(), nf = static_cast<size_t>(m.n_flat);
        r.temps = m.temps;
        r.gamma_normal.resize(nt * na);
        r.gamma_umklapp.resize(nt * na);
        r.gamma_extra.resize(nt * na);
        for (size_t t = 0; t < nt; ++t)
            for (size_t a = 0; a < na; ++a) {
                const size_t f = t * nf + static_cast<size_t>(m.active_flat[a]);
                r.gamma_normal[t * na + a] = m.gamma_normal[f];
                r.gamma_umklapp[t * na + a] = m.gamma_umklapp[f];
                r.gamma_extra[t * na + a] = m.gamma_extra[f];
            }
        r.wavevector.resize(3 * na);
        for (size_t a = 0; a < na; ++a)
            for (int k = 0; k < 3; ++k) r.wavevector[3 * a + k] = checked.wavevector[a][k];
        auto& v = p->nl_view;
        v.n_temps = static_cast<int32_t>(nt);
        v.temps = r.temps.data();
        v.gamma_normal = r.gamma_normal.data();
        v.gamma_umklapp = r.gamma_umklapp.data();
        v.gamma_extra = r.gamma_extra.data();
        v.wavevector = r.wavevector.data();
    }
    p->neq_background.resize(m.n_active);
    p->omega_active.resize(m.n_active);
    p->velocity_active.resize(3 * static_cast<size_t>(m.n_active));
    for (int a = 0; a < m.n_active; ++a) {
        const int f = m.active_flat[a];
        p->neq_background[a] = m.bose(c.background_temperature, f);
        p->omega_active[a] = m.omega_flat[f];
        for (int k = 0; k < 3; ++k) p->velocity_active[3 * a + k] = m.velocity_flat[f][k];
    }
    if (layered) build_layers(*p);
    // Rough tables on every 'R' facet, or on every reflecting facet with strict
    // elastic tables at T_bg under conservative transport
    // (src/solver/RoughBoundaryInitialization.cpp:9-25).
    std::vector<bool> enabled(d.facets.size());
    for (size_t f = 0; f < d.facets.size(); ++f) {
        const auto& fc = d.facets[f];
        const bool reflecting = fc.condition != 'T' && !(fc.condition == 'P' && fc.periodic_pair >= 0);
        enabled[f] = conservative ? reflecting : fc.condition == 'R';
    }
    p->rough = build_rough_tables(d, m, enabled, conservative ? c.background_temperature : 0.0, p->facet_to_rough);
    if (layered) layered_rough_tables(*p, enabled, conservative ? c.background_temperature : 0.0);
    if (layered) layered_reservoirs(*p);
    else p->reservoirs = build_reservoir_tables(d, m, p->particle_density, c.time_step, p->facet_to_reservoir);
    if (linearized) {
        // Linearized reservoir occupations (src/solver/ReservoirInjector.cpp:104-109),
        // each material's block of the bank from its own table.
        std::vector<const ModeTables*> banks{&m};
        std::vector<int> begin{0};
        if (layered) {
            banks = p->layer_modes;
            begin.assign(p->layers.mode_begin.begin(), p->layers.mode_begin.end() - 1);
        }
        for (auto& r : p->reservoirs)
            for (size_t k = 0; k < banks.size(); ++k) {
                const ModeTables& t = *banks[k];
                for (int a = 0; a < t.n_active; ++a) {
                    const int f = t.active_flat[a];
                    const double hw = kHbarEvPs * t.omega_flat[f];
                    r.occupation[begin[k] + a] = hw > 0.0 ? t.bose(c.background_temperature, f) +
                                                                t.heat_capacity(c.background_temperature, f) *
                                                                    (r.temperature - c.background_temperature) / hw
                                                          : 0.0;
                }
            }
    }

    // Prescribed heat source (src/solver/HeatSourceDriving.cpp:8-11).
    p->heat = layered ? build_heat_source(c, d, p->layer_modes) : build_heat_source(c, d, m);
    if (p->heat.enabled) {
        auto& h = p->heat_view;
        h.n_cells = d.cell_count();
        h.rate = p->heat.rates.data();
        h.mode_eligible = p->heat.mode_eligible.data();
        h.square = p->heat.square ? 1 : 0;
        h.start_ps = p->heat.start;
        h.end_ps = p->heat.end;
        h.period_ps = p->heat.period;
        h.on_duration_ps = p->heat.on_duration;
        h.duty_cycle = p->heat.duty_cycle;
        h.amplitude = p->heat.amplitude;
        h.spectrum = p->heat.spectrum;
        h.mode_weight = p->heat.mode_weight.data();
        h.frequency_center_thz = p->heat.frequency_center;
        h.frequency_sigma_thz = p->heat.frequency_sigma;
    }

    // C-ABI views.
    for (size_t f = 0; f < d.facets.size(); ++f) {
        pmc_facet v{};
        for (int k = 0; k < 3; ++k) {
            v.normal[k] = d.facets[f].normal[k];
            v.shift[k] = d.facets[f].shift[k];
        }
        v.condition = d.facets[f].condition;
        v.rough_table = p->facet_to_rough[f];
        v.reservoir = p->facet_to_reservoir[f];
        v.periodic_pair = d.facets[f].periodic_pair;
        p->facets.push_back(v);
    }
    auto& mv = p->mode_view;
    mv.n_active = m.n_active;
    mv.n_flat = m.n_flat;
    mv.active_flat = m.active_flat.data();
    if (layered) { // the global bank: flat index = global active index
        const auto& L = p->layers;
        mv.n_active = static_cast<int>(p->omega_active.size());
        mv.n_flat = mv.n_active;
        mv.active_flat = L.active_flat.data();
    }
    mv.omega = p->omega_active.data();
    mv.velocity = p->velocity_active.data();
    mv.neq_background = p->neq_background.data();
    mv.n_temps = static_cast<int>(m.temps.size());
    mv.temps = m.temps.data();
    mv.gamma_total = m.gamma_total.data();
    mv.energy_weight = p->energy_weight;
    mv.background_energy_density = p->background_energy_density;
    mv.zero_point_density = m.zero_point_density;
    mv.lut_n = static_cast<int>(m.lut_t.size());
    mv.lut_t = m.lut_t.data();
    mv.lut_e = m.lut_e.data();
    mv.energy_normalization = m.energy_normalization;
    if (layered) {
        auto& L = p->layers;
        mv.gamma_total = L.gamma_total.data();
        mv.lut_n = static_cast<int>(L.lut_t.size());
        mv.lut_t = L.lut_t.data();
        mv.lut_e = L.lut_e.data();
        auto& lv = p->layers_view;
        lv.count = L.count;
        lv.axis = c.interface_axis;
        lv.planes = L.planes.data();
        lv.mode_begin = L.mode_begin.data();
        lv.cell_material = L.cell_material.data();
        lv.background_energy_density = L.background_energy_density.data();
        lv.lut_begin = L.lut_begin.data();
        lv.spatial_volume = L.spatial_volume.data();
        lv.band_width = 2.0 * M_PI * c.interface_bin_thz;
        lv.bank_begin = L.bank_begin.data();
        lv.bank_mode = L.bank_mode.data();
        lv.bank_omega = L.bank_omega.data();
        lv.bank_prefix = L.bank_prefix.data();
        lv.amm_fraction = L.amm_fraction;
        lv.amm_block = L.amm_fraction > 0 ? L.amm_block.data() : nullptr;
        lv.amm_reflected = L.amm_fraction > 0 ? L.amm_reflected.data() : nullptr;
        lv.amm_edge_begin = L.amm_fraction > 0 ? L.amm_edge_begin.data() : nullptr;
        lv.amm_edge_mode = L.amm_fraction > 0 && !L.amm_edge_mode.empty() ? L.amm_edge_mode.data() : nullptr;
        lv.amm_edge_prob = L.amm_fraction > 0 && !L.amm_edge_prob.empty() ? L.amm_edge_prob.data() : nullptr;
        lv.facet_rough = L.facet_rough.empty() ? nullptr : L.facet_rough.data();
        lv.heat_capacity_density = L.heat_capacity_density.empty() ? nullptr : L.heat_capacity_density.data();
    }
    auto& g = p->geometry_view;
    g.kind = 0;
    for (int k = 0; k < 3; ++k) {
        g.bmin[k] = d.bmin[k];
        g.bmax[k] = d.bmax[k];
        g.grid[k] = d.grid[k];
        g.inv_cell[k] = d.inv_cell[k];
    }
    g.n_facets = static_cast<int>(p->facets.size());
    g.facets = p->facets.data();
    for (int k = 0; k < 6; ++k) g.box_plane_facet[k] = d.box_plane_facet[k];
    g.voxel_to_cell = d.voxel_to_cell.data();
    g.n_cells = d.cell_count();
    g.cell_volume = d.cell_volume.data();
    g.mesh = nullptr;
    if (!d.is_box) {
        // Mesh model: faces, face and simplex hierarchies, star tetrahedra.
        const SurfaceMesh& sm = d.mesh;
        auto& a = p->mesh;
        const int nf = static_cast<int>(sm.faces.size());
        static std::mutex mesh_lock;
        static std::map<std::string, Problem::MeshArrays> mesh_cache; // per surface, reused across problems
        bool cached = false;
        {
            std::lock_guard<std::mutex> hold(mesh_lock);
            auto it = mesh_cache.find(d.surface_key);
            if (!d.surface_key.empty() && it != mesh_cache.end()) {
                a = it->second;
                cached = true;
            }
        }
        if (!cached) {
        std::vector<std::array<Vec3, 2>> boxes(nf);
        for (const auto& v : sm.vertices) a.vertices.insert(a.vertices.end(), v.begin(), v.end());
        for (int f = 0; f < nf; ++f) {
            for (int k = 0; k < 3; ++k) {
                a.faces.push_back(sm.faces[f][k]);
                a.face_normal.push_back(sm.face_normal[f][k]);
            }
            a.face_k.push_back(sm.face_k[f]);
            for (int k = 0; k < 3; ++k) a.face_bounds.push_back(sm.face_bounds[f][0][k]);
            for (int k = 0; k < 3; ++k) a.face_bounds.push_back(sm.face_bounds[f][1][k]);
            a.face_facet.push_back(sm.face_to_facet[f]);
            boxes[f] = sm.face_bounds[f];
        }
        a.face_bvh = build_bvh(boxes, sm.face_centroid, 4);
        const auto& star = sm.simplices();
        std::vector<std::array<Vec3, 2>> sboxes;
        std::vector<Vec3> scent;
        for (const auto& s : star) {
            for (const auto& v : s.v) a.simplex_vertices.insert(a.simplex_vertices.end(), v.begin(), v.end());
            for (const auto& r : s.inv) a.simplex_inverse.insert(a.simplex_inverse.end(), r.begin(), r.end());
            for (const auto& b : s.bounds) a.simplex_bounds.insert(a.simplex_bounds.end(), b.begin(), b.end());
            a.simplex_volume.push_back(s.volume);
            sboxes.push_back(s.bounds);
            scent.push_back({0.5 * (s.bounds[0][0] + s.bounds[1][0]), 0.5 * (s.bounds[0][1] + s.bounds[1][1]),
                             0.5 * (s.bounds[0][2] + s.bounds[1][2])});
        }
        a.simplex_bvh = build_bvh(sboxes, scent, 4);
        // 4M voxels of 1-byte bounds in quarter-angstrom units (0 to 63.75 A).
        a.field = build_distance_field(sm, d.bmin, d.bmax, 4L << 20, 0.25);
        if (!d.surface_key.empty()) {
            std::lock_guard<std::mutex> hold(mesh_lock);
            if (mesh_cache.size() >= 8) mesh_cache.erase(mesh_cache.begin());
            mesh_cache[d.surface_key] = a;
        }
        }
        const auto& star = sm.simplices();
        auto& mv = p->mesh_view;
        mv.n_vertices = static_cast<int>(sm.vertices.size());
        mv.vertices = a.vertices.data();
        mv.n_faces = nf;
        mv.faces = a.faces.data();
        mv.face_normal = a.face_normal.data();
        mv.face_k = a.face_k.data();
        mv.face_bounds = a.face_bounds.data();
        mv.face_facet = a.face_facet.data();
        mv.n_face_nodes = static_cast<int>(a.face_bvh.nodes.size());
        mv.face_nodes = a.face_bvh.nodes.data();
        mv.face_order = a.face_bvh.order.data();
        mv.n_simplices = static_cast<int>(star.size());
        mv.simplex_vertices = a.simplex_vertices.data();
        mv.simplex_inverse = a.simplex_inverse.data();
        mv.simplex_bounds = a.simplex_bounds.data();
        mv.simplex_volume = a.simplex_volume.data();
        mv.n_simplex_nodes = static_cast<int>(a.simplex_bvh.nodes.size());
        mv.simplex_nodes = a.simplex_bvh.nodes.data();
        mv.simplex_order = a.simplex_bvh.order.data();
        for (int k = 0; k < 3; ++k) mv.field_dims[k] = a.field.dims[k];
        mv.field_h = a.field.h;
        mv.field_unit = a.field.unit;
        mv.field = a.field.bound.data();
        mv.field_inside = a.field.inside.data();
        g.kind = 1;
        g.mesh = &p->mesh_view;
    }
    for (const auto& r : p->rough) {
        pmc_rough_table v{};
        v.specularity = r.specularity.data();
        v.spec_match = r.spec_match.data();
        v.diffuse_begin = r.diffuse_begin.data();
        v.diffuse_end = r.diffuse_end.data();
        v.n_outgoing = static_cast<int>(r.outgoing_sorted.size());
        v.outgoing_active = r.outgoing_active.data();
        v.outgoing_sorted = r.outgoing_sorted.data();
        v.residual_prefix = r.residual_prefix.data();
        v.n_roulette = static_cast<int>(r.roulette_active.size());
        v.roulette_active = r.roulette_active.data();
        v.roulette_cdf = r.roulette_cdf.data();
        v.strict_elastic = r.strict_elastic ? 1 : 0;
        v.mode_begin = r.mode_begin;
        v.mode_count = r.mode_count;
        p->rough_views.push_back(v);
    }
    double expected = 0;
    for (const auto& r : p->reservoirs) {
        pmc_reservoir v{};
        v.facet = r.facet;
        v.temperature = r.temperature;
        for (int k = 0; k < 3; ++k) v.normal[k] = r.normal[k];
        v.expected_count = r.expected;
        v.initial_count = r.initial_count;
        v.cdf = r.cdf.data();
        v.alias_prob = r.alias_prob.data();
        v.alias_index = r.alias_index.data();
        v.occupation = r.occupation.data();
        v.n_faces = static_cast<int>(r.face_cdf.size());
        v.face_cdf = r.face_cdf.data();
        v.triangles = r.triangles.data();
        p->reservoir_views.push_back(v);
        expected += r.expected;
    }
    auto& s = p->setup;
    s.capacity = static_cast<int64_t>(std::ceil(static_cast<double>(n) * opt.capacity_factor + 64.0 * expected + 4096.0));
    s.time_step_ps = c.time_step;
    s.push_distance_a = 1e-10 * std::max(c.time_step, 1.0);
    s.seed = c.random_seed;
    s.instance = 0;
    s.background_temperature = c.background_temperature;
    s.lifetime_local = c.lifetime_local ? 1 : 0;
    s.lifetime_temperature = c.lifetime_temperature;
    // Fixed incoming flux under conservative transport and for layered materials,
    // else one-to-one refill (src/solver/ReservoirInitialization.cpp:30-33; the run
    // summary's reservoir_gen names only the conservative case, Diagnostics.cpp:102-104).
    s.refill_policy = conservative || layered ? 1 : 0;
    s.mode_sampling = opt.mode_sampling;
    s.temperature_inversion = 0;
    s.conservative_walls = conservative ? 1 : 0;

    // Initial temperatures (src/solver/Initialization.cpp:181-265).
    auto& in = p->init;
    in.count = n;
    // Initial mode excitation (src/solver/Initialization.cpp:146-178): skipped for a
    // multiplier within 1e-15 of 1.
    in.excite = c.has_mode_excitation && c.excite_multiplier > 0.0 && std::abs(c.excite_multiplier - 1.0) >= 1e-15;
    in.excite_fmin_thz = c.excite_fmin;
    in.excite_fmax_thz = c.excite_fmax;
    in.excite_multiplier = c.excite_multiplier;
    double tmin = 299.0, tmax = 301.0;
    bool any = false;
    for (const auto& r : p->reservoirs) {
        if (!std::isfinite(r.temperature)) continue;
        tmin = any ? std::min(tmin, r.temperature) : r.temperature;
        tmax = any ? std::max(tmax, r.temperature) : r.temperature;
        any = true;
    }
    if (tmin > tmax) std::swap(tmin, tmax);
    if (!c.initial_linear) {
        in.temperature_mode = 0;
        in.uniform_temperature = c.initial_uniform_temperature;
    } else {
        int cold = -1, hot = -1;
        double ct = std::numeric_limits<double>::infinity(), ht = -std::numeric_limits<double>::infinity();
        for (int f : d.reservoir_facets()) {
            const double t = d.facets[f].value;
            if (!std::isfinite(t)) continue;
            if (t < ct) { ct = t; cold = f; }
            if (t > ht) { ht = t; hot = f; }
        }
        in.temperature_mode = 0;
        in.uniform_temperature = 0.5 * (tmin + tmax);
        if (cold >= 0 && hot >= 0 && ht > ct + 1e-12) {
            const auto& pc = d.facets[cold].centroid;
            const auto& ph = d.facets[hot].centroid;
            const double ax[3] = {ph[0] - pc[0], ph[1] - pc[1], ph[2] - pc[2]};
            if (ax[0] * ax[0] + ax[1] * ax[1] + ax[2] * ax[2] > 1e-20) {
                in.temperature_mode = 1;
                for (int k = 0; k < 3; ++k) {
                    in.cold_point[k] = pc[k];
                    in.hot_point[k] = ph[k];
                }
                in.cold_temperature = ct;
                in.hot_temperature = ht;
            }
        }
    }
    return p;
}

MonteCarloSolver::MonteCarloSolver(std::unique_ptr<Problem> problem, std::string folder, std::ostream* log)
    : problem_(std::move(problem)), folder_(std::move(folder)), log_(log) {
    auto& p = *problem_;
    throw_if_failed(pmc_create(&p.setup, &ctx_));
    throw_if_failed(pmc_upload_material(ctx_, &p.mode_view));
    throw_if_failed(pmc_upload_geometry(ctx_, &p.geometry_view));
    throw_if_failed(pmc_upload_boundary_tables(ctx_, static_cast<int>(p.rough_views.size()), p.rough_views.data(),
                                               static_cast<int>(p.reservoir_views.size()), p.reservoir_views.data()));
    if (p.linearized) throw_if_failed(pmc_upload_callaway(ctx_, &p.callaway_view));
    if (p.nonlinear) throw_if_failed(pmc_upload_callaway_nonlinear(ctx_, &p.nl_view));
    if (p.layer_modes.size() > 1) throw_if_failed(pmc_upload_layers(ctx_, &p.layers_view));
    if (p.heat.enabled) throw_if_failed(pmc_upload_heat_source(ctx_, &p.heat_view));
    throw_if_failed(pmc_initialize_carriers(ctx_, &p.init));
    if (!folder_.empty()) {
        write_domain_summary(p.domain, p.config, folder_, static_cast<int>(p.mesh.simplex_volume.size()));
        write_convergence_header();
    }
    if (log_) {
        *log_ << "MonteCarloSolver initialized: particle_count=" << p.init.count << ", time_step=" << p.config.time_step
              << ", convergence_write_interval=" << p.config.convergence_write_interval
              << ", density=" << p.particle_density << '\n';
    }
    record_observables();
}

MonteCarloSolver::~MonteCarloSolver() { pmc_destroy(ctx_); }

void MonteCarloSolver::throw_if_failed(int rc) const {
    if (rc == PMC_OK) return;
    pmc_error e{};
    if (ctx_) pmc_last_error(ctx_, &e);
    std::string msg = e.message[0] ? e.message : pmc_status_string(rc);
    throw std::runtime_error(msg);
}

long long MonteCarloSolver::particle_count() const { return ledger().live_carriers; }

pmc_ledger MonteCarloSolver::ledger() const {
    pmc_ledger l{};
    pmc_read_ledger(ctx_, &l);
    return l;
}

std::vector<double> MonteCarloSolver::cell_temperatures() const {
    std::vector<double> t(problem_->domain.cell_count());
    pmc_read_cells(ctx_, t.data(), nullptr);
    return t;
}

void MonteCarloSolver::run_timestep() { run_steps(1); }

void MonteCarloSolver::run_steps(int n) {
    const int interval = problem_->config.convergence_write_interval;
    while (n > 0) {
        int next = step_ == 0 ? 1 : ((step_ / interval) + 1) * interval;
        const int k = std::min(n, next - step_);
        const auto t0 = std::chrono::steady_clock::now();
        throw_if_failed(pmc_step(ctx_, k));
        step_seconds_ += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        step_ += k;
        n -= k;
        if (step_ == 1 || step_ % interval == 0) record_observables();
    }
}

void MonteCarloSolver::record_observables() {
    // MonteCarloSolver::update_heat_flux_and_conductivity (src/solver/Observables.cpp:38-142).
    const auto& p = *problem_;
    const auto& c = p.config;
    const auto& d = p.domain;
    const int nc = d.cell_count();
    cell_T_ = cell_temperatures();
    std::vector<double> q(3 * static_cast<size_t>(nc));
    throw_if_failed(pmc_read_heat_flux(ctx_, q.data()));
    flux_.assign(nc, {0, 0, 0});
    for (int i = 0; i < nc; ++i) flux_[i] = {q[3 * i], q[3 * i + 1], q[3 * i + 2]};
    int axis = 0;
    if (c.transport_axis >= 0) axis = c.transport_axis;
    else if (c.grid[1] > c.grid[0] && c.grid[1] >= c.grid[2]) axis = 1;
    else if (c.grid[2] > c.grid[0] && c.grid[2] > c.grid[1]) axis = 2;
    double phi = 0, vol = 0;
    for (int i = 0; i < nc; ++i) {
        phi += flux_[i][axis] * d.cell_volume[i];
        vol += d.cell_volume[i];
    }
    mean_flux_ = phi / vol;
    kappa_eff_ = kappa_int_ = 0;
    if (c.compute_kappa) {
        double tl = cell_T_.front(), tr = cell_T_.back();
        const auto rf = d.reservoir_facets();
        if (rf.size() >= 2) {
            int fl = rf.front(), fr = rf.front();
            for (int f : rf) {
                if (d.facets[f].centroid[axis] < d.facets[fl].centroid[axis]) fl = f;
                if (d.facets[f].centroid[axis] > d.facets[fr].centroid[axis]) fr = f;
            }
            tl = d.facets[fl].value;
            tr = d.facets[fr].value;
        }
        const double L = std::abs(d.bmax[axis] - d.bmin[axis]) * 1e-10;
        const double grad_end = std::abs(L) > 1e-24 ? (tr - tl) / L : 0.0;
        kappa_eff_ = std::abs(grad_end) > 1e-18 ? -mean_flux_ / grad_end : 0.0;
        double sx = 0, st = 0, sxx = 0, sxt = 0;
        int nfit = 0;
        for (int i = 0; i < nc; ++i) {
            const double x = d.cell_center[i][axis] * 1e-10, t = cell_T_[i];
            sx += x;
            st += t;
            sxx += x * x;
            sxt += x * t;
            ++nfit;
        }
        double grad_fit = 0;
        const double window = std::abs(cell_T_.back() - cell_T_.front());
        if (nfit >= 2) {
            const double den = static_cast<double>(nfit) * sxx - sx * sx;
            if (std::abs(den) > 1e-30 && window > 1e-6) grad_fit = (static_cast<double>(nfit) * sxt - sx * st) / den;
        }
        kappa_int_ = std::abs(grad_fit) > 1e-18 ? -mean_flux_ / grad_fit : 0.0;
    }
    if (!folder_.empty()) append_convergence_row();
}

void MonteCarloSolver::write_convergence_header() {
    // ResultWriter::format_convergence_header (src/solver/ResultWriter.cpp:82-98).
    std::ofstream out(std::filesystem::path(folder_) / "convergence.txt", std::ios::trunc);
    if (!out) throw std::runtime_error("Cannot open output file: " + folder_ + "/convergence.txt");
    out << "# timestep time_ps";
    for (int i = 0; i < problem_->domain.cell_count(); ++i) out << " T_" << (i + 1);
    out << " heatflux kappa_int kappa_eff particle_count absorbed injected recovered net"
           " reservoir_absorbed_energy_ev reservoir_injected_energy_ev"
           " hs_occupation_updates hs_injected_energy_ev lifetime_energy_residual_ev"
           " energy_balance_residual_ev total_thermal_energy_ev"
           " hs_prescribed_total_ev hs_pending_total_ev hs_ledger_residual_ev\n";
    if (problem_->config.write_cell_heat_flux) {
        std::ofstream f(std::filesystem::path(folder_) / "grid_heat_flux.csv", std::ios::trunc);
        f << "timestep,time_ps,cell_index,temperature_K,qx_W_m2,qy_W_m2,qz_W_m2\n";
    }
}

void MonteCarloSolver::append_convergence_row() {
    // ResultWriter::format_convergence (src/solver/ResultWriter.cpp:100-119).
    const pmc_ledger l = ledger();
    const double time = static_cast<double>(step_) * problem_->config.time_step;
    std::ofstream out(std::filesystem::path(folder_) / "convergence.txt", std::ios::app);
    if (!out) throw std::runtime_error("Cannot open output file: " + folder_ + "/convergence.txt");
    out << step_ << " " << time;
    for (double t : cell_T_) out << " " << t;
    out << std::setprecision(12);
    const double prescribed = source_prescribed_ev(l);
    out << " " << mean_flux_ << " " << kappa_int_ << " " << kappa_eff_ << " " << l.live_carriers << " "
        << l.step_absorbed << " " << l.step_injected << " " << 0 << " " << (l.step_injected - l.step_absorbed) << " "
        << l.step_absorbed_energy_ev << " " << l.step_injected_energy_ev << " " << l.step_source_updates << " "
        << l.step_source_energy_ev << " " << l.step_lifetime_residual_ev << " " << l.step_balance_residual_ev << " "
        << l.thermal_energy_ev << " " << prescribed << " " << l.source_pending_ev << " "
        << (prescribed - l.total_source_energy_ev - l.source_pending_ev) << '\n';
    if (problem_->heat.enabled) write_source_ledger(l);
    if (problem_->config.write_cell_heat_flux) {
        std::ofstream f(std::filesystem::path(folder_) / "grid_heat_flux.csv", std::ios::app);
        f << std::setprecision(17);
        for (size_t i = 0; i < flux_.size(); ++i)
            f << step_ << ',' << time << ',' << i << ',' << cell_T_[i] << ',' << flux_[i][0] << ',' << flux_[i][1]
              << ',' << flux_[i][2] << '\n';
    }
}

// Prescribed energy so far: every cell accrues its rate times the integrated
// time factor (PrescribedHeatSource::accrue, src/solver/PrescribedHeatSource.cpp:232-252).
double MonteCarloSolver::source_prescribed_ev(const pmc_ledger& l) const {
    if (!problem_->heat.enabled) return 0.0;
    double rate = 0.0;
    for (double r : problem_->heat.rates) rate += r;
    return rate * l.source_time_ps;
}

void MonteCarloSolver::write_source_ledger(const pmc_ledger& l) const {
    // heat_source_ledger.txt (ResultWriter::format_source_ledger, src/solver/ResultWriter.cpp:343-354).
    const auto& p = *problem_;
    const int nc = p.domain.cell_count();
    std::vector<double> pending(nc);
    throw_if_failed(pmc_read_source_pending(ctx_, pending.data()));
    std::ofstream out(std::filesystem::path(folder_) / "heat_source_ledger.txt", std::ios::trunc);
    if (!out) throw std::runtime_error("Cannot open output file: " + folder_ + "/heat_source_ledger.txt");
    static const char* kSpectra[] = {"thermal", "weighted", "gaussian"};
    out << std::setprecision(17) << "# prescribed_lattice_power; not an electron-phonon calculation\n"
        << "# spectrum=" << kSpectra[problem_->config.heat_spectrum] << "; pending energy is not yet represented by phonons\n"
        << "# cell x_A y_A z_A material base_power_W prescribed_eV deposited_eV pending_eV\n";
    const auto& cm = p.layers.cell_material;
    for (int c = 0; c < nc; ++c) {
        const auto& x = p.domain.cell_center[c];
        const double rate = p.heat.rates[c];
        const double prescribed = rate * l.source_time_ps;
        out << c << " " << x[0] << " " << x[1] << " " << x[2] << " " << (cm.empty() ? 0 : cm[c]) << " "
            << rate / 6.241509074e6 << " " << prescribed << " " << (prescribed - pending[c]) << " " << pending[c]
            << '\n';
    }
}

void MonteCarloSolver::append_source_summary(std::ostream& out, const pmc_ledger& l) const {
    // [local_heat_source_occupation_injection] (ResultWriter::format_source_summary,
    // src/solver/ResultWriter.cpp:121-170).
    const SimConfig& c = problem_->config;
    const HeatSource& h = problem_->heat;
    const double prescribed = source_prescribed_ev(l);
    out << "\n[local_heat_source_occupation_injection]\n";
    out << "enabled = " << (h.enabled ? "true" : "false") << '\n';
    out << "profile = " << (c.heat_profile == HeatProfile::Gaussian ? "gaussian" : "uniform") << '\n';
    if (c.heat_total_power) out << "specified_total_power_w = " << *c.heat_total_power << '\n';
    else out << "power_density_wm3 = " << c.heat_power_density << '\n';
    out << "model = prescribed_lattice_power\n";
    out << "electron_transport = false\n";
    out << "empty_cell_policy = defer_locally\n";
    static const char* kSpectra[] = {"thermal", "weighted", "gaussian"};
    out << "spectrum = " << kSpectra[c.heat_spectrum] << '\n';
    if (c.heat_spectrum == 2) {
        out << "frequency_center_thz = " << c.heat_frequency_center.value_or(0) << '\n';
        out << "frequency_sigma_thz = " << c.heat_frequency_sigma.value_or(0) << '\n';
        out << "spectral_normalization = existing_carrier_weights_within_each_cell\n";
        out << "spectral_weight = exp(-0.5*((f-f_center)/sigma_f)^2); energy_per_mode\n";
    }
    out << "power_input = " << (c.heat_total_power ? "total_power" : "power_density") << '\n';
    out << "base_total_power_w = " << (h.enabled ? h.base_power_w() : 0.0) << '\n';
    out << "frequency_min_thz = " << c.heat_frequency_min << '\n';
    out << "frequency_max_thz = " << c.heat_frequency_max << '\n';
    out << "branches_zero_based = [";
    for (size_t i = 0; i < c.heat_branches.size(); ++i) out << (i ? ", " : "") << c.heat_branches[i];
    out << "]\nbranch_weights_per_sampled_mode = [";
    for (size_t i = 0; i < c.heat_branch_weights.size(); ++i) out << (i ? ", " : "") << c.heat_branch_weights[i];
    out << "]\n";
    out << "prescribed_energy_total_ev = " << prescribed << '\n';
    out << "pending_energy_total_ev = " << l.source_pending_ev << '\n';
    out << "deposition_complete = " << (l.source_pending_ev == 0.0 ? "true" : "false") << '\n';
    out << "source_ledger_residual_ev = " << (prescribed - l.total_source_energy_ev - l.source_pending_ev) << '\n';
    out << "time_profile = " << (c.heat_time_square ? "square" : "constant") << '\n';
    out << "time_start_ps = " << c.heat_time_start << '\n';
    out << "time_end_ps = " << c.heat_time_end << '\n';
    out << "period_ps = " << c.heat_period << '\n';
    out << "on_duration_ps = " << c.heat_on_duration << '\n';
    out << "duty_cycle = " << c.heat_duty_cycle << '\n';
    out << "amplitude = " << c.heat_amplitude << '\n';
    out << "occupation_updates_total = " << l.source_updates << '\n';
    out << "injected_energy_total_ev = " << l.total_source_energy_ev << '\n';
    out << "occupation_updates_last_step = " << l.step_source_updates << '\n';
    out << "injected_energy_last_step_ev = " << l.step_source_energy_ev << '\n';
    if (h.enabled && !folder_.empty()) write_source_ledger(l);
}

void MonteCarloSolver::append_profile_summary(std::ostream& out) const {
    // [performance] / [rough_boundary_diagnostics] keys of
    // ResultWriter::format_run_summary (src/solver/ResultWriter.cpp:172-341) plus a [gpu] section.
    const auto& p = *problem_;
    const pmc_ledger l = ledger();
    out << std::setprecision(17);
    out << "\n[performance]\n";
    out << "openmp_enabled = false\n";
    out << "openmp_thread_count = 1\n";
    out << "random_seed = " << problem_->config.random_seed << '\n';
    out << "profile_timers_enabled = " << (problem_->config.profile_timers ? "true" : "false") << '\n';
    out << "collision_cache_corrections_total = 0\n";
    out << "collision_cache_failures_total = 0\n";
    out << "transport_failure_policy = strict\n";
    out << "excessive_collision_recoveries_total = 0\n";
    const double den = std::max(1.0, static_cast<double>(l.rough_events));
    out << "\n[rough_boundary_diagnostics]\n";
    out << "events_total = " << l.rough_events << '\n';
    out << "specular_selected = " << l.rough_specular << '\n';
    out << "diffuse_selected = " << l.rough_diffuse << '\n';
    out << "specular_ratio = " << static_cast<double>(l.rough_specular) / den << '\n';
    out << "diffuse_ratio = " << static_cast<double>(l.rough_diffuse) / den << '\n';
    // [material_interface_diagnostics] (src/solver/ResultWriter.cpp:240-289).
    const SimConfig& c = problem_->config;
    const auto& L = problem_->layers;
    const int materials = std::max<int>(1, static_cast<int>(c.material_folders.size()));
    const double iden = std::max(1.0, static_cast<double>(l.interface_events));
    out << "\n[material_interface_diagnostics]\n";
    out << "material_count = " << materials << '\n';
    out << "model = "
        << (c.interface_model == "amm" ? "acoustic_mismatch" : c.interface_model == "mmm" ? "mixed_mismatch"
                                                                                           : "diffuse_mismatch")
        << '\n';
    if (materials > 1) {
        out << "frequency_bin_thz = " << c.interface_bin_thz << '\n';
        out << "amm_fraction = " << L.amm_fraction << '\n';
        if (L.amm_fraction > 0) {
            out << "amm_approximation = dispersive_scalar_acoustic_no_mode_conversion\n";
            out << "amm_optical_treatment = specular_reflection\n";
            out << "amm_kernel = reciprocal_linearized_qcell_overlap_v2\n";
            out << "amm_bin_role = candidate_search_only\n";
            out << "amm_channel_flux_units = inverse_Angstrom2_inverse_ps\n";
            for (size_t i = 0; i < L.amm_audit.size(); ++i) {
                const AmmAudit& a = L.amm_audit[i];
                const std::string prefix = "interface_" + std::to_string(i) + "_";
                out << prefix << "geometric_overlap_flux = " << a.geometric_overlap_flux << '\n';
                out << prefix << "capacity_removed_flux = " << a.capacity_removed_flux << '\n';
                out << prefix << "maximum_capacity_ratio = " << a.maximum_capacity_ratio << '\n';
                for (int side = 0; side < 2; ++side) {
                    const std::string key = prefix + "side_" + std::to_string(side) + "_";
                    out << key << "incident_acoustic_flux = " << a.incident_acoustic_flux[side] << '\n';
                    out << key << "transmitted_flux = " << a.transmitted_flux[side] << '\n';
                    out << key << "no_overlap_flux = " << a.no_overlap_flux[side] << '\n';
                }
            }
            out << "parallel_wavevector_bin_inv_a = " << c.interface_parallel_bin << '\n';
            for (size_t m = 0; m < c.mass_densities.size(); ++m)
                out << "material_" << m << "_mass_density_kg_m3 = " << c.mass_densities[m] << '\n';
        }
        out << "carrier_sampling = equal_physical_state_weight\n";
        out << "fixed_flux_reservoirs = true\n";
        for (size_t k = 0; k < L.spatial_volume.size(); ++k)
            out << "material_" << k << "_carrier_volume_a3 = " << L.spatial_volume[k] << '\n';
    }
    out << "events_total = " << l.interface_events << '\n';
    out << "transmitted_total = " << l.interface_transmitted << '\n';
    out << "reflected_total = " << (l.interface_events - l.interface_transmitted) << '\n';
    out << "transmission_ratio = " << static_cast<double>(l.interface_transmitted) / iden << '\n';
    out << "reflection_ratio = " << static_cast<double>(l.interface_events - l.interface_transmitted) / iden << '\n';
    // src/solver/ResultWriter.cpp:290-310: strict transport never loses a carrier,
    // so no recovery happens; the device checks containment every 100 steps.
    out << "\n[escaped_particle_recovery]\n";
    out << "check_interval = 100\n";
    out << "policy = strict\n";
    out << "recovery_events = 0\n";
    out << "recovered_particles_total = 0\n";
    out << "\n[reservoir_balance_diagnostics]\n";
    out << "reservoir_gen = " << (p.conservative ? "fixed_incoming_flux" : "one_to_one") << '\n';
    out << "absorbed_particles_total = " << l.total_absorbed << '\n';
    out << "injected_particles_total = " << l.total_injected << '\n';
    out << "recovered_particles_total = 0\n";
    out << "net_particles_total = " << (l.total_injected - l.total_absorbed) << '\n';
    out << "absorbed_deviational_energy_total_ev = " << l.total_absorbed_energy_ev << '\n';
    out << "injected_deviational_energy_total_ev = " << l.total_injected_energy_ev << '\n';
    out << "\n[energy_weighting]\n";
    out << "initial_particle_count = " << p.init.count << '\n';
    out << "current_particle_count = " << l.live_carriers << '\n';
    out << "particle_spatial_weight_a3 = " << p.particle_volume << '\n';
    out << "background_reference_temperature_k = " << c.background_temperature << '\n';
    out << "total_thermal_energy_ev = " << l.thermal_energy_ev << '\n';
    out << "lifetime_energy_residual_total_ev = " << l.total_lifetime_residual_ev << '\n';
    out << "energy_balance_residual_total_ev = " << l.total_balance_residual_ev << '\n';
    append_source_summary(out, l);
    out << "\n[gpu]\n";
    out << "device_steps = " << l.step << '\n';
    out << "carrier_steps = " << l.carrier_steps << '\n';
    out << "step_loop_seconds = " << step_seconds_ << '\n';
    out << "particle_steps_per_second = " << (step_seconds_ > 0 ? static_cast<double>(l.carrier_steps) / step_seconds_ : 0.0)
        << '\n';
    out << "rta_pseudo_temperature_evaluations = " << l.newton_evals << '\n';
    out << "rta_exact_carrier_sums = " << l.newton_passes << '\n';
    out << "source_exact_carrier_sums = " << l.source_exact_passes << '\n';
    out << "total_energy_balance_residual_ev = " << l.total_balance_residual_ev << '\n';
}

} // namespace pmc
