// C ABI of the host facade (include/pmc_solver.h).
#include "pmc_solver.h"
#include "solver.hpp"

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <exception>
#include <fstream>
#include <memory>
#include <string>

struct pmcs_problem {
    std::unique_ptr<pmc::Problem> owned;
    const pmc::Problem* view = nullptr; // non-owning view held by a solver
    const pmc::Problem* operator->() const { return owned ? owned.get() : view; }
};

struct pmcs_solver {
    std::unique_ptr<pmc::MonteCarloSolver> s;
    pmcs_problem problem_view;
};

namespace {
template <class F>
int guarded(char* err, int32_t errlen, F&& f) {
    try {
        f();
        return 0;
    } catch (const std::exception& e) {
        if (err && errlen > 0) {
            std::strncpy(err, e.what(), static_cast<size_t>(errlen - 1));
            err[errlen - 1] = '\0';
        }
        return 1;
    }
}

void fill_material(const pmc::RawMaterial& m, double lookup_dt, pmcs_material_view* out) {
    out->volume_a3 = m.volume_a3;
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) out->reciprocal[3 * r + c] = m.reciprocal[r][c];
    out->qpoints = m.qpoints;
    out->branches = m.branches;
    for (int k = 0; k < 3; ++k) out->mesh[k] = m.mesh[k];
    out->frequency_thz = m.frequency_thz.data();
    out->qpoint_fractions = m.qpoint_fractions.front().data();
    out->group_velocity = m.group_velocity.front().data();
    out->ntemps = static_cast<int32_t>(m.temperatures.size());
    out->temperatures = m.temperatures.data();
    out->gamma = m.gamma.data();
    out->gamma_normal = m.gamma_normal.empty() ? nullptr : m.gamma_normal.data();
    out->gamma_umklapp = m.gamma_umklapp.empty() ? nullptr : m.gamma_umklapp.data();
    out->gamma_isotope = m.gamma_isotope.empty() ? nullptr : m.gamma_isotope.data();
    out->temperature_lookup_dt = lookup_dt;
}
} // namespace

extern "C" {

int pmcs_problem_from_toml(const char* path, int32_t mode_sampling, pmcs_problem** out, char* err, int32_t errlen) {
    return guarded(err, errlen, [&] {
        pmc::ProblemOptions o;
        o.mode_sampling = mode_sampling;
        auto* p = new pmcs_problem{pmc::build_problem(pmc::parse_toml_file(path), o), nullptr};
        *out = p;
    });
}

int pmcs_problem_from_text(const char* text, const char* dir, int32_t mode_sampling, pmcs_problem** out, char* err,
                           int32_t errlen) {
    return guarded(err, errlen, [&] {
        pmc::ProblemOptions o;
        o.mode_sampling = mode_sampling;
        auto* p = new pmcs_problem{pmc::build_problem(pmc::parse_toml_text(text, dir ? dir : "."), o), nullptr};
        *out = p;
    });
}

void pmcs_problem_destroy(pmcs_problem* p) { delete p; }

int pmcs_problem_views(const pmcs_problem* pr, pmc_setup* setup, pmc_mode_table* modes, pmc_geometry* geometry,
                       pmc_init* init, int32_t* n_rough, const pmc_rough_table** rough, int32_t* n_res,
                       const pmc_reservoir** res) {
    if (!pr) return PMC_ERR_ARGUMENT;
    const auto& p = *(pr->operator->());
    if (setup) *setup = p.setup;
    if (modes) *modes = p.mode_view;
    if (geometry) *geometry = p.geometry_view;
    if (init) *init = p.init;
    if (n_rough) *n_rough = static_cast<int32_t>(p.rough_views.size());
    if (rough) *rough = p.rough_views.data();
    if (n_res) *n_res = static_cast<int32_t>(p.reservoir_views.size());
    if (res) *res = p.reservoir_views.data();
    return PMC_OK;
}

int pmcs_problem_heat_source(const pmcs_problem* pr, pmc_heat_source* out) {
    if (!pr) return -1;
    const auto& p = *(pr->operator->());
    if (out) *out = p.heat_view;
    return p.heat.enabled ? 1 : 0;
}

int pmcs_write_finfet_stl(const char* path, const double substrate_nm[3], const double fin_nm[3],
                          const double step_nm[3], char* err, int errlen) {
    try {
        pmc::FinfetSpec spec;
        for (int a = 0; a < 3; ++a) {
            spec.substrate[a] = substrate_nm[a];
            spec.fin[a] = fin_nm[a];
            spec.step[a] = step_nm[a];
        }
        const auto mesh = pmc::finfet_mesh(spec);
        pmc::write_binary_stl(path, mesh, "phonomc_amd procedural FinFET");
        return static_cast<int>(mesh.faces.size());
    } catch (const std::exception& e) {
        if (err && errlen > 0) std::snprintf(err, static_cast<size_t>(errlen), "%s", e.what());
        return -1;
    }
}

int pmcs_problem_callaway(const pmcs_problem* pr, pmc_callaway_table* out) {
    if (!pr) return -1;
    const auto& p = *(pr->operator->());
    if (out) *out = p.callaway_view;
    return p.linearized ? 1 : 0;
}

int pmcs_problem_callaway_nonlinear(const pmcs_problem* pr, pmc_callaway_rates* out) {
    if (!pr) return -1;
    const auto& p = *(pr->operator->());
    if (out) *out = p.nl_view;
    return p.nonlinear ? 1 : 0;
}

int pmcs_problem_amm_audit(const pmcs_problem* pr, int32_t face, double* out) {
    if (!pr || !out) return -1;
    const auto& audits = (*pr)->layers.amm_audit;
    if (face < 0 || face >= static_cast<int32_t>(audits.size())) return 0;
    const pmc::AmmAudit& a = audits[static_cast<size_t>(face)];
    const double v[9] = {a.incident_acoustic_flux[0], a.incident_acoustic_flux[1], a.transmitted_flux[0],
                         a.transmitted_flux[1], a.no_overlap_flux[0], a.no_overlap_flux[1],
                         a.geometric_overlap_flux, a.capacity_removed_flux, a.maximum_capacity_ratio};
    std::copy(v, v + 9, out);
    return 1;
}

int pmcs_problem_material(const pmcs_problem* pr, pmcs_material_view* out) {
    if (!pr || !out) return PMC_ERR_ARGUMENT;
    fill_material((*pr)->raw, (*pr)->config.temperature_lookup_dt, out);
    return PMC_OK;
}

int pmcs_problem_layers(const pmcs_problem* pr, pmc_layers* out) {
    if (!pr) return -1;
    const auto& p = *(pr->operator->());
    if (out) *out = p.layers_view;
    return p.layer_modes.size() > 1 ? 1 : 0;
}

int pmcs_problem_layer_material(const pmcs_problem* pr, int32_t layer, pmcs_material_view* out) {
    if (!pr || !out) return PMC_ERR_ARGUMENT;
    const auto& p = *(pr->operator->());
    if (layer < 0 || layer >= static_cast<int>(p.layer_raw.size())) return PMC_ERR_ARGUMENT;
    fill_material(*p.layer_raw[layer], p.config.temperature_lookup_dt, out);
    return PMC_OK;
}

int pmcs_problem_scalars(const pmcs_problem* pr, double* out) {
    if (!pr || !out) return PMC_ERR_ARGUMENT;
    const auto& p = *(pr->operator->());
    out[0] = p.particle_volume;
    out[1] = p.particle_density;
    out[2] = p.energy_weight;
    out[3] = p.background_energy_density;
    out[4] = p.modes.zero_point_density;
    out[5] = p.modes.energy_normalization;
    out[6] = p.domain.volume;
    out[7] = p.domain.cell_count();
    out[8] = static_cast<double>(p.init.count);
    out[9] = p.setup.push_distance_a;
    return PMC_OK;
}

double pmcs_problem_bose(const pmcs_problem* pr, double t, int32_t flat) { return (*pr)->modes.bose(t, flat); }

double pmcs_problem_relaxation_weight(const pmcs_problem* pr, double t, int32_t flat, double dt) {
    return (*pr)->modes.relaxation_weight(t, flat, dt);
}

double pmcs_problem_temperature_from_energy(const pmcs_problem* pr, double e) {
    return (*pr)->modes.temperature_from_energy_density(e);
}

double pmcs_problem_energy_from_temperature(const pmcs_problem* pr, double t) {
    return (*pr)->modes.energy_density_from_temperature(t);
}

int pmcs_problem_facets(const pmcs_problem* pr, int32_t* n, double* normals, double* centroids, double* areas,
                        char* conditions, double* shifts, double* values, int32_t* face_counts) {
    if (!pr || !n) return PMC_ERR_ARGUMENT;
    const auto& d = (*pr)->domain;
    *n = static_cast<int32_t>(d.facets.size());
    for (size_t f = 0; f < d.facets.size(); ++f) {
        for (int k = 0; k < 3; ++k) {
            if (normals) normals[3 * f + k] = d.facets[f].normal[k];
            if (centroids) centroids[3 * f + k] = d.facets[f].centroid[k];
            if (shifts) shifts[3 * f + k] = d.facets[f].shift[k];
        }
        if (areas) areas[f] = d.facets[f].area;
        if (conditions) conditions[f] = d.facets[f].condition;
        if (values) values[f] = d.facets[f].value;
        if (face_counts) face_counts[f] = static_cast<int32_t>(d.facets[f].faces.size());
    }
    return PMC_OK;
}

int pmcs_problem_write_summary(const pmcs_problem* pr, const char* folder, char* err, int32_t errlen) {
    if (!pr || !folder) return PMC_ERR_ARGUMENT;
    try {
        pmc::write_domain_summary((*pr)->domain, (*pr)->config, folder,
                                  static_cast<int>((*pr)->mesh.simplex_volume.size()));
    } catch (const std::exception& e) {
        if (err && errlen > 0) std::snprintf(err, static_cast<size_t>(errlen), "%s", e.what());
        return PMC_ERR_ARGUMENT;
    }
    return PMC_OK;
}

int pmcs_problem_cells(const pmcs_problem* pr, double* centers, double* volumes) {
    if (!pr) return PMC_ERR_ARGUMENT;
    const auto& d = (*pr)->domain;
    for (int c = 0; c < d.cell_count(); ++c) {
        for (int k = 0; k < 3; ++k)
            if (centers) centers[3 * c + k] = d.cell_center[c][k];
        if (volumes) volumes[c] = d.cell_volume[c];
    }
    return PMC_OK;
}

int pmcs_solver_create(pmcs_problem* pr, const char* folder, pmcs_solver** out, char* err, int32_t errlen) {
    return guarded(err, errlen, [&] {
        if (!pr || !pr->owned) throw std::invalid_argument("null problem");
        auto owned = std::move(pr->owned);
        delete pr;
        auto* s = new pmcs_solver{std::make_unique<pmc::MonteCarloSolver>(std::move(owned), folder ? folder : ""), {}};
        s->problem_view.view = &s->s->problem();
        *out = s;
    });
}

int pmcs_solver_run(pmcs_solver* s, int32_t steps, char* err, int32_t errlen) {
    return guarded(err, errlen, [&] { s->s->run_steps(steps); });
}

int32_t pmcs_solver_timestep(const pmcs_solver* s) { return s->s->current_timestep(); }

pmc_ctx* pmcs_solver_context(pmcs_solver* s) { return s->s->context(); }

const pmcs_problem* pmcs_solver_problem(const pmcs_solver* s) { return s ? &s->problem_view : nullptr; }

int pmcs_solver_observables(pmcs_solver* s, double* kappa_eff, double* kappa_int, double* mean_flux, double* cell_T) {
    if (!s) return PMC_ERR_ARGUMENT;
    if (kappa_eff) *kappa_eff = s->s->kappa_eff();
    if (kappa_int) *kappa_int = s->s->kappa_int();
    if (mean_flux) *mean_flux = s->s->average_heat_flux();
    if (cell_T) {
        const auto t = s->s->cell_temperatures();
        std::memcpy(cell_T, t.data(), t.size() * sizeof(double));
    }
    return PMC_OK;
}

int pmcs_solver_append_summary(pmcs_solver* s, const char* path) {
    if (!s || !path) return PMC_ERR_ARGUMENT;
    std::ofstream out(path, std::ios::app);
    if (!out) return PMC_ERR_ARGUMENT;
    s->s->append_profile_summary(out);
    return out ? PMC_OK : PMC_ERR_ARGUMENT;
}

void pmcs_solver_destroy(pmcs_solver* s) { delete s; }

int pmcs_synthetic_material(int32_t mesh, int32_t callaway, pmcs_material_view* out, void** handle) {
    try {
        pmc::SyntheticSpec spec;
        spec.mesh = mesh;
        spec.callaway_split = callaway != 0;
        auto* m = new pmc::RawMaterial(pmc::make_synthetic_material(spec));
        fill_material(*m, 0.1, out);
        *handle = m;
        return PMC_OK;
    } catch (const std::exception&) {
        return PMC_ERR_ARGUMENT;
    }
}

void pmcs_material_free(void* handle) { delete static_cast<pmc::RawMaterial*>(handle); }

} // extern "C"
