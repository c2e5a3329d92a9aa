// C entry points into the reference sources (see ref_harness.h).  Test
// infrastructure: builds reference objects from in-memory inputs and calls
// their public methods unchanged.
#include "ref_harness.h"

#include "MonteCarloSolver.h"
#include "PhononMaterial.h"
#include "SimulationConfig.h"
#include "solver/PhononSourceOperator.h"
#include "solver/PrescribedHeatSource.h"
#include "SimulationDomain.h"
#include "material/MaterialData.h"
#include "solver/BackgroundCache.h"
#include "solver/BoundaryInteraction.h"
#include "solver/GridState.h"
#include "solver/GridStatistics.h"
#include "solver/ParticleStorage.h"
#include "solver/ReservoirSampler.h"
#include "solver/RoughBoundarySampler.h"
#include "solver/RtaCollisionOperator.h"
#include "solver/CarrierCollisionOperator.h"
#include "solver/NonlinearCallawayCollision.h"
#include "solver/InterfaceInteraction.h"
#include "solver/InterfaceModeSampler.h"
#include "solver/AcousticInterfaceSampler.h"

#include <chrono>
#include <cstring>
#include <exception>
#include <functional>
#include <memory>
#include <random>
#include <sstream>
#include <string>
#include <vector>

#ifdef PHONOMC_USE_OPENMP
#include <omp.h>
#endif

namespace {
void copy_error(const std::exception& e, char* err, int errlen) {
    if (err && errlen > 0) {
        std::strncpy(err, e.what(), static_cast<size_t>(errlen - 1));
        err[errlen - 1] = '\0';
    }
}

template <class F>
int guarded(char* err, int errlen, F&& body) {
    try {
        body();
        return 0;
    } catch (const std::exception& e) {
        copy_error(e, err, errlen);
        return 1;
    }
}

phonomc::MaterialData to_material_data(const ref_material_arrays& a) {
    phonomc::MaterialData d;
    d.lattice.volume_a3 = a.volume_a3;
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) d.lattice.reciprocal[r][c] = a.reciprocal[3 * r + c];
    d.qpoints = a.qpoints;
    d.branches = a.branches;
    d.mesh = {a.mesh[0], a.mesh[1], a.mesh[2]};
    const size_t nm = static_cast<size_t>(a.qpoints) * a.branches;
    d.frequency_thz.assign(a.frequency_thz, a.frequency_thz + nm);
    d.qpoint_fractions.resize(a.qpoints);
    for (int q = 0; q < a.qpoints; ++q)
        d.qpoint_fractions[q] = {a.qpoint_fractions[3 * q], a.qpoint_fractions[3 * q + 1],
                                 a.qpoint_fractions[3 * q + 2]};
    d.group_velocity.resize(nm);
    for (size_t m = 0; m < nm; ++m)
        d.group_velocity[m] = {a.group_velocity[3 * m], a.group_velocity[3 * m + 1],
                               a.group_velocity[3 * m + 2]};
    d.temperatures.assign(a.temperatures, a.temperatures + a.ntemps);
    d.gamma.assign(a.gamma, a.gamma + nm * a.ntemps);
    if (a.gamma_normal && a.gamma_umklapp) {
        d.gamma_normal.assign(a.gamma_normal, a.gamma_normal + nm * a.ntemps);
        d.gamma_umklapp.assign(a.gamma_umklapp, a.gamma_umklapp + nm * a.ntemps);
    }
    if (a.gamma_isotope) d.gamma_resistive["gamma_isotope"].assign(a.gamma_isotope, a.gamma_isotope + nm);
    return d;
}

struct Domain {
    SimulationConfig config;
    std::unique_ptr<SimulationDomain> domain;
};

PhononMaterial& mat(void* p) { return *static_cast<PhononMaterial*>(p); }
Domain& dom(void* p) { return *static_cast<Domain*>(p); }

// Rough sampler configured the way MonteCarloSolver::initialize_rough_boundary_scattering
// does for non-conservative transport.
phonomc::RoughBoundarySampler make_rough(const SimulationDomain& g, const PhononMaterial& m) {
    std::vector<phonomc::RoughFacetSpec> facets;
    for (int f = 0; f < g.mesh().facet_count(); ++f) {
        facets.push_back({g.is_rough_facet(f), g.mesh().facet_normals()[f],
                          std::max(0.0, g.roughness_for_facet(f, 0.0))});
    }
    phonomc::RoughBoundarySampler s;
    s.configure(facets, {&m}, nullptr, false, 0.0);
    return s;
}

// Rough sampler configured the way MonteCarloSolver::initialize_rough_boundary_scattering
// does under conservative transport (src/solver/RoughBoundaryInitialization.cpp:9-25):
// every reflecting facet (not a reservoir, not a paired periodic face) gets a
// strict-elastic table built at the equilibrium temperature.
phonomc::RoughBoundarySampler make_rough_strict(const SimulationDomain& g, const PhononMaterial& m, double t_bg) {
    std::vector<phonomc::RoughFacetSpec> facets;
    for (int f = 0; f < g.mesh().facet_count(); ++f) {
        const char c = g.facet_boundary_condition(f);
        const bool reflecting = c != 'T' && !(c == 'P' && g.has_periodic_pair(f));
        facets.push_back({reflecting, g.mesh().facet_normals()[f], std::max(0.0, g.roughness_for_facet(f, 0.0))});
    }
    phonomc::RoughBoundarySampler s;
    s.configure(facets, {&m}, nullptr, true, t_bg);
    return s;
}

double canonical(std::mt19937_64& rng) {
    return std::uniform_real_distribution<double>(0.0, 1.0)(rng);
}
} // namespace

extern "C" {

int ref_material_create(const ref_material_arrays* arrays, void** out, char* err, int errlen) {
    return guarded(err, errlen, [&] {
        *out = new PhononMaterial(to_material_data(*arrays), arrays->temperature_lookup_dt);
    });
}

int ref_material_create_stationary(const ref_material_arrays* arrays, int include_stationary, void** out, char* err,
                                   int errlen) {
    return guarded(err, errlen, [&] {
        *out = new PhononMaterial(to_material_data(*arrays), arrays->temperature_lookup_dt, include_stationary != 0);
    });
}

int ref_callaway_data(void* material, double temperature, double* omega, double* capacity, double* wavevector,
                      double* normal_rate, double* resistive_rate, char* err, int errlen) {
    return guarded(err, errlen, [&] {
        const auto d = mat(material).callaway_data(temperature);
        for (size_t a = 0; a < d.omega.size(); ++a) {
            omega[a] = d.omega[a];
            capacity[a] = d.capacity[a];
            for (int k = 0; k < 3; ++k) wavevector[3 * a + k] = d.wavevector[a][k];
            normal_rate[a] = d.normal_rate[a];
            resistive_rate[a] = d.resistive_rate[a];
        }
    });
}

int ref_callaway_apply(void* material, double t_bg, int ncells, int n, const int* active, const int* cell,
                       double* occupation, double dt, double particle_volume, double* residual_ev, char* err,
                       int errlen) {
    return guarded(err, errlen, [&] {
        const auto& m = mat(material);
        phonomc::ParticleStorage p;
        p.reset(n);
        for (int i = 0; i < n; ++i) {
            p.modes[i] = m.active_mode_at(active[i]);
            p.grid_ids[i] = cell[i];
            p.occupation[i] = occupation[i];
            p.temperatures[i] = t_bg;
            p.material_ids[i] = 0;
        }
        p.material_spatial_volumes_a3 = {particle_volume};
        phonomc::CellParticleIndex index;
        index.ensure(p, ncells);
        std::vector<int> materials(ncells, 0);
        std::vector<double> temps(ncells, t_bg);
        phonomc::CarrierCollisionOperator op;
        op.configure({&m}, t_bg);
        *residual_ev = op.apply(p, {&m}, {index.counts(), materials, temps, &index}, dt, particle_volume);
        for (int i = 0; i < n; ++i) occupation[i] = p.occupation[i];
    });
}

int ref_nl_callaway_evolve(int n, const double* omega, const double* wavevector, const double* normal_rate,
                           const double* resistive_rate, const double* occupation, double temperature, double dt,
                           double* out, char* err, int errlen) {
    return guarded(err, errlen, [&] {
        phonomc::CallawayData d;
        d.omega.assign(omega, omega + n);
        d.normal_rate.assign(normal_rate, normal_rate + n);
        d.resistive_rate.assign(resistive_rate, resistive_rate + n);
        d.wavevector.resize(static_cast<size_t>(n));
        for (int i = 0; i < n; ++i)
            for (int k = 0; k < 3; ++k) d.wavevector[static_cast<size_t>(i)][static_cast<size_t>(k)] = wavevector[3 * i + k];
        const auto r = phonomc::nonlinear_callaway_evolve(d, std::vector<double>(occupation, occupation + n),
                                                          temperature, dt);
        std::copy(r.begin(), r.end(), out);
    });
}

int ref_nl_callaway_rates(void* material, double temperature, int n, const int* active, double* normal_rate,
                          double* resistive_rate, char* err, int errlen) {
    return guarded(err, errlen, [&] {
        const auto& m = mat(material);
        for (int i = 0; i < n; ++i) {
            const auto r = m.mode_callaway_rates(temperature, m.active_mode_at(active[i]));
            normal_rate[i] = r[0];
            resistive_rate[i] = r[1];
        }
    });
}

int ref_nl_callaway_apply(void* material, double t_bg, int ncells, const double* cell_temperature, int local,
                          double lifetime_temperature, int n, const int* active, const int* cell, double* occupation,
                          double dt, double particle_volume, double* residual_ev, char* err, int errlen) {
    return guarded(err, errlen, [&] {
        const auto& m = mat(material);
        phonomc::ParticleStorage p;
        p.reset(n);
        for (int i = 0; i < n; ++i) {
            p.modes[i] = m.active_mode_at(active[i]);
            p.grid_ids[i] = cell[i];
            p.occupation[i] = occupation[i];
            p.temperatures[i] = cell_temperature[cell[i]];
            p.material_ids[i] = 0;
        }
        p.material_spatial_volumes_a3 = {particle_volume};
        phonomc::CellParticleIndex index;
        index.ensure(p, ncells);
        std::vector<int> materials(ncells, 0);
        std::vector<double> temps(cell_temperature, cell_temperature + ncells);
        phonomc::NonlinearCallawayCollision op;
        op.configure({&m}, t_bg);
        *residual_ev = op.apply(p, {&m}, {index.counts(), materials, temps, &index},
                                {dt, particle_volume, local != 0, lifetime_temperature});
        for (int i = 0; i < n; ++i) occupation[i] = p.occupation[i];
    });
}

void ref_material_destroy(void* material) { delete static_cast<PhononMaterial*>(material); }

int ref_material_active_count(void* material) { return mat(material).active_mode_count(); }

void ref_material_active_modes(void* material, int* q, int* b) {
    const auto& list = mat(material).active_mode_list();
    for (size_t i = 0; i < list.size(); ++i) {
        q[i] = list[i][0];
        b[i] = list[i][1];
    }
}

double ref_bose_occupation(void* material, double t, int q, int b) {
    return mat(material).bose_occupation(t, {q, b});
}

double ref_relaxation_weight(void* material, double t, int q, int b, double dt) {
    return mat(material).mode_relaxation_weight(t, {q, b}, dt);
}

double ref_temperature_from_energy_density(void* material, double e) {
    return mat(material).temperature_from_energy_density(e);
}

double ref_energy_density_from_temperature(void* material, double t) {
    return mat(material).energy_density_from_temperature(t);
}

double ref_crystal_energy_density(void* material, double t) {
    return mat(material).crystal_energy_density(t);
}

double ref_zero_point_energy_density(void* material) {
    return mat(material).zero_point_energy_density();
}

double ref_energy_density_normalization(void* material) {
    return mat(material).energy_density_normalization();
}

int ref_domain_create(const char* toml_path, void** out, char* err, int errlen) {
    return guarded(err, errlen, [&] {
        auto d = std::make_unique<Domain>();
        d->config = load_simulation_config(toml_path);
        d->config.output_folder.clear();
        d->domain = std::make_unique<SimulationDomain>(d->config);
        *out = d.release();
    });
}

int ref_domain_summary(const char* toml_path, const char* folder, char* err, int errlen) {
    return guarded(err, errlen, [&] {
        SimulationConfig config = load_simulation_config(toml_path);
        config.output_folder = folder;
        SimulationDomain domain(config); // the constructor writes the summary (src/SimulationDomain.cpp:86)
    });
}

void ref_domain_destroy(void* domain) { delete static_cast<Domain*>(domain); }

int ref_domain_facet_count(void* domain) { return dom(domain).domain->mesh().facet_count(); }

int ref_domain_grid_count(void* domain) { return dom(domain).domain->grid_count(); }

void ref_domain_facets(void* domain, double* normals, double* centroids, double* areas,
                       char* conditions, double* shifts, double* roughness, int* rough,
                       int* face_counts) {
    const auto& g = *dom(domain).domain;
    const auto& m = g.mesh();
    for (int f = 0; f < m.facet_count(); ++f) {
        for (int a = 0; a < 3; ++a) {
            normals[3 * f + a] = m.facet_normals()[f][a];
            centroids[3 * f + a] = m.facet_centroids()[f][a];
            shifts[3 * f + a] = g.periodic_shift_for_facet(f)[a];
        }
        areas[f] = m.facet_areas()[f];
        conditions[f] = g.facet_boundary_condition(f);
        roughness[f] = g.roughness_for_facet(f, 0.0);
        rough[f] = g.is_rough_facet(f) ? 1 : 0;
        face_counts[f] = static_cast<int>(m.facets()[f].faces.size());
    }
}

void ref_domain_grid(void* domain, double* centers, double* volumes) {
    const auto& g = *dom(domain).domain;
    for (int c = 0; c < g.grid_count(); ++c) {
        for (int a = 0; a < 3; ++a) centers[3 * c + a] = g.grid_centers()[c][a];
        volumes[c] = g.grid_volumes()[c];
    }
}

void ref_domain_bounds(void* domain, double* bmin, double* bmax, double* volume) {
    const auto& g = *dom(domain).domain;
    for (int a = 0; a < 3; ++a) {
        bmin[a] = g.bounds_min()[a];
        bmax[a] = g.bounds_max()[a];
    }
    *volume = g.volume();
}

void ref_domain_trace(void* domain, int n, const double* pos, const double* vel, double* point,
                      double* t, int* facet) {
    const auto& g = *dom(domain).domain;
    for (int i = 0; i < n; ++i) {
        auto [cp, th, f] = g.trace_boundary_intersection({pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]},
                                                         {vel[3 * i], vel[3 * i + 1], vel[3 * i + 2]});
        for (int a = 0; a < 3; ++a) point[3 * i + a] = cp[a];
        t[i] = th;
        facet[i] = f;
    }
}

void ref_domain_grid_index(void* domain, int n, const double* pos, int* cell) {
    const auto& g = *dom(domain).domain;
    for (int i = 0; i < n; ++i) cell[i] = g.fast_grid_index({pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]});
}

void ref_domain_contains(void* domain, int n, const double* pos, int* inside) {
    const auto& g = *dom(domain).domain;
    for (int i = 0; i < n; ++i) inside[i] = g.contains_point({pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]}) ? 1 : 0;
}

void ref_domain_nearest_facet(void* domain, int n, const double* pos, int* facet) {
    const auto& g = *dom(domain).domain;
    for (int i = 0; i < n; ++i) facet[i] = g.mesh().nearest_facet({pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]});
}

int ref_domain_sample_surface(void* domain, int facet, int n, uint64_t seed, double* points,
                              double* uniforms, char* err, int errlen) {
    return guarded(err, errlen, [&] {
        const auto& g = *dom(domain).domain;
        std::mt19937_64 rng(seed);
        std::mt19937_64 probe = rng;
        const auto pts = g.mesh().sample_surface_points(n, {facet}, rng);
        for (int i = 0; i < n; ++i) {
            for (int a = 0; a < 3; ++a) points[3 * i + a] = pts[i][a];
            for (int k = 0; k < 3; ++k) uniforms[3 * i + k] = canonical(probe);
        }
    });
}

int ref_rough_sample(void* domain, void* material, int facet, double reference_temperature,
                     int n, const int* in_active, const double* in_occupation,
                     const uint64_t* seeds, int* out_active, double* out_occupation,
                     double* uniforms, char* err, int errlen) {
    return guarded(err, errlen, [&] {
        const auto& g = *dom(domain).domain;
        const auto& m = mat(material);
        const auto sampler = make_rough(g, m);
        const int table = sampler.table_index(0, facet);
        for (int i = 0; i < n; ++i) {
            std::mt19937_64 rng(seeds[i]);
            std::mt19937_64 probe = rng;
            uniforms[2 * i] = canonical(probe);
            uniforms[2 * i + 1] = canonical(probe);
            const auto r = sampler.sample(m, table, m.active_mode_at(in_active[i]), in_occupation[i],
                                          reference_temperature, rng, false);
            out_active[i] = m.active_index_for_mode(r.mode);
            out_occupation[i] = r.occupation;
        }
    });
}

int ref_rough_sample_strict(void* domain, void* material, int facet, double t_bg, int n, const int* in_active,
                            const double* in_occupation, const uint64_t* seeds, int* out_active,
                            double* out_occupation, double* uniforms, char* err, int errlen) {
    return guarded(err, errlen, [&] {
        const auto& g = *dom(domain).domain;
        const auto& m = mat(material);
        const auto sampler = make_rough_strict(g, m, t_bg);
        const int table = sampler.table_index(0, facet);
        if (table < 0) throw std::invalid_argument("facet has no strict-elastic table");
        for (int i = 0; i < n; ++i) {
            std::mt19937_64 rng(seeds[i]);
            std::mt19937_64 probe = rng;
            uniforms[2 * i] = canonical(probe);
            uniforms[2 * i + 1] = canonical(probe);
            // Signed weights, as interact_boundary passes under conservative
            // transport (src/solver/BoundaryInteraction.cpp:90-113).
            const auto r = sampler.sample(m, table, m.active_mode_at(in_active[i]), in_occupation[i], t_bg, rng, true);
            out_active[i] = m.active_index_for_mode(r.mode);
            out_occupation[i] = r.occupation;
        }
    });
}

int ref_interact_boundary(void* domain, void* material, int facet, double reference_temperature,
                          double energy_weight, int n, const int* in_active,
                          const double* in_velocity, const double* in_occupation,
                          const uint64_t* seeds, int* action, int* out_active,
                          double* out_velocity, double* out_occupation, double* absorbed_ev,
                          double* shift, char* err, int errlen) {
    return guarded(err, errlen, [&] {
        const auto& g = *dom(domain).domain;
        const auto& m = mat(material);
        const auto sampler = make_rough(g, m);
        // Hit record assembled as TransportOperator::process_boundary_collision does.
        phonomc::BoundaryHit hit;
        hit.valid = facet >= 0 && facet < g.mesh().facet_count();
        hit.condition = g.facet_boundary_condition(facet);
        hit.outward_normal = g.mesh().facet_normals()[facet];
        hit.has_periodic_pair = g.has_periodic_pair(facet);
        if (hit.has_periodic_pair) hit.periodic_shift = g.periodic_shift_for_facet(facet);
        hit.rough_table = sampler.table_index(0, facet);
        if (hit.condition == 'R' && hit.rough_table < 0) hit.roughness = g.roughness_for_facet(facet, 0.0);
        for (int i = 0; i < n; ++i) {
            const auto mode = m.active_mode_at(in_active[i]);
            const phonomc::BoundaryState incoming{
                mode, 0, {in_velocity[3 * i], in_velocity[3 * i + 1], in_velocity[3 * i + 2]},
                m.mode_angular_frequency(mode), in_occupation[i]};
            std::mt19937_64 rng(seeds[i]);
            const auto r = phonomc::interact_boundary(incoming, hit, {false, reference_temperature}, m,
                                                      energy_weight, sampler, rng);
            action[i] = static_cast<int>(r.action);
            out_active[i] = m.active_index_for_mode(r.state.mode);
            for (int a = 0; a < 3; ++a) {
                out_velocity[3 * i + a] = r.state.velocity[a];
                shift[3 * i + a] = r.position_shift[a];
            }
            out_occupation[i] = r.state.occupation;
            absorbed_ev[i] = r.absorbed_energy_ev;
        }
    });
}

int ref_reservoir(void* domain, void* material, int facet, double particle_density,
                  double time_step, uint64_t seed, int n, int* initial_count,
                  int* out_active, double* out_time, double* uniforms, char* err, int errlen) {
    return guarded(err, errlen, [&] {
        const auto& g = *dom(domain).domain;
        const auto& m = mat(material);
        phonomc::ReservoirSampler sampler;
        sampler.configure({{facet, 0, g.reservoir_value_for_facet(facet, 300.0),
                            g.mesh().facet_areas()[facet], g.mesh().facet_normals()[facet]}},
                          {&m}, g.mesh().facet_count(), particle_density, time_step);
        *initial_count = sampler.initial_emission_count(0);
        std::mt19937_64 rng(seed);
        std::mt19937_64 probe = rng;
        const auto draws = sampler.sample(0, n, rng);
        for (int i = 0; i < n; ++i) {
            out_active[i] = static_cast<int>(draws[i].active_index);
            out_time[i] = draws[i].time_ps;
            uniforms[2 * i] = canonical(probe);
            uniforms[2 * i + 1] = canonical(probe);
        }
    });
}

int ref_refresh_temperatures(void* material, int ncells, const double* volumes, double t_bg,
                             double particle_volume, int n, const int* active,
                             const int* cell, const double* occupation, double* cell_energy,
                             double* cell_temperature, char* err, int errlen) {
    return guarded(err, errlen, [&] {
        const auto& m = mat(material);
        std::vector<phonomc::BackgroundCache> background;
        background.emplace_back(m, t_bg, particle_volume, false);
        phonomc::ParticleStorage p;
        p.reset(n);
        for (int i = 0; i < n; ++i) {
            p.modes[i] = m.active_mode_at(active[i]);
            p.grid_ids[i] = cell[i];
            p.occupation[i] = occupation[i];
            p.material_ids[i] = 0;
        }
        phonomc::GridState grid;
        grid.material_ids.assign(ncells, 0);
        std::vector<double> vol(volumes, volumes + ncells);
        phonomc::GridStatistics stats;
        stats.refresh_temperatures(p, grid, vol, {&m}, background, false);
        for (int c = 0; c < ncells; ++c) {
            cell_energy[c] = grid.energy_density[c];
            cell_temperature[c] = grid.temperatures[c];
        }
    });
}

int ref_rta_apply(void* material, int ncells, const double* cell_temperature, int n,
                  const int* active, const int* cell, const double* carrier_temperature,
                  double* occupation, double dt, double particle_volume, int local_lifetime,
                  double lifetime_temperature, double* residual_ev, char* err, int errlen) {
    return guarded(err, errlen, [&] {
        const auto& m = mat(material);
        phonomc::ParticleStorage p;
        p.reset(n);
        for (int i = 0; i < n; ++i) {
            p.modes[i] = m.active_mode_at(active[i]);
            p.grid_ids[i] = cell[i];
            p.occupation[i] = occupation[i];
            p.temperatures[i] = carrier_temperature[i];
            p.material_ids[i] = 0;
        }
        phonomc::CellParticleIndex index;
        index.ensure(p, ncells);
        std::vector<int> materials(ncells, 0);
        std::vector<double> temps(cell_temperature, cell_temperature + ncells);
        phonomc::RtaCollisionOperator rta;
        *residual_ev = rta.apply(p, {&m}, {index.counts(), materials, temps, &index},
                                 {dt, particle_volume, local_lifetime != 0, lifetime_temperature});
        for (int i = 0; i < n; ++i) occupation[i] = p.occupation[i];
    });
}

int ref_heat_source_rates(void* domain, double* rates, char* err, int errlen) {
    return guarded(err, errlen, [&] {
        phonomc::PrescribedHeatSource hs;
        hs.configure(dom(domain).config, *dom(domain).domain);
        if (!hs.enabled()) throw std::runtime_error("input enables no heat source");
        const auto& r = hs.cell_rates_evps();
        std::copy(r.begin(), r.end(), rates);
    });
}

int ref_source_apply(void* domain, void* material, double integrated_time_ps, const double* cell_temperature,
                     int n, const int* active, const int* cell, double* occupation, double particle_volume,
                     double t_bg, double* deposited_ev, long long* updates, char* err, int errlen) {
    return guarded(err, errlen, [&] {
        const auto& d = dom(domain);
        const auto& m = mat(material);
        const int ncells = d.domain->grid_count();
        phonomc::PrescribedHeatSource hs;
        hs.configure(d.config, *d.domain);
        hs.accrue(integrated_time_ps);
        phonomc::PhononSourceOperator op;
        op.configure(d.config, {&m});
        phonomc::ParticleStorage p;
        p.reset(n);
        for (int i = 0; i < n; ++i) {
            p.modes[i] = m.active_mode_at(active[i]);
            p.grid_ids[i] = cell[i];
            p.occupation[i] = occupation[i];
            p.material_ids[i] = 0;
            p.alive[i] = 1;
        }
        phonomc::CellParticleIndex index;
        index.ensure(p, ncells);
        std::vector<int> materials(ncells, 0);
        std::vector<double> temps(cell_temperature, cell_temperature + ncells);
        const auto r = op.apply(p, {&m}, {materials, temps, &index}, hs, particle_volume, t_bg);
        *deposited_ev = r.energy_ev;
        *updates = r.occupation_updates;
        for (int i = 0; i < n; ++i) occupation[i] = p.occupation[i];
    });
}

int ref_run(const char* toml_path, const ref_material_arrays* arrays, int warmup, int steps, int threads,
            double* seconds, long long* particle_steps, char* err, int errlen) {
    long long interface[2];
    return ref_run_materials(toml_path, arrays, 1, warmup, steps, threads, seconds, particle_steps, interface, nullptr,
                             0, err, errlen);
}

int ref_run_materials(const char* toml_path, const ref_material_arrays* arrays, int n_materials, int warmup,
                      int steps, int threads, double* seconds, long long* particle_steps, long long* interface,
                      char* summary, int summary_len, char* err, int errlen) {
    return guarded(err, errlen, [&] {
#ifdef PHONOMC_USE_OPENMP
        if (threads > 0) omp_set_num_threads(threads);
#else
        (void)threads;
#endif
        SimulationConfig config = load_simulation_config(toml_path);
        SimulationDomain domain(config);
        // Constructor flags of MaterialFactory (src/material/MaterialFactory.cpp:51-55).
        std::vector<std::unique_ptr<PhononMaterial>> storage;
        std::vector<std::reference_wrapper<const PhononMaterial>> materials;
        for (int k = 0; k < n_materials; ++k) {
            storage.push_back(std::make_unique<PhononMaterial>(
                to_material_data(arrays[k]), config.temperature_lookup_dt,
                uses_conservative_transport(config) || config.heat_source_spectrum == HeatSourceSpectrum::Gaussian));
            materials.push_back(std::cref(*storage.back()));
        }
        MonteCarloSolver solver(config, domain, materials);
        for (int s = 0; s < warmup; ++s) solver.run_timestep();
        long long carried = 0;
        const auto begin = std::chrono::steady_clock::now();
        for (int s = 0; s < steps; ++s) {
            carried += solver.particle_count();
            solver.run_timestep();
        }
        const auto end = std::chrono::steady_clock::now();
        *seconds = std::chrono::duration<double>(end - begin).count();
        *particle_steps = carried;
        interface[0] = solver.interface_event_count();
        interface[1] = solver.interface_transmitted_count();
        if (summary && summary_len > 0) {
            std::ostringstream text;
            solver.append_profile_summary(text);
            const std::string t = text.str();
            const size_t k = std::min(t.size(), static_cast<size_t>(summary_len - 1));
            std::memcpy(summary, t.data(), k);
            summary[k] = '\0';
        }
    });
}

int ref_amm_rows(const ref_material_arrays* arrays, int n_materials, int axis, const double* densities,
                 double frequency_bin_thz, double parallel_bin_inv_a, int face, int side, int* reflected,
                 int* edge_begin, int* edge_active, double* edge_probability, int edge_capacity, double* audit,
                 char* err, int errlen) {
    return guarded(err, errlen, [&] {
        std::vector<std::unique_ptr<PhononMaterial>> storage;
        std::vector<const PhononMaterial*> mats;
        for (int k = 0; k < n_materials; ++k) {
            storage.push_back(std::make_unique<PhononMaterial>(to_material_data(arrays[k]),
                                                               arrays[k].temperature_lookup_dt));
            mats.push_back(storage.back().get());
        }
        phonomc::AcousticInterfaceSampler amm;
        amm.configure(mats, axis, std::vector<double>(densities, densities + n_materials), frequency_bin_thz,
                      parallel_bin_inv_a);
        const int source = face + side, destination = face + 1 - side;
        const int n = mats[static_cast<size_t>(source)]->active_mode_count();
        int e = 0;
        edge_begin[0] = 0;
        for (int i = 0; i < n; ++i) {
            const auto& r = amm.row(source, destination, i);
            reflected[i] = r.reflected;
            for (const auto& edge : r.transmitted) {
                if (e >= edge_capacity) throw std::runtime_error("edge capacity exceeded");
                edge_active[e] = edge.active;
                edge_probability[e] = edge.probability;
                ++e;
            }
            edge_begin[i + 1] = e;
        }
        const auto& a = amm.audit(face);
        const double values[9] = {a.incident_acoustic_flux[0], a.incident_acoustic_flux[1], a.transmitted_flux[0],
                                  a.transmitted_flux[1], a.no_overlap_flux[0], a.no_overlap_flux[1],
                                  a.geometric_overlap_flux, a.capacity_removed_flux, a.maximum_capacity_ratio};
        std::copy(values, values + 9, audit);
    });
}

int ref_interface_model(const ref_material_arrays* arrays, int n_materials, int axis, int model,
                        double amm_fraction, const double* densities, double parallel_bin_inv_a,
                        double frequency_bin_thz, double reference_temperature, const double* energy_weights, int n,
                        const int* material, const int* active, const double* velocity, const double* occupation,
                        const int* next_material, const uint64_t* seeds, int* action, int* out_material,
                        int* out_active, double* out_occupation, double* out_velocity, double* uniforms, char* err,
                        int errlen) {
    return guarded(err, errlen, [&] {
        std::vector<std::unique_ptr<PhononMaterial>> storage;
        std::vector<const PhononMaterial*> mats;
        for (int k = 0; k < n_materials; ++k) {
            storage.push_back(std::make_unique<PhononMaterial>(to_material_data(arrays[k]),
                                                               arrays[k].temperature_lookup_dt));
            mats.push_back(storage.back().get());
        }
        phonomc::InterfaceInteractionConfig cfg;
        cfg.axis = axis;
        cfg.enabled = true;
        cfg.linearized = false;
        cfg.model = model == 1 ? InterfaceModel::AcousticMismatch
                    : model == 2 ? InterfaceModel::MixedMismatch
                                 : InterfaceModel::DiffuseMismatch;
        cfg.frequency_bin_thz = frequency_bin_thz;
        cfg.amm_fraction = amm_fraction;
        cfg.parallel_bin_inv_a = parallel_bin_inv_a;
        cfg.mass_densities_kg_m3.assign(densities, densities + n_materials);
        cfg.reference_temperature = reference_temperature;
        phonomc::InterfaceInteraction interaction;
        interaction.configure(cfg, mats, std::vector<double>(energy_weights, energy_weights + n_materials));
        phonomc::InterfaceModeSampler bands;
        bands.configure(mats, axis);
        const double band_width = 2.0 * std::acos(-1.0) * frequency_bin_thz;
        const double p_amm = model == 1 ? 1.0 : model == 2 ? amm_fraction : 0.0;
        for (int i = 0; i < n; ++i) {
            const PhononMaterial& src = *mats.at(static_cast<size_t>(material[i]));
            phonomc::BoundaryState in;
            in.material = material[i];
            in.mode = src.active_mode_at(active[i]);
            for (int a = 0; a < 3; ++a) in.velocity[static_cast<size_t>(a)] = velocity[3 * i + a];
            in.omega = src.mode_angular_frequency(in.mode);
            in.occupation = occupation[i];
            // The draws the reference makes, in order, from a copy of its engine:
            // the channel draw under mixed mismatch, then the AMM row draw or the
            // DMM draws (a band draw only when the band carries flux, the choice
            // only when the transmission band does); 0.5 marks a draw not made.
            std::mt19937_64 rng(seeds[i]);
            std::mt19937_64 probe = rng;
            double* u = uniforms + 4 * i;
            for (int k = 0; k < 4; ++k) u[k] = 0.5;
            int base = 0;
            bool amm = p_amm >= 1;
            if (p_amm > 0 && p_amm < 1) {
                u[0] = canonical(probe);
                amm = u[0] < p_amm;
                base = 1;
            }
            if (amm) {
                if (!(std::abs(in.velocity[static_cast<size_t>(axis)]) < 1e-7)) u[base] = canonical(probe);
            } else {
                const int inc = in.velocity[static_cast<size_t>(axis)] > 0.0 ? 1 : 0;
                double fr = 0.0, ft = 0.0;
                std::mt19937_64 scratch(0);
                bands.sample_band(material[i], 1 - inc, in.omega, band_width, scratch, fr);
                bands.sample_band(next_material[i], inc, in.omega, band_width, scratch, ft);
                u[base] = fr > 0.0 ? canonical(probe) : 0.5;
                u[base + 1] = ft > 0.0 ? canonical(probe) : 0.5;
                u[base + 2] = ft > 0.0 ? canonical(probe) : 0.5;
            }
            try {
                const auto r = interaction.interact(in, next_material[i], rng);
                if (r.action == phonomc::BoundaryAction::Transmit) action[i] = 1;
                else if (r.state.material == in.material && r.state.mode == in.mode) action[i] = -1;
                else action[i] = 0;
                out_material[i] = r.state.material;
                out_active[i] = mats.at(static_cast<size_t>(r.state.material))->active_index_for_mode(r.state.mode);
                out_occupation[i] = r.state.occupation;
                for (int a = 0; a < 3; ++a) out_velocity[3 * i + a] = r.state.velocity[static_cast<size_t>(a)];
            } catch (const std::exception&) {
                action[i] = -2;
                out_material[i] = material[i];
                out_active[i] = active[i];
                out_occupation[i] = occupation[i];
                for (int a = 0; a < 3; ++a) out_velocity[3 * i + a] = velocity[3 * i + a];
            }
        }
    });
}

int ref_interface_dmm(const ref_material_arrays* arrays, int n_materials, int axis, double frequency_bin_thz,
                      double reference_temperature, const double* energy_weights, int n, const int* material,
                      const int* active, const double* velocity, const double* occupation,
                      const int* next_material, const uint64_t* seeds, int* action, int* out_material,
                      int* out_active, double* out_occupation, double* out_velocity, double* uniforms, char* err,
                      int errlen) {
    return guarded(err, errlen, [&] {
        std::vector<std::unique_ptr<PhononMaterial>> storage;
        std::vector<const PhononMaterial*> mats;
        for (int k = 0; k < n_materials; ++k) {
            storage.push_back(std::make_unique<PhononMaterial>(to_material_data(arrays[k]),
                                                               arrays[k].temperature_lookup_dt));
            mats.push_back(storage.back().get());
        }
        phonomc::InterfaceInteractionConfig cfg;
        cfg.axis = axis;
        cfg.enabled = true;
        cfg.linearized = false;
        cfg.model = InterfaceModel::DiffuseMismatch;
        cfg.frequency_bin_thz = frequency_bin_thz;
        cfg.reference_temperature = reference_temperature;
        phonomc::InterfaceInteraction interaction;
        interaction.configure(cfg, mats, std::vector<double>(energy_weights, energy_weights + n_materials));
        // A second sampler on the same tables tells which draws an event makes
        // (a band draw only when the band carries flux, the choice only when
        // the transmission band does).
        phonomc::InterfaceModeSampler bands;
        bands.configure(mats, axis);
        const double band_width = 2.0 * std::acos(-1.0) * frequency_bin_thz;
        for (int i = 0; i < n; ++i) {
            const PhononMaterial& src = *mats.at(static_cast<size_t>(material[i]));
            phonomc::BoundaryState in;
            in.material = material[i];
            in.mode = src.active_mode_at(active[i]);
            for (int a = 0; a < 3; ++a) in.velocity[static_cast<size_t>(a)] = velocity[3 * i + a];
            in.omega = src.mode_angular_frequency(in.mode);
            in.occupation = occupation[i];
            const int inc = in.velocity[static_cast<size_t>(axis)] > 0.0 ? 1 : 0;
            double fr = 0.0, ft = 0.0;
            std::mt19937_64 scratch(0);
            bands.sample_band(material[i], 1 - inc, in.omega, band_width, scratch, fr);
            bands.sample_band(next_material[i], inc, in.omega, band_width, scratch, ft);
            std::mt19937_64 rng(seeds[i]);
            std::mt19937_64 probe = rng;
            double* u = uniforms + 3 * i;
            u[0] = fr > 0.0 ? canonical(probe) : 0.5;
            u[1] = ft > 0.0 ? canonical(probe) : 0.5;
            u[2] = ft > 0.0 ? canonical(probe) : 0.5;
            const auto r = interaction.interact(in, next_material[i], rng);
            if (r.action == phonomc::BoundaryAction::Transmit) action[i] = 1;
            else if (r.state.material == in.material && r.state.mode == in.mode) action[i] = -1;
            else action[i] = 0;
            out_material[i] = r.state.material;
            out_active[i] = mats.at(static_cast<size_t>(r.state.material))->active_index_for_mode(r.state.mode);
            out_occupation[i] = r.state.occupation;
            for (int a = 0; a < 3; ++a) out_velocity[3 * i + a] = r.state.velocity[static_cast<size_t>(a)];
        }
    });
}
} // extern "C"
