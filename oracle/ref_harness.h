/* C entry points into the reference PhonoMC sources compiled by oracle/ref.mk.
 *
 * Test infrastructure only: tests/golden/make_golden.py and the CPU baseline leg
 * of bench.py load oracle/_ref/libphonomc_ref.so through these functions to
 * produce known-answer vectors and reference timings.  Nothing in the product
 * links or loads it.
 *
 * Every function returns 0 on success; on failure it returns nonzero and copies
 * the reference's exception text into err (errlen bytes, NUL terminated). */
#pragma once
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* In-memory phonon table in the layout of the reference's MaterialData
 * (include/material/MaterialData.h:19-33): q-major/branch-minor; gamma is
 * temperature-major; frequencies in THz, velocities in Angstrom/ps. */
typedef struct ref_material_arrays {
    double volume_a3;
    double reciprocal[9]; /* row-major, angular (2*pi included), 1/Angstrom */
    int qpoints, branches;
    int mesh[3];
    const double* frequency_thz;    /* qpoints*branches */
    const double* qpoint_fractions; /* qpoints*3 */
    const double* group_velocity;   /* qpoints*branches*3 */
    int ntemps;
    const double* temperatures;   /* ntemps */
    const double* gamma;          /* ntemps*qpoints*branches */
    const double* gamma_normal;   /* optional, same shape as gamma */
    const double* gamma_umklapp;  /* optional, same shape as gamma */
    const double* gamma_isotope;  /* optional, qpoints*branches */
    double temperature_lookup_dt; /* K */
} ref_material_arrays;

int ref_material_create(const ref_material_arrays* arrays, void** out, char* err, int errlen);
/* With the constructor flag MaterialFactory passes under conservative transport
 * (src/material/MaterialFactory.cpp:51-55): stationary modes join the active bank. */
int ref_material_create_stationary(const ref_material_arrays* arrays, int include_stationary, void** out, char* err,
                                   int errlen);
/* PhononMaterial::callaway_data(T) (src/material/PhononMaterialCollisionData.cpp:32-165)
 * by active index: omega, capacity, wavevector (3 per mode), normal and resistive rates. */
int ref_callaway_data(void* material, double temperature, double* omega, double* capacity, double* wavevector,
                      double* normal_rate, double* resistive_rate, char* err, int errlen);
/* CarrierCollisionOperator::configure(T_bg) + apply on a carrier set: occupations
 * updated in place, *residual_ev = the returned lifetime residual. */
int ref_callaway_apply(void* material, double t_bg, int ncells, int n, const int* active, const int* cell,
                       double* occupation, double dt, double particle_volume, double* residual_ev, char* err,
                       int errlen);
/* nonlinear_callaway_evolve (src/solver/NonlinearCallawayCollision.cpp:251-285) on one cell. */
int ref_nl_callaway_evolve(int n, const double* omega, const double* wavevector, const double* normal_rate,
                           const double* resistive_rate, const double* occupation, double temperature, double dt,
                           double* out, char* err, int errlen);
/* PhononMaterial::mode_callaway_rates (src/material/PhononMaterialCollisionData.cpp:167-196). */
int ref_nl_callaway_rates(void* material, double temperature, int n, const int* active, double* normal_rate,
                          double* resistive_rate, char* err, int errlen);
/* NonlinearCallawayCollision::configure(T_bg) + apply (:251-344) on a carrier set. */
int ref_nl_callaway_apply(void* material, double t_bg, int ncells, const double* cell_temperature, int local,
                          double lifetime_temperature, int n, const int* active, const int* cell, double* occupation,
                          double dt, double particle_volume, double* residual_ev, char* err, int errlen);
void ref_material_destroy(void* material);
int ref_material_active_count(void* material);
/* Active-mode list in the reference's sampling order (src/PhononMaterial.cpp:136-156). */
void ref_material_active_modes(void* material, int* q, int* b);
double ref_bose_occupation(void* material, double temperature, int q, int b);
double ref_relaxation_weight(void* material, double temperature, int q, int b, double dt);
double ref_temperature_from_energy_density(void* material, double e);
double ref_energy_density_from_temperature(void* material, double t);
double ref_crystal_energy_density(void* material, double t);
double ref_zero_point_energy_density(void* material);
double ref_energy_density_normalization(void* material);

/* SimulationConfig + SimulationDomain built from a TOML file by the reference parser. */
int ref_domain_create(const char* toml_path, void** out, char* err, int errlen);
/* SimulationDomain::write_domain_summary (src/geometry/DomainOutput.cpp:12-241) of the
 * input with its output folder replaced by `folder`: summary.txt and grid_centers.csv. */
int ref_domain_summary(const char* toml_path, const char* folder, char* err, int errlen);
void ref_domain_destroy(void* domain);
int ref_domain_facet_count(void* domain);
int ref_domain_grid_count(void* domain);
/* Per facet: normal[3], centroid[3], area, condition char, periodic shift[3], roughness,
 * rough flag, face count. */
void ref_domain_facets(void* domain, double* normals, double* centroids, double* areas,
                       char* conditions, double* shifts, double* roughness, int* rough,
                       int* face_counts);
void ref_domain_grid(void* domain, double* centers, double* volumes);
void ref_domain_bounds(void* domain, double* bmin, double* bmax, double* volume);
void ref_domain_trace(void* domain, int n, const double* pos, const double* vel, double* point,
                      double* t, int* facet);
void ref_domain_grid_index(void* domain, int n, const double* pos, int* cell);
/* SimulationDomain::contains_point and SurfaceMesh::nearest_facet per point. */
void ref_domain_contains(void* domain, int n, const double* pos, int* inside);
void ref_domain_nearest_facet(void* domain, int n, const double* pos, int* facet);
/* sample_surface_points on one facet (src/geometry/SurfaceSampling.cpp:20-73) with an
 * mt19937_64 seeded by seed; also returns the three canonical uniforms it consumed per
 * point (face pick, r1, r2) so a restatement can be fed the same values. */
int ref_domain_sample_surface(void* domain, int facet, int n, uint64_t seed, double* points,
                              double* uniforms, char* err, int errlen);

/* RoughBoundarySampler configured as MonteCarloSolver does for non-conservative
 * transport (src/solver/RoughBoundaryInitialization.cpp:9-25), sampled once per
 * input with mt19937_64(seed[i]).  Returns out mode (active index), out occupation
 * and the first two canonical uniforms of that engine. */
int ref_rough_sample(void* domain, void* material, int facet, double reference_temperature,
                     int n, const int* in_active, const double* in_occupation,
                     const uint64_t* seeds, int* out_active, double* out_occupation,
                     double* uniforms, char* err, int errlen);
/* RoughBoundarySampler as configured under conservative (linearized Callaway)
 * transport (src/solver/RoughBoundaryInitialization.cpp:9-25): strict-elastic
 * tables at t_bg on every reflecting facet, sampled with signed weights
 * (src/solver/BoundaryInteraction.cpp:90-113).  Same outputs as ref_rough_sample. */
int ref_rough_sample_strict(void* domain, void* material, int facet, double t_bg, int n, const int* in_active,
                            const double* in_occupation, const uint64_t* seeds, int* out_active,
                            double* out_occupation, double* uniforms, char* err, int errlen);
/* interact_boundary (src/solver/BoundaryInteraction.cpp:53-162) for one facet. */
int ref_interact_boundary(void* domain, void* material, int facet, double reference_temperature,
                          double energy_weight, int n, const int* in_active,
                          const double* in_velocity, const double* in_occupation,
                          const uint64_t* seeds, int* action, int* out_active,
                          double* out_velocity, double* out_occupation, double* absorbed_ev,
                          double* shift, char* err, int errlen);

/* ReservoirSampler for one reservoir facet: rounded initial count and n draws with
 * mt19937_64(seed) (src/solver/ReservoirSampler.cpp:17-146); uniforms holds the two
 * canonical values (mode, time) consumed per draw. */
int ref_reservoir(void* domain, void* material, int facet, double particle_density,
                  double time_step, uint64_t seed, int n, int* initial_count,
                  int* out_active, double* out_time, double* uniforms, char* err, int errlen);

/* GridStatistics::refresh_temperatures on a supplied carrier set. */
int ref_refresh_temperatures(void* material, int ncells, const double* volumes, double t_bg,
                             double particle_volume, int n, const int* active,
                             const int* cell, const double* occupation, double* cell_energy,
                             double* cell_temperature, char* err, int errlen);

/* RtaCollisionOperator::apply on a supplied carrier set; occupation is updated in
 * place and the residual returned through residual_ev. */
/* PrescribedHeatSource::configure cell rates (eV/ps) for the input the domain was built from. */
int ref_heat_source_rates(void* domain, double* rates, char* err, int errlen);
/* PrescribedHeatSource::accrue(integrated_time) + PhononSourceOperator::apply on a carrier set. */
int ref_source_apply(void* domain, void* material, double integrated_time_ps, const double* cell_temperature,
                     int n, const int* active, const int* cell, double* occupation, double particle_volume,
                     double t_bg, double* deposited_ev, long long* updates, char* err, int errlen);
int ref_rta_apply(void* material, int ncells, const double* cell_temperature, int n,
                  const int* active, const int* cell, const double* carrier_temperature,
                  double* occupation, double dt, double particle_volume, int local_lifetime,
                  double lifetime_temperature, double* residual_ev, char* err, int errlen);

/* Whole reference run: TOML input, in-memory material, `warmup` untimed then
 * `steps` timed calls of MonteCarloSolver::run_timestep.  The reference writes
 * its usual result folder (the TOML's io.output_folder).  seconds = wall time
 * of the timed step loop only; particle_steps counts its live carriers. */
int ref_run(const char* toml_path, const ref_material_arrays* arrays, int warmup, int steps, int threads,
            double* seconds, long long* particle_steps, char* err, int errlen);

/* The same over several in-memory materials (materials.folders, in order);
 * interface[2] = MonteCarloSolver::interface_event_count and
 * interface_transmitted_count after the last step; summary (if not null) =
 * MonteCarloSolver::append_profile_summary's text, truncated to summary_len - 1. */
int ref_run_materials(const char* toml_path, const ref_material_arrays* arrays, int n_materials, int warmup,
                      int steps, int threads, double* seconds, long long* particle_steps, long long* interface,
                      char* summary, int summary_len, char* err, int errlen);

/* InterfaceInteraction::interact (DMM) on n carriers: material[i], active
 * index active[i] of that material, velocity[3i..], occupation[i], toward
 * next_material[i], with std::mt19937_64(seeds[i]).  Interface settings as
 * MonteCarloSolver::initialize_interface_mode_banks sets them (axis, bin,
 * reference temperature, per-material energy weights).  Outputs: action (1
 * transmitted, 0 reflected into a new mode, -1 axis mirror of the same mode),
 * material, active index, occupation and velocity after the event, and the
 * uniforms on [0, 1) the event consumed in the order (reflection band,
 * transmission band, transmit choice); 0.5 where the event drew none. */
/* AcousticInterfaceSampler::configure + row (src/solver/AcousticInterfaceSampler.cpp:27-455):
 * the rows of one interface side by the source material's active index (reflected
 * partner, transmitted edges as CSR) and the interface's audit (incident, transmitted,
 * no-overlap flux per side, overlap, capacity removed, largest capacity ratio). */
int ref_amm_rows(const ref_material_arrays* arrays, int n_materials, int axis, const double* densities,
                 double frequency_bin_thz, double parallel_bin_inv_a, int face, int side, int* reflected,
                 int* edge_begin, int* edge_active, double* edge_probability, int edge_capacity, double* audit,
                 char* err, int errlen);
/* InterfaceInteraction::interact for model 0 DMM, 1 AMM, 2 MMM (fraction amm_fraction)
 * on n carriers: outcomes as ref_interface_dmm (action -2: the reference threw) and
 * the draws it made, four per carrier in order (0.5 where none was made). */
int ref_interface_model(const ref_material_arrays* arrays, int n_materials, int axis, int model,
                        double amm_fraction, const double* densities, double parallel_bin_inv_a,
                        double frequency_bin_thz, double reference_temperature, const double* energy_weights, int n,
                        const int* material, const int* active, const double* velocity, const double* occupation,
                        const int* next_material, const uint64_t* seeds, int* action, int* out_material,
                        int* out_active, double* out_occupation, double* out_velocity, double* uniforms, char* err,
                        int errlen);
int ref_interface_dmm(const ref_material_arrays* arrays, int n_materials, int axis, double frequency_bin_thz,
                      double reference_temperature, const double* energy_weights, int n, const int* material,
                      const int* active, const double* velocity, const double* occupation,
                      const int* next_material, const uint64_t* seeds, int* action, int* out_material,
                      int* out_active, double* out_occupation, double* out_velocity, double* uniforms, char* err,
                      int errlen);

#ifdef __cplusplus
}
#endif
