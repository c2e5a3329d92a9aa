/* pmc_solver.h -- C ABI of the host facade (TOML input -> tables -> device step).
 *
 * pmcs_problem holds everything the host builds once from a TOML input: the
 * material tables, domain, rough and reservoir tables and the pmc_capi.h views
 * of them.  pmcs_solver wraps the MonteCarloSolver-like facade that drives the
 * device step and writes the reference's result files.
 *
 * Replaces: load_simulation_config (src/SimulationConfig.cpp:57-65),
 * SimulationDomain(cfg) (src/SimulationDomain.cpp:73-90), PhononMaterial
 * construction (src/PhononMaterial.cpp:31-46) and the MonteCarloSolver facade
 * (include/MonteCarloSolver.h:42-103) as driven by main (src/main.cpp:86-130). */
#ifndef PMC_SOLVER_H
#define PMC_SOLVER_H
#include "pmc_capi.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct pmcs_problem pmcs_problem;
typedef struct pmcs_solver pmcs_solver;

/* Raw phonon table in MaterialData layout (include/material/MaterialData.h:19-33). */
typedef struct pmcs_material_view {
    double volume_a3;
    double reciprocal[9];
    int32_t qpoints, branches;
    int32_t mesh[3];
    const double* frequency_thz;
    const double* qpoint_fractions;
    const double* group_velocity;
    int32_t ntemps;
    const double* temperatures;
    const double* gamma;
    const double* gamma_normal;
    const double* gamma_umklapp;
    const double* gamma_isotope;
    double temperature_lookup_dt;
} pmcs_material_view;

/* Host scalars of a problem: [0] particle volume A^3, [1] particle density 1/A^3,
 * [2] energy weight, [3] background energy density, [4] zero-point density,
 * [5] energy normalization, [6] domain volume, [7] cell count, [8] carriers,
 * [9] push distance A. */
enum { PMCS_SCALARS = 10 };

int pmcs_problem_from_toml(const char* path, int32_t mode_sampling, pmcs_problem** out, char* err, int32_t errlen);
int pmcs_problem_from_text(const char* text, const char* input_directory, int32_t mode_sampling, pmcs_problem** out,
                           char* err, int32_t errlen);
void pmcs_problem_destroy(pmcs_problem* problem);
int pmcs_problem_views(const pmcs_problem* problem, pmc_setup* setup, pmc_mode_table* modes, pmc_geometry* geometry,
                       pmc_init* init, int32_t* n_rough, const pmc_rough_table** rough, int32_t* n_reservoirs,
                       const pmc_reservoir** reservoirs);
int pmcs_problem_material(const pmcs_problem* problem, pmcs_material_view* out);
/* Heat-source view (valid while the problem lives): 1 if the input enables one, 0 if not. */
int pmcs_problem_heat_source(const pmcs_problem* problem, pmc_heat_source* out);
/* Linearized Callaway view (valid while the problem lives): 1 if the input selects
 * scattering.model = 'callaway' (linearized), 0 if not. */
int pmcs_problem_callaway(const pmcs_problem* problem, pmc_callaway_table* out);
/* Nonlinear Callaway rate tables (valid while the problem lives): 1 if the input
 * selects scattering.callaway_formulation = 'nonlinear', 0 if not. */
int pmcs_problem_callaway_nonlinear(const pmcs_problem* problem, pmc_callaway_rates* out);
/* Layered materials: fills *out and returns 1 (0 for one material). */
int pmcs_problem_layers(const pmcs_problem* problem, pmc_layers* out);
/* Audit of the AMM tables of interface `face` (AcousticInterfaceSampler::Audit,
 * include/solver/AcousticInterfaceSampler.h:12-17): incident acoustic flux,
 * transmitted flux and no-overlap flux of sides 0 and 1, geometric overlap flux,
 * capacity-removed flux, largest capacity ratio.  1 when filled, 0 without AMM. */
int pmcs_problem_amm_audit(const pmcs_problem* problem, int32_t face, double* out);
/* Raw table of layer `layer` (materials.folders order). */
int pmcs_problem_layer_material(const pmcs_problem* problem, int32_t layer, pmcs_material_view* out);
/* Writes the procedural FinFET surface (a substrate box with one fin centred on
 * its top, lattice faces as triangles; sizes and lattice step in nm) as a binary
 * STL; returns the triangle count, or -1 with err filled. */
int pmcs_write_finfet_stl(const char* path, const double substrate_nm[3], const double fin_nm[3],
                          const double step_nm[3], char* err, int errlen);
int pmcs_problem_scalars(const pmcs_problem* problem, double* out);
/* Derived material queries, for checks against the reference's PhononMaterial. */
double pmcs_problem_bose(const pmcs_problem* problem, double temperature, int32_t flat_mode);
double pmcs_problem_relaxation_weight(const pmcs_problem* problem, double temperature, int32_t flat_mode, double dt);
double pmcs_problem_temperature_from_energy(const pmcs_problem* problem, double energy_density);
double pmcs_problem_energy_from_temperature(const pmcs_problem* problem, double temperature);
/* Facet table: per facet normal[3], centroid[3], area, condition, shift[3], value, face count. */
int pmcs_problem_facets(const pmcs_problem* problem, int32_t* n, double* normals, double* centroids, double* areas,
                        char* conditions, double* shifts, double* values, int32_t* face_counts);
int pmcs_problem_cells(const pmcs_problem* problem, double* centers, double* volumes);
/* summary.txt ([input], [geometry], [grid], [boundary], [files]) and
 * grid_centers.csv of the problem in `folder`, as the solver writes them at
 * start-up (SimulationDomain::write_domain_summary, src/geometry/DomainOutput.cpp:12-241);
 * host only. */
int pmcs_problem_write_summary(const pmcs_problem* problem, const char* folder, char* err, int32_t errlen);

/* Builds a solver (takes ownership of the problem); output_folder may be NULL for no files. */
int pmcs_solver_create(pmcs_problem* problem, const char* output_folder, pmcs_solver** out, char* err, int32_t errlen);
int pmcs_solver_run(pmcs_solver* solver, int32_t steps, char* err, int32_t errlen);
int32_t pmcs_solver_timestep(const pmcs_solver* solver);
pmc_ctx* pmcs_solver_context(pmcs_solver* solver);
const pmcs_problem* pmcs_solver_problem(const pmcs_solver* solver);
int pmcs_solver_observables(pmcs_solver* solver, double* kappa_eff, double* kappa_int, double* mean_flux,
                            double* cell_temperature);
int pmcs_solver_append_summary(pmcs_solver* solver, const char* path);
void pmcs_solver_destroy(pmcs_solver* solver);

/* Seeded synthetic Si-like table (simple cubic, mesh^3 q grid, 6 branches);
 * the arrays stay valid until pmcs_material_free(handle). */
int pmcs_synthetic_material(int32_t mesh, int32_t callaway_split, pmcs_material_view* out, void** handle);
void pmcs_material_free(void* handle);

#ifdef __cplusplus
}
#endif
#endif
