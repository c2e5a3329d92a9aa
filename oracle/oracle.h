/* oracle.h -- CPU restatement (FP64, OpenMP) of the reference's deviational
 * Monte Carlo step, used as the checker of the HIP path.
 *
 * Test infrastructure only: tests/, __graft_entry__.smoke() and the CPU
 * baseline leg of bench.py load oracle/_build/liboracle.so.  The product never
 * links or calls it.
 *
 * Functions restate the reference (lyushisyan/PhonoMC) file:line cited at each
 * definition in oracle.cpp.  Random draws come from Philox4x32-10 with the same
 * counter layout as the device path (id, step, draw, stream), so a GPU run and
 * an oracle run started from the same population see the same uniforms. */
#ifndef PMC_ORACLE_H
#define PMC_ORACLE_H
#include <stdint.h>

#include "pmc_capi.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Philox4x32-10 block function. */
void orc_philox(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0, uint32_t k1, uint32_t out[4]);
double orc_u53(uint32_t lo, uint32_t hi);

/* Analytic box trace (src/SimulationDomain.cpp:342-450): returns facet or -1. */
int orc_trace_box(const pmc_geometry* g, const double pos[3], const double vel[3], double point[3], double* t);
/* fast_grid_index (src/SimulationDomain.cpp:92-126). */
int orc_grid_index(const pmc_geometry* g, const double pos[3]);
/* Mesh geometry (kind 1): SurfaceMesh::trace_boundary_intersection, contains_point,
 * nearest_facet and sample_volume_points on Philox draws (oracle.cpp). */
int orc_trace_mesh(const pmc_geometry* g, const double pos[3], const double vel[3], double point[3], double* t);
int orc_contains_mesh(const pmc_geometry* g, const double pos[3]);
int orc_nearest_facet(const pmc_geometry* g, const double pos[3]);
int orc_sample_volume(const pmc_geometry* g, uint32_t k0, uint32_t k1, uint32_t id, double point[3], double* u_mode);

/* RoughBoundarySampler::sample (src/solver/RoughBoundarySampler.cpp:459-533) fed
 * explicit uniforms u[0..2] in draw order; returns the number consumed. */
int orc_rough_sample(const pmc_mode_table* m, const pmc_rough_table* r, int in_active, double in_occupation,
                     const double* u, int* out_active, double* out_occupation, int* kind);
/* The same with signed_weights (RoughBoundarySampler::sample's last argument):
 * mapped occupations may be negative under linearized transport. */
int orc_rough_sample_signed(const pmc_mode_table* m, const pmc_rough_table* r, int in_active, double in_occupation,
                            const double* u, int* out_active, double* out_occupation, int* kind, int signed_weights);

/* Linearized Callaway collisions of a carrier set (CarrierCollisionOperator::apply,
 * src/solver/CarrierCollisionOperator.cpp:38-142, with CallawayOperator,
 * src/material/CallawayOperator.cpp:9-201, in long double): occupations updated in
 * place, returns the lifetime residual (eV); status 0, or -9 work limit, -10 no
 * convergence, -11 invalid table. */
double orc_callaway_apply(const pmc_mode_table* m, const pmc_callaway_table* t, int ncells, int64_t n,
                          const int32_t* active, const int32_t* cell, double* occupation, double dt,
                          double energy_weight, int* status);

/* Reservoir mode/time draw (src/solver/ReservoirSampler.cpp:127-146 for
 * sampling = 1; Walker alias for sampling = 0) and surface point
 * (src/geometry/SurfaceSampling.cpp:54-72). */
int orc_reservoir_mode(const pmc_reservoir* r, int n_active, double u, int sampling);
void orc_surface_point(const pmc_reservoir* r, double u_face, double r1, double r2, double point[3]);

/* GridStatistics::refresh_temperatures (src/solver/GridStatistics.cpp:95-235). */
void orc_refresh(const pmc_mode_table* m, int ncells, const double* volume, int64_t n, const int32_t* active,
                 const int32_t* cell, const double* occupation, double* energy_density, double* temperature);

/* RtaCollisionOperator::apply (src/solver/RtaCollisionOperator.cpp:13-189);
 * occupation updated in place, pseudo-temperatures written per cell, residual returned. */
double orc_rta(const pmc_mode_table* m, int ncells, const double* cell_temperature, int64_t n,
               const int32_t* active, const int32_t* cell, double* occupation, double dt, double energy_weight,
               int local_lifetime, double lifetime_temperature, double* pseudo_temperature);

/* Whole-step simulator. */
typedef struct orc_sim orc_sim;
int orc_create(const pmc_setup* setup, const pmc_mode_table* modes, const pmc_geometry* geometry, int32_t n_rough,
               const pmc_rough_table* rough, int32_t n_res, const pmc_reservoir* res, orc_sim** out);
/* Same population as pmc_initialize_carriers (positions held in FP64). */
int orc_initialize(orc_sim* sim, const pmc_init* init);
/* Explicit population: FP32 positions as uploaded to the device, uid = index. */
int orc_seed(orc_sim* sim, int64_t n, const float* position, const int32_t* mode, const double* occupation);
int orc_step(orc_sim* sim, int32_t nsteps);
int orc_cells(orc_sim* sim, double* temperature, double* energy_density);
int orc_heat_flux(orc_sim* sim, double* flux);
int orc_ledger(orc_sim* sim, pmc_ledger* ledger);
/* Thermal-spectrum deposit of the pending per-cell budgets (src/solver/PhononSourceOperator.cpp:120-341):
 * a pseudo-temperature per cell, or, with capacity (per active mode heat capacity at T_bg, linearized
 * transport), increments proportional to the mode heat capacity. */
int orc_source_deposit(const pmc_mode_table* m, int ncells, double* pending, const double* cell_T, double t_bg,
                       const int32_t* elig, int64_t n, const int32_t* active, const int32_t* cell, double* occ,
                       const double* capacity, double* energy, int64_t* updates);
/* Prescribed heat source, thermal spectrum (src/solver/PhononSourceOperator.cpp:120-341). */
/* orc_source_deposit with the source spectrum of pmc_heat_source: 0 thermal,
 * 1 weighted (mode_weight per active mode), 2 gaussian (center, sigma in THz). */
int orc_source_deposit_spectrum(const pmc_mode_table* m, int ncells, double* pending, const double* cell_T,
                                double t_bg, const int32_t* elig, int64_t n, const int32_t* active,
                                const int32_t* cell, double* occ, const double* capacity, int spectrum,
                                const double* mode_weight, double center_thz, double sigma_thz, double* energy_out,
                                int64_t* updates_out);
int orc_set_heat_source(orc_sim* sim, const pmc_heat_source* source);
/* Layered materials with DMM interfaces (src/solver/LayeredMaterialLayout.cpp,
 * src/solver/InterfaceInteraction.cpp); before orc_initialize. */
int orc_set_layers(orc_sim* sim, const pmc_layers* layers);
/* One DMM interaction of a carrier of global mode `active` (material by mode range)
 * with velocity vel and occupation occ toward material dest, on the uniforms u[3].
 * Returns the action (1 transmit, 0 reflect into a new mode, -1 mirror) and the
 * outgoing mode, occupation and velocity. */
/* InterfaceInteraction::interact for any interface model (DMM, AMM, MMM) on the
 * draws u in order (the channel draw first under MMM, then the AMM row draw or
 * the three DMM draws); returns 1 transmitted, 0 reflected into a new mode,
 * -1 mirrored, -2 no AMM row for the mode in this direction. */
int orc_interface(orc_sim* sim, int active, const double vel[3], double occ, int dest, const double u[4],
                  int* out_active, double* out_occ, double out_vel[3]);
int orc_interface_dmm(orc_sim* sim, int active, const double vel[3], double occ, int dest, const double u[3],
                      int* out_active, double* out_occ, double out_vel[3]);
/* Linearized Callaway transport (before initialization): linearized initial
 * occupations, linear cell temperature, conservative reflections and Callaway
 * collisions; the reservoir occupations and refill policy come from the setup
 * and reservoir tables. */
int orc_set_callaway(orc_sim* sim, const pmc_callaway_table* table);
/* Nonlinear Callaway transport (before initialization): Bose occupations,
 * lookup-table cell temperatures, conservative reflections and the nonlinear
 * N/R substeps. */
int orc_set_callaway_nonlinear(orc_sim* sim, const pmc_callaway_rates* rates);
/* nonlinear_callaway_evolve (src/solver/NonlinearCallawayCollision.cpp:251-285) on
 * one cell of n carriers (rates in 1/ps, wavevector [n*3]); returns 0, or the
 * reason it failed: 1 no representable energy constraint, 2 singular moment
 * Jacobian, 3 line search failed, 4 no convergence, 5 invalid occupation,
 * 6 invalid state or mode. */
int orc_nl_callaway_evolve(int n, const double* omega, const double* wavevector, const double* normal_rate,
                           const double* resistive_rate, const double* occupation, double T, double dt, double* out);
/* PhononMaterial::mode_callaway_rates at T for one active mode; 7 when T is
 * outside the table. */
int orc_nl_callaway_rates(const pmc_mode_table* m, const pmc_callaway_rates* r, double T, int active, double* normal,
                          double* resistive);
/* NonlinearCallawayCollision::apply over all cells; returns the residual (eV),
 * *status as orc_nl_callaway_evolve or 7. */
double orc_nl_callaway_apply(const pmc_mode_table* m, const pmc_callaway_rates* r, int ncells,
                             const double* cell_temperature, int lifetime_local, double lifetime_temperature,
                             int64_t n, const int32_t* active, const int32_t* cell, double* occ, double dt,
                             double energy_weight, int* status);
/* Live carriers sorted by uid. */
int64_t orc_carrier_count(orc_sim* sim);
int orc_carriers(orc_sim* sim, double* position, int32_t* mode, double* occupation, int32_t* cell, uint32_t* uid);
void orc_set_threads(int threads);
const char* orc_last_error(orc_sim* sim);
void orc_destroy(orc_sim* sim);

#ifdef __cplusplus
}
#endif
#endif
