/* pmc_capi.h -- C ABI of the MI355X deviational Monte Carlo step.
 *
 * One context owns every device buffer of one simulation instance on the
 * current HIP device: structure-of-arrays carriers, mode/rate tables, geometry,
 * boundary tables, cell tallies and ledgers.  The host hands over immutable
 * tables once; pmc_step() then advances whole time steps on the device.
 *
 * Each entry point replaces a reference (lyushisyan/PhonoMC) interface, cited
 * as file:line.  All functions return PMC_OK (0) or an error code; after a
 * failed step the context is latched and every later call returns
 * PMC_ERR_LATCHED, like the reference's "Cannot resume a failed timestep"
 * guard (src/MonteCarloSolver.cpp:349-354).  Not reentrant: one host thread
 * per context. */
#ifndef PMC_CAPI_H
#define PMC_CAPI_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum pmc_status {
    PMC_OK = 0,
    PMC_ERR_ARGUMENT = 1,       /* invalid argument or table shape */
    PMC_ERR_DEVICE = 2,         /* HIP runtime failure */
    PMC_ERR_EXCESS_EVENTS = 3,  /* a carrier needed more than 64 events in one step */
    PMC_ERR_BOUNDARY = 4,       /* boundary intersection / outgoing mode failed */
    PMC_ERR_ESCAPED = 5,        /* carrier found outside the domain */
    PMC_ERR_INVALID_STATE = 6,  /* non-finite occupation or energy */
    PMC_ERR_CAPACITY = 7,       /* carrier slots exhausted */
    PMC_ERR_LATCHED = 8,        /* an earlier step failed; create a new context */
    PMC_ERR_UNSUPPORTED = 9     /* feature outside this build */
};

typedef struct pmc_ctx pmc_ctx;

/* Step controls; replaces the constructor arguments MonteCarloSolver reads from
 * SimulationConfig (src/MonteCarloSolver.cpp:50-207, include/SimulationConfig.h:42-146). */
typedef struct pmc_setup {
    int64_t capacity;              /* carrier slots to allocate (>= initial count) */
    double time_step_ps;
    double push_distance_a;        /* 1e-10*max(dt,1) in the reference (src/MonteCarloSolver.cpp:74) */
    uint64_t seed;                 /* Philox key (with instance) */
    uint32_t instance;
    double background_temperature; /* deviational reference T_bg */
    int32_t lifetime_local;        /* 1: rates at cell temperature; 0: at lifetime_temperature */
    double lifetime_temperature;
    int32_t refill_policy;         /* 0 one-to-one (previous absorbed count), 1 fixed incoming flux */
    int32_t mode_sampling;         /* 0 alias table, 1 inverse CDF (reference index order) */
    int32_t temperature_inversion; /* 0 lookup table (reference), 1 per-cell Newton on the exact
                                    * Bose-Einstein crystal energy (single-material RTA transport;
                                    * pmc_upload_callaway, pmc_upload_layers and batches reject it) */
    int32_t conservative_walls;    /* 1: conservative transport with Bose occupations
                                    * (simulation.conservative_boundaries = true, include/SimulationConfig.h:153-156):
                                    * every reflecting facet reflects through its strict-elastic table with
                                    * signed occupations (src/solver/BoundaryInteraction.cpp:90-113).
                                    * pmc_upload_callaway and pmc_upload_callaway_nonlinear imply it. */
    int32_t reserved[7];
} pmc_setup;

/* Material and background tables by active-mode index; replaces PhononMaterial
 * queries (include/PhononMaterial.h:14-125) and BackgroundCache
 * (include/solver/BackgroundCache.h:11-55). */
typedef struct pmc_mode_table {
    int32_t n_active;
    int32_t n_flat;                /* qpoints*branches */
    const int32_t* active_flat;    /* [n_active] flat mode index */
    const double* omega;           /* [n_active] rad/ps */
    const double* velocity;        /* [n_active*3] Angstrom/ps */
    const double* neq_background;  /* [n_active] Bose occupation at T_bg */
    int32_t n_temps;
    const double* temps;           /* [n_temps] K */
    const double* gamma_total;     /* [n_temps*n_flat] THz linewidth (RTA total) */
    double energy_weight;          /* M_active*V_particle/max(1,Nq*Vuc) */
    double background_energy_density;
    double zero_point_density;
    int32_t lut_n;
    const double* lut_t;           /* [lut_n] K */
    const double* lut_e;           /* [lut_n] eV/A^3, strictly increasing */
    /* max(1, Nq*V_uc) (A^3): the crystal energy density is sum over active modes
     * of hw*n_BE(T) / energy_normalization + zero_point_density
     * (PhononMaterial::crystal_energy_density, src/material/Thermodynamics.cpp:151-163).
     * Read by the Newton temperature inversion only. */
    double energy_normalization;
} pmc_mode_table;

/* One surface facet; replaces SimulationDomain facet queries
 * (include/SimulationDomain.h:58-66). */
typedef struct pmc_facet {
    double normal[3];              /* outward unit normal */
    double shift[3];               /* periodic translation */
    int32_t condition;             /* 'T', 'P' or 'R' */
    int32_t rough_table;           /* index or -1 */
    int32_t reservoir;             /* index or -1 */
    int32_t periodic_pair;         /* facet or -1 */
} pmc_facet;

/* Node of a bounding-volume hierarchy over triangles or tetrahedra.  Boxes are
 * FP32, rounded outward.  An interior node (count == 0) has its left child next
 * in the array and its right child at `index`; a leaf holds `count` items
 * order[index .. index+count). */
typedef struct pmc_bvh_node {
    float bmin[3], bmax[3];
    int32_t index;
    int32_t count;
} pmc_bvh_node;

/* Closed triangle surface of a mesh model (geometry kind 1); replaces
 * SurfaceMesh::trace_boundary_intersection (src/SurfaceMesh.cpp:1164-1270),
 * contains_point (:1126-1151, with the star tetrahedra of :783-817) and
 * sample_volume_points (src/geometry/SurfaceSampling.cpp:113-155). */
typedef struct pmc_mesh {
    int32_t n_vertices;
    const double* vertices;        /* [n_vertices*3] Angstrom */
    int32_t n_faces;
    const int32_t* faces;          /* [n_faces*3] outward winding */
    const double* face_normal;     /* [n_faces*3] outward unit normal */
    const double* face_k;          /* [n_faces] plane n.x + k = 0 */
    const double* face_bounds;     /* [n_faces*6] min x,y,z then max x,y,z */
    const int32_t* face_facet;     /* [n_faces] facet of each face */
    int32_t n_face_nodes;
    const pmc_bvh_node* face_nodes;
    const int32_t* face_order;     /* [n_faces] */
    int32_t n_simplices;           /* 0: membership by ray casting alone */
    const double* simplex_vertices;/* [n_simplices*12] shared apex, then the face's three vertices */
    const double* simplex_inverse; /* [n_simplices*9] row-major inverse of [a-o | b-o | c-o] */
    const double* simplex_bounds;  /* [n_simplices*6] */
    const double* simplex_volume;  /* [n_simplices] */
    int32_t n_simplex_nodes;
    const pmc_bvh_node* simplex_nodes;
    const int32_t* simplex_order;  /* [n_simplices] */
    /* Distance lower bounds on a voxel grid from bmin with edge field_h:
     * field[(i*dims[1] + j)*dims[2] + k] * field_unit <= the distance from any
     * point of voxel (i,j,k) to every face (255: at least 255 * field_unit).
     * Optional (field == NULL): a carrier whose flight is shorter than the
     * bound skips the face search. */
    int32_t field_dims[3];
    double field_h, field_unit;
    const uint8_t* field;
    /* Per voxel with a positive bound: 1 inside the surface, 0 outside
     * (contains_point of one voxel centre per connected region). */
    const uint8_t* field_inside;
} pmc_mesh;

/* Geometry and cell grid; replaces SimulationDomain::trace_boundary_intersection
 * (src/SimulationDomain.cpp:342-450) and fast_grid_index (:92-126). */
typedef struct pmc_geometry {
    int32_t kind;                  /* 0: analytic axis-aligned box, 1: triangle mesh */
    double bmin[3], bmax[3];
    int32_t n_facets;
    const pmc_facet* facets;
    int32_t box_plane_facet[6];    /* xmin,xmax,ymin,ymax,zmin,zmax -> facet */
    int32_t grid[3];
    double inv_cell[3];
    const int32_t* voxel_to_cell;  /* [grid product] */
    int32_t n_cells;
    const double* cell_volume;     /* [n_cells] A^3 */
    const pmc_mesh* mesh;          /* kind 1 only */
} pmc_geometry;

/* Rough-facet reflection table; replaces RoughBoundarySampler
 * (src/solver/RoughBoundarySampler.cpp:77-533). */
typedef struct pmc_rough_table {
    const double* specularity;     /* [n_active] */
    const int32_t* spec_match;     /* [n_active] */
    const int32_t* diffuse_begin;  /* [n_active] */
    const int32_t* diffuse_end;    /* [n_active] */
    int32_t n_outgoing;
    const int32_t* outgoing_active;/* [n_outgoing] active order */
    const int32_t* outgoing_sorted;/* [n_outgoing] frequency order */
    const double* residual_prefix; /* [n_outgoing+1] */
    int32_t n_roulette;
    const int32_t* roulette_active;
    const double* roulette_cdf;
    int32_t strict_elastic;
    /* The material's active-mode range for layered materials (the arrays are by
     * global active index; a row outside the range has no specular partner or
     * window); mode_count 0: every active mode.  The last-resort draw over any
     * mode (RoughBoundarySampler::sample_diffuse, :451-455) stays in this range. */
    int32_t mode_begin;
    int32_t mode_count;
} pmc_rough_table;

/* Thermal reservoir on one facet; replaces ReservoirSampler/ReservoirInjector
 * (src/solver/ReservoirSampler.cpp:17-146, src/solver/ReservoirInjector.cpp:62-132). */
typedef struct pmc_reservoir {
    int32_t facet;
    double temperature;
    double normal[3];
    double expected_count;
    int32_t initial_count;
    const double* cdf;             /* [n_active] inward-flux CDF */
    const double* alias_prob;      /* [n_active] */
    const int32_t* alias_index;    /* [n_active] */
    const double* occupation;      /* [n_active] Bose occupation at the reservoir T */
    int32_t n_faces;
    const double* face_cdf;        /* [n_faces] */
    const double* triangles;       /* [n_faces*9] */
} pmc_reservoir;

/* Layered materials along one axis with diffuse-mismatch (DMM) interfaces;
 * replaces LayeredMaterialLayout (src/solver/LayeredMaterialLayout.cpp:11-115),
 * InterfaceModeSampler (src/solver/InterfaceModeSampler.cpp:11-110) and the DMM
 * branch of InterfaceInteraction::interact (src/solver/InterfaceInteraction.cpp:47-153).
 * The materials' active modes form one global bank (pmc_mode_table indexes it);
 * every material shares the energy weight and carriers are denser where the
 * phase-space density is (src/MonteCarloSolver.cpp:81-100). */
typedef struct pmc_layers {
    int32_t count;                 /* materials, in layer order along the axis */
    int32_t axis;                  /* 0 x, 1 y, 2 z */
    const double* planes;          /* [count-1] interface positions (A), increasing */
    const int32_t* mode_begin;     /* [count+1] global active-mode range of each material */
    const int32_t* cell_material;  /* [n_cells] */
    const double* background_energy_density; /* [count] eV/A^3 at T_bg */
    const int32_t* lut_begin;      /* [count+1] rows of each material in pmc_mode_table.lut_t / lut_e */
    const double* spatial_volume;  /* [count] A^3 of space per carrier */
    double band_width;             /* DMM frequency band, rad/ps */
    const int32_t* bank_begin;     /* [2*count+1] banks by (material, direction: 0 toward -axis, 1 toward +axis) */
    const int32_t* bank_mode;      /* global active modes, each bank sorted by frequency */
    const double* bank_omega;      /* [bank_begin[2*count]] */
    const double* bank_prefix;     /* bank b: n_b+1 cumulative |v_axis| / normalization from bank_begin[b] + b */
    /* Acoustic (AMM) and mixed (MMM) mismatch (InterfaceInteraction::interact,
     * src/solver/InterfaceInteraction.cpp:46-153; AcousticInterfaceSampler,
     * src/solver/AcousticInterfaceSampler.cpp:27-472).  amm_fraction 0: DMM only
     * (the AMM arrays may be NULL); 1: AMM only; in between: an event takes AMM
     * with this probability.  AMM rows by interface f and side s (0: material f
     * toward +axis, 1: material f+1 toward -axis): block 2f+s starts at
     * amm_block[2f+s] and holds one entry per active mode of the source material
     * (local index) in amm_reflected and one more in amm_edge_begin. */
    double amm_fraction;
    const int32_t* amm_block;      /* [2*(count-1)+1] */
    const int32_t* amm_reflected;  /* global active mode of the mirror partner; -1: no incoming row */
    const int32_t* amm_edge_begin; /* row r's transmitted edges are [amm_edge_begin[r], amm_edge_begin[r+1]) */
    const int32_t* amm_edge_mode;  /* global active mode transmitted into */
    const double* amm_edge_prob;   /* its probability */
    /* Reflection tables by material (RoughBoundarySampler::table_index,
     * src/solver/RoughBoundarySampler.cpp:68-74, used at
     * src/solver/TransportOperator.cpp:302): facet f reflects a carrier of
     * material k through table facet_rough[k * n_facets + f] of
     * pmc_upload_boundary_tables (-1: no table).  Those tables hold global active
     * indices with mode_begin / mode_count set to the material's range.  NULL:
     * every material uses pmc_facet.rough_table. */
    const int32_t* facet_rough;
    /* Linearized transport (pmc_upload_callaway before this call): each material's
     * heat-capacity density C_bg, eV/(K A^3); cell temperatures are
     * T_bg + (e - E_bg,k) / C_bg,k with background_energy_density holding each
     * material's exact crystal energy density (BackgroundCache,
     * include/solver/BackgroundCache.h:13-30).  NULL otherwise. */
    const double* heat_capacity_density;
} pmc_layers;

/* Linearized Callaway scattering about the fixed reference T_bg; replaces
 * CarrierCollisionOperator::apply (src/solver/CarrierCollisionOperator.cpp:38-142)
 * with CallawayOperator (src/material/CallawayOperator.cpp:62-201) on the
 * CallawayData of PhononMaterial::callaway_data
 * (src/material/PhononMaterialCollisionData.cpp:32-165).  Uploading it switches
 * the context to the reference's linearized, conservative transport
 * (include/SimulationConfig.h:148-156): linearized initial occupations
 * n_eq(T_bg) + C(T_bg)(T - T_bg)/hw (src/material/Thermodynamics.cpp:27-37),
 * cell temperature T_bg + (e - E_bg)/C_bg (src/solver/GridStatistics.cpp:193-225),
 * discrete-mode reflection with signed occupations on every reflecting facet
 * (src/solver/BoundaryInteraction.cpp:90-113).  The caller supplies linearized
 * reservoir occupations in pmc_reservoir.occupation, strict-elastic rough tables
 * and pmc_setup.refill_policy = 1 (fixed incoming flux).  All arrays by active
 * index, evaluated at T_bg.  Call after pmc_upload_material and
 * pmc_upload_geometry and before the population is created. */
typedef struct pmc_callaway_table {
    const double* capacity;        /* [n_active] mode heat capacity, eV/K */
    const double* wavevector;      /* [n_active*3] Cartesian q, 1/Angstrom */
    const double* normal_rate;     /* [n_active] 4 pi gamma_N, 1/ps */
    const double* resistive_rate;  /* [n_active] 4 pi (gamma_U + isotope/impurity/defect), 1/ps */
    double heat_capacity_density;  /* C_bg, eV/(K Angstrom^3) (BackgroundCache.h:24-27) */
    double background_energy_density; /* E_bg: exact crystal energy density at T_bg */
} pmc_callaway_table;

/* Nonlinear Callaway scattering (scattering.callaway_formulation = 'nonlinear');
 * replaces NonlinearCallawayCollision::apply
 * (src/solver/NonlinearCallawayCollision.cpp:287-344) and nonlinear_callaway_evolve
 * (:251-285): per cell, N(dt/2) R(dt) N(dt/2) exponential BGK substeps toward the
 * displaced Bose target whose a-weighted energy (and, for N, crystal momentum)
 * moments match the carriers' (channel, :25-248), rates from
 * PhononMaterial::mode_callaway_rates (src/material/PhononMaterialCollisionData.cpp:167-196)
 * at the cell temperature (lifetime_local) or pmc_setup.lifetime_temperature.
 * Uploading it switches the context to the reference's nonlinear conservative
 * transport (include/SimulationConfig.h:148-156): Bose occupations, lookup-table
 * cell temperatures, discrete-mode reflection with the strict-elastic tables on
 * every reflecting facet.  The caller supplies Bose reservoir occupations,
 * strict-elastic rough tables at T_bg and pmc_setup.refill_policy = 1.  Tables
 * by (temperature sample, active index); raw linewidths, the 4 pi factor is the
 * device's.  Call after pmc_upload_material and pmc_upload_geometry and before
 * the population is created. */
typedef struct pmc_callaway_rates {
    int32_t n_temps;
    const double* temps;           /* [n_temps] K, increasing */
    const double* gamma_normal;    /* [n_temps*n_active] gamma_N */
    const double* gamma_umklapp;   /* [n_temps*n_active] gamma_U */
    const double* gamma_extra;     /* [n_temps*n_active] isotope/impurity/defect linewidths */
    const double* wavevector;      /* [n_active*3] Cartesian q, 1/Angstrom */
} pmc_callaway_rates;

/* Per-step and cumulative ledgers; replaces BoundaryLedger/EnergyLedger
 * (src/solver/BoundaryLedger.cpp:30-141, src/solver/EnergyLedger.cpp:58-90). */
typedef struct pmc_ledger {
    int64_t step;
    int64_t live_carriers;
    int64_t slots;
    int64_t step_absorbed, step_injected;
    double step_absorbed_energy_ev, step_injected_energy_ev;
    double step_lifetime_residual_ev, step_balance_residual_ev;
    int64_t total_absorbed, total_injected;
    double total_absorbed_energy_ev, total_injected_energy_ev;
    double total_lifetime_residual_ev, total_balance_residual_ev;
    double thermal_energy_ev;
    int64_t carrier_steps;         /* sum of live carriers at step start */
    int64_t rough_events, rough_specular, rough_diffuse;
    int64_t newton_passes;         /* exact carrier sums of the pseudo-temperature solve */
    int64_t newton_evals;          /* all evaluations of the pseudo-temperature function */
    double step_source_energy_ev;  /* heat-source energy deposited this step (EnergyLedger::record_source) */
    double total_source_energy_ev;
    int64_t source_updates;        /* occupation updates by the heat source */
    int64_t source_exact_passes;   /* exact carrier sums of the source pseudo-temperature solve */
    int64_t interface_events;      /* material-interface crossings handled (layered materials) */
    int64_t interface_transmitted; /* of which transmitted */
    /* Heat-source ledger (PrescribedHeatSource::accrue / deposit_cell,
     * src/solver/PrescribedHeatSource.cpp:232-275): updates this step, the
     * integrated time factor (prescribed energy of a cell = its rate times
     * this) and the energy still pending in cells without eligible carriers. */
    int64_t step_source_updates;
    double source_time_ps;
    double source_pending_ev;
} pmc_ledger;

typedef struct pmc_error {
    int32_t code;
    int64_t carrier;               /* first offending slot or -1 */
    float position[3];
    int32_t facet;
    char message[256];
} pmc_error;

/* Prescribed lattice heat source with the thermal spectrum; replaces
 * PrescribedHeatSource (src/solver/PrescribedHeatSource.cpp:22-252), the time
 * profile of MonteCarloSolver::local_heat_source_integrated_time_factor
 * (src/solver/HeatSourceDriving.cpp:14-46) and PhononSourceOperator::apply
 * (src/solver/PhononSourceOperator.cpp:120-341).  Applied as two half-step
 * deposits around transport and collisions (src/MonteCarloSolver.cpp:366-377). */
typedef struct pmc_heat_source {
    int32_t n_cells;
    const double* rate;            /* [n_cells] eV/ps */
    const int32_t* mode_eligible;  /* [n_active] 1: the spectrum deposits into the mode */
    int32_t square;                /* 0 constant, 1 square wave */
    double start_ps, end_ps;       /* end < 0: no end */
    double period_ps, on_duration_ps, duty_cycle, amplitude;
    /* Spectrum (PhononSourceOperator, src/solver/PhononSourceOperator.cpp:20-118,
     * :186-220): 0 thermal (pseudo-temperature, or heat capacities under
     * linearized transport); 1 weighted: a cell's eligible carriers share its
     * energy in proportion to mode_weight (the branch weight over the largest);
     * 2 gaussian: in proportion to exp(mode_weight) with
     * mode_weight = -((f - f_c) / sigma)^2 / 2, f = omega / 2 pi in THz, normalized
     * per cell about the cell's largest weight.  mode_weight may be NULL for 0. */
    int32_t spectrum;
    const double* mode_weight;     /* [n_active] */
    double frequency_center_thz, frequency_sigma_thz; /* gaussian */
} pmc_heat_source;

/* Initial population (src/solver/Initialization.cpp:32-275). */
typedef struct pmc_init {
    int64_t count;
    int32_t temperature_mode;      /* 0 uniform, 1 linear cold->hot, 2 per-cell field */
    double uniform_temperature;
    double cold_point[3], hot_point[3];
    double cold_temperature, hot_temperature;
    const double* cell_temperature;/* [n_cells] for mode 2 (warm start) */
    /* Initial targeted modal over-occupation (MonteCarloSolver::apply_initial_mode_excitation,
     * src/solver/Initialization.cpp:146-178): with excite = 1 the initial occupation
     * of every carrier whose frequency omega / 2 pi lies in [excite_fmin_thz,
     * excite_fmax_thz] (a negative bound: none) is multiplied by excite_multiplier. */
    int32_t excite;
    double excite_fmin_thz, excite_fmax_thz, excite_multiplier;
} pmc_init;

int pmc_create(const pmc_setup* setup, pmc_ctx** out);
int pmc_upload_material(pmc_ctx* ctx, const pmc_mode_table* table);
int pmc_upload_geometry(pmc_ctx* ctx, const pmc_geometry* geometry);
int pmc_upload_boundary_tables(pmc_ctx* ctx, int32_t n_rough, const pmc_rough_table* rough,
                               int32_t n_reservoirs, const pmc_reservoir* reservoirs);
/* Enables the heat source (after material and geometry). */
int pmc_upload_callaway(pmc_ctx* ctx, const pmc_callaway_table* table);
/* Nonlinear Callaway (pmc_callaway_rates); single contexts only, batches reject it. */
int pmc_upload_callaway_nonlinear(pmc_ctx* ctx, const pmc_callaway_rates* rates);
int pmc_upload_heat_source(pmc_ctx* ctx, const pmc_heat_source* source);
/* Samples the initial population on the device (Initialization.cpp:48-113) and
 * performs the initial temperature refresh (Observables.cpp:12-21). */
int pmc_initialize_carriers(pmc_ctx* ctx, const pmc_init* init);
/* Uploads an explicit population instead (positions FP32, active-mode index, occupation). */
int pmc_seed_carriers(pmc_ctx* ctx, int64_t n, const float* position, const int32_t* mode,
                      const double* occupation);
/* Runs nsteps calls of the time step (MonteCarloSolver::run_timestep,
 * src/MonteCarloSolver.cpp:348-398) without host synchronisation between steps. */
int pmc_step(pmc_ctx* ctx, int32_t nsteps);
/* Cell temperatures and energy densities of the latest refresh. */
int pmc_read_cells(pmc_ctx* ctx, double* temperature, double* energy_density);
/* Heat-flux tally (src/solver/GridStatistics.cpp:269-347), W/m^2, [n_cells*3]. */
int pmc_read_heat_flux(pmc_ctx* ctx, double* flux);
/* Heat-source energy still pending per cell (eV, [n_cells]): the budget of a
 * cell without eligible carriers (PrescribedHeatSource::cell_pending_ev,
 * src/solver/PrescribedHeatSource.cpp:232-262).  Zeros without a source. */
int pmc_read_source_pending(pmc_ctx* ctx, double* pending);
int pmc_read_ledger(pmc_ctx* ctx, pmc_ledger* ledger);
/* Carrier state for tests and checkpoints: slots in [0, *n) (dead slots have mode -1).
 * On input *n is the capacity (slots) of the arrays given; with every array NULL
 * it is ignored.  On output *n is the number of slots in use; PMC_ERR_ARGUMENT
 * when that exceeds the capacity (nothing is copied). */
int pmc_read_carriers(pmc_ctx* ctx, int64_t* n, float* position, int32_t* mode, double* occupation,
                      int32_t* cell);
/* Philox stream ids of slots [0, n) (initial carriers 0..N-1, injected ones numbered on);
 * n may not exceed the slots in use. */
int pmc_read_uids(pmc_ctx* ctx, int64_t n, uint32_t* uid);
int pmc_last_error(pmc_ctx* ctx, pmc_error* error);

/* Evaluation strategy of the RTA pseudo-temperature solve
 * (src/solver/RtaCollisionOperator.cpp:91-165).  The reference sums
 * hw*a*(n - n_BE(T)) over a cell's carriers for every trial T.  With the
 * series on, one pass stores per cell the exact value at the cell temperature
 * T0 plus Taylor terms of sum hw*a*n_BE(T) in u = T0/T - 1; trial
 * temperatures within |u| <= series_radius whose truncation estimate stays
 * below 1e-3 of the convergence tolerance use the series, all others the exact
 * carrier sum.  series_on = 0 reproduces the reference's exact sums only. */
typedef struct pmc_tuning {
    int32_t series_on;             /* default 1 */
    double series_radius;          /* default 0.01 */
    int32_t reserved[6];
} pmc_tuning;
int pmc_set_tuning(pmc_ctx* ctx, const pmc_tuning* tuning);
int pmc_get_tuning(pmc_ctx* ctx, pmc_tuning* tuning);
/* Kernel classes of the per-kernel timers. */
enum pmc_kernel_class {
    PMC_KERNEL_BEGIN = 0,      /* step set-up (+ containment check every 100 steps) */
    PMC_KERNEL_TRANSPORT = 1,  /* free flight of carriers that reach no wall, cell counts */
    PMC_KERNEL_EVENTS = 2,     /* carriers with wall events (boundary conditions, rough walls) */
    PMC_KERNEL_INJECT = 3,     /* reservoir injection and remaining flight */
    PMC_KERNEL_REFRESH = 4,    /* cell offsets, cell temperatures, per-cell constants, ledgers */
    PMC_KERNEL_SORT = 5,       /* cell-order compaction of the live carriers + cell energy */
    PMC_KERNEL_MOMENTS = 6,    /* per-chunk pseudo-temperature sums */
    PMC_KERNEL_NEWTON = 7,     /* per-cell pseudo-temperature solve */
    PMC_KERNEL_RELAX = 8,      /* per-chunk relaxation */
    PMC_KERNEL_SOURCE = 9,     /* heat-source half-step deposits */
    PMC_KERNEL_OUTPUT = 10,    /* observables outside the step (heat flux) */
    PMC_KERNEL_STEP = 11,      /* whole steps */
    PMC_KERNEL_CALLAWAY = 12,  /* per-cell linearized Callaway collisions */
    PMC_KERNEL_CLASSES = 13
};
/* Accumulated kernel time (ms) and launch counts per class since the last reset,
 * measured with HIP events on the context's stream while timing is enabled. */
int pmc_kernel_times(pmc_ctx* ctx, double* total_ms, int64_t* launches, int32_t n);
int pmc_reset_kernel_times(pmc_ctx* ctx);
/* Enables per-kernel event timing for later pmc_step calls (off by default). */
int pmc_set_timing(pmc_ctx* ctx, int32_t enabled);
void pmc_destroy(pmc_ctx* ctx);
const char* pmc_status_string(int32_t code);

/* Launch parameters of the step that do not change its results: how many
 * event blocks share one transport block's list of carriers with wall events,
 * and the workgroup size of the per-cell Callaway kernel.  No counterpart in the
 * reference (OpenMP loops, src/solver/TransportOperator.cpp:113-132 and
 * src/solver/CarrierCollisionOperator.cpp:38-142). */
typedef struct pmc_launch_config {
    int32_t event_split;       /* 1, 2, 4 or 8 */
    int32_t callaway_threads;  /* 256, 384 or 1024 (a cell holds at most 8 carriers per thread) */
} pmc_launch_config;
int pmc_get_launch_config(pmc_ctx* ctx, pmc_launch_config* out);
int pmc_set_launch_config(pmc_ctx* ctx, const pmc_launch_config* config);
/* Autotuner: times `steps` steps per candidate of each parameter with HIP events
 * on the context's stream and keeps the fastest (Callaway workgroup sizes only
 * where the fullest cell fits with a 1.5x margin).  The steps are real steps of
 * the simulation; the parameters change no result beyond FP64 summation order.
 * report_ms[k] receives the mean kernel time of candidate k in the order
 * event_split 1, 2, 4, 8, then callaway_threads 256, 384, 1024 (0 when not tried). */
int pmc_autotune(pmc_ctx* ctx, int32_t steps, pmc_launch_config* chosen, double* report_ms, int32_t report_len);

/* Uploads layered materials (after pmc_upload_boundary_tables, before the
 * population is created).  RTA only; no rough facets and no heat source. */
int pmc_upload_layers(pmc_ctx* ctx, const pmc_layers* layers);

/* Batched stepping of independent instances.  No counterpart in the reference,
 * which runs one MonteCarloSolver per process (src/main.cpp:86-102); each step
 * of the batch is the reference's run_timestep (src/MonteCarloSolver.cpp:348-398)
 * applied to every instance.  A batch launches each kernel of the step once
 * over all instances (per-instance block ranges) and replays the step as a
 * captured hipGraph.  The contexts keep their own tables, carriers, ledgers
 * and errors and stay readable with pmc_read_*; results equal those of
 * stepping each context alone.  Every context must be initialized, at the
 * same step, with the same geometry kind and collision model, without a heat
 * source; tuning changes after pmc_batch_create are not seen by the batch.
 * While a batch exists, step its contexts only through it. */
typedef struct pmc_batch pmc_batch;
int pmc_batch_create(pmc_ctx* const* contexts, int32_t n, pmc_batch** out);
/* Advances every instance by nsteps.  On a failure the failing context is
 * latched with its message (pmc_last_error) and the batch returns its code;
 * pmc_batch_failed names the first failed instance. */
int pmc_batch_step(pmc_batch* batch, int32_t nsteps);
int pmc_batch_failed(const pmc_batch* batch);
/* 1 (default): replay the captured step graph; 0: launch kernel by kernel. */
int pmc_batch_set_graph(pmc_batch* batch, int32_t enabled);
void pmc_batch_destroy(pmc_batch* batch);

#ifdef __cplusplus
}
#endif
#endif
