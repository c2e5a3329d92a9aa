// Device-side data layout of one Monte Carlo context (internal header shared by
// the HIP translation units).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pmc_dev {

constexpr double kHbar = 6.582119569e-4;      // eV*ps
constexpr double kBoltz = 8.617333262145e-5;  // eV/K
constexpr double kPi = 3.14159265358979323846;
constexpr double kFluxUnit = 1.602176634e13;  // eV/(ps*A^2) -> W/m^2

// Carrier mode word: bits 0..27 active-mode index, bits 28..30 velocity sign
// flips (x,y,z) from mirror reflections, -1 marks an absorbed (dead) slot.
constexpr int kModeMask = (1 << 28) - 1;
constexpr int kFlipShift = 28;

// Philox counter word 3 tags one random stream per use.
enum Stream : uint32_t { kStreamTransport = 1, kStreamInject = 2, kStreamInit = 3, kStreamFlux = 4 };

constexpr int kMaxReservoirs = 16;
constexpr int kMaxFacetsDev = 64;

// Carrier passes: each wave walks kRowsPerWave rows of 64 consecutive slots,
// loading kUnroll rows ahead of their use.
constexpr int kRowsPerWave = 16;
constexpr int kSortRows = 8;   // rows per wave of the cell sort (values held in registers)
constexpr int kUnroll = 4;
constexpr int kWavesPerBlock = 4;
constexpr int kSlotsPerBlock = 64 * kRowsPerWave * kWavesPerBlock;
constexpr int kInjectBlocks = 64;
constexpr int kMaxLayers = 8;     // materials of a layered domain
// Flight-kernel variants: axis-aligned box, triangle mesh, box with material
// interfaces.  Each compiles only its own boundary code.
constexpr int kGeomBox = 0, kGeomMesh = 1, kGeomLayered = 2;

// Collision kernels work on chunks of up to kChunk consecutive carriers of one
// cell in the cell-sorted buffer (one block each) and write per-chunk partial
// sums; per-cell results add the chunks of a cell in chunk order.
constexpr int kChunk = 2048;
constexpr int kChunkThreads = 256;
constexpr int kChunkWaves = kChunkThreads / 64;

// RTA pseudo-temperature sums per cell:
//   D      = sum hw a (n - n_BE(T0))                 exact value at the cell temperature T0
//   M_k    = sum hw a (-x0)^k P_k(n_BE(T0)) / k!    Taylor terms of sum hw a n_BE(T) in u = T0/T - 1
//   S0     = sum hw a n                              f(0)
//   ESCALE = sum hw a max(1, n)
//   UHIGH  = sum hw a (n - max(0, 1/x_high - 1/2))   upper bound of f(high) (n_BE(x) > 1/x - 1/2)
//   BAD    = carriers whose x0 is too small for the series
constexpr int kSeriesOrder = 6;
enum AccCol : int {
    kAccD = 0,
    kAccM1 = 1, // .. kAccM1 + kSeriesOrder - 1
    kAccS0 = 1 + kSeriesOrder,
    kAccEscale,
    kAccUHigh,
    kAccBad,
    kAccCols
};

// Per-block ledger partials (one row per block, summed by the refresh kernel).
enum Partial : int {
    kPAbsorbedE = 0,
    kPInjectedE = 1,
    kPAbsorbedN = 2,
    kPInjectedN = 3,
    kPRoughEvents = 4,
    kPRoughSpecular = 5,
    kPRoughDiffuse = 6,
    kPIfaceEvents = 7,
    kPIfaceTransmitted = 8,
    kPAbsorbedRes = 9, // + reservoir index
    kPartialCols = 9 + kMaxReservoirs
};
// Heat-source sums per chunk: Taylor terms of sum hw n_BE(T) about the base
// temperature (no constant term: f(T_b) = 0 by construction), eligible count,
// carriers the series cannot take; then the increment pass's applied energy and
// largest-energy carrier (energy, uid, slot).
constexpr int kSrcM1 = 0;
constexpr int kSrcCount = kSeriesOrder;
constexpr int kSrcBad = kSeriesOrder + 1;
constexpr int kSrcCols = kSeriesOrder + 2;
constexpr int kSrcApplied = 0, kSrcMaxE = 1, kSrcMaxUid = 2, kSrcMaxSlot = 3, kSrcApplyCols = 4;
constexpr int kSrcCellCols = 4; // base T, source T, target (raw eV), deposit (1) / keep pending (0)

// Cell tally columns: deviational energy, lifetime residual, live carriers.
constexpr int kTallyStride = 4;
enum TallyCol : int { kTallyEnergy = 0, kTallyCount = 1, kTallyResidual = 2 };

struct FacetDev {
    double n[3];
    double shift[3];
    int cond, rough, res, pair;
    int axis; // axis of the normal of an axis-aligned facet, else -1 (mirror reflections need one)
};

// Threaded bounding-volume hierarchy node: boxes padded by the mesh tolerance at
// upload, nodes in depth-first order.  An interior node (count 0) has its left
// child next and `index` = the first node after its subtree, where a traversal
// that misses the box continues; a leaf holds items index .. index+count-1.
struct BvhDev {
    float bmin[3], bmax[3];
    int index, count;
};

struct RoughDev {
    const double* spec;
    const int* match;
    const int* dbeg;
    const int* dend;
    int n_out;
    const int* out_active;
    const int* out_sorted;
    const double* prefix;
    int n_roul;
    const int* roul_active;
    const double* roul_cdf;
    int strict;
    int mode_lo, mode_n; // active modes of the table's material (last-resort draw)
};

struct ResDev {
    int facet;
    double temperature;
    double n[3];
    int initial;
    double expected;
    const double* cdf;
    const double* aprob;
    const int* aidx;
    const double* occ;
    int nfaces;
    const double* face_cdf;
    const double* tri;
};

// Scalars that live on the device across kernels of one step.
struct DevState {
    int n_slots;          // slots in use (live + dead)
    int n_live;
    int n_slots_start;    // slots before injection this step
    int step;             // completed steps
    double elapsed;       // simulated time (ps), summed dt by dt like the reference's elapsed_time_
    unsigned next_uid;
    int emit_total;
    int emit_count[kMaxReservoirs];     // emissions this step per reservoir
    int absorbed_now[kMaxReservoirs];   // absorptions this step per reservoir
    long long step_absorbed, step_injected;
    double step_absorbed_e, step_injected_e;
    double step_residual_e, step_balance_e;
    long long total_absorbed, total_injected;
    double total_absorbed_e, total_injected_e, total_residual_e, total_balance_e;
    double thermal_e;
    long long carrier_steps;
    long long rough_events, rough_specular, rough_diffuse;
    long long iface_events, iface_transmitted;
    double step_source_e, total_source_e;   // heat-source energy deposited (physical eV)
    long long source_updates;               // occupation updates by the heat source
    long long step_source_updates;          // of which this step
    double source_time;                     // integrated source time factor (ps): prescribed = total rate * this
    long long source_passes;                // exact carrier sums of the source pseudo-temperature solve
    long long newton_passes;  // exact carrier sums of the pseudo-temperature solve
    long long newton_evals;   // all evaluations (series or exact)
    int cw_big;               // length of the large-cell Callaway list this step
    int error_code;
    long long error_slot;
    float error_pos[3];
    int error_facet;
};

// Per-cell RTA constants of one step: temperature of the refresh (series
// centre), initial bracket top, rate-table rows/weight and the pseudo-temperature.
struct CellRta {
    double t0, high, gt, tr, inv_t0;
    int glo, ghi;
};

struct View {
    // carriers (structure of arrays)
    float* px;
    float* py;
    float* pz;
    int* mode;
    double* occ;
    int* cell;
    unsigned* uid;
    int capacity;
    // mode tables
    int na, nflat;
    const int* active_flat;
    const double* omega;
    const float4* vel;      // x,y,z,|v| FP32
    const double* veld;     // FP64 velocity, 3 per mode
    const double* neq;
    const double2* mtab;    // (omega, n_eq(T_bg)) per active mode
    int ntemps;
    const double* temps;
    const double2* gamma2;  // [ntemps][na]: (gamma(T_r), gamma(T_r+1)) of each active mode
    double w, e_bg, zp;
    double e_norm;          // max(1, Nq*Vuc): crystal energy density = sum hw n / e_norm + zp
    int lut_n;
    const double* lut_t;
    const double* lut_e;
    // geometry
    double bmin[3], bmax[3];
    float fbmin[3], fbmax[3];
    int plane_facet[6];
    int nfacets;
    const FacetDev* facets;
    int grid[3];
    double inv_cell[3];
    const int* voxel;
    int ncells;
    const double* cell_vol;
    // triangle mesh (geom_kind 1); faces and star tetrahedra in BVH leaf order
    int geom_kind;
    double mesh_tol;          // FP32 positions within this distance of a wall or face are on it (A)
    int nfaces;
    const double4* face_plane; // outward normal, k: n.x + k = 0
    const double* face_tri;    // 9 per face
    const double* face_box;    // 6 per face: min, max
    const int* face_facet;
    const BvhDev* face_nodes;
    int n_face_nodes;
    const unsigned char* field; // distance lower bounds (pmc_mesh field), or null
    const unsigned char* field_inside; // side of each voxel with a positive bound
    int field_dims[3];
    float field_inv_h, field_unit;
    int event_split;            // k_events blocks per transport block
    int nsimplices;
    const double* simplex_v;   // 12 per tetrahedron: apex, face vertices
    const double* simplex_inv; // 9 per tetrahedron
    const double* simplex_box; // 6 per tetrahedron
    const double* simplex_cdf; // cumulative volume (A^3), original order
    const int* simplex_pos;    // original index -> BVH order
    const BvhDev* simplex_nodes;
    int n_simplex_nodes;
    // layered materials (layers > 1): planes along layer_axis, material of each
    // cell and of each global mode range, DMM banks by (material, direction)
    int layers;
    int layer_axis;
    float layer_plane[kMaxLayers - 1];
    int mode_begin[kMaxLayers + 1];
    const int* cell_mat;
    const int* layer_rough; // [layers * nfacets] reflection table by (material, facet); null: facets' own
    double e_bg_mat[kMaxLayers];
    double cbg_mat[kMaxLayers]; // linearized transport: heat-capacity density of each material
    int lut_begin[kMaxLayers + 1];
    double layer_accept[kMaxLayers]; // initial positions: smallest spatial volume / this material's
    double band;                     // DMM frequency band, rad/ps
    int bank_begin[2 * kMaxLayers + 1];
    const int* bank_mode;
    const double* bank_omega;
    const double* bank_prefix;
    // Per bank, band_count + 1 positions: lower_bound of j * band in the bank's
    // sorted frequencies, j = 0 .. band_count (every frequency is below
    // band_count * band), so a band's modes are [start[j], start[j + 1]).
    const int* band_start;
    int band_count;
    // Acoustic (AMM) / mixed (MMM) mismatch: probability of the AMM channel per
    // event (0 DMM, 1 AMM), and the AMM rows by (interface, side) block:
    // row amm_block[2f+s] + local source mode, its mirror partner and its
    // transmitted edges [amm_ebeg[r], amm_ebeg[r + 1]).
    double amm_fraction;
    int amm_block[2 * (kMaxLayers - 1) + 1];
    const int* amm_refl;
    const int* amm_ebeg;
    const int* amm_emode;
    const double* amm_eprob;
    // boundary tables
    const RoughDev* rough;
    const ResDev* res;
    int nres;
    // cells
    double* cell_tally;   // kTallyStride per cell
    double* partials;     // block rows of pcols values
    int pcols;            // columns written per row: kPAbsorbedRes + nres (<= kPartialCols)
    int partial_rows;     // allocated: transport blocks x kMaxEventSplit + injection blocks
    double* cell_e;       // energy density of the latest refresh
    double* cell_T;
    double* cell_tx;      // Newton temperature of the post-transport energy (inversion 1)
    double* cell_flux;    // 3 per cell
    int* cell_count;      // live carriers per cell after transport and injection
    int* cell_start;      // [ncells+1] first slot of each cell in the sorted buffer
    int* cell_cursor;     // scatter cursors
    CellRta* crta;
    int* chunk_start;     // [ncells+1] first chunk of each cell
    int* chunk_cell;      // [max_chunks] cell of each chunk
    int max_chunks;
    double* chunk_acc;    // kAccCols RTA sums per chunk
    double* chunk_out;    // 2 per chunk: energy after relaxation, lifetime residual
    double* aw;           // relaxation weight of each sorted carrier (moments -> relax); source increments
    // prescribed heat source (thermal spectrum)
    int hs_on;
    double t_bg;
    const double* hs_rate;       // eV/ps per cell
    const int* hs_elig;          // per active mode
    double* hs_pending;          // eV per cell
    double* hs_acc;              // kSrcCols per chunk
    double* hs_cell;             // kSrcCellCols per cell: T_b, T_s, target, deposit flag
    double* chunk_src;           // energy change of the second half-step deposit per chunk (tally units)
    // spectrum: 0 thermal, 1 weighted (hs_w = branch weight), 2 gaussian (hs_w = log weight)
    int hs_spectrum;
    const double* hs_w;
    // time profile of the source (pmc_heat_source): constant or square wave
    int hs_square;
    double hs_start, hs_end, hs_period, hs_on_duration, hs_duty, hs_amplitude;
    int* events;          // per transport block: slots with a wall event this step
    int* block_events;    // per transport block: number of such slots
    DevState* st;
    // destination buffer of the cell sort
    float* o_px;
    float* o_py;
    float* o_pz;
    int* o_mode;
    double* o_occ;
    int* o_cell;
    unsigned* o_uid;
    // controls
    double dt;
    float dtf;
    double push;
    uint32_t key0, key1;
    int lifetime_local;
    double lifetime_T;
    int sampling;
    int inversion;
    int series_on;        // evaluate pseudo-temperature sums from the per-cell series
    double series_radius; // largest |T0/T - 1| at which the series is tried
    int refill_fixed;     // reservoirs emit floor(expected) + Bernoulli(frac) carriers per step
    // Transport policy (include/SimulationConfig.h:148-156): discrete-mode
    // reflection through the strict-elastic tables on every reflecting facet
    // (conservative), occupations and cell temperatures linearized about T_bg
    // (linearized).  Linearized Callaway sets both; nonlinear Callaway and
    // conservative RTA set conservative only.
    int conservative;
    int linearized;
    // linearized Callaway collisions (fixed reference T_bg)
    int callaway;
    const double* cw_cap;   // mode heat capacity at T_bg, eV/K
    const double* cw_q;     // Cartesian wavevector, 3 per mode, 1/A
    const double* cw_gn;    // normal rate 1/ps
    const double* cw_gr;    // resistive rate 1/ps
    double* cw_dir;         // 4 per slot: explicit normal-gain directions of cells that need them
    double* cw_cell;        // kCwRec per cell: projector, rates, moments of the Callaway setup
    const double4* cw_pack; // 2 per mode: (g_N, g_R, B0, B1), (B2, B3, hw, 0); B = sqrt(C) (1, q / w)
    double cw_cbg;          // heat-capacity density at T_bg, eV/(K A^3)
    int cw_threads;         // k_callaway workgroup size: 256, 384 or kCwThreads
    int* cw_big;            // cells of more than kCwWaveMax carriers this step (count in DevState)
    // initial mode excitation (pmc_init.excite): occupation factor in a frequency band
    int ex_on;
    double ex_fmin, ex_fmax, ex_mult; // THz (< 0: no bound)
    // nonlinear Callaway collisions (rates at the cell or lifetime temperature)
    int nl_callaway;
    int nl_nt;              // rate temperature samples
    const double* nl_temps;
    const double* nl_gn;    // [nl_nt * na] gamma_N
    const double* nl_gu;    // [nl_nt * na] gamma_U
    const double* nl_ge;    // [nl_nt * na] isotope/impurity/defect
    const double* nl_q;     // Cartesian wavevector, 3 per mode, 1/A
    double* nl_scratch;     // 7 per slot (structure of arrays): a, metric, n0, 4 directions
};

// Transport blocks for a slot capacity (a multiple of 8, one per XCD round).
__host__ __device__ inline int wave_blocks_for(int capacity) {
    const int b = (capacity + kSlotsPerBlock - 1) / kSlotsPerBlock;
    return (b + 7) & ~7;
}

// Event blocks: event_split per transport block (1 to kMaxEventSplit).
constexpr int kMaxEventSplit = 8;
__host__ __device__ inline int event_rows_for(const View& v) { return wave_blocks_for(v.capacity) * v.event_split; }

// Linearized Callaway kernels.  A cell of at most kCwWaveMax carriers is one
// wavefront's work (kCwTinyPer or kCwWavePer carriers per lane in registers,
// reductions by cross-lane shuffles); a larger cell goes on a per-step list to
// one workgroup of 256, 384 or kCwThreads threads, up to kCwPer carriers per
// thread.
constexpr int kCwThreads = 1024, kCwWaves = kCwThreads / 64, kCwPer = 8;
constexpr int kCwTinyPer = 2, kCwWavePer = 8;
constexpr int kCwTinyMax = 64 * kCwTinyPer, kCwWaveMax = 64 * kCwWavePer;
constexpr int kCwBigBlocks = 2048; // workgroups that walk a context's list of large cells
constexpr int kCwRec = 32; // doubles of the per-cell Callaway record


} // namespace pmc_dev
