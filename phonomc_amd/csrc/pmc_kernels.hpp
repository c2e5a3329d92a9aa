// Launchers for the step kernels (defined in pmc_kernels.hip).
#pragma once
#include "pmc_capi.h"
#include "pmc_device.hpp"

namespace pmc_launch {
// Kernel classes of a batched step, in launch order.
enum BatchKind {
    kBatchBegin,
    kBatchContainment,
    kBatchTransport,
    kBatchEvents,
    kBatchInject,
    kBatchRefresh1,
    kBatchSort,
    kBatchCellPrep,
    kBatchMoments,
    kBatchNewton,
    kBatchRelax,
    kBatchCallawayWave,     // one wavefront per cell of at most kCwWaveMax carriers
    kBatchCallaway,         // the listed larger cells
    kBatchRefresh2,
    kBatchSourceSums0,      // heat-source half step before transport
    kBatchSourceSolve0,
    kBatchSourceIncrements,
    kBatchSourceApply0,
    kBatchSourceSums1,      // heat-source half step after collisions
    kBatchSourceSolve1,
    kBatchSourceApply1,
    kBatchKinds
};
int batch_blocks(const pmc_dev::View& v, int kind);
void batch_launch(int kind, bool mesh, const pmc_dev::View* views, const int* offs, int n, int total, hipStream_t s);
void batch_errors(const pmc_dev::View* views, int n, int* out, hipStream_t s);
int wave_blocks(int capacity);
int event_split(int capacity);
int partial_rows(const pmc_dev::View& v);
void begin(const pmc_dev::View& v, hipStream_t s);
void transport(const pmc_dev::View& v, hipStream_t s);
void inject(const pmc_dev::View& v, hipStream_t s);
void refresh(const pmc_dev::View& v, hipStream_t s, int phase);
void events(const pmc_dev::View& v, hipStream_t s);
void sort(const pmc_dev::View& v, hipStream_t s);
void cell_prep(const pmc_dev::View& v, hipStream_t s);
void invert(const pmc_dev::View& v, hipStream_t s, int source);
void moments(const pmc_dev::View& v, hipStream_t s);
void newton(const pmc_dev::View& v, hipStream_t s);
void relax(const pmc_dev::View& v, hipStream_t s);
void callaway(const pmc_dev::View& v, hipStream_t s);
void nlcallaway(const pmc_dev::View& v, hipStream_t s);
void source(const pmc_dev::View& v, hipStream_t s, int half);
void clear_cells(const pmc_dev::View& v, hipStream_t s);
void flux(const pmc_dev::View& v, hipStream_t s);
void containment(const pmc_dev::View& v, hipStream_t s);
void init(const pmc_dev::View& v, hipStream_t s, int count, int tmode, double t0, const double* cold,
          const double* hot, double tcold, double thot, const double* field);
void count_all(const pmc_dev::View& v, hipStream_t s);
} // namespace pmc_launch
