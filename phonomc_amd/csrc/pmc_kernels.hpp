// Launchers for the step kernels (defined in pmc_kernels.hip).
#pragma once
#include "pmc_capi.h"
#include "pmc_device.hpp"

namespace pmc_launch {
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
void moments(const pmc_dev::View& v, hipStream_t s);
void newton(const pmc_dev::View& v, hipStream_t s);
void relax(const pmc_dev::View& v, hipStream_t s);
void callaway(const pmc_dev::View& v, hipStream_t s);
void source(const pmc_dev::View& v, hipStream_t s, double factor, int half);
void clear_cells(const pmc_dev::View& v, hipStream_t s);
void flux(const pmc_dev::View& v, hipStream_t s);
void containment(const pmc_dev::View& v, hipStream_t s);
void init(const pmc_dev::View& v, hipStream_t s, int count, int tmode, double t0, const double* cold,
          const double* hot, double tcold, double thot, const double* field);
void count_all(const pmc_dev::View& v, hipStream_t s);
} // namespace pmc_launch
