// Host build of the device's nonlinear Callaway cell routine
// (phonomc_amd/csrc/pmc_nlcallaway.hpp) for CPU checks against the oracle and
// the reference.  The executor emulates one wavefront: 64 lanes stride over the
// cell's slots, and the lane sums combine in the order of the device's DPP
// reduction (xor 1, xor 2, half-row mirror, row mirror, then rows 0, 16, 32, 48).
#include "pmc_nlcallaway.hpp"

#include <array>
#include <cstdint>
#include <vector>

namespace {
struct HostWave {
    template <int K, class F>
    void map_reduce(int begin, int end, double (&acc)[K], F&& f) {
        std::vector<std::array<double, K>> lane(64);
        for (auto& l : lane) l.fill(0.0);
        for (int l = 0; l < 64; ++l) {
            double part[K];
            for (int k = 0; k < K; ++k) part[k] = 0.0;
            for (int i = begin + l; i < end; i += 64) f(i, part);
            for (int k = 0; k < K; ++k) lane[l][k] = part[k];
        }
        auto stage = [&](auto partner) {
            const auto old = lane;
            for (int l = 0; l < 64; ++l)
                for (int k = 0; k < K; ++k) lane[l][k] = old[l][k] + old[partner(l)][k];
        };
        stage([](int l) { return l ^ 1; });
        stage([](int l) { return l ^ 2; });
        stage([](int l) { return (l & ~7) | (7 - (l & 7)); });
        stage([](int l) { return (l & ~15) | (15 - (l & 15)); });
        for (int k = 0; k < K; ++k) acc[k] = ((lane[0][k] + lane[16][k]) + lane[32][k]) + lane[48][k];
    }
};
} // namespace

extern "C" {

// One cell of n carriers in slots 0..n-1: mode[i] active indices, occ updated in
// place; omega_neq holds (omega, n_BE(T_bg)) per active mode; out2 = energy
// after, lifetime residual (eV).  Returns the pmc_nl::Status.
int nlh_cell(int n, const int* mode, double* occ, const double* omega_neq, const double* q, int n_active, int n_temps,
             const double* temps, const double* gn, const double* gu, const double* ge, double T, double rate_T,
             double dt, double energy_weight, double* out2) {
    std::vector<double> scratch(7 * static_cast<size_t>(n > 0 ? n : 1));
    pmc_nl::Data D;
    D.mode = mode;
    D.mode_mask = (1 << 28) - 1;
    D.occ = occ;
    D.omega = omega_neq;
    D.neq_bg = omega_neq + 1;
    D.omega_stride = 2;
    D.q = q;
    D.n_active = n_active;
    D.n_temps = n_temps;
    D.temps = temps;
    D.gn = gn;
    D.gu = gu;
    D.ge = ge;
    D.a = scratch.data();
    D.metric = scratch.data() + n;
    D.n0 = scratch.data() + 2 * static_cast<size_t>(n);
    D.b = scratch.data() + 3 * static_cast<size_t>(n);
    D.bstride = n;
    D.energy_weight = energy_weight;
    HostWave x;
    double out[2];
    const int rc = pmc_nl::cell(x, D, 0, n, T, rate_T, dt, out);
    out2[0] = out[0];
    out2[1] = out[1];
    return rc;
}

// Rates of active modes at T (rate_bracket, rate_of); returns the status.
int nlh_rates(int n, const int* active, int n_active, int n_temps, const double* temps, const double* gn,
              const double* gu, const double* ge, double T, double* normal, double* resistive) {
    pmc_nl::Data D{};
    D.n_active = n_active;
    D.n_temps = n_temps;
    D.temps = temps;
    pmc_nl::RateAt r;
    const int rc = pmc_nl::rate_bracket(D, T, r);
    if (rc) return rc;
    for (int i = 0; i < n; ++i) {
        normal[i] = pmc_nl::rate_of(D, gn, r, active[i]);
        resistive[i] = pmc_nl::rate_of(D, gu, r, active[i]) + pmc_nl::rate_of(D, ge, r, active[i]);
    }
    return 0;
}
}
