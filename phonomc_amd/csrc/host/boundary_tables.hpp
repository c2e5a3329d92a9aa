// Immutable boundary sampling tables built once on the host:
//  - rough/specular reflection tables per rough facet
//    (src/solver/RoughBoundarySampler.cpp:77-368, include/solver/BoundaryFluxSampling.h:16-89);
//  - thermal-reservoir inward-flux tables per reservoir facet
//    (src/solver/ReservoirSampler.cpp:17-101), with a Walker alias table beside the CDF.
#pragma once
#include "domain.hpp"
#include "material.hpp"

#include <vector>

namespace pmc {

struct RoughTable {
    int facet = -1;
    bool strict_elastic = false;
    int mode_begin = 0, mode_count = 0; // a layered material's global active range (0 count: all modes)
    std::vector<double> specularity;      // per active mode
    std::vector<int> spec_match;          // per active mode, -1 = none
    std::vector<int> diffuse_begin, diffuse_end;
    std::vector<int> outgoing_active;     // active order
    std::vector<int> outgoing_sorted;     // frequency order (std::sort, as the reference)
    std::vector<double> residual_prefix;  // outgoing_sorted.size() + 1
    std::vector<int> roulette_active;
    std::vector<double> roulette_cdf;
};

struct ReservoirTable {
    int facet = -1;
    double temperature = 300;
    Vec3 normal{};
    double area = 0;
    double expected = 0;
    int initial_count = 0;
    std::vector<double> cdf;         // normalized inward-flux CDF over active modes
    std::vector<double> alias_prob;  // Walker alias table of the same PMF
    std::vector<int> alias_index;
    std::vector<double> occupation;  // Bose occupation at the reservoir temperature
    std::vector<double> face_cdf;    // area CDF over the facet's triangles (last = 1)
    std::vector<double> triangles;   // 9 doubles per triangle, facet face order
};

// Rough tables for every facet flagged rough; equilibrium_temperature > 0 selects
// the strict-elastic (conservative) construction.  facet_to_table has one entry per facet.
std::vector<RoughTable> build_rough_tables(const Domain& d, const ModeTables& m, const std::vector<bool>& enabled,
                                           double equilibrium_temperature, std::vector<int>& facet_to_table);

std::vector<ReservoirTable> build_reservoir_tables(const Domain& d, const ModeTables& m, double particle_density,
                                                   double time_step, std::vector<int>& facet_to_reservoir);

// Walker/Vose alias table of a nonnegative weight vector (prob in [0,1], alias index).
void build_alias(const std::vector<double>& weights, std::vector<double>& prob, std::vector<int>& alias);

} // namespace pmc
