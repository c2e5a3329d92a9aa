// Acoustic-mismatch (AMM) interface tables of a layered film: the scalar,
// dispersive acoustic mismatch of AcousticInterfaceSampler::configure
// (src/solver/AcousticInterfaceSampler.cpp:27-437).  Every q point stands for a
// finite reciprocal-space cell; linearized in (frequency, k_a, k_b) it becomes a
// parallelepiped, and the transmission channel between an incident acoustic
// mode on one side and one on the other is the overlap of their cells weighted
// by the acoustic transmission 4 Z1c1 Z2c2 / (Z1c1 + Z2c2)^2, with both
// marginals capped at the modes' own flux.  Optical modes and modes without a
// channel reflect into their mirror partner (k_axis -> -k_axis).
#pragma once
#include "material.hpp"

#include <array>
#include <vector>

namespace pmc {

struct AmmAudit {
    std::array<double, 2> incident_acoustic_flux{}, transmitted_flux{}, no_overlap_flux{};
    double geometric_overlap_flux = 0;
    double capacity_removed_flux = 0;
    double maximum_capacity_ratio = 1;
};

// Rows of one interface: side 0 holds the modes of the lower material moving
// toward +axis, side 1 those of the upper material moving toward -axis, by the
// source material's active index.  A row's transmitted edges name active modes
// of the other material (already the outgoing states) with probabilities; the
// remainder reflects into `reflected` (-1: the mode does not cross this way).
struct AmmInterface {
    std::array<std::vector<int>, 2> reflected;
    std::array<std::vector<int>, 2> edge_begin; // n + 1 offsets into the edges
    std::array<std::vector<int>, 2> edge_mode;
    std::array<std::vector<double>, 2> edge_prob;
    AmmAudit audit;
};

// One table per interface between consecutive materials.  axis 0..2, densities
// in kg/m^3 (one per material), frequency bin in THz and parallel-wavevector
// bin in 1/Angstrom (candidate search only).  Throws the reference's messages.
std::vector<AmmInterface> build_amm_tables(const std::vector<const ModeTables*>& modes,
                                           const std::vector<const RawMaterial*>& raws, int axis,
                                           const std::vector<double>& densities, double frequency_bin_thz,
                                           double parallel_bin_inv_a);

// Scalar acoustic transmission between impedances z at incidence cosines c
// (AcousticInterfaceSampler::transmission, :17-25).
double amm_transmission(double z1, double z2, double c1, double c2);

} // namespace pmc
