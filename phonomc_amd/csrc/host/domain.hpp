// Simulation domain: surface facets with their boundary conditions, periodic
// pairing and the cell grid (src/SimulationDomain.cpp:73-827,
// src/SurfaceMesh.cpp:232-760).
#pragma once
#include "config.hpp"
#include "mesh.hpp"

#include <array>
#include <string>
#include <vector>

namespace pmc {

using Vec3 = std::array<double, 3>;

struct Facet {
    Vec3 normal{}, centroid{};
    double area = 0;
    std::vector<int> faces; // indices into Domain::faces, in grouping order
    char condition = 'R';
    double value = 0;       // reservoir temperature (K) or roughness (Angstrom)
    int periodic_pair = -1;
    Vec3 shift{};
};

struct Domain {
    bool is_box = true;
    std::string surface_key;               // identity of a mesh model's surface (file, size, time, merge flag)
    SurfaceMesh mesh;                      // surface of a mesh model (empty for the box)
    Vec3 bmin{}, bmax{};
    double volume = 0;
    std::vector<Vec3> vertices;
    std::vector<std::array<int, 3>> faces; // outward winding
    std::vector<double> face_area;
    std::vector<Vec3> face_normal;
    std::vector<Facet> facets;
    std::array<int, 6> box_plane_facet{-1, -1, -1, -1, -1, -1}; // xmin,xmax,ymin,ymax,zmin,zmax

    std::array<int, 3> grid{1, 1, 1};
    Vec3 inv_cell{};
    std::vector<int> voxel_to_cell;
    std::vector<Vec3> cell_center;
    std::vector<double> cell_volume;

    int cell_count() const { return static_cast<int>(cell_center.size()); }
    int fast_grid_index(const Vec3& p) const;
    std::vector<int> reservoir_facets() const;
    std::vector<int> facets_in_region(const std::array<double, 6>& box) const;
};

Domain build_domain(const SimConfig& config);

// summary.txt geometry sections and grid_centers.csv (src/geometry/DomainOutput.cpp:12-241).
// n_simplices: star tetrahedra of a mesh model (0 for a box).
void write_domain_summary(const Domain& d, const SimConfig& c, const std::string& folder, int n_simplices = 0);

} // namespace pmc
