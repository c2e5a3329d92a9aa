// Closed triangle surfaces: the STL/OBJ reader (src/geometry/MeshReader.cpp),
// the derived data of the reference's SurfaceMesh (orientation, face metrics,
// coplanar facet groups, enclosed volume; src/SurfaceMesh.cpp:232-760), its
// ray-cast containment (:1100-1124) and closest-hit trace (:1164-1312), plus the
// device BVH and a procedural FinFET generator.
#pragma once
#include "pmc_capi.h"

#include <array>
#include <cstdint>
#include <string>
#include <vector>

namespace pmc {

using Vec3 = std::array<double, 3>;
using Tri = std::array<int, 3>;

struct MeshData {
    std::vector<Vec3> vertices;
    std::vector<Tri> faces;
};

// .stl (binary or ASCII, vertices merged on a 1e-10 lattice) or .obj, units as stored.
MeshData read_mesh_file(const std::string& path);
void write_binary_stl(const std::string& path, const MeshData& mesh, const std::string& name);

struct MeshFacet {
    std::vector<int> faces;          // grouping order
    Vec3 normal{}, centroid{};
    double area = 0;
    std::vector<int> boundary_edges; // sorted edge ids used once inside the facet
};

struct SurfaceMesh {
    bool merge_coplanar = true;
    std::vector<Vec3> vertices;
    std::vector<Tri> faces;
    std::vector<std::array<int, 2>> edges;
    std::vector<std::vector<int>> edge_faces;
    std::vector<std::array<int, 3>> face_edges;
    std::vector<Vec3> face_normal, face_centroid;
    std::vector<double> face_area, face_k;
    std::vector<std::array<Vec3, 2>> face_bounds;
    std::vector<std::array<int, 2>> face_adjacency;
    std::vector<int> face_to_facet;
    std::vector<MeshFacet> facets;
    Vec3 bmin{}, bmax{};
    double volume = 0;

    void set(std::vector<Vec3> v, std::vector<Tri> f);
    void shift_to_origin();
    // Odd number of distinct crossings of the +x ray.
    bool contains(const Vec3& p) const;
    // Membership as the reference decides it without an external Delaunay tool:
    // star tetrahedra from the vertex mean (kept when their volumes add up to the
    // enclosed volume within 1e-3), inclusive to 1e-10 in barycentric
    // coordinates, then the ray cast (src/SurfaceMesh.cpp:771-816, 1126-1152).
    bool contains_point(const Vec3& p) const;
    // Sorted boundary-edge lengths of a facet (periodic matching).
    std::vector<double> boundary_edge_lengths(int facet) const;

    // Star tetrahedra (apex, then a face's three vertices) with their bounds,
    // inverse edge matrices and volumes; empty when they do not add up to the
    // enclosed volume (src/SurfaceMesh.cpp:773-817).
    struct Simplex {
        std::array<Vec3, 4> v;
        std::array<Vec3, 2> bounds;
        std::array<std::array<double, 3>, 3> inv;
        double volume = 0;
    };
    const std::vector<Simplex>& simplices() const;

private:
    mutable std::vector<Simplex> simplices_;
    mutable bool simplices_ready_ = false;
    void build_star() const;
    void rebuild();
    void bounding_box();
    void edge_list();
    void orient();
    void face_metrics();
    void neighbors();
    void facet_groups();
    void enclosed_volume();
};

bool ray_triangle(const Vec3& o, const Vec3& d, const Vec3& a, const Vec3& b, const Vec3& c, double& t);
Vec3 closest_point_on_triangle(const Vec3& p, const Vec3& a, const Vec3& b, const Vec3& c);

// Bounding-volume hierarchy (pmc_bvh_node layout): binned surface-area
// heuristic over item centroids, nodes in depth-first order with the left child
// next, boxes rounded outward to FP32.
struct Bvh {
    std::vector<pmc_bvh_node> nodes;
    std::vector<int32_t> order;
    int depth = 0;
};
Bvh build_bvh(const std::vector<std::array<Vec3, 2>>& boxes, const std::vector<Vec3>& centroids, int leaf);

// Lower bounds of the distance to the surface on a voxel grid over [lo, hi]
// (pmc_mesh field layout): at most max_voxels voxels, bounds quantized down to
// multiples of unit and saturated at 255 units.
struct DistanceField {
    std::array<int, 3> dims{1, 1, 1};
    double h = 0, unit = 0;
    std::vector<uint8_t> bound;
    // Voxels with a positive bound do not touch the surface, so each connected
    // region of them lies on one side: 1 inside, 0 outside (by contains_point of
    // one centre per region).
    std::vector<uint8_t> inside;
};
DistanceField build_distance_field(const SurfaceMesh& mesh, const Vec3& lo, const Vec3& hi, long max_voxels,
                                   double unit);

// Procedural FinFET on a lattice: a substrate box with one fin on its top face,
// every exposed lattice face split into two triangles (watertight, manifold).
struct FinfetSpec {
    double substrate[3] = {200, 100, 50}; // nm
    double fin[3] = {10, 40, 100};        // nm, centred on the substrate top
    double step[3] = {5, 10, 5};          // lattice spacing, nm (about 4k triangles)
};
MeshData finfet_mesh(const FinfetSpec& spec);

} // namespace pmc
