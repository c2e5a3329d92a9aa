#include "domain.hpp"

#include <algorithm>
#include <cmath>
#include <filesystem>
#include <fstream>
#include <limits>
#include <sstream>
#include <stdexcept>

namespace pmc {
namespace {
Vec3 sub(const Vec3& a, const Vec3& b) { return {a[0] - b[0], a[1] - b[1], a[2] - b[2]}; }
Vec3 add(const Vec3& a, const Vec3& b) { return {a[0] + b[0], a[1] + b[1], a[2] + b[2]}; }
Vec3 mul(const Vec3& a, double s) { return {a[0] * s, a[1] * s, a[2] * s}; }
double dot(const Vec3& a, const Vec3& b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
Vec3 cross(const Vec3& a, const Vec3& b) {
    return {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
}
double norm(const Vec3& a) { return std::sqrt(dot(a, a)); }

// Axis-aligned box as 12 triangles grouped into 6 facets in plane order
// xmin, xmax, ymin, ymax, zmin, zmax -- the vertex list, triangle list and
// coplanar grouping order of src/SimulationDomain.cpp:140-168 and
// src/SurfaceMesh.cpp:581-722, with inward-wound triangles flipped the way
// orient_closed_surface leaves a closed convex shell (swap of vertices 1 and 2).
void build_box_mesh(Domain& d, const Vec3& size) {
    d.vertices = {{0, 0, 0}, {0, 0, 1}, {0, 1, 1}, {0, 1, 0}, {1, 0, 0}, {1, 0, 1}, {1, 1, 1}, {1, 1, 0}};
    for (auto& v : d.vertices)
        for (int a = 0; a < 3; ++a) v[a] *= size[a];
    d.faces = {{0, 1, 2}, {0, 2, 3}, {4, 5, 6}, {4, 6, 7}, {0, 4, 5}, {0, 5, 1},
               {3, 7, 6}, {3, 6, 2}, {0, 4, 7}, {0, 7, 3}, {1, 5, 6}, {1, 6, 2}};
    const Vec3 center = mul(size, 0.5);
    d.face_area.resize(d.faces.size());
    d.face_normal.resize(d.faces.size());
    std::vector<Vec3> face_centroid(d.faces.size());
    for (size_t i = 0; i < d.faces.size(); ++i) {
        auto& f = d.faces[i];
        Vec3 a = d.vertices[f[0]], b = d.vertices[f[1]], c = d.vertices[f[2]];
        Vec3 n = cross(sub(b, a), sub(c, a));
        const Vec3 cen = mul(add(add(a, b), c), 1.0 / 3.0);
        if (dot(n, sub(cen, center)) < 0) {
            std::swap(f[1], f[2]);
            b = d.vertices[f[1]];
            c = d.vertices[f[2]];
            n = cross(sub(b, a), sub(c, a));
        }
        const double area2 = norm(n);
        d.face_normal[i] = area2 > 0 ? mul(n, 1.0 / area2) : Vec3{0, 0, 0};
        d.face_area[i] = area2 * 0.5;
        face_centroid[i] = {(a[0] + b[0] + c[0]) / 3.0, (a[1] + b[1] + c[1]) / 3.0, (a[2] + b[2] + c[2]) / 3.0};
    }
    d.facets.clear();
    for (int k = 0; k < 6; ++k) {
        Facet f;
        f.faces = {2 * k, 2 * k + 1};
        Vec3 nsum{0, 0, 0}, csum{0, 0, 0};
        double asum = 0;
        for (int fi : f.faces) {
            nsum = add(nsum, d.face_normal[fi]);
            csum = add(csum, mul(face_centroid[fi], d.face_area[fi]));
            asum += d.face_area[fi];
        }
        const double nn = norm(nsum);
        f.normal = nn > 0 ? mul(nsum, 1.0 / nn) : Vec3{0, 0, 0};
        f.area = asum;
        f.centroid = asum > 0 ? mul(csum, 1.0 / asum) : csum;
        d.facets.push_back(f);
    }
    d.bmin = {0, 0, 0};
    d.bmax = size;
    long double v6 = 0;
    for (const auto& f : d.faces)
        v6 += static_cast<long double>(dot(d.vertices[f[0]], cross(d.vertices[f[1]], d.vertices[f[2]])));
    d.volume = std::abs(static_cast<double>(v6 / 6.0L));
    for (int k = 0; k < 6; ++k) d.box_plane_facet[k] = k;
}
} // namespace

std::vector<int> Domain::facets_in_region(const std::array<double, 6>& box) const {
    // src/SimulationDomain.cpp:452-472: relative box on facet centroids.
    const Vec3 ext = sub(bmax, bmin);
    const double tol = 1e-12 * std::max({1.0, std::abs(ext[0]), std::abs(ext[1]), std::abs(ext[2])});
    Vec3 lo, hi;
    for (int a = 0; a < 3; ++a) {
        lo[a] = bmin[a] + std::min(box[a], box[a + 3]) * ext[a];
        hi[a] = bmin[a] + std::max(box[a], box[a + 3]) * ext[a];
    }
    std::vector<int> out;
    for (int f = 0; f < static_cast<int>(facets.size()); ++f) {
        const auto& c = facets[f].centroid;
        bool inside = true;
        for (int a = 0; a < 3; ++a) inside = inside && c[a] >= lo[a] - tol && c[a] <= hi[a] + tol;
        if (inside) out.push_back(f);
    }
    return out;
}

std::vector<int> Domain::reservoir_facets() const {
    std::vector<int> out;
    for (int f = 0; f < static_cast<int>(facets.size()); ++f)
        if (facets[f].condition == 'T') out.push_back(f);
    return out;
}

int Domain::fast_grid_index(const Vec3& p) const {
    // src/SimulationDomain.cpp:92-126.
    const Vec3 ext = sub(bmax, bmin);
    auto axis_index = [&](int axis, int n) {
        if (n <= 1 || ext[axis] <= 0.0) return 0;
        const double rel = (p[axis] - bmin[axis]) * inv_cell[axis];
        int idx = static_cast<int>(std::floor(rel));
        return std::clamp(idx, 0, n - 1);
    };
    const int cell = (axis_index(0, grid[0]) * grid[1] + axis_index(1, grid[1])) * grid[2] + axis_index(2, grid[2]);
    return voxel_to_cell[cell];
}

Domain build_domain(const SimConfig& c) {
    Domain d;
    if (c.model != "box")
        throw std::runtime_error("Mesh geometry '" + c.model + "' requires the mesh domain builder (not a box).");
    if (c.sizes.size() != 3 || std::any_of(c.sizes.begin(), c.sizes.end(), [](double s) { return !std::isfinite(s) || s <= 0; }))
        throw std::runtime_error("Box requires three finite positive sizes.");
    build_box_mesh(d, {c.sizes[0], c.sizes[1], c.sizes[2]});
    d.is_box = true;

    // Boundary conditions (src/SimulationDomain.cpp:474-515): unselected facets are R, eta = 0.
    if (c.boundary_conditions.size() != c.boundary_boxes.size() || c.boundary_values.size() != c.boundary_boxes.size())
        throw std::runtime_error("boundary_position, boundary_conditions, and boundary_values must have identical lengths.");
    for (auto& f : d.facets) {
        f.condition = 'R';
        f.value = 0.0;
    }
    for (size_t i = 0; i < c.boundary_boxes.size(); ++i) {
        const char cond = c.boundary_conditions[i];
        const double value = cond == 'R' ? c.boundary_values[i] * 10.0 : c.boundary_values[i];
        if (!std::isfinite(value) || value < 0 || (cond == 'P' && value != 0))
            throw std::runtime_error("Boundary values must be finite, nonnegative, and zero for periodic boundaries.");
        const auto matched = d.facets_in_region(c.boundary_boxes[i]);
        if (matched.empty())
            throw std::runtime_error("boundary_position region " + std::to_string(i) + " did not match any mesh facet.");
        for (int f : matched) {
            d.facets[f].condition = cond;
            d.facets[f].value = value;
        }
    }
    // Periodic pairs (src/SimulationDomain.cpp:518-623); one facet per region on a box.
    if (c.periodic_pairs.size() % 2 != 0) throw std::runtime_error("periodic_pair must have even number of regions.");
    for (size_t i = 0; i < c.periodic_pairs.size(); i += 2) {
        const auto ma = d.facets_in_region(c.periodic_pairs[i]);
        const auto mb = d.facets_in_region(c.periodic_pairs[i + 1]);
        if (ma.empty() || mb.empty()) throw std::runtime_error("Connected facet mapping failed for periodic_pair regions.");
        if (ma.size() != mb.size()) throw std::runtime_error("periodic_pair region pair must match the same number of facets.");
        if (ma.size() != 1) throw std::runtime_error("Multi-facet periodic regions require the mesh domain builder.");
        const int a = ma[0], b = mb[0];
        if (d.facets[a].condition != 'P' || d.facets[b].condition != 'P')
            throw std::runtime_error("Connected facets must both be periodic ('P').");
        if ((d.facets[a].periodic_pair >= 0 && d.facets[a].periodic_pair != b) ||
            (d.facets[b].periodic_pair >= 0 && d.facets[b].periodic_pair != a))
            throw std::runtime_error("Facet already paired with a different periodic connection.");
        if (dot(d.facets[a].normal, d.facets[b].normal) > -0.98)
            throw std::runtime_error("Connected facets normals do not match opposite direction.");
        const double aa = d.facets[a].area, ab = d.facets[b].area;
        if (std::abs(aa - ab) / std::max({1.0, aa, ab}) > 1e-3) throw std::runtime_error("Connected facets area mismatch.");
        d.facets[a].periodic_pair = b;
        d.facets[b].periodic_pair = a;
        d.facets[a].shift = sub(d.facets[b].centroid, d.facets[a].centroid);
        d.facets[b].shift = sub(d.facets[a].centroid, d.facets[b].centroid);
    }
    for (int f = 0; f < static_cast<int>(d.facets.size()); ++f)
        if (d.facets[f].condition == 'P' && d.facets[f].periodic_pair < 0)
            throw std::runtime_error("Periodic boundary facet " + std::to_string(f) + " is not assigned to a periodic pair.");

    // Grid (src/SimulationDomain.cpp:626-766); every box voxel is a cell.
    const int nx = c.grid[0], ny = c.grid[1], nz = c.grid[2];
    if (nx <= 0 || ny <= 0 || nz <= 0) throw std::runtime_error("Grid sizes must be positive.");
    d.grid = {nx, ny, nz};
    const Vec3 ext = sub(d.bmax, d.bmin);
    for (int a = 0; a < 3; ++a) d.inv_cell[a] = ext[a] > 0.0 ? static_cast<double>(d.grid[a]) / ext[a] : 0.0;
    d.voxel_to_cell.assign(static_cast<size_t>(nx) * ny * nz, -1);
    for (int ix = 0; ix < nx; ++ix)
        for (int iy = 0; iy < ny; ++iy)
            for (int iz = 0; iz < nz; ++iz) {
                const Vec3 p{d.bmin[0] + (static_cast<double>(ix) + 0.5) * ext[0] / static_cast<double>(nx),
                             d.bmin[1] + (static_cast<double>(iy) + 0.5) * ext[1] / static_cast<double>(ny),
                             d.bmin[2] + (static_cast<double>(iz) + 0.5) * ext[2] / static_cast<double>(nz)};
                d.voxel_to_cell[(ix * ny + iy) * nz + iz] = static_cast<int>(d.cell_center.size());
                d.cell_center.push_back(p);
            }
    d.cell_volume.assign(d.cell_center.size(), d.volume / static_cast<double>(d.cell_center.size()));
    return d;
}

void write_domain_summary(const Domain& d, const SimConfig& c, const std::string& folder) {
    namespace fs = std::filesystem;
    fs::create_directories(folder);
    std::ofstream out(fs::path(folder) / "summary.txt");
    auto join = [](const std::vector<double>& v) {
        std::ostringstream o;
        o << "[";
        for (size_t i = 0; i < v.size(); ++i) o << (i ? ", " : "") << v[i];
        o << "]";
        return o.str();
    };
    out << "[input]\n";
    out << "model = " << c.model << '\n';
    out << "sizes_nm = [" << c.sizes[0] / 10.0 << ", " << c.sizes[1] / 10.0 << ", " << c.sizes[2] / 10.0 << "]\n";
    out << "sizes_angstrom = [" << c.sizes[0] << ", " << c.sizes[1] << ", " << c.sizes[2] << "]\n";
    out << "particle_count = " << c.particle_count << '\n';
    out << "time_step = " << c.time_step << '\n';
    out << "iterations = " << c.iterations << '\n';
    out << "convergence_write_interval = " << c.convergence_write_interval << '\n';
    out << "random_seed = " << c.random_seed << '\n';
    out << "temperature_lookup_dt = " << c.temperature_lookup_dt << '\n';
    out << "background_temperature = " << c.background_temperature << '\n';
    out << "callaway_formulation = " << (c.callaway_nonlinear ? "nonlinear" : "linearized") << '\n';
    out << "conservative_boundaries = " << (c.conservative_boundaries ? "true" : "false") << '\n';
    out << "lifetime_temperature = ";
    if (c.lifetime_local) out << "local";
    else out << c.lifetime_temperature;
    out << '\n';
    if (c.transport_axis >= 0) out << "transport_axis = " << "xyz"[c.transport_axis] << '\n';
    out << "compute_kappa = " << (c.compute_kappa ? "true" : "false") << '\n';
    out << "profile_timers = " << (c.profile_timers ? "true" : "false") << '\n';
    out << "progress_temperature_summary_only = " << (c.progress_temperature_summary_only ? "true" : "false") << '\n';
    out << "merge_coplanar_facets = " << (c.merge_coplanar_facets ? "true" : "false") << '\n';
    out << "initial_temperature = ";
    if (c.initial_linear) out << "linear";
    else out << c.initial_uniform_temperature;
    out << '\n';
    out << "grid_xyz = [" << c.grid[0] << ", " << c.grid[1] << ", " << c.grid[2] << "]\n";
    out << "boundary_conditions = ";
    for (size_t i = 0; i < c.boundary_conditions.size(); ++i) {
        const char b = c.boundary_conditions[i];
        out << (i ? " " : "") << (b == 'T' ? "T" : b == 'P' ? "P" : "R");
    }
    out << '\n';
    out << "boundary_values = " << join(c.boundary_values) << '\n';
    out << "boundary_position_selector_mode = " << (c.boundary_boxes.empty() ? "none" : "region") << '\n';
    out << "boundary_position_selector_count = " << c.boundary_boxes.size() << '\n';
    out << "periodic_pair_region_count = " << c.periodic_pairs.size() << '\n';
    out << "material_folder = " << c.material_folder << '\n';
    out << "output_folder = " << folder << '\n';
    out << "heat_source_enabled = " << (c.heat_source_enabled ? "true" : "false") << '\n';
    out << "heat_source_profile = " << (c.heat_profile == HeatProfile::Gaussian ? "gaussian" : "uniform") << '\n';
    out << "heat_source_power_density = " << c.heat_power_density << '\n';
    out << '\n';
    out << "[geometry]\n";
    out << "volume = " << d.volume << '\n';
    out << "bounds_min = [" << d.bmin[0] << ", " << d.bmin[1] << ", " << d.bmin[2] << "]\n";
    out << "bounds_max = [" << d.bmax[0] << ", " << d.bmax[1] << ", " << d.bmax[2] << "]\n";
    out << "n_faces = " << d.faces.size() << '\n';
    out << "n_facets = " << d.facets.size() << '\n';
    out << "n_simplices = 0\n\n";
    out << "[grid]\n";
    out << "count = " << d.cell_count() << '\n';
    out << "lookup = voxel_to_active_grid_o1\n\n";
    int reservoirs = 0, rough = 0, pairs = 0;
    for (int f = 0; f < static_cast<int>(d.facets.size()); ++f) {
        reservoirs += d.facets[f].condition == 'T';
        rough += d.facets[f].condition == 'R';
        pairs += d.facets[f].periodic_pair > f;
    }
    out << "[boundary]\n";
    out << "reservoir_facets = " << reservoirs << '\n';
    out << "rough_facets = " << rough << '\n';
    out << "periodic_pairs = " << pairs << '\n';
    out << '\n';
    out << "[files]\n";
    out << "grid_centers_csv = " << (fs::path(folder) / "grid_centers.csv").string() << '\n';
    std::ofstream centers(fs::path(folder) / "grid_centers.csv");
    centers << "index,x_nm,y_nm,z_nm,volume_nm3\n";
    for (size_t i = 0; i < d.cell_center.size(); ++i) {
        const auto& p = d.cell_center[i];
        centers << i << "," << (p[0] / 10.0) << "," << (p[1] / 10.0) << "," << (p[2] / 10.0) << ","
                << (d.cell_volume[i] / 1000.0) << "\n";
    }
}

} // namespace pmc
