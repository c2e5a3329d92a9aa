#include "domain.hpp"

#include <algorithm>
#include <cmath>
#include <deque>
#include <filesystem>
#include <fstream>
#include <iomanip>
#include <limits>
#include <map>
#include <mutex>
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

namespace {
// Mesh model (src/SimulationDomain.cpp:129-200): file resolved against the
// working directory, then the input directory; STL coordinates are nm and
// become Angstrom; the surface is shifted so its bounding box starts at 0.
void build_mesh_domain(Domain& d, const SimConfig& c) {
    namespace fs = std::filesystem;
    fs::path path(c.model);
    if (!path.is_absolute()) {
        const fs::path p1 = fs::current_path() / path, p2 = fs::path(c.input_directory) / path;
        path = fs::exists(p1) ? p1 : p2;
    }
    path = path.lexically_normal();
    if (!fs::exists(path)) throw std::runtime_error("Model file does not exist: " + path.string());
    MeshData data = read_mesh_file(path.string());
    std::string ext = path.extension().string();
    for (auto& ch : ext) ch = static_cast<char>(std::tolower(static_cast<unsigned char>(ch)));
    if (ext == ".stl")
        for (auto& v : data.vertices)
            for (auto& x : v) x *= 10.0;
    d.mesh.merge_coplanar = c.merge_coplanar_facets;
    d.mesh.set(std::move(data.vertices), std::move(data.faces));
    d.mesh.shift_to_origin();
    d.is_box = false;
    d.vertices = d.mesh.vertices;
    d.faces = d.mesh.faces;
    d.face_area = d.mesh.face_area;
    d.face_normal = d.mesh.face_normal;
    d.bmin = d.mesh.bmin;
    d.bmax = d.mesh.bmax;
    d.volume = d.mesh.volume;
    d.facets.clear();
    for (const auto& mf : d.mesh.facets) {
        Facet f;
        f.normal = mf.normal;
        f.centroid = mf.centroid;
        f.area = mf.area;
        f.faces = mf.faces;
        d.facets.push_back(f);
    }
    d.box_plane_facet = {-1, -1, -1, -1, -1, -1};
}
} // namespace

namespace {
// Mesh surfaces and their voxel grids, reused across problems in one process:
// parameter sweeps (batches of one geometry with different sources, reservoir
// temperatures or seeds) would otherwise re-read the surface and re-test every
// voxel centre for each instance.  Keyed by the resolved file (path, size,
// modification time) and the coplanar-merge flag, the grid also by its sizes.
struct GridCells {
    Vec3 inv_cell{};
    std::vector<int> voxel_to_cell;
    std::vector<Vec3> cell_center;
};
std::mutex cache_lock;
std::map<std::string, Domain> surface_cache;
std::map<std::string, GridCells> grid_cache;
constexpr size_t kCacheEntries = 8;

std::string surface_key(const SimConfig& c) {
    namespace fs = std::filesystem;
    fs::path path(c.model);
    if (!path.is_absolute()) {
        const fs::path p1 = fs::current_path() / path, p2 = fs::path(c.input_directory) / path;
        path = fs::exists(p1) ? p1 : p2;
    }
    std::error_code ec;
    const auto size = fs::file_size(path, ec);
    const auto stamp = fs::last_write_time(path, ec).time_since_epoch().count();
    std::ostringstream k;
    k << fs::absolute(path.lexically_normal()).string() << '|' << size << '|' << stamp << '|'
      << c.merge_coplanar_facets;
    return k.str();
}
} // namespace

Domain build_domain(const SimConfig& c) {
    Domain d;
    std::string key;
    if (c.model != "box") {
        key = surface_key(c);
        bool hit = false;
        {
            std::lock_guard<std::mutex> hold(cache_lock);
            auto it = surface_cache.find(key);
            if (it != surface_cache.end()) {
                d = it->second;
                hit = true;
            }
        }
        if (!hit) {
            build_mesh_domain(d, c);
            d.surface_key = key;
            std::lock_guard<std::mutex> hold(cache_lock);
            if (surface_cache.size() >= kCacheEntries) surface_cache.erase(surface_cache.begin());
            surface_cache[key] = d;
        }
    } else {
        if (c.sizes.size() != 3 ||
            std::any_of(c.sizes.begin(), c.sizes.end(), [](double s) { return !std::isfinite(s) || s <= 0; }))
            throw std::runtime_error("Box requires three finite positive sizes.");
        build_box_mesh(d, {c.sizes[0], c.sizes[1], c.sizes[2]});
        d.is_box = true;
    }

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
    // Periodic pairs (src/SimulationDomain.cpp:518-623): one facet per region, or
    // nearest centroid after the mean shift of the two regions.
    if (c.periodic_pairs.size() % 2 != 0) throw std::runtime_error("periodic_pair must have even number of regions.");
    for (size_t i = 0; i < c.periodic_pairs.size(); i += 2) {
        const auto ma = d.facets_in_region(c.periodic_pairs[i]);
        const auto mb = d.facets_in_region(c.periodic_pairs[i + 1]);
        if (ma.empty() || mb.empty()) throw std::runtime_error("Connected facet mapping failed for periodic_pair regions.");
        if (ma.size() != mb.size()) throw std::runtime_error("periodic_pair region pair must match the same number of facets.");
        std::vector<std::pair<int, int>> pairs;
        if (ma.size() == 1) {
            pairs.push_back({ma[0], mb[0]});
        } else {
            Vec3 mean_a{0, 0, 0}, mean_b{0, 0, 0};
            for (int f : ma) mean_a = add(mean_a, d.facets[f].centroid);
            for (int f : mb) mean_b = add(mean_b, d.facets[f].centroid);
            mean_a = mul(mean_a, 1.0 / static_cast<double>(ma.size()));
            mean_b = mul(mean_b, 1.0 / static_cast<double>(mb.size()));
            const Vec3 guess = sub(mean_b, mean_a);
            std::vector<int> left = mb;
            for (int fa : ma) {
                const Vec3 target = add(d.facets[fa].centroid, guess);
                double best = std::numeric_limits<double>::infinity();
                int bj = -1;
                for (int j = 0; j < static_cast<int>(left.size()); ++j) {
                    const Vec3 diff = sub(d.facets[left[j]].centroid, target);
                    const double d2 = dot(diff, diff);
                    if (d2 < best) {
                        best = d2;
                        bj = j;
                    }
                }
                if (bj < 0) throw std::runtime_error("Failed to build one-to-one mapping for periodic_pair facets.");
                pairs.push_back({fa, left[bj]});
                left.erase(left.begin() + bj);
            }
        }
        for (const auto& [a, b] : pairs) {
        if (d.facets[a].condition != 'P' || d.facets[b].condition != 'P')
            throw std::runtime_error("Connected facets must both be periodic ('P').");
        if ((d.facets[a].periodic_pair >= 0 && d.facets[a].periodic_pair != b) ||
            (d.facets[b].periodic_pair >= 0 && d.facets[b].periodic_pair != a))
            throw std::runtime_error("Facet already paired with a different periodic connection.");
        if (dot(d.facets[a].normal, d.facets[b].normal) > -0.98)
            throw std::runtime_error("Connected facets normals do not match opposite direction.");
        const double aa = d.facets[a].area, ab = d.facets[b].area;
        if (std::abs(aa - ab) / std::max({1.0, aa, ab}) > 1e-3) throw std::runtime_error("Connected facets area mismatch.");
        if (!d.is_box) {
            const auto la = d.mesh.boundary_edge_lengths(a), lb = d.mesh.boundary_edge_lengths(b);
            if (la.size() != lb.size()) throw std::runtime_error("Connected facets boundary edge count mismatch.");
            for (size_t k = 0; k < la.size(); ++k)
                if (std::abs(la[k] - lb[k]) / std::max({1.0, la[k], lb[k]}) > 1e-3)
                    throw std::runtime_error("Connected facets boundary edge-length spectrum mismatch.");
        }
        d.facets[a].periodic_pair = b;
        d.facets[b].periodic_pair = a;
        d.facets[a].shift = sub(d.facets[b].centroid, d.facets[a].centroid);
        d.facets[b].shift = sub(d.facets[a].centroid, d.facets[b].centroid);
        }
    }
    for (int f = 0; f < static_cast<int>(d.facets.size()); ++f)
        if (d.facets[f].condition == 'P' && d.facets[f].periodic_pair < 0)
            throw std::runtime_error("Periodic boundary facet " + std::to_string(f) + " is not assigned to a periodic pair.");

    // Grid (src/SimulationDomain.cpp:626-766): voxels whose centre lies in the
    // solid become cells (every voxel of a box); the others take the cell of the
    // nearest such voxel by a multi-source six-neighbour breadth-first fill.
    const int nx = c.grid[0], ny = c.grid[1], nz = c.grid[2];
    if (nx <= 0 || ny <= 0 || nz <= 0) throw std::runtime_error("Grid sizes must be positive.");
    d.grid = {nx, ny, nz};
    const std::string grid_key = key.empty() ? std::string() : key + "|" + std::to_string(nx) + "x" +
                                                                 std::to_string(ny) + "x" + std::to_string(nz);
    if (!grid_key.empty()) {
        std::lock_guard<std::mutex> hold(cache_lock);
        auto it = grid_cache.find(grid_key);
        if (it != grid_cache.end()) {
            d.inv_cell = it->second.inv_cell;
            d.voxel_to_cell = it->second.voxel_to_cell;
            d.cell_center = it->second.cell_center;
            d.cell_volume.assign(d.cell_center.size(), d.volume / static_cast<double>(d.cell_center.size()));
            return d;
        }
    }
    const Vec3 ext = sub(d.bmax, d.bmin);
    for (int a = 0; a < 3; ++a) d.inv_cell[a] = ext[a] > 0.0 ? static_cast<double>(d.grid[a]) / ext[a] : 0.0;
    d.voxel_to_cell.assign(static_cast<size_t>(nx) * ny * nz, -1);
    for (int ix = 0; ix < nx; ++ix)
        for (int iy = 0; iy < ny; ++iy)
            for (int iz = 0; iz < nz; ++iz) {
                const Vec3 p{d.bmin[0] + (static_cast<double>(ix) + 0.5) * ext[0] / static_cast<double>(nx),
                             d.bmin[1] + (static_cast<double>(iy) + 0.5) * ext[1] / static_cast<double>(ny),
                             d.bmin[2] + (static_cast<double>(iz) + 0.5) * ext[2] / static_cast<double>(nz)};
                if (d.is_box || d.mesh.contains_point(p)) {
                    d.voxel_to_cell[(ix * ny + iy) * nz + iz] = static_cast<int>(d.cell_center.size());
                    d.cell_center.push_back(p);
                }
            }
    if (d.cell_center.empty()) throw std::runtime_error("No valid grid centers found inside mesh.");
    std::deque<int> queue;
    for (int v = 0; v < static_cast<int>(d.voxel_to_cell.size()); ++v)
        if (d.voxel_to_cell[v] >= 0) queue.push_back(v);
    const int yz = ny * nz;
    while (!queue.empty()) {
        const int v = queue.front();
        queue.pop_front();
        const int ix = v / yz, iy = (v % yz) / nz, iz = v % nz;
        const int src = d.voxel_to_cell[v];
        const int nbr[6][3] = {{ix - 1, iy, iz}, {ix + 1, iy, iz}, {ix, iy - 1, iz},
                               {ix, iy + 1, iz}, {ix, iy, iz - 1}, {ix, iy, iz + 1}};
        for (const auto& q : nbr) {
            if (q[0] < 0 || q[0] >= nx || q[1] < 0 || q[1] >= ny || q[2] < 0 || q[2] >= nz) continue;
            const int w = (q[0] * ny + q[1]) * nz + q[2];
            if (d.voxel_to_cell[w] >= 0) continue;
            d.voxel_to_cell[w] = src;
            queue.push_back(w);
        }
    }
    d.cell_volume.assign(d.cell_center.size(), d.volume / static_cast<double>(d.cell_center.size()));
    if (!grid_key.empty()) {
        std::lock_guard<std::mutex> hold(cache_lock);
        if (grid_cache.size() >= kCacheEntries) grid_cache.erase(grid_cache.begin());
        grid_cache[grid_key] = GridCells{d.inv_cell, d.voxel_to_cell, d.cell_center};
    }
    return d;
}

void write_domain_summary(const Domain& d, const SimConfig& c, const std::string& folder, int n_simplices) {
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
    // Linearized Callaway transport is conservative transport (uses_conservative_transport,
    // include/SimulationConfig.h:153-156).
    const bool conservative = c.collision == CollisionModel::Callaway || c.callaway_nonlinear || c.conservative_boundaries;
    out << "conservative_boundaries = " << (conservative ? "true" : "false") << '\n';
    if (conservative) {
        out << "collision_model = " << (c.collision == CollisionModel::Callaway ? "callaway" : "rta") << '\n';
        out << "collision_reference_temperature = " << c.background_temperature << '\n';
        out << "reservoir_generation = fixed_incoming_flux\n";
    }
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
    // Selector tokens as the reference keeps them: the mode, then each box's six
    // numbers at 17 significant digits (ConfigParsing.cpp flatten_boxes).
    auto tokens = [](const std::vector<std::array<double, 6>>& boxes) {
        if (boxes.empty()) return std::string("[]");
        std::ostringstream o;
        o << "[relative_box";
        for (const auto& b : boxes)
            for (double x : b) {
                std::ostringstream t;
                t << std::setprecision(17) << x;
                o << ", " << t.str();
            }
        o << "]";
        return o.str();
    };
    auto strings = [](const std::vector<std::string>& v) {
        std::ostringstream o;
        o << "[";
        for (size_t i = 0; i < v.size(); ++i) o << (i ? ", " : "") << v[i];
        o << "]";
        return o.str();
    };
    auto scaled = [&](const std::vector<double>& v, double f) {
        std::vector<double> w(v);
        for (double& x : w) x *= f;
        return join(w);
    };
    out << "boundary_position_selector_mode = " << (c.boundary_boxes.empty() ? "none" : "region") << '\n';
    out << "boundary_position_selector_count = " << c.boundary_boxes.size() << '\n';
    out << "boundary_position_tokens = " << tokens(c.boundary_boxes) << '\n';
    out << "periodic_pair_region_count = " << c.periodic_pairs.size() << '\n';
    out << "periodic_pair_tokens = " << tokens(c.periodic_pairs) << '\n';
    out << "material_folder = " << c.material_folder << '\n';
    out << "material_folders = " << strings(c.material_folders) << '\n';
    out << "material_interface_axis = " << "xyz"[c.interface_axis] << '\n';
    out << "material_interface_positions_nm = " << scaled(c.interface_positions, 0.1) << '\n';
    out << "material_interface_model = "
        << (c.interface_model == "amm" ? "acoustic_mismatch" : c.interface_model == "mmm" ? "mixed_mismatch"
                                                                                           : "diffuse_mismatch")
        << '\n';
    out << "material_interface_amm_fraction = " << (c.interface_model == "amm" ? 1.0 : c.interface_amm_fraction) << '\n';
    out << "material_interface_parallel_bin_inv_a = " << c.interface_parallel_bin << '\n';
    out << "output_folder = " << folder << '\n';
    out << "heat_source_enabled = " << (c.heat_source_enabled ? "true" : "false") << '\n';
    out << "heat_source_profile = " << (c.heat_profile == HeatProfile::Gaussian ? "gaussian" : "uniform") << '\n';
    out << "heat_source_min_nm = " << scaled(c.heat_min, 0.1) << '\n';
    out << "heat_source_max_nm = " << scaled(c.heat_max, 0.1) << '\n';
    out << "heat_source_center_nm = " << scaled(c.heat_center, 0.1) << '\n';
    out << "heat_source_centers_nm = " << scaled(c.heat_centers, 0.1) << '\n';
    out << "heat_source_sigma_nm = " << scaled(c.heat_sigma, 0.1) << '\n';
    out << "heat_source_half_width_nm = " << scaled(c.heat_half_width, 0.1) << '\n';
    out << "heat_source_min_angstrom = " << join(c.heat_min) << '\n';
    out << "heat_source_max_angstrom = " << join(c.heat_max) << '\n';
    out << "heat_source_center_angstrom = " << join(c.heat_center) << '\n';
    out << "heat_source_centers_angstrom = " << join(c.heat_centers) << '\n';
    out << "heat_source_sigma_angstrom = " << join(c.heat_sigma) << '\n';
    out << "heat_source_half_width_angstrom = " << join(c.heat_half_width) << '\n';
    out << "heat_source_power_density = " << c.heat_power_density << '\n';
    out << "heat_source_power_densities = " << join(c.heat_power_densities) << '\n';
    out << "heat_source_frequency_min_THz = " << c.heat_frequency_min << '\n';
    out << "heat_source_frequency_max_THz = " << c.heat_frequency_max << '\n';
    out << "heat_source_time_profile = " << (c.heat_time_square ? "square" : "constant") << '\n';
    out << "mode_excitation_enabled = " << (c.has_mode_excitation ? "true" : "false") << '\n';
    out << "mode_excitation_frequency_min_THz = " << c.excite_fmin << '\n';
    out << "mode_excitation_frequency_max_THz = " << c.excite_fmax << '\n';
    out << "mode_excitation_occupation_multiplier = " << c.excite_multiplier << '\n';
    out << '\n';
    out << "[geometry]\n";
    out << "volume = " << d.volume << '\n';
    out << "bounds_min = [" << d.bmin[0] << ", " << d.bmin[1] << ", " << d.bmin[2] << "]\n";
    out << "bounds_max = [" << d.bmax[0] << ", " << d.bmax[1] << ", " << d.bmax[2] << "]\n";
    out << "n_faces = " << d.faces.size() << '\n';
    out << "n_facets = " << d.facets.size() << '\n';
    out << "n_simplices = " << n_simplices << "\n\n";
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
