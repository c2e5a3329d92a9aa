#include "mesh.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <functional>
#include <limits>
#include <map>
#include <numeric>
#include <queue>
#include <sstream>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <unordered_set>

namespace pmc {
namespace {
Vec3 add(const Vec3& a, const Vec3& b) { return {a[0] + b[0], a[1] + b[1], a[2] + b[2]}; }
Vec3 sub(const Vec3& a, const Vec3& b) { return {a[0] - b[0], a[1] - b[1], a[2] - b[2]}; }
Vec3 mul(const Vec3& a, double s) { return {a[0] * s, a[1] * s, a[2] * s}; }
double dot(const Vec3& a, const Vec3& b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
Vec3 cross(const Vec3& a, const Vec3& b) {
    return {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
}
double norm(const Vec3& a) { return std::sqrt(dot(a, a)); }
Vec3 vmin(const Vec3& a, const Vec3& b) { return {std::min(a[0], b[0]), std::min(a[1], b[1]), std::min(a[2], b[2])}; }
Vec3 vmax(const Vec3& a, const Vec3& b) { return {std::max(a[0], b[0]), std::max(a[1], b[1]), std::max(a[2], b[2])}; }

std::string lower(std::string s) {
    for (auto& c : s) c = static_cast<char>(std::tolower(static_cast<unsigned char>(c)));
    return s;
}

// Vertex merge on a 1e-10 lattice, first occurrence keeps its index
// (src/geometry/MeshReader.cpp:41-64).
struct VertexMerge {
    std::unordered_map<std::string, int> index;
    int add(const Vec3& p, std::vector<Vec3>& vertices) {
        const double h = 1e-10;
        const std::string key = std::to_string(std::llround(p[0] / h)) + ":" + std::to_string(std::llround(p[1] / h)) +
                                ":" + std::to_string(std::llround(p[2] / h));
        const auto it = index.find(key);
        if (it != index.end()) return it->second;
        const int i = static_cast<int>(vertices.size());
        vertices.push_back(p);
        index.emplace(key, i);
        return i;
    }
};

MeshData read_binary_stl(const std::string& path) {
    std::ifstream in(path, std::ios::binary);
    if (!in) throw std::runtime_error("Failed to open STL file: " + path);
    char header[80];
    in.read(header, 80);
    uint32_t n = 0;
    in.read(reinterpret_cast<char*>(&n), 4);
    MeshData m;
    VertexMerge merge;
    for (uint32_t i = 0; i < n; ++i) {
        float rec[12];
        uint16_t attr = 0;
        in.read(reinterpret_cast<char*>(rec), sizeof rec);
        in.read(reinterpret_cast<char*>(&attr), 2);
        Tri t;
        for (int k = 0; k < 3; ++k)
            t[k] = merge.add({rec[3 + 3 * k], rec[4 + 3 * k], rec[5 + 3 * k]}, m.vertices);
        m.faces.push_back(t);
    }
    if (m.vertices.empty() || m.faces.empty()) throw std::runtime_error("Binary STL mesh has no vertices/faces.");
    return m;
}

MeshData read_ascii_stl(const std::string& path) {
    std::ifstream in(path);
    if (!in) throw std::runtime_error("Failed to open STL file: " + path);
    MeshData m;
    VertexMerge merge;
    std::vector<int> tri;
    std::string line;
    while (std::getline(in, line)) {
        std::istringstream iss(line);
        std::string tag;
        if (!(iss >> tag) || tag != "vertex") continue;
        double x = 0, y = 0, z = 0;
        iss >> x >> y >> z;
        tri.push_back(merge.add({x, y, z}, m.vertices));
        if (tri.size() == 3) {
            m.faces.push_back({tri[0], tri[1], tri[2]});
            tri.clear();
        }
    }
    if (m.vertices.empty() || m.faces.empty()) throw std::runtime_error("ASCII STL mesh has no vertices/faces.");
    return m;
}

MeshData read_obj(const std::string& path) {
    std::ifstream in(path);
    if (!in) throw std::runtime_error("Failed to open OBJ file: " + path);
    MeshData m;
    std::string line;
    while (std::getline(in, line)) {
        std::istringstream iss(line);
        std::string tag;
        if (!(iss >> tag) || tag[0] == '#') continue;
        if (tag == "v") {
            double x = 0, y = 0, z = 0;
            iss >> x >> y >> z;
            m.vertices.push_back({x, y, z});
        } else if (tag == "f") {
            std::vector<int> poly;
            std::string tok;
            while (iss >> tok) {
                const auto slash = tok.find('/');
                const std::string s = slash == std::string::npos ? tok : tok.substr(0, slash);
                if (s.empty()) continue;
                int idx = std::stoi(s);
                idx = idx < 0 ? static_cast<int>(m.vertices.size()) + idx : idx - 1;
                if (idx < 0 || idx >= static_cast<int>(m.vertices.size()))
                    throw std::runtime_error("OBJ face index out of range.");
                poly.push_back(idx);
            }
            for (size_t i = 1; i + 1 < poly.size(); ++i) m.faces.push_back({poly[0], poly[i], poly[i + 1]});
        }
    }
    if (m.vertices.empty() || m.faces.empty()) throw std::runtime_error("OBJ mesh has no vertices/faces.");
    return m;
}

bool binary_stl(const std::string& path) {
    // Size rule first: a binary file may also begin with "solid".
    std::ifstream in(path, std::ios::binary);
    if (!in) return false;
    std::string first(5, '\0');
    in.read(&first[0], 5);
    if (!in) return false;
    in.seekg(0, std::ios::end);
    const auto size = static_cast<uint64_t>(in.tellg());
    if (size < 84) return false;
    in.seekg(80, std::ios::beg);
    uint32_t n = 0;
    in.read(reinterpret_cast<char*>(&n), 4);
    if (84u + static_cast<uint64_t>(n) * 50u == size) return true;
    return lower(first) != "solid";
}
} // namespace

MeshData read_mesh_file(const std::string& path) {
    const auto dot_pos = path.find_last_of('.');
    const std::string ext = dot_pos == std::string::npos ? "" : lower(path.substr(dot_pos));
    if (ext == ".obj") return read_obj(path);
    if (ext == ".stl") return binary_stl(path) ? read_binary_stl(path) : read_ascii_stl(path);
    throw std::runtime_error("Unsupported mesh extension: " + ext + " (expected .stl or .obj)");
}

void write_binary_stl(const std::string& path, const MeshData& m, const std::string& name) {
    std::ofstream out(path, std::ios::binary);
    if (!out) throw std::runtime_error("Cannot write STL file: " + path);
    char header[80];
    std::memset(header, ' ', sizeof header);
    std::memcpy(header, name.data(), std::min<size_t>(name.size(), sizeof header));
    out.write(header, 80);
    const uint32_t n = static_cast<uint32_t>(m.faces.size());
    out.write(reinterpret_cast<const char*>(&n), 4);
    for (const auto& f : m.faces) {
        const Vec3 &a = m.vertices[f[0]], &b = m.vertices[f[1]], &c = m.vertices[f[2]];
        Vec3 nrm = cross(sub(b, a), sub(c, a));
        const double l = norm(nrm);
        if (l > 0) nrm = mul(nrm, 1.0 / l);
        float rec[12] = {static_cast<float>(nrm[0]), static_cast<float>(nrm[1]), static_cast<float>(nrm[2])};
        for (int k = 0; k < 3; ++k)
            for (int j = 0; j < 3; ++j) rec[3 + 3 * k + j] = static_cast<float>(m.vertices[f[k]][j]);
        out.write(reinterpret_cast<const char*>(rec), sizeof rec);
        const uint16_t attr = 0;
        out.write(reinterpret_cast<const char*>(&attr), 2);
    }
}

bool ray_triangle(const Vec3& o, const Vec3& d, const Vec3& a, const Vec3& b, const Vec3& c, double& t) {
    // Moller-Trumbore with the reference's 1e-10 tolerances (src/SurfaceMesh.cpp:162-186).
    const double eps = 1e-10;
    const Vec3 e1 = sub(b, a), e2 = sub(c, a);
    const Vec3 p = cross(d, e2);
    const double det = dot(e1, p);
    if (std::abs(det) < eps) return false;
    const double inv = 1.0 / det;
    const Vec3 s = sub(o, a);
    const double u = dot(s, p) * inv;
    if (u < -eps || u > 1.0 + eps) return false;
    const Vec3 q = cross(s, e1);
    const double v = dot(d, q) * inv;
    if (v < -eps || (u + v) > 1.0 + eps) return false;
    t = dot(e2, q) * inv;
    return t > eps;
}

Vec3 closest_point_on_triangle(const Vec3& p, const Vec3& a, const Vec3& b, const Vec3& c) {
    // Voronoi-region closest point (Ericson, Real-Time Collision Detection, 5.1.5),
    // in the case order of src/SurfaceMesh.cpp:110-158.
    const Vec3 ab = sub(b, a), ac = sub(c, a), ap = sub(p, a);
    const double d1 = dot(ab, ap), d2 = dot(ac, ap);
    if (d1 <= 0.0 && d2 <= 0.0) return a;
    const Vec3 bp = sub(p, b);
    const double d3 = dot(ab, bp), d4 = dot(ac, bp);
    if (d3 >= 0.0 && d4 <= d3) return b;
    const double vc = d1 * d4 - d3 * d2;
    if (vc <= 0.0 && d1 >= 0.0 && d3 <= 0.0) return add(a, mul(ab, d1 / (d1 - d3)));
    const Vec3 cp = sub(p, c);
    const double d5 = dot(ab, cp), d6 = dot(ac, cp);
    if (d6 >= 0.0 && d5 <= d6) return c;
    const double vb = d5 * d2 - d1 * d6;
    if (vb <= 0.0 && d2 >= 0.0 && d6 <= 0.0) return add(a, mul(ac, d2 / (d2 - d6)));
    const double va = d3 * d6 - d5 * d4;
    if (va <= 0.0 && (d4 - d3) >= 0.0 && (d5 - d6) >= 0.0)
        return add(b, mul(sub(c, b), (d4 - d3) / ((d4 - d3) + (d5 - d6))));
    const double den = 1.0 / (va + vb + vc);
    return add(add(a, mul(ab, vb * den)), mul(ac, vc * den));
}

// ---------------------------------------------------------------- SurfaceMesh
void SurfaceMesh::set(std::vector<Vec3> v, std::vector<Tri> f) {
    vertices = std::move(v);
    faces = std::move(f);
    if (vertices.empty() || faces.empty()) throw std::runtime_error("SurfaceMesh requires non-empty vertices and faces.");
    rebuild();
}

void SurfaceMesh::shift_to_origin() {
    bounding_box();
    const Vec3 lo = bmin;
    for (auto& v : vertices) v = sub(v, lo);
    rebuild();
}

void SurfaceMesh::rebuild() {
    // Order of src/SurfaceMesh.cpp:263-275.
    simplices_.clear();
    simplices_ready_ = false;
    bounding_box();
    edge_list();
    orient();
    face_metrics();
    neighbors();
    facet_groups();
    enclosed_volume();
}

void SurfaceMesh::bounding_box() {
    bmin = bmax = vertices.front();
    for (const auto& v : vertices) {
        bmin = vmin(bmin, v);
        bmax = vmax(bmax, v);
    }
}

void SurfaceMesh::edge_list() {
    // Edge ids in order of first appearance over (v0,v1), (v0,v2), (v1,v2) of each face.
    std::map<std::pair<int, int>, int> ids;
    edges.clear();
    edge_faces.clear();
    face_edges.assign(faces.size(), {0, 0, 0});
    for (size_t i = 0; i < faces.size(); ++i) {
        const auto& f = faces[i];
        const std::pair<int, int> pairs[3] = {std::minmax(f[0], f[1]), std::minmax(f[0], f[2]),
                                              std::minmax(f[1], f[2])};
        for (int j = 0; j < 3; ++j) {
            const auto ins = ids.emplace(pairs[j], static_cast<int>(edges.size()));
            if (ins.second) {
                edges.push_back({pairs[j].first, pairs[j].second});
                edge_faces.emplace_back();
            }
            face_edges[i][j] = ins.first->second;
            edge_faces[ins.first->second].push_back(static_cast<int>(i));
        }
    }
}

void SurfaceMesh::orient() {
    // Consistent winding over the face adjacency graph, then each connected shell
    // turned to positive signed volume (src/SurfaceMesh.cpp:277-413).  Surfaces
    // with an edge not shared by exactly two faces are oriented face by face from
    // ray-cast membership on both sides of each face.
    for (const auto& attached : edge_faces) {
        if (attached.size() == 2) continue;
        for (auto& f : faces) {
            const Vec3 a = vertices[f[0]], b = vertices[f[1]], c = vertices[f[2]];
            const Vec3 raw = cross(sub(b, a), sub(c, a));
            const double len = norm(raw);
            if (!(len > 0.0)) throw std::runtime_error("Degenerate triangle in CAD surface.");
            const Vec3 n = mul(raw, 1.0 / len);
            const Vec3 cen = mul(add(add(a, b), c), 1.0 / 3.0);
            const double edge = std::min({norm(sub(b, a)), norm(sub(c, a)), norm(sub(c, b))});
            bool resolved = false;
            for (double factor : {1e-8, 1e-7, 1e-6}) {
                const double eps = std::max(1e-8, edge * factor);
                const bool plus = contains(add(cen, mul(n, eps)));
                const bool minus = contains(add(cen, mul(n, -eps)));
                if (plus == minus) continue;
                if (plus) std::swap(f[1], f[2]);
                resolved = true;
                break;
            }
            if (!resolved)
                throw std::runtime_error("Cannot identify interior/exterior sides of CAD surface; check closure and "
                                         "overlapping faces.");
        }
        return;
    }
    auto direction = [&](int face, int u, int v) {
        const Tri& f = faces[face];
        for (int k = 0; k < 3; ++k) {
            const int a = f[k], b = f[(k + 1) % 3];
            if (a == u && b == v) return 1;
            if (a == v && b == u) return -1;
        }
        return 0;
    };
    std::vector<std::vector<std::pair<int, int>>> adj(faces.size());
    for (size_t e = 0; e < edge_faces.size(); ++e) {
        const int a = edge_faces[e][0], b = edge_faces[e][1];
        const int da = direction(a, edges[e][0], edges[e][1]), db = direction(b, edges[e][0], edges[e][1]);
        if (da == 0 || db == 0) throw std::runtime_error("Surface mesh edge-to-face topology is inconsistent.");
        adj[a].push_back({b, -da * db});
        adj[b].push_back({a, -da * db});
    }
    std::vector<int> sign(faces.size(), 0);
    std::vector<std::vector<int>> shells;
    for (int root = 0; root < static_cast<int>(faces.size()); ++root) {
        if (sign[root] != 0) continue;
        shells.emplace_back();
        std::queue<int> q;
        sign[root] = 1;
        q.push(root);
        while (!q.empty()) {
            const int cur = q.front();
            q.pop();
            shells.back().push_back(cur);
            for (const auto& [other, rel] : adj[cur]) {
                const int want = sign[cur] * rel;
                if (sign[other] == 0) {
                    sign[other] = want;
                    q.push(other);
                } else if (sign[other] != want) {
                    throw std::runtime_error("Surface mesh is non-orientable or has inconsistent face winding.");
                }
            }
        }
    }
    for (const auto& shell : shells) {
        long double v6 = 0.0L;
        for (int fi : shell) {
            const Tri& f = faces[fi];
            v6 += static_cast<long double>(sign[fi]) *
                  static_cast<long double>(dot(vertices[f[0]], cross(vertices[f[1]], vertices[f[2]])));
        }
        if (v6 < 0.0L)
            for (int fi : shell) sign[fi] = -sign[fi];
    }
    for (size_t i = 0; i < faces.size(); ++i)
        if (sign[i] < 0) std::swap(faces[i][1], faces[i][2]);
}

void SurfaceMesh::face_metrics() {
    const size_t n = faces.size();
    face_normal.assign(n, {0, 0, 0});
    face_centroid.assign(n, {0, 0, 0});
    face_area.assign(n, 0.0);
    face_k.assign(n, 0.0);
    face_bounds.assign(n, {Vec3{0, 0, 0}, Vec3{0, 0, 0}});
    for (size_t i = 0; i < n; ++i) {
        const Vec3 &a = vertices[faces[i][0]], &b = vertices[faces[i][1]], &c = vertices[faces[i][2]];
        const Vec3 nrm = cross(sub(b, a), sub(c, a));
        const double area2 = norm(nrm);
        face_normal[i] = area2 > 0.0 ? mul(nrm, 1.0 / area2) : Vec3{0, 0, 0};
        face_area[i] = area2 * 0.5;
        face_centroid[i] = {(a[0] + b[0] + c[0]) / 3.0, (a[1] + b[1] + c[1]) / 3.0, (a[2] + b[2] + c[2]) / 3.0};
        face_k[i] = -dot(face_normal[i], a);
        face_bounds[i] = {vmin(vmin(a, b), c), vmax(vmax(a, b), c)};
    }
}

void SurfaceMesh::neighbors() {
    face_adjacency.clear();
    for (const auto& ef : edge_faces)
        for (size_t i = 0; i < ef.size(); ++i)
            for (size_t j = i + 1; j < ef.size(); ++j)
                face_adjacency.push_back({std::min(ef[i], ef[j]), std::max(ef[i], ef[j])});
    std::sort(face_adjacency.begin(), face_adjacency.end());
    face_adjacency.erase(std::unique(face_adjacency.begin(), face_adjacency.end()), face_adjacency.end());
}

void SurfaceMesh::facet_groups() {
    // Faces bucketed by plane (normal and offset rounded to 1e-6) in order of
    // first appearance; with merging, each bucket splits into connected
    // components (breadth-first over face adjacency, seeds in bucket order)
    // (src/SurfaceMesh.cpp:581-722).
    face_to_facet.assign(faces.size(), -1);
    facets.clear();
    auto key_of = [](const Vec3& n, double d) {
        auto q = [](double x) { return std::to_string(static_cast<long long>(std::llround(x * 1e6))); };
        return q(n[0]) + ":" + q(n[1]) + ":" + q(n[2]) + ":" + q(d);
    };
    std::unordered_map<std::string, int> bucket_of;
    std::vector<std::vector<int>> buckets;
    for (size_t i = 0; i < faces.size(); ++i) {
        const auto ins = bucket_of.emplace(key_of(face_normal[i], dot(face_normal[i], vertices[faces[i][0]])),
                                           static_cast<int>(buckets.size()));
        if (ins.second) buckets.emplace_back();
        buckets[ins.first->second].push_back(static_cast<int>(i));
    }
    auto emit = [&](const std::vector<int>& members) {
        MeshFacet f;
        Vec3 ns{0, 0, 0}, cs{0, 0, 0};
        double as = 0;
        f.faces = members;
        for (int fi : members) {
            ns = add(ns, face_normal[fi]);
            cs = add(cs, mul(face_centroid[fi], face_area[fi]));
            as += face_area[fi];
        }
        const double l = norm(ns);
        f.normal = l > 0 ? mul(ns, 1.0 / l) : Vec3{0, 0, 0};
        f.area = as;
        f.centroid = as > 0 ? mul(cs, 1.0 / as) : cs;
        for (int fi : members) face_to_facet[fi] = static_cast<int>(facets.size());
        facets.push_back(std::move(f));
    };
    if (!merge_coplanar) {
        for (int fi = 0; fi < static_cast<int>(faces.size()); ++fi) emit({fi});
    } else {
        std::vector<std::vector<int>> nb(faces.size());
        for (const auto& a : face_adjacency) {
            nb[a[0]].push_back(a[1]);
            nb[a[1]].push_back(a[0]);
        }
        std::vector<char> seen(faces.size(), 0);
        for (const auto& bucket : buckets) {
            const std::unordered_set<int> in(bucket.begin(), bucket.end());
            for (int seed : bucket) {
                if (seen[seed]) continue;
                std::vector<int> comp;
                std::queue<int> q;
                seen[seed] = 1;
                q.push(seed);
                while (!q.empty()) {
                    const int cur = q.front();
                    q.pop();
                    comp.push_back(cur);
                    for (int o : nb[cur])
                        if (!seen[o] && in.count(o)) {
                            seen[o] = 1;
                            q.push(o);
                        }
                }
                emit(comp);
            }
        }
    }
    for (auto& f : facets) {
        std::unordered_map<int, int> count;
        for (int fi : f.faces)
            for (int e : face_edges[fi]) count[e] += 1;
        for (const auto& [e, n] : count)
            if (n == 1) f.boundary_edges.push_back(e);
        std::sort(f.boundary_edges.begin(), f.boundary_edges.end());
    }
}

void SurfaceMesh::enclosed_volume() {
    long double v6 = 0.0L;
    for (const auto& f : faces)
        v6 += static_cast<long double>(dot(vertices[f[0]], cross(vertices[f[1]], vertices[f[2]])));
    volume = std::abs(static_cast<double>(v6 / 6.0L));
    if (!std::isfinite(volume) || volume <= 1e-16)
        throw std::runtime_error("Surface mesh encloses no finite positive volume.");
}

bool SurfaceMesh::contains(const Vec3& p) const {
    // +x ray; crossings closer than 1e-9 count once (src/SurfaceMesh.cpp:1100-1124).
    std::vector<double> hits;
    for (const auto& f : faces) {
        double t = 0;
        if (ray_triangle(p, {1.0, 0.0, 0.0}, vertices[f[0]], vertices[f[1]], vertices[f[2]], t)) hits.push_back(t);
    }
    if (hits.empty()) return false;
    std::sort(hits.begin(), hits.end());
    int unique = 1;
    for (size_t i = 1; i < hits.size(); ++i)
        if (std::abs(hits[i] - hits[i - 1]) > 1e-9) ++unique;
    return unique % 2 == 1;
}

void SurfaceMesh::build_star() const {
    simplices_.clear();
    Vec3 center{0, 0, 0};
    for (const auto& v : vertices) center = add(center, v);
    center = mul(center, 1.0 / static_cast<double>(vertices.size()));
    if (!contains(center)) center = mul(add(bmin, bmax), 0.5);
    double total = 0.0;
    for (const auto& f : faces) {
        const Vec3 a = vertices[f[0]], b = vertices[f[1]], c = vertices[f[2]];
        const double vol = std::abs(dot(sub(a, center), cross(sub(b, center), sub(c, center)))) / 6.0;
        if (vol <= 1e-16) continue;
        // Inverse of [a-c, b-c, c-c] columns by cofactors (adjugate / determinant).
        const double m[3][3] = {{a[0] - center[0], b[0] - center[0], c[0] - center[0]},
                                {a[1] - center[1], b[1] - center[1], c[1] - center[1]},
                                {a[2] - center[2], b[2] - center[2], c[2] - center[2]}};
        const double A = m[1][1] * m[2][2] - m[1][2] * m[2][1];
        const double B = -(m[1][0] * m[2][2] - m[1][2] * m[2][0]);
        const double C = m[1][0] * m[2][1] - m[1][1] * m[2][0];
        const double D = -(m[0][1] * m[2][2] - m[0][2] * m[2][1]);
        const double E = m[0][0] * m[2][2] - m[0][2] * m[2][0];
        const double F = -(m[0][0] * m[2][1] - m[0][1] * m[2][0]);
        const double G = m[0][1] * m[1][2] - m[0][2] * m[1][1];
        const double H = -(m[0][0] * m[1][2] - m[0][2] * m[1][0]);
        const double I = m[0][0] * m[1][1] - m[0][1] * m[1][0];
        const double det = m[0][0] * A + m[0][1] * B + m[0][2] * C;
        if (std::abs(det) < 1e-14) continue;
        const double id = 1.0 / det;
        Simplex s;
        s.v = {center, a, b, c};
        s.volume = vol;
        s.bounds = {vmin(vmin(center, a), vmin(b, c)), vmax(vmax(center, a), vmax(b, c))};
        s.inv = {{{A * id, D * id, G * id}, {B * id, E * id, H * id}, {C * id, F * id, I * id}}};
        simplices_.push_back(s);
        total += vol;
    }
    const double rel = std::abs(total - volume) / std::max(volume, 1e-18);
    if (!std::isfinite(rel) || rel > 1e-3) simplices_.clear();
    simplices_ready_ = true;
}

const std::vector<SurfaceMesh::Simplex>& SurfaceMesh::simplices() const {
    if (!simplices_ready_) build_star();
    return simplices_;
}

bool SurfaceMesh::contains_point(const Vec3& p) const {
    if (!simplices_ready_) build_star();
    const double tol = 1e-10;
    for (const auto& s : simplices_) {
        if (p[0] < s.bounds[0][0] - tol || p[1] < s.bounds[0][1] - tol || p[2] < s.bounds[0][2] - tol ||
            p[0] > s.bounds[1][0] + tol || p[1] > s.bounds[1][1] + tol || p[2] > s.bounds[1][2] + tol)
            continue;
        const Vec3 q = sub(p, s.v[0]);
        const double l1 = s.inv[0][0] * q[0] + s.inv[0][1] * q[1] + s.inv[0][2] * q[2];
        const double l2 = s.inv[1][0] * q[0] + s.inv[1][1] * q[1] + s.inv[1][2] * q[2];
        const double l3 = s.inv[2][0] * q[0] + s.inv[2][1] * q[1] + s.inv[2][2] * q[2];
        const double l0 = 1.0 - l1 - l2 - l3;
        if (l0 >= -tol && l1 >= -tol && l2 >= -tol && l3 >= -tol && l0 <= 1.0 + tol && l1 <= 1.0 + tol &&
            l2 <= 1.0 + tol && l3 <= 1.0 + tol)
            return true;
    }
    return contains(p);
}

std::vector<double> SurfaceMesh::boundary_edge_lengths(int facet) const {
    std::vector<double> out;
    for (int e : facets[facet].boundary_edges) out.push_back(norm(sub(vertices[edges[e][0]], vertices[edges[e][1]])));
    std::sort(out.begin(), out.end());
    return out;
}

// ---------------------------------------------------------------- distance field
DistanceField build_distance_field(const SurfaceMesh& m, const Vec3& lo, const Vec3& hi, long max_voxels,
                                   double unit) {
    DistanceField f;
    f.unit = unit;
    const Vec3 ext = sub(hi, lo);
    const double big = std::max({ext[0], ext[1], ext[2], 1e-9});
    double prod = 1.0;
    for (int a = 0; a < 3; ++a) prod *= std::max(ext[a], 1e-3 * big);
    f.h = std::max(std::cbrt(prod / static_cast<double>(max_voxels)), 1e-6 * big);
    for (int a = 0; a < 3; ++a) f.dims[a] = std::max(1, static_cast<int>(std::ceil(ext[a] / f.h)));
    while (static_cast<long>(f.dims[0]) * f.dims[1] * f.dims[2] > max_voxels) {
        f.h *= 1.05;
        for (int a = 0; a < 3; ++a) f.dims[a] = std::max(1, static_cast<int>(std::ceil(ext[a] / f.h)));
    }
    const double half_diag = 0.5 * f.h * std::sqrt(3.0);
    const double cap = 255.0 * unit, reach = cap + half_diag;
    const int nx = f.dims[0], ny = f.dims[1], nz = f.dims[2];
    f.bound.assign(static_cast<size_t>(nx) * ny * nz, 255);
    const int nf = static_cast<int>(m.faces.size());
    // Voxel range of each face's box grown by `reach`; a voxel no face range
    // covers is at least `reach` from every face, so its bound is at least cap.
    std::vector<std::array<int, 6>> range(nf);
    for (int t = 0; t < nf; ++t)
        for (int a = 0; a < 3; ++a) {
            const double b0 = (m.face_bounds[t][0][a] - reach - lo[a]) / f.h - 0.5;
            const double b1 = (m.face_bounds[t][1][a] + reach - lo[a]) / f.h - 0.5;
            range[t][a] = std::clamp(static_cast<int>(std::ceil(b0)), 0, f.dims[a] - 1);
            range[t][3 + a] = std::clamp(static_cast<int>(std::floor(b1)), -1, f.dims[a] - 1);
        }
#pragma omp parallel for schedule(dynamic, 1)
    for (int i = 0; i < nx; ++i) { // one x slab per iteration: no two threads write one voxel
        for (int t = 0; t < nf; ++t) {
            if (i < range[t][0] || i > range[t][3]) continue;
            const auto& tri = m.faces[t];
            const Vec3 &a = m.vertices[tri[0]], &b = m.vertices[tri[1]], &c = m.vertices[tri[2]];
            for (int j = range[t][1]; j <= range[t][4]; ++j)
                for (int k = range[t][2]; k <= range[t][5]; ++k) {
                    const Vec3 centre{lo[0] + (i + 0.5) * f.h, lo[1] + (j + 0.5) * f.h, lo[2] + (k + 0.5) * f.h};
                    const double d = norm(sub(closest_point_on_triangle(centre, a, b, c), centre)) - half_diag;
                    const double q = std::floor(std::max(0.0, d) / unit);
                    uint8_t& cell = f.bound[(static_cast<size_t>(i) * ny + j) * nz + k];
                    if (q < cell) cell = static_cast<uint8_t>(q);
                }
        }
    }
    // Side of each face-connected region of positive-bound voxels.
    f.inside.assign(f.bound.size(), 0);
    std::vector<uint8_t> seen(f.bound.size(), 0);
    std::vector<size_t> queue;
    for (size_t start = 0; start < f.bound.size(); ++start) {
        if (f.bound[start] == 0 || seen[start]) continue;
        const int si = static_cast<int>(start / (static_cast<size_t>(ny) * nz));
        const int sj = static_cast<int>((start / nz) % ny), sk = static_cast<int>(start % nz);
        const Vec3 centre{lo[0] + (si + 0.5) * f.h, lo[1] + (sj + 0.5) * f.h, lo[2] + (sk + 0.5) * f.h};
        const uint8_t side = m.contains_point(centre) ? 1 : 0;
        queue.assign(1, start);
        seen[start] = 1;
        while (!queue.empty()) {
            const size_t v = queue.back();
            queue.pop_back();
            f.inside[v] = side;
            const int i = static_cast<int>(v / (static_cast<size_t>(ny) * nz));
            const int j = static_cast<int>((v / nz) % ny), k = static_cast<int>(v % nz);
            const int nb[6][3] = {{i - 1, j, k}, {i + 1, j, k}, {i, j - 1, k}, {i, j + 1, k}, {i, j, k - 1}, {i, j, k + 1}};
            for (const auto& n : nb) {
                if (n[0] < 0 || n[1] < 0 || n[2] < 0 || n[0] >= nx || n[1] >= ny || n[2] >= nz) continue;
                const size_t w = (static_cast<size_t>(n[0]) * ny + n[1]) * nz + n[2];
                if (f.bound[w] == 0 || seen[w]) continue;
                seen[w] = 1;
                queue.push_back(w);
            }
        }
    }
    return f;
}

// ---------------------------------------------------------------- device BVH
Bvh build_bvh(const std::vector<std::array<Vec3, 2>>& boxes, const std::vector<Vec3>& centroids, int leaf) {
    constexpr int kBins = 12;
    Bvh bvh;
    const int n = static_cast<int>(boxes.size());
    bvh.order.resize(n);
    std::iota(bvh.order.begin(), bvh.order.end(), 0);
    if (n == 0) return bvh;
    auto area = [](const Vec3& lo, const Vec3& hi) {
        const Vec3 e = sub(hi, lo);
        return 2.0 * (e[0] * e[1] + e[1] * e[2] + e[2] * e[0]);
    };
    std::function<int(int, int, int)> build = [&](int begin, int end, int depth) -> int {
        bvh.depth = std::max(bvh.depth, depth);
        const int idx = static_cast<int>(bvh.nodes.size());
        bvh.nodes.push_back({});
        Vec3 lo = boxes[bvh.order[begin]][0], hi = boxes[bvh.order[begin]][1];
        Vec3 clo = centroids[bvh.order[begin]], chi = clo;
        for (int i = begin + 1; i < end; ++i) {
            const int f = bvh.order[i];
            lo = vmin(lo, boxes[f][0]);
            hi = vmax(hi, boxes[f][1]);
            clo = vmin(clo, centroids[f]);
            chi = vmax(chi, centroids[f]);
        }
        pmc_bvh_node node{};
        for (int a = 0; a < 3; ++a) {
            // Rounded outward so the FP32 box contains the FP64 one.
            node.bmin[a] = std::nextafter(static_cast<float>(lo[a]), -INFINITY);
            node.bmax[a] = std::nextafter(static_cast<float>(hi[a]), INFINITY);
        }
        const int count = end - begin;
        int axis = -1, split = -1;
        auto bin_of = [&](int f, int a) {
            const double ext = chi[a] - clo[a];
            return std::clamp(static_cast<int>((centroids[f][a] - clo[a]) / ext * kBins), 0, kBins - 1);
        };
        if (count > leaf) {
            double best = std::numeric_limits<double>::infinity();
            for (int a = 0; a < 3; ++a) {
                if (chi[a] - clo[a] <= 1e-12) continue;
                int bin_n[kBins] = {};
                Vec3 bin_lo[kBins], bin_hi[kBins];
                for (int i = begin; i < end; ++i) {
                    const int f = bvh.order[i];
                    const int b = bin_of(f, a);
                    bin_lo[b] = bin_n[b] ? vmin(bin_lo[b], boxes[f][0]) : boxes[f][0];
                    bin_hi[b] = bin_n[b] ? vmax(bin_hi[b], boxes[f][1]) : boxes[f][1];
                    ++bin_n[b];
                }
                for (int s = 1; s < kBins; ++s) {
                    int nl = 0, nr = 0;
                    Vec3 llo{}, lhi{}, rlo{}, rhi{};
                    for (int b = 0; b < kBins; ++b) {
                        if (!bin_n[b]) continue;
                        if (b < s) {
                            llo = nl ? vmin(llo, bin_lo[b]) : bin_lo[b];
                            lhi = nl ? vmax(lhi, bin_hi[b]) : bin_hi[b];
                            nl += bin_n[b];
                        } else {
                            rlo = nr ? vmin(rlo, bin_lo[b]) : bin_lo[b];
                            rhi = nr ? vmax(rhi, bin_hi[b]) : bin_hi[b];
                            nr += bin_n[b];
                        }
                    }
                    if (!nl || !nr) continue;
                    const double cost = nl * area(llo, lhi) + nr * area(rlo, rhi);
                    if (cost < best) {
                        best = cost;
                        axis = a;
                        split = s;
                    }
                }
            }
        }
        if (axis < 0) {
            // Leaf; a set of coincident centroids larger than `leaf` stays whole.
            node.index = begin;
            node.count = count;
            bvh.nodes[idx] = node;
            return idx;
        }
        const auto mid_it = std::stable_partition(bvh.order.begin() + begin, bvh.order.begin() + end,
                                                  [&](int f) { return bin_of(f, axis) < split; });
        const int mid = static_cast<int>(mid_it - bvh.order.begin());
        build(begin, mid, depth + 1); // left child is idx + 1
        node.index = build(mid, end, depth + 1);
        node.count = 0;
        bvh.nodes[idx] = node;
        return idx;
    };
    build(0, n, 1);
    return bvh;
}

// ---------------------------------------------------------------- FinFET
MeshData finfet_mesh(const FinfetSpec& s) {
    // Solid = substrate box [0,S] plus the fin box centred on the substrate top;
    // lattice coordinates every step along each axis; each lattice face between a
    // solid and an empty (or outside) cell becomes two outward-wound triangles.
    const double fx0 = 0.5 * (s.substrate[0] - s.fin[0]), fy0 = 0.5 * (s.substrate[1] - s.fin[1]);
    const double top[3] = {s.substrate[0], s.substrate[1], s.substrate[2] + s.fin[2]};
    int n[3];
    for (int a = 0; a < 3; ++a) {
        n[a] = static_cast<int>(std::llround(top[a] / s.step[a]));
        if (std::abs(n[a] * s.step[a] - top[a]) > 1e-9 * top[a])
            throw std::runtime_error("FinFET sizes must be multiples of the lattice step.");
    }
    auto coord = [&](int a, int i) { return i * s.step[a]; };
    auto solid = [&](int i, int j, int k) {
        if (i < 0 || j < 0 || k < 0 || i >= n[0] || j >= n[1] || k >= n[2]) return false;
        const double x = coord(0, i) + 0.5 * s.step[0], y = coord(1, j) + 0.5 * s.step[1],
                     z = coord(2, k) + 0.5 * s.step[2];
        if (z < s.substrate[2]) return true;
        return x > fx0 && x < fx0 + s.fin[0] && y > fy0 && y < fy0 + s.fin[1];
    };
    MeshData m;
    std::map<std::array<int, 3>, int> vid;
    auto vertex = [&](int i, int j, int k) {
        const std::array<int, 3> key{i, j, k};
        const auto it = vid.find(key);
        if (it != vid.end()) return it->second;
        const int id = static_cast<int>(m.vertices.size());
        m.vertices.push_back({coord(0, i), coord(1, j), coord(2, k)});
        vid.emplace(key, id);
        return id;
    };
    for (int i = 0; i < n[0]; ++i)
        for (int j = 0; j < n[1]; ++j)
            for (int k = 0; k < n[2]; ++k) {
                if (!solid(i, j, k)) continue;
                const int c[3] = {i, j, k};
                for (int a = 0; a < 3; ++a)
                    for (int side = 0; side < 2; ++side) {
                        int nb[3] = {i, j, k};
                        nb[a] += side ? 1 : -1;
                        if (solid(nb[0], nb[1], nb[2])) continue;
                        // Quad on the plane a = c[a] + side, corners ordered
                        // counter-clockwise seen from the outward side.
                        const int u = (a + 1) % 3, v = (a + 2) % 3;
                        int p[4][3];
                        const int du[4] = {0, 1, 1, 0}, dv[4] = {0, 0, 1, 1};
                        for (int q = 0; q < 4; ++q) {
                            p[q][a] = c[a] + side;
                            p[q][u] = c[u] + du[q];
                            p[q][v] = c[v] + dv[q];
                        }
                        int ids[4];
                        for (int q = 0; q < 4; ++q) ids[q] = vertex(p[q][0], p[q][1], p[q][2]);
                        if (side) {
                            m.faces.push_back({ids[0], ids[1], ids[2]});
                            m.faces.push_back({ids[0], ids[2], ids[3]});
                        } else {
                            m.faces.push_back({ids[0], ids[2], ids[1]});
                            m.faces.push_back({ids[0], ids[3], ids[2]});
                        }
                    }
            }
    return m;
}

} // namespace pmc
