// CPU restatement of the reference's deviational Monte Carlo step (see oracle.h).
// Checker only: FP64 throughout, positions kept in FP64, the collision cache and
// push/nudge placement of TransportOperator kept as the reference has them.
#include "oracle.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <limits>
#include <numeric>
#include <stdexcept>
#include <string>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {
constexpr double kHbar = 6.582119569e-4;
constexpr double kBoltz = 8.617333262145e-5;
constexpr double kPi = 3.14159265358979323846;
using V3 = std::array<double, 3>;

V3 add(const V3& a, const V3& b) { return {a[0] + b[0], a[1] + b[1], a[2] + b[2]}; }
V3 sub(const V3& a, const V3& b) { return {a[0] - b[0], a[1] - b[1], a[2] - b[2]}; }
V3 mul(const V3& a, double s) { return {a[0] * s, a[1] * s, a[2] * s}; }
double dot(const V3& a, const V3& b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
double norm(const V3& a) { return std::sqrt(dot(a, a)); }

// src/material/Thermodynamics.cpp:71-82.
double bose(double t, double w) {
    if (t <= 0.0 || w <= 0.0) return 0.0;
    const double x = kHbar * w / (kBoltz * t);
    if (x > 700.0) return 0.0;
    return 1.0 / std::max(std::exp(x) - 1.0, 1e-12);
}

double lerp(double x0, double x1, double y0, double y1, double x) {
    if (std::abs(x1 - x0) <= 1e-18) return y0;
    const double t = (x - x0) / (x1 - x0);
    return y0 * (1.0 - t) + y1 * t;
}

// src/material/Thermodynamics.cpp:41-69 (clamped, upper_bound).
double interp(const double* xs, const double* ys, int n, double x) {
    if (n <= 0) return 0.0;
    if (n == 1) return ys[0];
    if (x <= xs[0]) return ys[0];
    if (x >= xs[n - 1]) return ys[n - 1];
    const int hi = static_cast<int>(std::upper_bound(xs, xs + n, x) - xs);
    return lerp(xs[hi - 1], xs[hi], ys[hi - 1], ys[hi], x);
}

// src/material/Thermodynamics.cpp:95-133.
double relaxation_weight(const pmc_mode_table* m, double t, int active, double dt) {
    const int flat = m->active_flat[active];
    const size_t nm = static_cast<size_t>(m->n_flat);
    const int nt = m->n_temps;
    double g;
    if (t <= m->temps[0]) g = m->gamma_total[flat];
    else if (t >= m->temps[nt - 1]) g = m->gamma_total[(nt - 1) * nm + flat];
    else {
        const int hi = static_cast<int>(std::upper_bound(m->temps, m->temps + nt, t) - m->temps);
        g = lerp(m->temps[hi - 1], m->temps[hi], m->gamma_total[(hi - 1) * nm + flat], m->gamma_total[hi * nm + flat], t);
    }
    if (g == 0.0 || dt == 0.0) return 0.0;
    return -std::expm1(-(g * dt) * (4.0 * kPi));
}

struct Uniforms {
    uint32_t uid, step, ev, k0, k1;
    int used = 0;
    double u[2] = {0, 0};
    double next() {
        if ((used & 1) == 0) {
            uint32_t w[4];
            orc_philox(uid, step, 2 * ev + (used >> 1), 1u, k0, k1, w);
            u[0] = orc_u53(w[0], w[1]);
            u[1] = orc_u53(w[2], w[3]);
        }
        return u[used++ & 1];
    }
};
} // namespace

extern "C" {

void orc_philox(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0, uint32_t k1, uint32_t out[4]) {
    // Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11): round multipliers and
    // Weyl key increments of the published algorithm.
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = static_cast<uint64_t>(0xD2511F53u) * c0;
        const uint64_t p1 = static_cast<uint64_t>(0xCD9E8D57u) * c2;
        const uint32_t hi0 = static_cast<uint32_t>(p0 >> 32), lo0 = static_cast<uint32_t>(p0);
        const uint32_t hi1 = static_cast<uint32_t>(p1 >> 32), lo1 = static_cast<uint32_t>(p1);
        const uint32_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
        c0 = n0;
        c1 = lo1;
        c2 = n2;
        c3 = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0;
    out[1] = c1;
    out[2] = c2;
    out[3] = c3;
}

double orc_u53(uint32_t lo, uint32_t hi) {
    const uint64_t x = ((static_cast<uint64_t>(hi) << 32) | lo) >> 11;
    return static_cast<double>(x) * 0x1.0p-53;
}

int orc_trace_box(const pmc_geometry* g, const double pos[3], const double vel[3], double point[3], double* t) {
    // SimulationDomain::trace_boundary_intersection, analytic box branch.
    const double inf = std::numeric_limits<double>::infinity();
    const double velocity_eps = 1e-18, time_eps = 1e-14;
    double ext[3];
    for (int a = 0; a < 3; ++a) ext[a] = g->bmax[a] - g->bmin[a];
    const double scale = std::max({1.0, std::abs(ext[0]), std::abs(ext[1]), std::abs(ext[2])});
    const double tol = std::max(1e-9, 64.0 * std::numeric_limits<double>::epsilon() * scale);
    for (int a = 0; a < 3; ++a) {
        point[a] = pos[a];
        if (!std::isfinite(pos[a]) || pos[a] < g->bmin[a] - tol || pos[a] > g->bmax[a] + tol) {
            *t = inf;
            return -1;
        }
    }
    double best = inf;
    int plane = -1;
    for (int a = 0; a < 3; ++a) {
        const double c = vel[a];
        if (!std::isfinite(c) || std::abs(c) <= velocity_eps) continue;
        const int side = c > 0.0 ? 1 : 0;
        const double target = side ? g->bmax[a] : g->bmin[a];
        const double ht = (target - pos[a]) / c;
        if (std::isfinite(ht) && ht > time_eps && ht < best) {
            best = ht;
            plane = 2 * a + side;
        }
    }
    if (plane < 0) {
        *t = inf;
        return -1;
    }
    double col[3];
    for (int a = 0; a < 3; ++a) col[a] = pos[a] + vel[a] * best;
    const int ha = plane / 2;
    col[ha] = (plane % 2 == 0) ? g->bmin[ha] : g->bmax[ha];
    for (int a = 0; a < 3; ++a) {
        if (!std::isfinite(col[a]) || col[a] < g->bmin[a] - tol || col[a] > g->bmax[a] + tol) {
            *t = inf;
            return -1;
        }
        col[a] = std::clamp(col[a], g->bmin[a], g->bmax[a]);
    }
    for (int a = 0; a < 3; ++a) point[a] = col[a];
    *t = best;
    return g->box_plane_facet[plane];
}

int orc_grid_index(const pmc_geometry* g, const double p[3]) {
    int idx[3];
    for (int a = 0; a < 3; ++a) {
        const int n = g->grid[a];
        if (n <= 1 || (g->bmax[a] - g->bmin[a]) <= 0.0) {
            idx[a] = 0;
            continue;
        }
        int k = static_cast<int>(std::floor((p[a] - g->bmin[a]) * g->inv_cell[a]));
        idx[a] = std::clamp(k, 0, n - 1);
    }
    return g->voxel_to_cell[(idx[0] * g->grid[1] + idx[1]) * g->grid[2] + idx[2]];
}

namespace {
V3 cross(const V3& a, const V3& b) {
    return {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
}
V3 mesh_vertex(const pmc_mesh* M, int face, int k) {
    const int v = M->faces[3 * face + k];
    return {M->vertices[3 * v], M->vertices[3 * v + 1], M->vertices[3 * v + 2]};
}
// Moller-Trumbore with 1e-10 tolerances (src/SurfaceMesh.cpp:162-186).
bool ray_triangle(const V3& o, const V3& d, const V3& a, const V3& b, const V3& c, double& t) {
    const double eps = 1e-10;
    const V3 e1 = sub(b, a), e2 = sub(c, a);
    const V3 p = cross(d, e2);
    const double det = dot(e1, p);
    if (std::abs(det) < eps) return false;
    const double inv = 1.0 / det;
    const V3 s = sub(o, a);
    const double u = dot(s, p) * inv;
    if (u < -eps || u > 1.0 + eps) return false;
    const V3 q = cross(s, e1);
    const double v = dot(d, q) * inv;
    if (v < -eps || (u + v) > 1.0 + eps) return false;
    t = dot(e2, q) * inv;
    return t > eps;
}
// Voronoi-region closest point (src/SurfaceMesh.cpp:111-159).
V3 closest_point_on_triangle(const V3& p, const V3& a, const V3& b, const V3& c) {
    const V3 ab = sub(b, a), ac = sub(c, a), ap = sub(p, a);
    const double d1 = dot(ab, ap), d2 = dot(ac, ap);
    if (d1 <= 0.0 && d2 <= 0.0) return a;
    const V3 bp = sub(p, b);
    const double d3 = dot(ab, bp), d4 = dot(ac, bp);
    if (d3 >= 0.0 && d4 <= d3) return b;
    const double vc = d1 * d4 - d3 * d2;
    if (vc <= 0.0 && d1 >= 0.0 && d3 <= 0.0) return add(a, mul(ab, d1 / (d1 - d3)));
    const V3 cp = sub(p, c);
    const double d5 = dot(ab, cp), d6 = dot(ac, cp);
    if (d6 >= 0.0 && d5 <= d6) return c;
    const double vb = d5 * d2 - d1 * d6;
    if (vb <= 0.0 && d2 >= 0.0 && d6 <= 0.0) return add(a, mul(ac, d2 / (d2 - d6)));
    const double va = d3 * d6 - d5 * d4;
    if (va <= 0.0 && (d4 - d3) >= 0.0 && (d5 - d6) >= 0.0)
        return add(b, mul(sub(c, b), (d4 - d3) / ((d4 - d3) + (d5 - d6))));
    const double denom = 1.0 / (va + vb + vc);
    return add(add(a, mul(ab, vb * denom)), mul(ac, vc * denom));
}
// +x ray cast, crossings closer than 1e-9 counted once (src/SurfaceMesh.cpp:1100-1123).
bool ray_cast_inside(const pmc_mesh* M, const V3& p) {
    std::vector<double> hits;
    for (int f = 0; f < M->n_faces; ++f) {
        double t = 0;
        if (ray_triangle(p, {1.0, 0.0, 0.0}, mesh_vertex(M, f, 0), mesh_vertex(M, f, 1), mesh_vertex(M, f, 2), t))
            hits.push_back(t);
    }
    if (hits.empty()) return false;
    std::sort(hits.begin(), hits.end());
    int unique = 1;
    for (size_t i = 1; i < hits.size(); ++i)
        if (std::abs(hits[i] - hits[i - 1]) > 1e-9) ++unique;
    return unique % 2 == 1;
}
} // namespace

int orc_trace_mesh(const pmc_geometry* g, const double pos[3], const double vel[3], double point[3], double* t) {
    // SurfaceMesh::trace_boundary_intersection (src/SurfaceMesh.cpp:1164-1270): the
    // nearest face plane crossing with t > 1e-10 whose point lies within 1e-9 of the
    // face's box and 1e-7 of the triangle.  Faces are visited in index order with a
    // strict improvement test; the reference visits them in BVH order, which can
    // only differ on an exact tie in t between faces of two facets.
    const pmc_mesh* M = g->mesh;
    const double eps = 1e-10;
    const V3 x{pos[0], pos[1], pos[2]}, v{vel[0], vel[1], vel[2]};
    double best_t = std::numeric_limits<double>::infinity();
    int best = -1;
    V3 best_p = x;
    for (int f = 0; f < M->n_faces; ++f) {
        const V3 n{M->face_normal[3 * f], M->face_normal[3 * f + 1], M->face_normal[3 * f + 2]};
        const double denom = dot(v, n);
        if (std::abs(denom) < eps) continue;
        const double th = -(dot(x, n) + M->face_k[f]) / denom;
        if (th <= eps || th >= best_t) continue;
        const V3 p = add(x, mul(v, th));
        const double* bb = M->face_bounds + 6 * f;
        if (p[0] < bb[0] - 1e-9 || p[1] < bb[1] - 1e-9 || p[2] < bb[2] - 1e-9 || p[0] > bb[3] + 1e-9 ||
            p[1] > bb[4] + 1e-9 || p[2] > bb[5] + 1e-9)
            continue;
        const V3 cp = closest_point_on_triangle(p, mesh_vertex(M, f, 0), mesh_vertex(M, f, 1), mesh_vertex(M, f, 2));
        if (norm(sub(cp, p)) > 1e-7) continue;
        best_t = th;
        best_p = p;
        best = f;
    }
    for (int a = 0; a < 3; ++a) point[a] = best_p[a];
    if (best < 0) {
        *t = std::numeric_limits<double>::infinity();
        return -1;
    }
    *t = best_t;
    return M->face_facet[best];
}

int orc_contains_mesh(const pmc_geometry* g, const double pos[3]) {
    // SurfaceMesh::contains_point (src/SurfaceMesh.cpp:1126-1151): star tetrahedra,
    // inclusive to 1e-10 in barycentric coordinates, then the +x ray cast with
    // crossings closer than 1e-9 counted once (:1100-1123).
    const pmc_mesh* M = g->mesh;
    const V3 p{pos[0], pos[1], pos[2]};
    const double tol = 1e-10;
    for (int s = 0; s < M->n_simplices; ++s) {
        const double* b = M->simplex_bounds + 6 * s;
        if (p[0] < b[0] - tol || p[1] < b[1] - tol || p[2] < b[2] - tol || p[0] > b[3] + tol || p[1] > b[4] + tol ||
            p[2] > b[5] + tol)
            continue;
        const double* o = M->simplex_vertices + 12 * s;
        const double* A = M->simplex_inverse + 9 * s;
        const V3 q{p[0] - o[0], p[1] - o[1], p[2] - o[2]};
        const double l1 = A[0] * q[0] + A[1] * q[1] + A[2] * q[2];
        const double l2 = A[3] * q[0] + A[4] * q[1] + A[5] * q[2];
        const double l3 = A[6] * q[0] + A[7] * q[1] + A[8] * q[2];
        const double l0 = 1.0 - l1 - l2 - l3;
        if (l0 >= -tol && l1 >= -tol && l2 >= -tol && l3 >= -tol && l0 <= 1.0 + tol && l1 <= 1.0 + tol &&
            l2 <= 1.0 + tol && l3 <= 1.0 + tol)
            return 1;
    }
    return ray_cast_inside(M, p) ? 1 : 0;
}

int orc_nearest_facet(const pmc_geometry* g, const double pos[3]) {
    // SurfaceMesh::nearest_facet / nearest_face (src/SurfaceMesh.cpp:996-1022).
    const pmc_mesh* M = g->mesh;
    const V3 p{pos[0], pos[1], pos[2]};
    double best = std::numeric_limits<double>::infinity();
    int idx = -1;
    for (int f = 0; f < M->n_faces; ++f) {
        const V3 cp = closest_point_on_triangle(p, mesh_vertex(M, f, 0), mesh_vertex(M, f, 1), mesh_vertex(M, f, 2));
        const double d = norm(sub(p, cp));
        if (d < best) {
            best = d;
            idx = f;
        }
    }
    return idx < 0 ? -1 : M->face_facet[idx];
}

int orc_sample_volume(const pmc_geometry* g, uint32_t k0, uint32_t k1, uint32_t id, double point[3], double* u_mode) {
    // SurfaceMesh::sample_volume_points (src/geometry/SurfaceSampling.cpp:113-155) on
    // Philox draws: a star tetrahedron by volume, then Dirichlet(1,1,1,1) weights
    // from normalized exponentials.  Without a star decomposition, rejection in the
    // bounding box against the ray cast (:81-106), up to 256 trials per point.
    // Returns 0, or -1 when no trial was accepted.
    const pmc_mesh* M = g->mesh;
    uint32_t a[4], b[4], c[4];
    orc_philox(id, 0, 0, 3u, k0, k1, a);
    orc_philox(id, 0, 1, 3u, k0, k1, b);
    *u_mode = orc_u53(b[2], b[3]);
    if (M->n_simplices > 0) {
        orc_philox(id, 0, 2, 3u, k0, k1, c);
        double total = 0.0;
        for (int s = 0; s < M->n_simplices; ++s) total += M->simplex_volume[s];
        const double target = orc_u53(a[0], a[1]) * total;
        double acc = 0.0;
        int pick = M->n_simplices - 1;
        for (int s = 0; s < M->n_simplices; ++s) {
            acc += M->simplex_volume[s];
            if (acc > target) {
                pick = s;
                break;
            }
        }
        const double u[4] = {orc_u53(a[2], a[3]), orc_u53(b[0], b[1]), orc_u53(c[0], c[1]), orc_u53(c[2], c[3])};
        double w[4], sum = 0.0;
        for (int k = 0; k < 4; ++k) {
            w[k] = -std::log(std::max(u[k], 1e-15));
            sum += w[k];
        }
        const double* v = M->simplex_vertices + 12 * pick;
        for (int d = 0; d < 3; ++d) {
            double x = 0.0;
            for (int k = 0; k < 4; ++k) x += (w[k] / sum) * v[3 * k + d];
            point[d] = x;
        }
        return 0;
    }
    for (uint32_t trial = 0; trial < 256; ++trial) {
        orc_philox(id, 1 + trial, 0, 3u, k0, k1, a);
        orc_philox(id, 1 + trial, 1, 3u, k0, k1, b);
        const double u[3] = {orc_u53(a[0], a[1]), orc_u53(a[2], a[3]), orc_u53(b[0], b[1])};
        V3 p;
        for (int d = 0; d < 3; ++d) p[d] = g->bmin[d] + (g->bmax[d] - g->bmin[d]) * u[d];
        if (ray_cast_inside(M, p)) {
            for (int d = 0; d < 3; ++d) point[d] = p[d];
            return 0;
        }
    }
    return -1;
}

int orc_rough_sample(const pmc_mode_table* m, const pmc_rough_table* r, int in_ai, double in_occ, const double* u,
                     int* out_ai, double* out_occ, int* kind) {
    return orc_rough_sample_signed(m, r, in_ai, in_occ, u, out_ai, out_occ, kind, 0);
}

int orc_rough_sample_signed(const pmc_mode_table* m, const pmc_rough_table* r, int in_ai, double in_occ,
                            const double* u, int* out_ai, double* out_occ, int* kind, int signed_weights) {
    int used = 0;
    *out_ai = in_ai;
    *out_occ = in_occ;
    *kind = 0;
    const int na = m->n_active;
    const double p_spec = std::clamp(r->specularity[in_ai], 0.0, 1.0);
    const double w_in = m->omega[in_ai], neq_in = m->neq_background[in_ai];
    auto mapped_to = [&](int o) {
        const double mapped = m->neq_background[o] + w_in * (in_occ - neq_in) / std::max(1e-30, m->omega[o]);
        if (std::isfinite(mapped) && (signed_weights || mapped >= 0.0)) {
            *out_ai = o;
            *out_occ = mapped;
        }
    };
    if (u[used++] <= p_spec) {
        const int o = r->spec_match[in_ai];
        if (o >= 0 && o < na) {
            *kind = 1;
            mapped_to(o);
            return used;
        }
    }
    // sample_diffuse (src/solver/RoughBoundarySampler.cpp:373-456).
    int o = -1;
    const int begin = r->diffuse_begin[in_ai], end = r->diffuse_end[in_ai];
    bool chosen = false;
    if (begin >= 0 && end > begin && end <= r->n_outgoing) {
        const double uw = u[used++];
        // flux_weighted_window_index (include/solver/BoundaryFluxSampling.h:64-88).
        const double* prefix = r->residual_prefix;
        const double fb = prefix[begin], fe = prefix[end], wf = fe - fb;
        if (wf > 0.0 && std::isfinite(wf)) {
            const double uc = std::clamp(uw, 0.0, std::nextafter(1.0, 0.0));
            const double target = fb + uc * wf;
            const double* it = std::upper_bound(prefix + begin + 1, prefix + end + 1, target);
            const int pos = std::clamp(static_cast<int>(it - prefix) - 1, begin, end - 1);
            o = r->outgoing_sorted[pos];
            chosen = true;
        }
    }
    if (!chosen) {
        if (r->strict_elastic) return used;
        if (r->n_roulette > 0 && r->roulette_cdf[r->n_roulette - 1] > 0.0) {
            const double x = u[used++] * r->roulette_cdf[r->n_roulette - 1];
            size_t pos = static_cast<size_t>(std::lower_bound(r->roulette_cdf, r->roulette_cdf + r->n_roulette, x) -
                                             r->roulette_cdf);
            if (pos >= static_cast<size_t>(r->n_roulette)) pos = r->n_roulette - 1;
            o = r->roulette_active[pos];
        } else if (r->n_outgoing > 0) {
            const int k = std::min(static_cast<int>(u[used++] * r->n_outgoing), r->n_outgoing - 1);
            o = r->outgoing_active[k];
        } else {
            // any mode of the table's material (RoughBoundarySampler::sample_diffuse :451-455)
            const int lo = r->mode_count ? r->mode_begin : 0, cnt = r->mode_count ? r->mode_count : na;
            o = lo + std::min(static_cast<int>(u[used++] * cnt), cnt - 1);
        }
    }
    *kind = 2;
    mapped_to(o);
    return used;
}

int orc_reservoir_mode(const pmc_reservoir* r, int na, double u, int sampling) {
    if (sampling == 1) {
        const double* it = std::lower_bound(r->cdf, r->cdf + na, u);
        return it == r->cdf + na ? na - 1 : static_cast<int>(it - r->cdf);
    }
    const double x = u * na;
    int j = std::min(static_cast<int>(x), na - 1);
    return (x - j) < r->alias_prob[j] ? j : r->alias_index[j];
}

void orc_surface_point(const pmc_reservoir* r, double u_face, double r1, double r2, double point[3]) {
    int f = static_cast<int>(std::lower_bound(r->face_cdf, r->face_cdf + r->n_faces, u_face) - r->face_cdf);
    f = std::min(f, r->n_faces - 1);
    const double* tri = r->triangles + 9 * f;
    const double sr1 = std::sqrt(r1);
    const double w0 = 1.0 - sr1, w1 = sr1 * (1.0 - r2), w2 = sr1 * r2;
    for (int a = 0; a < 3; ++a) point[a] = w0 * tri[a] + w1 * tri[3 + a] + w2 * tri[6 + a];
}

void orc_refresh(const pmc_mode_table* m, int ncells, const double* volume, int64_t n, const int32_t* active,
                 const int32_t* cell, const double* occ, double* e, double* t) {
    for (int c = 0; c < ncells; ++c) e[c] = 0.0;
    for (int64_t i = 0; i < n; ++i) {
        const int a = active[i];
        const double energy = kHbar * m->omega[a] * (occ[i] - m->neq_background[a]);
        e[cell[i]] += energy * m->energy_weight;
    }
    for (int c = 0; c < ncells; ++c) {
        e[c] = e[c] / volume[c] + m->background_energy_density;
        t[c] = interp(m->lut_e, m->lut_t, m->lut_n, e[c]);
    }
}

double orc_rta(const pmc_mode_table* m, int ncells, const double* cell_T, int64_t n, const int32_t* active,
               const int32_t* cell, double* occ, double dt, double weight_e, int local, double lifetime_T,
               double* pseudo) {
    // Cell-ordered index, as CellParticleIndex::ensure builds it (src/solver/CellParticleIndex.cpp:14-44).
    std::vector<int64_t> offsets(ncells + 1, 0), order(n);
    for (int64_t i = 0; i < n; ++i) offsets[cell[i] + 1]++;
    for (int c = 0; c < ncells; ++c) offsets[c + 1] += offsets[c];
    std::vector<int64_t> cursor(offsets.begin(), offsets.end() - 1);
    for (int64_t i = 0; i < n; ++i) order[cursor[cell[i]]++] = i;
    std::vector<double> a(n);
    for (int64_t i = 0; i < n; ++i) a[i] = relaxation_weight(m, local ? cell_T[cell[i]] : lifetime_T, active[i], dt);
    double residual = 0.0;
    for (int c = 0; c < ncells; ++c) {
        const int64_t b = offsets[c], e = offsets[c + 1];
        if (pseudo) pseudo[c] = cell_T[c];
        if (e <= b) continue;
        auto evaluate = [&](double temperature, double* derivative) {
            double value = 0.0, slope = 0.0;
            const double T = std::max(0.0, temperature);
            for (int64_t k = b; k < e; ++k) {
                const int64_t i = order[k];
                const double w = std::max(0.0, m->omega[active[i]]);
                const double hw = kHbar * w;
                if (!(hw > 0.0) || !(a[i] > 0.0)) continue;
                const double neq = bose(T, w);
                value += hw * a[i] * (occ[i] - neq);
                if (T > 0.0) slope -= hw * a[i] * (hw * neq * (neq + 1.0) / (kBoltz * T * T));
            }
            if (derivative) *derivative = slope;
            return value;
        };
        double low = 0.0, high = std::max(1000.0, 2.0 * std::max(1.0, cell_T[c]));
        const double f_low = evaluate(low, nullptr);
        double f_high = evaluate(high, nullptr);
        for (int expand = 0; f_high > 0.0 && expand < 20; ++expand) {
            high *= 2.0;
            f_high = evaluate(high, nullptr);
        }
        double T = std::clamp(cell_T[c], low, high);
        if (f_low <= 0.0) T = low;
        else if (f_high >= 0.0) T = high;
        else {
            double escale = 0.0;
            for (int64_t k = b; k < e; ++k) {
                const int64_t i = order[k];
                const double hw = kHbar * std::max(0.0, m->omega[active[i]]);
                escale += hw * a[i] * std::max(1.0, occ[i]);
            }
            for (int iter = 0; iter < 40; ++iter) {
                double d = 0.0;
                const double v = evaluate(T, &d);
                if (std::abs(v) <= 1e-13 * std::max(1.0, escale)) break;
                if (v > 0.0) low = T;
                else high = T;
                double cand = 0.5 * (low + high);
                if (d < 0.0 && std::isfinite(d)) {
                    const double newton = T - v / d;
                    if (newton > low && newton < high && std::isfinite(newton)) cand = newton;
                }
                T = cand;
            }
        }
        if (pseudo) pseudo[c] = T;
        double before = 0.0, after = 0.0;
        for (int64_t k = b; k < e; ++k) {
            const int64_t i = order[k];
            const double w = std::max(0.0, m->omega[active[i]]);
            const double hw = kHbar * w;
            const double old = occ[i];
            const double nn = old + a[i] * (bose(T, w) - old);
            before += hw * old;
            after += hw * nn;
            occ[i] = std::max(0.0, nn);
        }
        residual += (after - before) * weight_e;
    }
    return residual;
}

// Thermal-spectrum branch of PhononSourceOperator::apply
// (src/solver/PhononSourceOperator.cpp:120-341): cells in order, each cell's
// eligible carriers in storage order; deposits every pending budget of a cell
// with eligible carriers and zeroes it.
int orc_source_deposit(const pmc_mode_table* m, int ncells, double* pending, const double* cell_T, double t_bg,
                       const int32_t* elig, int64_t n, const int32_t* active, const int32_t* cell, double* occ,
                       const double* capacity, double* energy_out, int64_t* updates_out) {
    return orc_source_deposit_spectrum(m, ncells, pending, cell_T, t_bg, elig, n, active, cell, occ, capacity, 0,
                                       nullptr, 0.0, 1.0, energy_out, updates_out);
}

int orc_source_deposit_spectrum(const pmc_mode_table* m, int ncells, double* pending, const double* cell_T,
                                double t_bg, const int32_t* elig, int64_t n, const int32_t* active,
                                const int32_t* cell, double* occ, const double* capacity, int spectrum,
                                const double* mode_weight, double center_thz, double sigma_thz, double* energy_out,
                                int64_t* updates_out) {
    *energy_out = 0.0;
    *updates_out = 0;
    double pending_total = 0.0;
    for (int c = 0; c < ncells; ++c) pending_total += pending[c];
    if (pending_total <= 0.0) return 0;
    std::vector<std::vector<int64_t>> members(ncells);
    for (int64_t i = 0; i < n; ++i) {
        const int c = cell[i];
        if (pending[c] <= 0.0) continue;
        const double hw = kHbar * m->omega[active[i]];
        if (std::isfinite(hw) && hw > 1e-14 && elig[active[i]]) members[c].push_back(i);
    }
    for (int c = 0; c < ncells; ++c) {
        if (pending[c] <= 0.0 || members[c].empty()) continue; // the budget stays in the cell
        const auto& idx = members[c];
        const double target = pending[c] / m->energy_weight;
        if (!std::isfinite(target)) return PMC_ERR_INVALID_STATE;
        const double tb = std::isfinite(cell_T[c]) && cell_T[c] >= 0.0 ? cell_T[c] : t_bg;
        std::vector<double> inc(idx.size());
        if (spectrum == 2) {
            // Gaussian in frequency (src/solver/PhononSourceOperator.cpp:186-206), normalized on
            // the cell's carriers about their largest log weight, in long double.
            std::vector<long double> lw(idx.size());
            long double peak = -std::numeric_limits<long double>::infinity();
            for (size_t k = 0; k < idx.size(); ++k) {
                const long double f = m->omega[active[idx[k]]] / 6.283185307179586476925286766559L;
                const long double z = (f - center_thz) / sigma_thz;
                lw[k] = -.5L * z * z;
                peak = std::max(peak, lw[k]);
            }
            long double sum = 0;
            for (auto& w : lw) {
                w = std::exp(w - peak);
                sum += w;
            }
            for (size_t k = 0; k < idx.size(); ++k)
                inc[k] = static_cast<double>(target * lw[k] / sum) / (kHbar * m->omega[active[idx[k]]]);
        } else if (spectrum == 1) {
            // Branch weights (:207-217): target * (w / sum w) / hw.
            double sum = 0.0;
            for (int64_t i : idx) sum += mode_weight[active[i]];
            for (size_t k = 0; k < idx.size(); ++k) {
                const int a = active[idx[k]];
                inc[k] = target * (mode_weight[a] / sum) / (kHbar * m->omega[a]);
            }
        } else if (capacity) {
            // Linearized transport: target * C_i / sum C / hw_i with the mode heat
            // capacities at T_bg (src/solver/PhononSourceOperator.cpp:221-236).
            long double total = 0;
            for (int64_t i : idx) total += capacity[active[i]];
            if (!(total > 0)) return PMC_ERR_INVALID_STATE;
            for (size_t k = 0; k < idx.size(); ++k) {
                const int a = active[idx[k]];
                inc[k] = static_cast<double>(target * capacity[a] / total) / (kHbar * m->omega[a]);
            }
        } else {
            auto energy = [&](double t, double* deriv) {
                double e = 0.0, slope = 0.0;
                for (int64_t i : idx) {
                    const double w = m->omega[active[i]];
                    const double hw = kHbar * std::max(0.0, w);
                    const double n0 = bose(tb, w), nn = bose(t, w);
                    e += hw * (nn - n0);
                    if (deriv && t > 0.0) slope += hw * (hw * nn * (nn + 1.0) / (kBoltz * t * t));
                }
                if (deriv) *deriv = slope;
                return e;
            };
            double low = std::max(0.0, tb);
            double high = std::max(low + 1.0, 1.05 * low + 1.0);
            double he = energy(high, nullptr);
            for (int x = 0; he < target && x < 60; ++x) {
                high = 2.0 * high + 1.0;
                he = energy(high, nullptr);
            }
            if (!(he >= target) || !std::isfinite(he)) return PMC_ERR_INVALID_STATE;
            double t = 0.5 * (low + high);
            const double tol = 1e-12 * std::max(1e-6, target);
            for (int iter = 0; iter < 60; ++iter) {
                double d = 0.0;
                const double r = energy(t, &d) - target;
                if (std::abs(r) <= tol) break;
                if (r > 0.0) high = t;
                else low = t;
                double cand = 0.5 * (low + high);
                if (d > 0.0 && std::isfinite(d)) {
                    const double nt = t - r / d;
                    if (nt > low && nt < high && std::isfinite(nt)) cand = nt;
                }
                t = cand;
            }
            for (size_t k = 0; k < idx.size(); ++k) {
                const double w = m->omega[active[idx[k]]];
                inc[k] = bose(t, w) - bose(tb, w);
            }
        }
        double applied = 0.0, largest = -1.0;
        size_t at = 0;
        for (size_t k = 0; k < idx.size(); ++k) {
            const double w = m->omega[active[idx[k]]];
            const double e = kHbar * w * inc[k];
            applied += e;
            if (e > largest) {
                largest = e;
                at = k;
            }
        }
        inc[at] += (target - applied) / (kHbar * m->omega[active[idx[at]]]);
        for (size_t k = 0; k < idx.size(); ++k) {
            const double next = occ[idx[k]] + inc[k];
            if (!std::isfinite(inc[k]) || !std::isfinite(next) || next < 0.0) return PMC_ERR_INVALID_STATE;
        }
        for (size_t k = 0; k < idx.size(); ++k) occ[idx[k]] += inc[k];
        *energy_out += pending[c];
        *updates_out += static_cast<int64_t>(idx.size());
        pending[c] = 0.0;
    }
    return 0;
}

} // extern "C"

// ---------------------------------------------------------------- linearized Callaway
// CallawayOperator (src/material/CallawayOperator.cpp:9-201) restated on
// explicit vectors in long double, as the reference computes it.
namespace {
using LVec = std::vector<long double>;

long double ldot(const LVec& a, const LVec& b) {
    long double s = 0;
    for (size_t i = 0; i < a.size(); ++i) s += a[i] * b[i];
    return s;
}
long double lnorm(const LVec& a) { return std::sqrt(ldot(a, a)); }

// gains (:28-64): sqrt(rate)-scaled basis, normalized, two-pass modified
// Gram-Schmidt, skipped below 1e-12, scaled by sqrt(rate) again.
std::vector<LVec> cw_gains(const std::vector<double>& rates, std::vector<LVec> basis) {
    std::vector<LVec> orth;
    for (auto v : basis) {
        for (size_t i = 0; i < v.size(); ++i) v[i] *= std::sqrt(static_cast<long double>(rates[i]));
        const long double original = lnorm(v);
        if (original == 0) continue;
        for (auto& x : v) x /= original;
        for (int pass = 0; pass < 2; ++pass)
            for (const auto& q : orth) {
                const long double proj = ldot(q, v);
                for (size_t i = 0; i < v.size(); ++i) v[i] -= proj * q[i];
            }
        const long double length = lnorm(v);
        if (length < 1e-12L) continue;
        for (auto& x : v) x /= length;
        orth.push_back(std::move(v));
    }
    for (auto& q : orth)
        for (size_t i = 0; i < q.size(); ++i) q[i] *= std::sqrt(static_cast<long double>(rates[i]));
    return orth;
}

struct CwOp {
    LVec scale, rate;
    std::vector<LVec> ngain, rgain;
    long double bound = 0;
};

// Constructor (:66-95).
bool cw_build(CwOp& op, const std::vector<double>& omega, const std::vector<double>& cap,
              const std::vector<V3>& q, const std::vector<double>& gn, const std::vector<double>& gr) {
    const size_t n = omega.size();
    op.scale.resize(n);
    op.rate.resize(n);
    std::vector<LVec> basis(4, LVec(n));
    for (size_t i = 0; i < n; ++i) {
        if (!std::isfinite(omega[i]) || omega[i] <= 0 || !std::isfinite(cap[i]) || cap[i] <= 0 ||
            !std::isfinite(gn[i]) || gn[i] < 0 || !std::isfinite(gr[i]) || gr[i] < 0)
            return false;
        op.scale[i] = std::sqrt(static_cast<long double>(cap[i]));
        basis[0][i] = op.scale[i];
        for (int a = 0; a < 3; ++a) basis[a + 1][i] = op.scale[i] * q[i][a] / omega[i];
        op.rate[i] = static_cast<long double>(gn[i]) + gr[i];
        op.bound = std::max(op.bound, op.rate[i]);
    }
    op.ngain = cw_gains(gn, basis);
    op.rgain = cw_gains(gr, {basis[0]});
    return true;
}

// action (:97-111).
LVec cw_action(const CwOp& op, const LVec& x) {
    LVec r(x.size());
    for (size_t i = 0; i < x.size(); ++i) r[i] = -op.rate[i] * x[i];
    for (const auto* gain : {&op.ngain, &op.rgain})
        for (const auto& u : *gain) {
            const long double c = ldot(u, x);
            for (size_t i = 0; i < x.size(); ++i) r[i] += u[i] * c;
        }
    return r;
}

// evolve (:135-201): 0, or -9 (work limit) / -10 (no convergence).
int cw_evolve(const CwOp& op, std::vector<double>& energy, double dt) {
    LVec x(energy.size());
    long double magnitude = 0;
    for (size_t i = 0; i < x.size(); ++i) {
        x[i] = energy[i] / op.scale[i];
        magnitude = std::max(magnitude, std::abs(x[i]));
    }
    if (magnitude == 0 || dt == 0 || op.bound == 0) return 0;
    const long double steps_ld = std::max(1.0L, std::ceil(dt * op.bound / 4.0L));
    if (!std::isfinite(steps_ld) || steps_ld > 100000 || steps_ld * x.size() > 200000000) return -9;
    const size_t steps = static_cast<size_t>(steps_ld);
    const long double h = static_cast<long double>(dt) / steps;
    const long double scaled = h * op.bound;
    for (auto& v : x) v /= magnitude;
    const long double tol = 2e-14L / steps;
    for (size_t st = 0; st < steps; ++st) {
        LVec sum = x, term = x;
        const long double reference = lnorm(x);
        bool converged = false;
        for (int k = 1; k <= 40; ++k) {
            term = cw_action(op, term);
            for (size_t i = 0; i < x.size(); ++i) {
                term[i] *= h / k;
                sum[i] += term[i];
            }
            if (scaled < k + 2) {
                const long double tail = lnorm(term) * scaled / (k + 1) / (1.0L - scaled / (k + 2));
                if (tail <= tol * reference) {
                    converged = true;
                    break;
                }
            }
        }
        if (!converged) return -10;
        x = std::move(sum);
    }
    for (size_t i = 0; i < x.size(); ++i) energy[i] = static_cast<double>(x[i] * magnitude * op.scale[i]);
    return 0;
}
} // namespace

extern "C" {

double orc_callaway_apply(const pmc_mode_table* m, const pmc_callaway_table* t, int ncells, int64_t n,
                          const int32_t* active, const int32_t* cell, double* occ, double dt, double energy_weight,
                          int* status) {
    // CarrierCollisionOperator::apply (src/solver/CarrierCollisionOperator.cpp:38-142):
    // each cell's carriers in index order are the quadrature; the new
    // occupations are staged and committed together.
    std::vector<std::vector<int64_t>> members(ncells);
    for (int64_t i = 0; i < n; ++i) members[cell[i]].push_back(i);
    std::vector<double> next(occ, occ + n);
    std::vector<long double> residual(ncells, 0);
    std::vector<int> err(ncells, 0);
#pragma omp parallel for schedule(dynamic, 1)
    for (int c = 0; c < ncells; ++c) {
        const auto& idx = members[c];
        if (idx.empty()) continue;
        std::vector<double> omega, cap, gn, gr, energy;
        std::vector<V3> q;
        for (int64_t i : idx) {
            const int a = active[i];
            omega.push_back(m->omega[a]);
            cap.push_back(t->capacity[a]);
            q.push_back({t->wavevector[3 * a], t->wavevector[3 * a + 1], t->wavevector[3 * a + 2]});
            gn.push_back(t->normal_rate[a]);
            gr.push_back(t->resistive_rate[a]);
            energy.push_back(6.582119569e-4 * m->omega[a] * (occ[i] - m->neq_background[a]));
        }
        if (dt == 0) continue;
        CwOp op;
        if (!cw_build(op, omega, cap, q, gn, gr)) {
            err[c] = -11;
            continue;
        }
        const int rc = cw_evolve(op, energy, dt);
        if (rc) {
            err[c] = rc;
            continue;
        }
        for (size_t j = 0; j < idx.size(); ++j) {
            const int64_t i = idx[j];
            const int a = active[i];
            const double hw = 6.582119569e-4 * m->omega[a];
            next[i] = m->neq_background[a] + energy[j] / hw;
            residual[c] += hw * (static_cast<long double>(next[i]) - occ[i]);
        }
        residual[c] *= energy_weight;
    }
    *status = 0;
    for (int e : err)
        if (e) {
            *status = e;
            return 0.0;
        }
    long double sum = 0;
    for (auto r : residual) sum += r;
    std::copy(next.begin(), next.end(), occ);
    return static_cast<double>(sum);
}

} // extern "C"

// ---------------------------------------------------------------- nonlinear Callaway
// nonlinear_callaway_evolve (src/solver/NonlinearCallawayCollision.cpp:251-285):
// three exponential BGK substeps N(dt/2) R(dt) N(dt/2), each relaxing toward the
// displaced Bose target n_BE(phi), phi = sum_j lambda_j b_j, whose a-weighted
// moments equal the carriers' (channel, :25-248).  Restated in long double, as
// the reference computes it.
namespace {
using LD = long double;

LD nl_bose(LD x) { return x > 700 ? 0 : 1 / std::expm1(x); } // :19-21

// One substep on the occupations n of one cell.  rate: the channel's rates,
// normal: energy plus crystal-momentum constraints (else energy only).
int nl_channel(std::vector<double>& n, const std::vector<double>& omega, const std::vector<V3>& q,
               const std::vector<double>& rate, double T, double dt, bool normal) {
    const size_t N = n.size();
    // weights a = 1 - exp(-rate dt) and the metric a n_eq (1 + n_eq) at the cell
    // temperature (:34-43); phi0 = hw / kT is the undisplaced equilibrium.
    std::vector<LD> a(N), metric(N), phi0(N);
    LD activity = 0;
    for (size_t i = 0; i < N; ++i) {
        a[i] = -std::expm1(-rate[i] * dt);
        activity += a[i];
        phi0[i] = kHbar * omega[i] / (kBoltz * T);
        const LD e = nl_bose(phi0[i]);
        metric[i] = a[i] * e * (1 + e);
    }
    if (activity == 0) return 0;
    // Constraint directions (energy phi0, then q_x, q_y, q_z for N), each
    // normalized, orthogonalized twice against the kept ones in the metric,
    // dropped when shorter than 1e-12 (:46-87).
    std::vector<std::vector<LD>> B;
    std::array<LD, 4> lam{};
    const int ndir = normal ? 4 : 1;
    for (int j = 0; j < ndir; ++j) {
        std::vector<LD> v(N);
        LD start = 0;
        for (size_t i = 0; i < N; ++i) {
            v[i] = j == 0 ? phi0[i] : static_cast<LD>(q[i][j - 1]);
            start += metric[i] * v[i] * v[i];
        }
        if (start == 0) continue;
        start = std::sqrt(start);
        for (auto& x : v) x /= start;
        for (int pass = 0; pass < 2; ++pass)
            for (const auto& b : B) {
                LD d = 0;
                for (size_t i = 0; i < N; ++i) d += metric[i] * v[i] * b[i];
                for (size_t i = 0; i < N; ++i) v[i] -= d * b[i];
            }
        LD len = 0;
        for (size_t i = 0; i < N; ++i) len += metric[i] * v[i] * v[i];
        len = std::sqrt(len);
        if (len < 1e-12L) continue;
        for (auto& x : v) x /= len;
        if (j == 0) lam[0] = start * len;
        B.push_back(std::move(v));
    }
    const int dim = static_cast<int>(B.size());
    if (dim == 0) return 1; // "no representable energy constraint"
    // Gradient scale per direction (:91-97).
    std::array<LD, 4> scale{};
    for (int j = 0; j < dim; ++j)
        for (size_t i = 0; i < N; ++i)
            scale[j] += std::abs(a[i] * B[j][i]) * std::max(static_cast<LD>(n[i]), nl_bose(phi0[i]));
    // Convex dual objective sum a (-log(1 - e^-phi) + n phi), its gradient
    // sum a b (n - n_BE(phi)) and Hessian sum a b b' n_BE (1 + n_BE) (:98-131);
    // infinite where some weighted phi is not positive and finite.
    std::array<LD, 4> g{};
    std::array<std::array<LD, 4>, 4> H{};
    std::vector<LD> target(N);
    auto objective = [&](const std::array<LD, 4>& x, bool hessian) -> LD {
        g = {};
        if (hessian) H = {};
        LD f = 0;
        for (size_t i = 0; i < N; ++i) {
            if (a[i] == 0) continue;
            LD phi = 0;
            for (int j = 0; j < dim; ++j) phi += x[j] * B[j][i];
            if (!(phi > 0) || !std::isfinite(phi)) return std::numeric_limits<LD>::infinity();
            const LD e = nl_bose(phi);
            if (hessian) target[i] = e;
            f += a[i] * (-std::log(-std::expm1(-phi)) + n[i] * phi);
            for (int j = 0; j < dim; ++j) {
                g[j] += a[i] * B[j][i] * (n[i] - e);
                if (hessian)
                    for (int k = 0; k < dim; ++k) H[j][k] += a[i] * B[j][i] * B[k][i] * e * (1 + e);
            }
        }
        return f;
    };
    auto gradient_error = [&]() {
        LD err = 0;
        for (int j = 0; j < dim; ++j) err = std::max(err, std::abs(g[j]) / std::max(1e-30L, scale[j]));
        return err;
    };
    // Damped Newton: 70 iterations at most, converged when every scaled
    // gradient component is below 2e-13 (:132-228).
    bool converged = false;
    for (int it = 0; it < 70 && !converged; ++it) {
        const LD f0 = objective(lam, true);
        const LD err = gradient_error();
        if (err < 2e-13L) {
            converged = true;
            break;
        }
        // H d = -g by Gaussian elimination with partial pivoting.
        std::array<LD, 4> d{};
        for (int j = 0; j < dim; ++j) d[j] = -g[j];
        auto M = H;
        for (int col = 0; col < dim; ++col) {
            int piv = col;
            for (int r = col + 1; r < dim; ++r)
                if (std::abs(M[r][col]) > std::abs(M[piv][col])) piv = r;
            if (!(std::abs(M[piv][col]) > 1e-30L)) return 2; // "Singular ... moment Jacobian"
            std::swap(M[col], M[piv]);
            std::swap(d[col], d[piv]);
            for (int r = col + 1; r < dim; ++r) {
                const LD f = M[r][col] / M[col][col];
                for (int k = col; k < dim; ++k) M[r][k] -= f * M[col][k];
                d[r] -= f * d[col];
            }
        }
        for (int r = dim - 1; r >= 0; --r) {
            for (int k = r + 1; k < dim; ++k) d[r] -= M[r][k] * d[k];
            d[r] /= M[r][r];
        }
        LD slope = 0;
        for (int j = 0; j < dim; ++j) slope += g[j] * d[j];
        // Backtracking: accept sufficient decrease or a 10 % smaller gradient error.
        bool accepted = false;
        for (LD t = 1; t > 1e-14L; t *= .5L) {
            std::array<LD, 4> trial = lam;
            for (int j = 0; j < dim; ++j) trial[j] += t * d[j];
            const LD f1 = objective(trial, false);
            const LD e1 = gradient_error();
            if (std::isfinite(f1) && (f1 <= f0 + 1e-4L * t * slope || e1 < err * .9L)) {
                lam = trial;
                accepted = true;
                break;
            }
        }
        if (!accepted) return 3; // "moment line search failed"
    }
    if (!converged) return 4; // "moment solve did not converge"
    for (size_t i = 0; i < N; ++i)
        if (a[i] > 0) {
            n[i] = static_cast<double>((1 - a[i]) * n[i] + a[i] * target[i]);
            if (!std::isfinite(n[i]) || n[i] < 0) return 5; // "occupation is invalid"
        }
    return 0;
}
} // namespace

extern "C" {

int orc_nl_callaway_evolve(int n, const double* omega, const double* wavevector, const double* normal_rate,
                           const double* resistive_rate, const double* occupation, double T, double dt, double* out) {
    // Argument checks of :254-275.
    if (n <= 0 || !std::isfinite(T) || T <= 0 || !std::isfinite(dt) || dt < 0) return 6;
    std::vector<double> w(omega, omega + n), gn(normal_rate, normal_rate + n), gr(resistive_rate, resistive_rate + n);
    std::vector<double> occ(occupation, occupation + n);
    std::vector<V3> q(n);
    for (int i = 0; i < n; ++i) {
        if (!std::isfinite(occ[i]) || occ[i] < 0 || !std::isfinite(w[i]) || w[i] <= 0 || !std::isfinite(gn[i]) ||
            gn[i] < 0 || !std::isfinite(gr[i]) || gr[i] < 0)
            return 6;
        for (int k = 0; k < 3; ++k) {
            q[i][k] = wavevector[3 * i + k];
            if (!std::isfinite(q[i][k])) return 6;
        }
    }
    if (dt > 0) {
        int rc = nl_channel(occ, w, q, gn, T, dt * .5, true);
        if (!rc) rc = nl_channel(occ, w, q, gr, T, dt, false);
        if (!rc) rc = nl_channel(occ, w, q, gn, T, dt * .5, true);
        if (rc) return rc;
    }
    std::copy(occ.begin(), occ.end(), out);
    return 0;
}

int orc_nl_callaway_rates(const pmc_mode_table* m, const pmc_callaway_rates* r, double T, int active, double* normal,
                          double* resistive) {
    // PhononMaterial::mode_callaway_rates (src/material/PhononMaterialCollisionData.cpp:167-196):
    // 4 pi times the linear interpolation between the bracketing samples (no clamping).
    const int nt = r->n_temps;
    if (!std::isfinite(T) || T <= 0 || nt <= 0 || T < r->temps[0] || T > r->temps[nt - 1]) return 7;
    const int hi = static_cast<int>(std::lower_bound(r->temps, r->temps + nt, T) - r->temps);
    const int lo = hi ? hi - 1 : hi;
    const double f = lo == hi ? 0 : (T - r->temps[lo]) / (r->temps[hi] - r->temps[lo]);
    const size_t na = static_cast<size_t>(m->n_active), a = static_cast<size_t>(active);
    auto rate = [&](const double* t) { return 4 * kPi * ((1 - f) * t[lo * na + a] + f * t[hi * na + a]); };
    *normal = rate(r->gamma_normal);
    *resistive = rate(r->gamma_umklapp) + rate(r->gamma_extra);
    return 0;
}

double orc_nl_callaway_apply(const pmc_mode_table* m, const pmc_callaway_rates* r, int ncells,
                             const double* cell_temperature, int lifetime_local, double lifetime_temperature,
                             int64_t n, const int32_t* active, const int32_t* cell, double* occ, double dt,
                             double energy_weight, int* status) {
    // NonlinearCallawayCollision::apply (src/solver/NonlinearCallawayCollision.cpp:287-344):
    // each cell's carriers in index order, rates at the cell temperature or the
    // fixed lifetime temperature, new occupations committed together, residual
    // sum hw (n' - n) times the energy weight.
    std::vector<std::vector<int64_t>> members(ncells);
    for (int64_t i = 0; i < n; ++i) members[cell[i]].push_back(i);
    std::vector<double> next(occ, occ + n);
    std::vector<LD> residual(ncells, 0);
    std::vector<int> err(ncells, 0);
#pragma omp parallel for schedule(dynamic, 1)
    for (int c = 0; c < ncells; ++c) {
        const auto& idx = members[c];
        if (idx.empty()) continue;
        const double T = cell_temperature[c], rT = lifetime_local ? T : lifetime_temperature;
        const size_t k = idx.size();
        std::vector<double> w(k), q(3 * k), gn(k), gr(k), n0(k), out(k);
        for (size_t j = 0; j < k; ++j) {
            const int a = active[idx[j]];
            if (orc_nl_callaway_rates(m, r, rT, a, &gn[j], &gr[j])) {
                err[c] = 7;
                break;
            }
            w[j] = m->omega[a];
            for (int d = 0; d < 3; ++d) q[3 * j + d] = r->wavevector[3 * a + d];
            n0[j] = occ[idx[j]];
        }
        if (err[c]) continue;
        const int rc = orc_nl_callaway_evolve(static_cast<int>(k), w.data(), q.data(), gn.data(), gr.data(),
                                              n0.data(), T, dt, out.data());
        if (rc) {
            err[c] = rc;
            continue;
        }
        for (size_t j = 0; j < k; ++j) {
            next[idx[j]] = out[j];
            residual[c] += kHbar * w[j] * (static_cast<LD>(out[j]) - n0[j]);
        }
        residual[c] *= energy_weight;
    }
    *status = 0;
    for (int e : err)
        if (e) {
            *status = e;
            return 0.0;
        }
    LD sum = 0;
    for (auto x : residual) sum += x;
    std::copy(next.begin(), next.end(), occ);
    return static_cast<double>(sum);
}

} // extern "C"

// ---------------------------------------------------------------- whole-step simulator
struct orc_sim {
    pmc_setup setup;
    pmc_mode_table m;
    pmc_geometry g;
    std::vector<pmc_facet> facets;
    std::vector<pmc_rough_table> rough;
    std::vector<int32_t> layer_rough; // [layers * facets] table by (material, facet), empty: facets' own
    std::vector<pmc_reservoir> res;
    uint32_t k0 = 0, k1 = 0;
    int step = 0;
    // carriers
    std::vector<V3> pos, vel, cpoint;
    std::vector<int> mode, cell, cfacet;
    std::vector<double> occ, ctime;
    std::vector<uint32_t> uid;
    std::vector<char> alive;
    uint32_t next_uid = 0;
    // cells
    std::vector<double> cell_e, cell_T, cell_flux;
    // ledgers
    std::vector<int> previous_absorbed;
    pmc_ledger ledger{};
    std::string error;
    // prescribed heat source (thermal spectrum)
    bool hs_on = false;
    pmc_heat_source hs{};
    std::vector<double> hs_rate, hs_pending;
    double source_time = 0.0;
    std::vector<int> hs_elig;
    std::vector<double> hs_weight; // spectrum weights per active mode (weighted spectrum)
    double elapsed = 0.0;
    // layered materials (DMM interfaces)
    int layers = 1, layer_axis = 0;
    double band = 0.0;
    std::vector<double> planes, e_bg_mat, accept, bank_omega, bank_prefix;
    std::vector<double> cbg_mat; // linearized transport: C_bg of each material
    std::vector<int> mode_begin, cell_mat, lut_begin, bank_begin, bank_mode;
    // acoustic / mixed mismatch (AMM rows in global active indices)
    double amm_fraction = 0.0;
    std::vector<int> amm_block, amm_refl, amm_ebeg, amm_emode;
    std::vector<double> amm_eprob;
    // Transport policy (include/SimulationConfig.h:148-156): conservative walls
    // (strict-elastic tables on every reflecting facet, signed occupations) and
    // linearized occupations; the collision operator is RTA, linearized
    // Callaway (callaway) or nonlinear Callaway (nl_callaway).
    bool conservative = false, linearized = false;
    bool callaway = false;
    pmc_callaway_table cw{};
    std::vector<double> cw_cap, cw_q, cw_gn, cw_gr;
    bool nl_callaway = false;
    pmc_callaway_rates nl{};
    std::vector<double> nl_t, nl_gn, nl_gu, nl_ge, nl_q;
};

namespace {
V3 table_velocity(const orc_sim& s, int a) {
    return {s.m.velocity[3 * a], s.m.velocity[3 * a + 1], s.m.velocity[3 * a + 2]};
}

// TransportOperator::update_collision_cache_single (src/solver/TransportOperator.cpp:186-250)
// for the analytic box: nudge candidates along the nearest facet normal and the velocity.
bool update_cache(orc_sim& s, int64_t i) {
    const double speed = norm(s.vel[i]);
    if (std::isfinite(speed) && speed <= 1e-18) {
        s.cpoint[i] = s.pos[i];
        s.cfacet[i] = -1;
        s.ctime[i] = std::numeric_limits<double>::infinity();
        return true;
    }
    const bool mesh = s.g.kind == 1;
    auto trace = [&](const double* x, double* cp, double* t) {
        return mesh ? orc_trace_mesh(&s.g, x, s.vel[i].data(), cp, t) : orc_trace_box(&s.g, x, s.vel[i].data(), cp, t);
    };
    double cp[3], t = 0;
    int f = trace(s.pos[i].data(), cp, &t);
    if (f < 0 || !std::isfinite(t)) {
        const V3 original = s.pos[i];
        double ext = 1.0;
        for (int a = 0; a < 3; ++a) ext = std::max(ext, std::abs(s.g.bmax[a] - s.g.bmin[a]));
        const double eps = std::max(s.setup.push_distance_a, 1e-10 * ext);
        int nearest = -1;
        if (mesh) {
            nearest = orc_nearest_facet(&s.g, original.data());
        } else { // nearest box plane
            int best = 0;
            double bd = std::numeric_limits<double>::infinity();
            for (int p = 0; p < 6; ++p) {
                const int a = p / 2;
                const double d = std::abs(original[a] - (p % 2 ? s.g.bmax[a] : s.g.bmin[a]));
                if (d < bd) {
                    bd = d;
                    best = p;
                }
            }
            nearest = s.g.box_plane_facet[best];
        }
        const V3 dir = mul(s.vel[i], 1.0 / speed);
        V3 cands[4] = {original, original, add(original, mul(dir, eps)), add(original, mul(dir, -eps))};
        if (nearest >= 0) {
            const pmc_facet& F = s.facets[nearest];
            const V3 n{F.normal[0], F.normal[1], F.normal[2]};
            cands[0] = add(original, mul(n, eps));
            cands[1] = add(original, mul(n, -eps));
        }
        for (const V3& c : cands) {
            bool inside = true;
            if (mesh) {
                inside = orc_contains_mesh(&s.g, c.data()) != 0;
            } else {
                for (int a = 0; a < 3; ++a) inside = inside && c[a] >= s.g.bmin[a] && c[a] <= s.g.bmax[a];
            }
            if (!inside) continue;
            f = trace(c.data(), cp, &t);
            if (f >= 0 && std::isfinite(t)) {
                s.pos[i] = c;
                break;
            }
        }
    }
    if (f < 0 || !std::isfinite(t)) {
        s.cfacet[i] = -1;
        s.ctime[i] = std::numeric_limits<double>::infinity();
        return false;
    }
    s.cpoint[i] = {cp[0], cp[1], cp[2]};
    s.cfacet[i] = f;
    s.ctime[i] = t / s.setup.time_step_ps;
    return true;
}

// advance() result of a failed discrete-mode reflection (conservative transport).
constexpr int kOrcModeMapping = 64;

struct StepTally {
    std::vector<int> absorbed_by_res;
    double absorbed_e = 0;
    long long absorbed = 0;
    long long rough_events = 0, rough_spec = 0, rough_diff = 0;
    long long iface_events = 0, iface_transmitted = 0;
};

// TransportOperator::advance_particle (src/solver/TransportOperator.cpp:405-466) with
// process_boundary_collision (:287-362) and interact_boundary
// (src/solver/BoundaryInteraction.cpp:53-162), non-conservative policy.
int layer_of_mode(const orc_sim& s, int mode) {
    int k = 0;
    while (k + 1 < s.layers && mode >= s.mode_begin[k + 1]) ++k;
    return k;
}

// InterfaceModeSampler::sample_band (src/solver/InterfaceModeSampler.cpp:72-109)
// with the uniform u for the flux draw.
int sample_band(const orc_sim& s, int material, int direction, double w, double u, double& flux) {
    const int b = 2 * material + direction;
    const int b0 = s.bank_begin[b], n = s.bank_begin[b + 1] - b0;
    const double* om = s.bank_omega.data() + b0;
    const double* pre = s.bank_prefix.data() + b0 + b;
    const double bin = std::floor(w / s.band);
    const double lo = bin * s.band, hi = (bin + 1) * s.band;
    const int begin = static_cast<int>(std::lower_bound(om, om + n, lo) - om);
    const int end = static_cast<int>(std::lower_bound(om, om + n, hi) - om);
    flux = end > begin ? pre[end] - pre[begin] : 0.0;
    if (!(flux > 0.0)) return -1;
    const double target = pre[begin] + u * flux;
    int pos = static_cast<int>(std::upper_bound(pre + begin + 1, pre + end + 1, target) - pre);
    pos = std::clamp(pos, begin + 1, end) - 1;
    return s.bank_mode[b0 + pos];
}

// InterfaceInteraction::interact, DMM branch (src/solver/InterfaceInteraction.cpp:47-153),
// on three uniforms: reflection band draw, transmission band draw, transmission
// test.  Returns 1 transmitted, 0 reflected into a new mode, -1 mirrored (same
// mode, velocity reversed along the axis).
// The energy-preserving remap into the selected mode, or a mirror along the
// axis when there is none or its occupation is invalid
// (src/solver/InterfaceInteraction.cpp:103-152, nonlinear occupations).
int interface_outcome(const orc_sim& s, int a, const V3& vel, double occ, int sel, bool transmit, int& out,
                      double& nocc, V3& nvel) {
    const int ax = s.layer_axis;
    const double w = s.m.omega[a];
    out = a;
    nocc = occ;
    nvel = vel;
    // Linearized transport (:100-105, :124-133): no outgoing mode, or one that does
    // not leave in the expected direction along the axis, fails (-3); signed
    // occupations are kept.
    if (s.linearized) {
        if (sel < 0) return -3;
        const double W = s.m.energy_weight;
        const double n = s.m.neq_background[sel] +
                         (W * w * (occ - s.m.neq_background[a])) / std::max(1e-300, W * s.m.omega[sel]);
        const double vs = table_velocity(s, sel)[ax];
        const bool out_up = transmit ? vel[ax] > 0.0 : !(vel[ax] > 0.0);
        if (!std::isfinite(n) || !std::isfinite(vs) || (out_up ? !(vs > 0.0) : !(vs < 0.0))) return -3;
        out = sel;
        nocc = n;
        nvel = table_velocity(s, sel);
        return transmit ? 1 : 0;
    }
    if (sel >= 0) {
        const double W = s.m.energy_weight;
        const double dev = occ - s.m.neq_background[a];
        const double n = s.m.neq_background[sel] + (W * w * dev) / std::max(1e-300, W * s.m.omega[sel]);
        if (std::isfinite(n) && n >= 0.0) {
            out = sel;
            nocc = n;
            nvel = table_velocity(s, sel);
            return transmit ? 1 : 0;
        }
    }
    nvel[ax] = -nvel[ax];
    return -1;
}

// Diffuse mismatch (:86-101): a band draw in the reflection bank of this
// material (u[0]), one in the transmission bank of the next (u[1]), the choice
// by the flux ratio (u[2]).
int dmm_core(const orc_sim& s, int a, const V3& vel, double occ, int dest, const double u[3], int& out, double& nocc,
             V3& nvel) {
    const int ax = s.layer_axis;
    const int mat = layer_of_mode(s, a);
    const double w = s.m.omega[a];
    const int inc = vel[ax] > 0.0 ? 1 : 0;
    double fr = 0.0, ft = 0.0;
    const int mr = sample_band(s, mat, 1 - inc, w, u[0], fr);
    const int mt = sample_band(s, dest, inc, w, u[1], ft);
    const double tr = ft > 0 ? ft / std::max(1e-300, fr + ft) : 0.0;
    const bool transmit = mt >= 0 && u[2] < tr;
    return interface_outcome(s, a, vel, occ, transmit ? mt : mr, transmit, out, nocc, nvel);
}

// InterfaceInteraction::interact (src/solver/InterfaceInteraction.cpp:46-153) on
// the draws u in the order they are made: the channel (mixed mismatch only),
// then the AMM row draw (AcousticInterfaceSampler::sample,
// src/solver/AcousticInterfaceSampler.cpp:457-471) or the three DMM draws.  A
// tangent carrier mirrors before any AMM draw.  Returns 1 transmitted, 0
// reflected into a new mode, -1 mirrored, -2 no AMM row for the mode in this
// direction ("AMM incoming state is tangent or unsupported.").
int interface_core(const orc_sim& s, int a, const V3& vel, double occ, int dest, const double u[4], int& out,
                   double& nocc, V3& nvel) {
    const bool mixed = s.amm_fraction > 0.0 && s.amm_fraction < 1.0;
    const bool amm = s.amm_fraction >= 1.0 || (mixed && u[0] < s.amm_fraction);
    const double* r = u + (mixed ? 1 : 0);
    if (!amm) return dmm_core(s, a, vel, occ, dest, r, out, nocc, nvel);
    const int ax = s.layer_axis, mat = layer_of_mode(s, a);
    out = a;
    nocc = occ;
    nvel = vel;
    if (std::abs(vel[ax]) < 1e-7) {
        nvel[ax] = -nvel[ax];
        return -1;
    }
    const int face = std::min(mat, dest), side = dest > mat ? 0 : 1;
    const int row = s.amm_block[2 * face + side] + (a - s.mode_begin[mat]);
    int sel = s.amm_refl[row];
    if (sel < 0) return -2;
    bool transmit = false;
    double x = r[0];
    for (int e = s.amm_ebeg[row]; e < s.amm_ebeg[row + 1]; ++e) {
        x -= s.amm_eprob[e];
        if (x < 0) {
            sel = s.amm_emode[e];
            transmit = true;
            break;
        }
    }
    return interface_outcome(s, a, vel, occ, sel, transmit, out, nocc, nvel);
}

bool interface_event(orc_sim& s, int64_t i, int dest, Uniforms& U, StepTally& tally) {
    const double u[4] = {U.next(), U.next(), U.next(), U.next()};
    int out = 0;
    double nocc = 0.0;
    V3 nvel{};
    const int action = interface_core(s, s.mode[i], s.vel[i], s.occ[i], dest, u, out, nocc, nvel);
    if (action == -2 || action == -3) return false;
    tally.iface_events++;
    if (action == 1) tally.iface_transmitted++;
    s.mode[i] = out;
    s.occ[i] = nocc;
    s.vel[i] = nvel;
    return true;
}

int advance(orc_sim& s, int64_t i, double remaining, StepTally& tally) {
    const double dt = s.setup.time_step_ps;
    const double push = s.setup.push_distance_a;
    Uniforms U;
    U.uid = s.uid[i];
    U.step = static_cast<uint32_t>(s.step);
    U.k0 = s.k0;
    U.k1 = s.k1;
    for (int ev = 0; remaining > 1e-14 && ev < 64; ++ev) {
        const double t_wall = s.ctime[i] * dt;
        // LayeredMaterialLayout::next_crossing (src/solver/LayeredMaterialLayout.cpp:33-60).
        double t_cross = std::numeric_limits<double>::infinity();
        int dest = -1;
        if (s.layers > 1) {
            const int mat = layer_of_mode(s, s.mode[i]);
            const double va = s.vel[i][s.layer_axis];
            double plane = 0.0;
            if (std::abs(va) > 1e-18) {
                if (va > 0.0 && mat + 1 < s.layers) {
                    plane = s.planes[mat];
                    dest = mat + 1;
                } else if (va < 0.0 && mat > 0) {
                    plane = s.planes[mat - 1];
                    dest = mat - 1;
                }
            }
            if (dest >= 0) {
                const double t = (plane - s.pos[i][s.layer_axis]) / va;
                if (std::isfinite(t) && t > 1e-14) t_cross = t;
                else dest = -1;
            }
        }
        const double t_event = std::min(t_wall, t_cross);
        if (!std::isfinite(t_event) || t_event > remaining) {
            s.pos[i] = add(s.pos[i], mul(s.vel[i], remaining));
            if (std::isfinite(s.ctime[i])) s.ctime[i] -= remaining / dt;
            return 0;
        }
        // The outer wall wins exact ties (TransportOperator.cpp:429-447).
        if (t_cross < t_wall) {
            remaining -= std::max(0.0, t_cross);
            s.pos[i] = add(s.pos[i], mul(s.vel[i], t_cross));
            U.ev = static_cast<uint32_t>(ev);
            U.used = 0;
            if (!interface_event(s, i, dest, U, tally)) return -PMC_ERR_BOUNDARY;
            s.pos[i] = add(s.pos[i], mul(s.vel[i], push / std::max(norm(s.vel[i]), 1e-12)));
            if (!update_cache(s, i)) return -PMC_ERR_BOUNDARY;
            continue;
        }
        remaining -= std::max(0.0, t_wall);
        const int facet = s.cfacet[i];
        s.pos[i] = s.cpoint[i];
        const pmc_facet& F = s.facets[facet];
        const V3 n{F.normal[0], F.normal[1], F.normal[2]};
        const int a = s.mode[i];
        if (F.condition == 'T') {
            const double dE = kHbar * s.m.omega[a] * (s.occ[i] - s.m.neq_background[a]);
            tally.absorbed_e += dE * s.m.energy_weight;
            tally.absorbed++;
            if (F.reservoir >= 0) tally.absorbed_by_res[F.reservoir]++;
            s.alive[i] = 0;
            return 1;
        }
        // Layered materials reflect through the table of the carrier's material
        // (RoughBoundarySampler::table_index, src/solver/TransportOperator.cpp:302).
        const int rtab = s.layer_rough.empty()
                             ? F.rough_table
                             : s.layer_rough[static_cast<size_t>(layer_of_mode(s, a)) * s.facets.size() + facet];
        if (F.condition == 'P' && F.periodic_pair >= 0) {
            s.pos[i] = add(s.pos[i], V3{F.shift[0], F.shift[1], F.shift[2]});
        } else if (rtab >= 0 && (F.condition == 'R' || s.conservative)) {
            // Conservative transport reflects every reflecting facet through its
            // table with signed occupations and needs an inward outgoing mode
            // (src/solver/BoundaryInteraction.cpp:90-113).
            U.ev = static_cast<uint32_t>(ev);
            U.used = 0;
            double u[4];
            for (int k = 0; k < 4; ++k) u[k] = U.next();
            int out, kind;
            double o;
            orc_rough_sample_signed(&s.m, &s.rough[rtab], a, s.occ[i], u, &out, &o, &kind,
                                    s.conservative ? 1 : 0);
            tally.rough_events++;
            if (kind == 1) tally.rough_spec++;
            if (kind == 2) tally.rough_diff++;
            V3 v = table_velocity(s, out);
            if (s.conservative) {
                if (!std::isfinite(o) || !(s.m.omega[out] > 0.0) || !(dot(v, n) < 0.0)) return -kOrcModeMapping;
            } else if (dot(v, n) > 0.0) {
                v = sub(v, mul(n, 2.0 * dot(v, n)));
            }
            s.mode[i] = out;
            s.vel[i] = v;
            s.occ[i] = o;
        } else if (s.conservative) {
            return -kOrcModeMapping;
        } else if (F.condition == 'R') {
            return -PMC_ERR_UNSUPPORTED;
        } else {
            s.vel[i] = sub(s.vel[i], mul(n, 2.0 * dot(s.vel[i], n)));
        }
        s.pos[i] = add(s.pos[i], mul(s.vel[i], push / std::max(norm(s.vel[i]), 1e-12)));
        if (!update_cache(s, i)) return -PMC_ERR_BOUNDARY;
        if (s.ctime[i] * dt < 1e-12) {
            s.pos[i] = add(s.pos[i], mul(s.vel[i], 1e-12));
            if (!update_cache(s, i)) return -PMC_ERR_BOUNDARY;
        }
    }
    return remaining > 1e-14 ? -PMC_ERR_EXCESS_EVENTS : 0;
}

void compact(orc_sim& s) {
    size_t w = 0;
    for (size_t r = 0; r < s.alive.size(); ++r) {
        if (!s.alive[r]) continue;
        if (w != r) {
            s.pos[w] = s.pos[r];
            s.vel[w] = s.vel[r];
            s.cpoint[w] = s.cpoint[r];
            s.mode[w] = s.mode[r];
            s.cell[w] = s.cell[r];
            s.cfacet[w] = s.cfacet[r];
            s.occ[w] = s.occ[r];
            s.ctime[w] = s.ctime[r];
            s.uid[w] = s.uid[r];
        }
        ++w;
    }
    for (auto* v : {&s.pos, &s.vel, &s.cpoint}) v->resize(w);
    for (auto* v : {&s.mode, &s.cell, &s.cfacet}) v->resize(w);
    for (auto* v : {&s.occ, &s.ctime}) v->resize(w);
    s.uid.resize(w);
    s.alive.assign(w, 1);
}

void refresh(orc_sim& s) {
    const int nc = s.g.n_cells;
    std::vector<int32_t> act(s.mode.begin(), s.mode.end());
    orc_refresh(&s.m, nc, s.g.cell_volume, static_cast<int64_t>(s.mode.size()), act.data(), s.cell.data(),
                s.occ.data(), s.cell_e.data(), s.cell_T.data());
    if (s.layers > 1) { // each cell with its material's background and lookup table
        std::vector<double> e(nc, 0.0);
        for (size_t i = 0; i < s.mode.size(); ++i) {
            const int a = s.mode[i];
            e[s.cell[i]] += kHbar * s.m.omega[a] * (s.occ[i] - s.m.neq_background[a]) * s.m.energy_weight;
        }
        for (int c = 0; c < nc; ++c) {
            const int k = s.cell_mat[c];
            s.cell_e[c] = e[c] / s.g.cell_volume[c] + s.e_bg_mat[k];
            s.cell_T[c] = interp(s.m.lut_e + s.lut_begin[k], s.m.lut_t + s.lut_begin[k],
                                 s.lut_begin[k + 1] - s.lut_begin[k], s.cell_e[c]);
        }
    }
    if (s.linearized) // T_bg + (e - E_bg) / C_bg (src/solver/GridStatistics.cpp:193-225), by material
        for (int c = 0; c < nc; ++c) {
            const bool lay = s.layers > 1;
            const double ebg = lay ? s.e_bg_mat[s.cell_mat[c]] : s.cw.background_energy_density;
            const double cbg = lay ? s.cbg_mat[s.cell_mat[c]] : s.cw.heat_capacity_density;
            s.cell_T[c] = s.setup.background_temperature + (s.cell_e[c] - ebg) / cbg;
        }
}

double thermal_energy(const orc_sim& s) {
    double total = 0.0;
    for (int c = 0; c < s.g.n_cells; ++c) total += (s.cell_e[c] - s.m.zero_point_density) * s.g.cell_volume[c];
    return total;
}

template <class T>
void set_error(orc_sim* s, const T& msg) {
    s->error = msg;
}

// MonteCarloSolver::local_heat_source_integrated_time_factor (src/solver/HeatSourceDriving.cpp:14-46).
double source_time_factor(const pmc_heat_source& h, double t0, double t1) {
    const double b = std::max(t0, h.start_ps);
    const double e = h.end_ps >= 0.0 ? std::min(t1, h.end_ps) : t1;
    if (e <= b) return 0.0;
    if (!h.square) return h.amplitude * (e - b);
    const double on = h.on_duration_ps < 0.0 ? h.duty_cycle * h.period_ps : h.on_duration_ps;
    if (on <= 0.0) return 0.0;
    auto on_time = [&](double rel) {
        if (!(rel > 0.0)) return 0.0;
        const double cycles = std::floor(rel / h.period_ps);
        return cycles * on + std::min(rel - cycles * h.period_ps, on);
    };
    return h.amplitude * std::max(0.0, on_time(e - h.start_ps) - on_time(b - h.start_ps));
}

// One heat-source application: PrescribedHeatSource::accrue
// (src/solver/PrescribedHeatSource.cpp:232-252) then orc_source_deposit.
int apply_source(orc_sim& s, double factor, double* energy_out, int64_t* updates_out) {
    *energy_out = 0.0;
    *updates_out = 0;
    s.source_time += factor; // every cell's prescribed energy is its rate times this sum
    double pending_total = 0.0;
    for (double p : s.hs_pending) pending_total += p;
    if (factor <= 0.0 && pending_total <= 0.0) return 0;
    const int nc = s.g.n_cells;
    for (int c = 0; c < nc; ++c) s.hs_pending[c] += s.hs_rate[c] * factor;
    std::vector<int32_t> act(s.mode.begin(), s.mode.end());
    return orc_source_deposit_spectrum(&s.m, nc, s.hs_pending.data(), s.cell_T.data(), s.setup.background_temperature,
                                       s.hs_elig.data(), static_cast<int64_t>(act.size()), act.data(), s.cell.data(),
                                       s.occ.data(), s.linearized ? s.cw_cap.data() : nullptr, s.hs.spectrum,
                                       s.hs_weight.empty() ? nullptr : s.hs_weight.data(), s.hs.frequency_center_thz,
                                       s.hs.frequency_sigma_thz, energy_out, updates_out);
}
} // namespace

extern "C" {

void orc_set_threads(int threads) {
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#else
    (void)threads;
#endif
}

int orc_create(const pmc_setup* setup, const pmc_mode_table* modes, const pmc_geometry* geometry, int32_t n_rough,
               const pmc_rough_table* rough, int32_t n_res, const pmc_reservoir* res, orc_sim** out) {
    if (!setup || !modes || !geometry || !out) return PMC_ERR_ARGUMENT;
    if (geometry->kind != 0 && !(geometry->kind == 1 && geometry->mesh)) return PMC_ERR_ARGUMENT;
    auto* s = new orc_sim();
    s->setup = *setup;
    s->m = *modes;
    s->g = *geometry;
    s->facets.assign(geometry->facets, geometry->facets + geometry->n_facets);
    s->g.facets = s->facets.data();
    s->rough.assign(rough, rough + n_rough);
    s->res.assign(res, res + n_res);
    s->conservative = setup->conservative_walls != 0;
    s->k0 = static_cast<uint32_t>(setup->seed & 0xffffffffu);
    s->k1 = static_cast<uint32_t>(setup->seed >> 32) ^ (setup->instance * 0x9E3779B9u);
    s->cell_e.assign(geometry->n_cells, 0.0);
    s->cell_T.assign(geometry->n_cells, 0.0);
    s->cell_flux.assign(3 * static_cast<size_t>(geometry->n_cells), 0.0);
    s->previous_absorbed.resize(n_res);
    for (int r = 0; r < n_res; ++r) s->previous_absorbed[r] = res[r].initial_count;
    *out = s;
    return PMC_OK;
}

static int finish_population(orc_sim* s) {
    const size_t n = s->mode.size();
    s->vel.resize(n);
    s->cpoint.resize(n);
    s->cfacet.assign(n, -1);
    s->ctime.assign(n, std::numeric_limits<double>::infinity());
    s->alive.assign(n, 1);
    s->cell.resize(n);
    for (size_t i = 0; i < n; ++i) {
        s->vel[i] = table_velocity(*s, s->mode[i]);
        s->cell[i] = orc_grid_index(&s->g, s->pos[i].data());
    }
    bool ok = true;
#pragma omp parallel for reduction(&& : ok)
    for (int64_t i = 0; i < static_cast<int64_t>(n); ++i) ok = update_cache(*s, i) && ok;
    if (!ok) {
        set_error(s, "Boundary intersection failed while updating collision cache.");
        return PMC_ERR_BOUNDARY;
    }
    s->next_uid = static_cast<uint32_t>(n);
    refresh(*s);
    s->ledger = pmc_ledger{};
    s->ledger.thermal_energy_ev = thermal_energy(*s);
    s->ledger.live_carriers = static_cast<int64_t>(n);
    s->ledger.slots = static_cast<int64_t>(n);
    return PMC_OK;
}

int orc_initialize(orc_sim* s, const pmc_init* init) {
    // Same draws as the device's initialization kernel.
    const int64_t n = init->count;
    s->pos.resize(n);
    s->mode.resize(n);
    s->occ.resize(n);
    s->uid.resize(n);
    const int na = s->m.n_active;
    for (int64_t i = 0; i < n; ++i) {
        uint32_t a[4], b[4];
        orc_philox(static_cast<uint32_t>(i), 0, 0, 3u, s->k0, s->k1, a);
        orc_philox(static_cast<uint32_t>(i), 0, 1, 3u, s->k0, s->k1, b);
        const double u[4] = {orc_u53(a[0], a[1]), orc_u53(a[2], a[3]), orc_u53(b[0], b[1]), orc_u53(b[2], b[3])};
        V3 p;
        for (int k = 0; k < 3; ++k) p[k] = s->g.bmin[k] + (s->g.bmax[k] - s->g.bmin[k]) * u[k];
        double u_mode = u[3];
        if (s->g.kind == 1 && orc_sample_volume(&s->g, s->k0, s->k1, static_cast<uint32_t>(i), p.data(), &u_mode)) {
            set_error(s, "sample_volume_points_naive: insufficient accepted samples.");
            return PMC_ERR_BOUNDARY;
        }
        int layer = 0;
        if (s->layers > 1) {
            // Density-weighted rejection (src/solver/Initialization.cpp:49-66), the
            // device's Philox words: (carrier, trial + 1, 0 and 1).
            bool accepted = false;
            for (uint32_t trial = 0; trial < 64 && !accepted; ++trial) {
                uint32_t ta[4], tb[4];
                orc_philox(static_cast<uint32_t>(i), 1 + trial, 0, 3u, s->k0, s->k1, ta);
                orc_philox(static_cast<uint32_t>(i), 1 + trial, 1, 3u, s->k0, s->k1, tb);
                const double q[3] = {orc_u53(ta[0], ta[1]), orc_u53(ta[2], ta[3]), orc_u53(tb[0], tb[1])};
                for (int k = 0; k < 3; ++k) p[k] = s->g.bmin[k] + (s->g.bmax[k] - s->g.bmin[k]) * q[k];
                const int ax = s->layer_axis;
                const float pa = std::min(std::max(static_cast<float>(p[ax]), static_cast<float>(s->g.bmin[ax])),
                                          static_cast<float>(s->g.bmax[ax]));
                layer = 0;
                while (layer + 1 < s->layers && pa >= static_cast<float>(s->planes[layer])) ++layer;
                accepted = orc_u53(tb[2], tb[3]) < s->accept[layer];
            }
            if (!accepted) {
                set_error(s, "sample_volume_points_naive: insufficient accepted samples.");
                return PMC_ERR_BOUNDARY;
            }
        }
        int m = std::min(static_cast<int>(u_mode * na), na - 1);
        if (s->layers > 1) {
            const int b0 = s->mode_begin[layer], nm = s->mode_begin[layer + 1] - b0;
            m = b0 + std::min(static_cast<int>(u_mode * nm), nm - 1);
        }
        double T = init->uniform_temperature;
        if (init->temperature_mode == 1) {
            const V3 ax{init->hot_point[0] - init->cold_point[0], init->hot_point[1] - init->cold_point[1],
                        init->hot_point[2] - init->cold_point[2]};
            const V3 rel{p[0] - init->cold_point[0], p[1] - init->cold_point[1], p[2] - init->cold_point[2]};
            const double sc = std::clamp(dot(rel, ax) / dot(ax, ax), 0.0, 1.0);
            T = init->cold_temperature + sc * (init->hot_temperature - init->cold_temperature);
        }
        V3 pf;
        for (int k = 0; k < 3; ++k)
            pf[k] = std::clamp(static_cast<double>(static_cast<float>(p[k])), static_cast<double>(static_cast<float>(s->g.bmin[k])),
                               static_cast<double>(static_cast<float>(s->g.bmax[k])));
        if (init->temperature_mode == 2) T = init->cell_temperature[orc_grid_index(&s->g, pf.data())];
        s->pos[i] = pf;
        s->mode[i] = m;
        if (s->linearized) { // PhononMaterial::linearized_bose_occupation (Thermodynamics.cpp:27-37)
            const double hw = kHbar * s->m.omega[m];
            s->occ[i] = hw > 0.0 ? s->m.neq_background[m] +
                                       s->cw_cap[m] * (T - s->setup.background_temperature) / hw
                                 : 0.0;
        } else {
            s->occ[i] = bose(T, s->m.omega[m]);
        }
        if (init->excite) { // apply_initial_mode_excitation (src/solver/Initialization.cpp:146-178)
            const double f = s->m.omega[m] / 6.283185307179586476925286766559;
            if (std::isfinite(f) && !(init->excite_fmin_thz >= 0.0 && f < init->excite_fmin_thz) &&
                !(init->excite_fmax_thz >= 0.0 && f > init->excite_fmax_thz))
                s->occ[i] *= init->excite_multiplier;
        }
        s->uid[i] = static_cast<uint32_t>(i);
    }
    return finish_population(s);
}

int orc_seed(orc_sim* s, int64_t n, const float* position, const int32_t* mode, const double* occupation) {
    s->pos.resize(n);
    s->mode.assign(mode, mode + n);
    s->occ.assign(occupation, occupation + n);
    s->uid.resize(n);
    for (int64_t i = 0; i < n; ++i) {
        s->pos[i] = {position[3 * i], position[3 * i + 1], position[3 * i + 2]};
        s->uid[i] = static_cast<uint32_t>(i);
    }
    return finish_population(s);
}

int orc_set_callaway(orc_sim* s, const pmc_callaway_table* t) {
    if (!s || !t || !t->capacity || !t->wavevector || !t->normal_rate || !t->resistive_rate) return PMC_ERR_ARGUMENT;
    const size_t na = static_cast<size_t>(s->m.n_active);
    s->cw_cap.assign(t->capacity, t->capacity + na);
    s->cw_q.assign(t->wavevector, t->wavevector + 3 * na);
    s->cw_gn.assign(t->normal_rate, t->normal_rate + na);
    s->cw_gr.assign(t->resistive_rate, t->resistive_rate + na);
    s->cw = *t;
    s->cw.capacity = s->cw_cap.data();
    s->cw.wavevector = s->cw_q.data();
    s->cw.normal_rate = s->cw_gn.data();
    s->cw.resistive_rate = s->cw_gr.data();
    s->callaway = true;
    s->conservative = true;
    s->linearized = true;
    return PMC_OK;
}

int orc_set_callaway_nonlinear(orc_sim* s, const pmc_callaway_rates* r) {
    if (!s || !r || r->n_temps <= 0 || !r->temps || !r->gamma_normal || !r->gamma_umklapp || !r->gamma_extra ||
        !r->wavevector)
        return PMC_ERR_ARGUMENT;
    const size_t na = static_cast<size_t>(s->m.n_active), nt = static_cast<size_t>(r->n_temps);
    s->nl_t.assign(r->temps, r->temps + nt);
    s->nl_gn.assign(r->gamma_normal, r->gamma_normal + nt * na);
    s->nl_gu.assign(r->gamma_umklapp, r->gamma_umklapp + nt * na);
    s->nl_ge.assign(r->gamma_extra, r->gamma_extra + nt * na);
    s->nl_q.assign(r->wavevector, r->wavevector + 3 * na);
    s->nl = *r;
    s->nl.temps = s->nl_t.data();
    s->nl.gamma_normal = s->nl_gn.data();
    s->nl.gamma_umklapp = s->nl_gu.data();
    s->nl.gamma_extra = s->nl_ge.data();
    s->nl.wavevector = s->nl_q.data();
    s->nl_callaway = true;
    s->conservative = true; // nonlinear occupations, conservative walls (include/SimulationConfig.h:153-156)
    return PMC_OK;
}

int orc_set_layers(orc_sim* s, const pmc_layers* L) {
    if (!s || !L || L->count < 2) return PMC_ERR_ARGUMENT;
    const int n = L->count;
    s->layers = n;
    s->layer_axis = L->axis;
    s->band = L->band_width;
    s->planes.assign(L->planes, L->planes + (n - 1));
    s->mode_begin.assign(L->mode_begin, L->mode_begin + n + 1);
    s->cell_mat.assign(L->cell_material, L->cell_material + s->g.n_cells);
    s->e_bg_mat.assign(L->background_energy_density, L->background_energy_density + n);
    if (s->linearized) {
        if (!L->heat_capacity_density) return PMC_ERR_UNSUPPORTED;
        s->cbg_mat.assign(L->heat_capacity_density, L->heat_capacity_density + n);
    } else {
        s->cbg_mat.clear();
    }
    s->lut_begin.assign(L->lut_begin, L->lut_begin + n + 1);
    s->bank_begin.assign(L->bank_begin, L->bank_begin + 2 * n + 1);
    const int nb = L->bank_begin[2 * n];
    s->bank_mode.assign(L->bank_mode, L->bank_mode + nb);
    s->bank_omega.assign(L->bank_omega, L->bank_omega + nb);
    s->bank_prefix.assign(L->bank_prefix, L->bank_prefix + nb + 2 * n);
    double vmin = L->spatial_volume[0];
    for (int k = 0; k < n; ++k) vmin = std::min(vmin, L->spatial_volume[k]);
    s->accept.resize(n);
    for (int k = 0; k < n; ++k) s->accept[k] = vmin / L->spatial_volume[k];
    if (L->facet_rough)
        s->layer_rough.assign(L->facet_rough, L->facet_rough + static_cast<size_t>(n) * s->facets.size());
    else
        s->layer_rough.clear();
    s->amm_fraction = L->amm_fraction;
    if (L->amm_fraction > 0.0) {
        s->amm_block.assign(L->amm_block, L->amm_block + 2 * (n - 1) + 1);
        const int rows = s->amm_block.back();
        s->amm_refl.assign(L->amm_reflected, L->amm_reflected + rows);
        s->amm_ebeg.assign(L->amm_edge_begin, L->amm_edge_begin + rows);
        const int edges = s->amm_ebeg.back();
        s->amm_emode.assign(L->amm_edge_mode, L->amm_edge_mode + edges);
        s->amm_eprob.assign(L->amm_edge_prob, L->amm_edge_prob + edges);
    }
    return PMC_OK;
}

int orc_interface(orc_sim* s, int active, const double vel[3], double occ, int dest, const double u[4],
                  int* out_active, double* out_occ, double out_vel[3]) {
    V3 nvel{};
    const int action = interface_core(*s, active, {vel[0], vel[1], vel[2]}, occ, dest, u, *out_active, *out_occ, nvel);
    for (int k = 0; k < 3; ++k) out_vel[k] = nvel[k];
    return action;
}

int orc_interface_dmm(orc_sim* s, int active, const double vel[3], double occ, int dest, const double u[3],
                      int* out_active, double* out_occ, double out_vel[3]) {
    V3 nvel{};
    const int action = dmm_core(*s, active, {vel[0], vel[1], vel[2]}, occ, dest, u, *out_active, *out_occ, nvel);
    for (int k = 0; k < 3; ++k) out_vel[k] = nvel[k];
    return action;
}

int orc_set_heat_source(orc_sim* s, const pmc_heat_source* h) {
    if (!s || !h || h->n_cells != s->g.n_cells) return PMC_ERR_ARGUMENT;
    s->hs = *h;
    s->hs_rate.assign(h->rate, h->rate + h->n_cells);
    s->hs_elig.assign(h->mode_eligible, h->mode_eligible + s->m.n_active);
    if (h->spectrum != 0 && !h->mode_weight) return PMC_ERR_ARGUMENT;
    if (h->mode_weight) s->hs_weight.assign(h->mode_weight, h->mode_weight + s->m.n_active);
    s->hs_pending.assign(h->n_cells, 0.0);
    s->hs_on = true;
    return PMC_OK;
}

int orc_step(orc_sim* s, int32_t nsteps) {
    const double dt = s->setup.time_step_ps;
    const int nres = static_cast<int>(s->res.size());
    for (int it = 0; it < nsteps; ++it) {
        pmc_ledger& L = s->ledger;
        if (s->g.kind == 1 && s->step % 100 == 0) {
            // TransportOperator::validate_containment (src/solver/TransportOperator.cpp:364-402).
            for (size_t i = 0; i < s->pos.size(); ++i) {
                if (orc_contains_mesh(&s->g, s->pos[i].data())) continue;
                set_error(s, "Particle " + std::to_string(i) + " escaped the domain.");
                return PMC_ERR_ESCAPED;
            }
        }
        L.carrier_steps += static_cast<int64_t>(s->mode.size());
        // First half of the symmetric source split (src/MonteCarloSolver.cpp:366-372).
        const double mid = s->elapsed + 0.5 * dt, end = s->elapsed + dt;
        double source_e = 0.0;
        int64_t source_n = 0;
        if (s->hs_on) {
            double e = 0.0;
            int64_t n = 0;
            if (apply_source(*s, source_time_factor(s->hs, s->elapsed, mid), &e, &n)) {
                set_error(s, "heat source failure");
                return PMC_ERR_INVALID_STATE;
            }
            source_e += e;
            source_n += n;
            if (e > 0.0) refresh(*s);
        }
        StepTally total;
        total.absorbed_by_res.assign(nres, 0);
        int err = 0;
        // Transport of existing carriers (src/MonteCarloSolver.cpp:301-346).
        const int64_t n0 = static_cast<int64_t>(s->mode.size());
#pragma omp parallel
        {
            StepTally local;
            local.absorbed_by_res.assign(nres, 0);
#pragma omp for schedule(dynamic, 64)
            for (int64_t i = 0; i < n0; ++i) {
                const int rc = advance(*s, i, dt, local);
                if (rc < 0) {
#pragma omp critical
                    err = err ? err : -rc;
                }
                if (s->alive[i]) s->cell[i] = orc_grid_index(&s->g, s->pos[i].data());
            }
#pragma omp critical
            {
                total.absorbed_e += local.absorbed_e;
                total.absorbed += local.absorbed;
                total.rough_events += local.rough_events;
                total.rough_spec += local.rough_spec;
                total.rough_diff += local.rough_diff;
                total.iface_events += local.iface_events;
                total.iface_transmitted += local.iface_transmitted;
                for (int r = 0; r < nres; ++r) total.absorbed_by_res[r] += local.absorbed_by_res[r];
            }
        }
        if (err) {
            set_error(s, err == kOrcModeMapping ? "Full-matrix boundary mode mapping failed while advancing particles."
                                                : "transport failure");
            return err == kOrcModeMapping ? PMC_ERR_BOUNDARY : err;
        }
        compact(*s);
        // Injection (src/solver/ReservoirInjector.cpp:62-132): one-to-one refill, or
        // floor(expected) + Bernoulli(frac) (ReservoirSampler.cpp:117-125) with the
        // device's Philox stream 4, counter (reservoir, step).
        std::vector<int> counts(s->previous_absorbed);
        if (s->setup.refill_policy == 1)
            for (int r = 0; r < nres; ++r) {
                const double expected = s->res[r].expected_count;
                int count = static_cast<int>(std::floor(expected));
                const double fraction = expected - count;
                if (fraction > 0.0) {
                    uint32_t w[4];
                    orc_philox(static_cast<uint32_t>(r), static_cast<uint32_t>(s->step), 0, 4u, s->k0, s->k1, w);
                    if (orc_u53(w[0], w[1]) < fraction) ++count;
                }
                counts[r] = count;
            }
        int emissions = 0;
        for (int c : counts) emissions += c;
        const int64_t base = static_cast<int64_t>(s->mode.size());
        s->pos.resize(base + emissions);
        s->vel.resize(base + emissions);
        s->cpoint.resize(base + emissions);
        s->mode.resize(base + emissions);
        s->cell.resize(base + emissions);
        s->cfacet.resize(base + emissions);
        s->occ.resize(base + emissions);
        s->ctime.resize(base + emissions);
        s->uid.resize(base + emissions);
        s->alive.resize(base + emissions, 1);
        std::vector<double> t_emit(emissions);
        double injected_e = 0.0;
        int k = 0;
        for (int r = 0; r < nres; ++r) {
            const pmc_reservoir& R = s->res[r];
            for (int j = 0; j < counts[r]; ++j, ++k) {
                uint32_t w0[4], w1[4], w2[4];
                orc_philox(static_cast<uint32_t>(k), static_cast<uint32_t>(s->step), 0, 2u, s->k0, s->k1, w0);
                orc_philox(static_cast<uint32_t>(k), static_cast<uint32_t>(s->step), 1, 2u, s->k0, s->k1, w1);
                orc_philox(static_cast<uint32_t>(k), static_cast<uint32_t>(s->step), 2, 2u, s->k0, s->k1, w2);
                const int m = orc_reservoir_mode(&R, s->m.n_active, orc_u53(w0[0], w0[1]), s->setup.mode_sampling);
                t_emit[k] = dt * orc_u53(w0[2], w0[3]);
                double pt[3];
                orc_surface_point(&R, orc_u53(w1[0], w1[1]), orc_u53(w1[2], w1[3]), orc_u53(w2[0], w2[1]), pt);
                V3 v = table_velocity(*s, m);
                const V3 n{R.normal[0], R.normal[1], R.normal[2]};
                if (dot(v, n) > 0.0) v = mul(v, -1.0);
                const double speed = std::max(1e-12, norm(v));
                const int64_t i = base + k;
                s->pos[i] = add(V3{pt[0], pt[1], pt[2]}, mul(v, s->setup.push_distance_a / speed));
                s->vel[i] = v;
                s->mode[i] = m;
                s->occ[i] = R.occupation[m];
                s->uid[i] = s->next_uid + static_cast<uint32_t>(k);
                s->alive[i] = 1;
                injected_e += kHbar * s->m.omega[m] * (s->occ[i] - s->m.neq_background[m]) * s->m.energy_weight;
            }
        }
        for (int64_t kk = 0; kk < emissions; ++kk)
            if (!update_cache(*s, base + kk)) {
                set_error(s, "Boundary intersection failed while updating collision cache.");
                return PMC_ERR_BOUNDARY;
            }
        {
            StepTally local;
            local.absorbed_by_res.assign(nres, 0);
            for (int64_t kk = 0; kk < emissions; ++kk) {
                const int64_t i = base + kk;
                const double remain = dt - t_emit[kk];
                if (remain > 1e-14) {
                    const int rc = advance(*s, i, remain, local);
                    if (rc < 0) {
                        set_error(s, "transport failure");
                        return -rc == kOrcModeMapping ? PMC_ERR_BOUNDARY : -rc;
                    }
                }
                if (s->alive[i]) s->cell[i] = orc_grid_index(&s->g, s->pos[i].data());
            }
            total.absorbed_e += local.absorbed_e;
            total.absorbed += local.absorbed;
            total.rough_events += local.rough_events;
            total.rough_spec += local.rough_spec;
            total.rough_diff += local.rough_diff;
            total.iface_events += local.iface_events;
            total.iface_transmitted += local.iface_transmitted;
            for (int r = 0; r < nres; ++r) total.absorbed_by_res[r] += local.absorbed_by_res[r];
        }
        compact(*s);
        s->next_uid += static_cast<uint32_t>(emissions);
        // Refresh, RTA, refresh (src/MonteCarloSolver.cpp:376-381).
        refresh(*s);
        std::vector<int32_t> act(s->mode.begin(), s->mode.end());
        double residual;
        if (s->callaway) {
            int status = 0;
            residual = orc_callaway_apply(&s->m, &s->cw, s->g.n_cells, static_cast<int64_t>(act.size()), act.data(),
                                          s->cell.data(), s->occ.data(), dt, s->m.energy_weight, &status);
            if (status) {
                set_error(s, "Callaway collision failure");
                return PMC_ERR_INVALID_STATE;
            }
        } else if (s->nl_callaway) {
            int status = 0;
            residual = orc_nl_callaway_apply(&s->m, &s->nl, s->g.n_cells, s->cell_T.data(), s->setup.lifetime_local,
                                             s->setup.lifetime_temperature, static_cast<int64_t>(act.size()),
                                             act.data(), s->cell.data(), s->occ.data(), dt, s->m.energy_weight,
                                             &status);
            if (status) {
                set_error(s, "nonlinear Callaway collision failure " + std::to_string(status));
                return PMC_ERR_INVALID_STATE;
            }
        } else {
            residual = orc_rta(&s->m, s->g.n_cells, s->cell_T.data(), static_cast<int64_t>(act.size()), act.data(),
                               s->cell.data(), s->occ.data(), dt, s->m.energy_weight, s->setup.lifetime_local,
                               s->setup.lifetime_temperature, nullptr);
        }
        if (s->hs_on) { // second half (src/MonteCarloSolver.cpp:376-378)
            double e = 0.0;
            int64_t n = 0;
            if (apply_source(*s, source_time_factor(s->hs, mid, end), &e, &n)) {
                set_error(s, "heat source failure");
                return PMC_ERR_INVALID_STATE;
            }
            source_e += e;
            source_n += n;
        }
        refresh(*s);
        s->elapsed += dt;
        // Ledgers (src/solver/EnergyLedger.cpp:58-90, src/solver/BoundaryLedger.cpp:86-141).
        const double energy = thermal_energy(*s);
        const double balance = energy - L.thermal_energy_ev - source_e - 0.0 - injected_e + total.absorbed_e - residual;
        L.step_source_energy_ev = source_e;
        L.total_source_energy_ev += source_e;
        L.source_updates += source_n;
        L.step_source_updates = source_n;
        L.source_time_ps = s->source_time;
        L.source_pending_ev = 0.0;
        for (double p : s->hs_pending) L.source_pending_ev += p;
        L.step_absorbed = total.absorbed;
        L.step_injected = emissions;
        L.step_absorbed_energy_ev = total.absorbed_e;
        L.step_injected_energy_ev = injected_e;
        L.step_lifetime_residual_ev = residual;
        L.step_balance_residual_ev = balance;
        L.total_absorbed += total.absorbed;
        L.total_injected += emissions;
        L.total_absorbed_energy_ev += total.absorbed_e;
        L.total_injected_energy_ev += injected_e;
        L.total_lifetime_residual_ev += residual;
        L.total_balance_residual_ev += balance;
        L.thermal_energy_ev = energy;
        L.rough_events += total.rough_events;
        L.rough_specular += total.rough_spec;
        L.rough_diffuse += total.rough_diff;
        L.interface_events += total.iface_events;
        L.interface_transmitted += total.iface_transmitted;
        L.live_carriers = static_cast<int64_t>(s->mode.size());
        L.slots = L.live_carriers;
        s->previous_absorbed = total.absorbed_by_res;
        s->step++;
        L.step = s->step;
    }
    return PMC_OK;
}

int orc_cells(orc_sim* s, double* temperature, double* energy) {
    for (int c = 0; c < s->g.n_cells; ++c) {
        if (temperature) temperature[c] = s->cell_T[c];
        if (energy) energy[c] = s->cell_e[c];
    }
    return PMC_OK;
}

int orc_heat_flux(orc_sim* s, double* flux) {
    const int nc = s->g.n_cells;
    std::vector<double> q(3 * static_cast<size_t>(nc), 0.0);
    for (size_t i = 0; i < s->mode.size(); ++i) {
        const int a = s->mode[i];
        const double e = kHbar * s->m.omega[a] * (s->occ[i] - s->m.neq_background[a]) * s->m.energy_weight;
        for (int k = 0; k < 3; ++k) q[3 * s->cell[i] + k] += s->vel[i][k] * e;
    }
    for (int c = 0; c < nc; ++c)
        for (int k = 0; k < 3; ++k) flux[3 * c + k] = q[3 * c + k] * (1.0 / s->g.cell_volume[c]) * 1.602176634e13;
    return PMC_OK;
}

int orc_ledger(orc_sim* s, pmc_ledger* out) {
    *out = s->ledger;
    return PMC_OK;
}

int64_t orc_carrier_count(orc_sim* s) { return static_cast<int64_t>(s->mode.size()); }

int orc_carriers(orc_sim* s, double* position, int32_t* mode, double* occupation, int32_t* cell, uint32_t* uid) {
    const size_t n = s->mode.size();
    std::vector<size_t> order(n);
    std::iota(order.begin(), order.end(), 0);
    std::sort(order.begin(), order.end(), [&](size_t a, size_t b) { return s->uid[a] < s->uid[b]; });
    for (size_t k = 0; k < n; ++k) {
        const size_t i = order[k];
        for (int a = 0; a < 3; ++a)
            if (position) position[3 * k + a] = s->pos[i][a];
        if (mode) mode[k] = s->mode[i];
        if (occupation) occupation[k] = s->occ[i];
        if (cell) cell[k] = s->cell[i];
        if (uid) uid[k] = s->uid[i];
    }
    return PMC_OK;
}

const char* orc_last_error(orc_sim* s) { return s->error.c_str(); }

void orc_destroy(orc_sim* s) { delete s; }

} // extern "C"
