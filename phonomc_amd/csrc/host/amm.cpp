// Acoustic-mismatch interface tables (see amm.hpp).  The channel geometry (the
// q cells, their clipped overlap and its quadrature) and the order in which
// channels are accumulated follow AcousticInterfaceSampler::configure
// (src/solver/AcousticInterfaceSampler.cpp:27-437) and channel::Cell
// (src/solver/ChannelCell.cpp:1-210), because the tables must equal the
// reference's: the quadrature points of an overlap depend on how it is cut
// into tetrahedra.
#include "amm.hpp"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <limits>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <unordered_map>

namespace pmc {
namespace {

using V3 = std::array<double, 3>;
using M3 = std::array<V3, 3>; // m[row][column]
using Loop = std::vector<V3>;  // vertices of a planar convex face, in order
using Solid = std::vector<Loop>;

constexpr double kTwoPi = 2.0 * 3.14159265358979323846;

V3 minus(V3 a, const V3& b) {
    for (int j = 0; j < 3; ++j) a[j] -= b[j];
    return a;
}
double dot3(const V3& a, const V3& b) {
    double s = 0;
    for (int j = 0; j < 3; ++j) s += a[j] * b[j];
    return s;
}
V3 cross3(const V3& a, const V3& b) {
    return {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
}
bool same_point(const V3& a, const V3& b) {
    const V3 d = minus(a, b);
    return dot3(d, d) < 1e-24;
}
V3 centroid(const Loop& pts) {
    V3 c{};
    const double n = static_cast<double>(pts.size());
    for (const V3& p : pts)
        for (int j = 0; j < 3; ++j) c[j] += p[j] / n;
    return c;
}
void push_new_tail(Loop& loop, const V3& p) {
    if (loop.empty() || !same_point(loop.back(), p)) loop.push_back(p);
}

// A parallelepiped centre + edge matrix (edge j = column j), with the dual rows
// that map a point to its coordinates along the edges, its bounding box and
// volume.
struct QCell {
    V3 center{}, lo{}, hi{};
    M3 edge{}, dual{};
    double volume = 0;

    QCell(const V3& c, const M3& e) : center(c), edge(e) {
        const V3 e0{e[0][0], e[1][0], e[2][0]}, e1{e[0][1], e[1][1], e[2][1]}, e2{e[0][2], e[1][2], e[2][2]};
        const double det = dot3(e0, cross3(e1, e2));
        volume = std::abs(det);
        if (!std::isfinite(volume) || volume <= 0) throw std::invalid_argument("Degenerate phonon channel cell");
        dual = {cross3(e1, e2), cross3(e2, e0), cross3(e0, e1)};
        for (auto& row : dual)
            for (double& x : row) x /= det;
        for (int i = 0; i < 3; ++i) {
            double r = 0;
            for (int j = 0; j < 3; ++j) r += .5 * std::abs(e[i][j]);
            lo[i] = center[i] - r;
            hi[i] = center[i] + r;
        }
    }
    V3 local(const V3& p) const {
        V3 y{};
        for (int i = 0; i < 3; ++i) y[i] = dot3(dual[i], minus(p, center));
        return y;
    }
    // Six faces: for each edge direction j, the side at -1/2 then +1/2; each
    // loop visits the other two edge directions (lower one first) as
    // (-,-), (+,-), (+,+), (-,+).
    Solid faces() const {
        std::array<V3, 8> v;
        for (int i = 0; i < 8; ++i) {
            v[i] = center;
            for (int d = 0; d < 3; ++d)
                for (int j = 0; j < 3; ++j) v[i][d] += edge[d][j] * ((i >> j) & 1 ? .5 : -.5);
        }
        Solid s;
        for (int j = 0; j < 3; ++j) {
            const int p = j == 0 ? 1 : 0, q = j == 2 ? 1 : 2; // the other two directions
            for (int side = 0; side < 2; ++side) {
                const int base = side << j;
                s.push_back({v[base], v[base | 1 << p], v[base | 1 << p | 1 << q], v[base | 1 << q]});
            }
        }
        return s;
    }
};

// Keeps the part of a convex solid where sign * n.(x - o) <= 1/2 (to 1e-12),
// closing the cut with a face ordered by angle about its centroid.
Solid clip_half(const Solid& s, const V3& n, const V3& o, double sign) {
    Solid out;
    Loop cap;
    const auto dist = [&](const V3& p) { return sign * dot3(n, minus(p, o)) - .5; };
    for (const Loop& f : s) {
        Loop kept;
        for (size_t i = 0; i < f.size(); ++i) {
            const V3& a = f[i];
            const V3& b = f[(i + 1) % f.size()];
            const double da = dist(a), db = dist(b);
            const bool in_a = da <= 1e-12, in_b = db <= 1e-12;
            if (in_a) push_new_tail(kept, a);
            if (in_a != in_b) {
                const double t = da / (da - db);
                V3 x{};
                for (int j = 0; j < 3; ++j) x[j] = a[j] + t * (b[j] - a[j]);
                push_new_tail(kept, x);
                if (std::none_of(cap.begin(), cap.end(), [&](const V3& c) { return same_point(c, x); }))
                    cap.push_back(x);
            }
        }
        if (kept.size() > 1 && same_point(kept.front(), kept.back())) kept.pop_back();
        if (kept.size() >= 3) out.push_back(std::move(kept));
    }
    if (cap.size() >= 3) {
        const V3 c = centroid(cap);
        V3 nn = n;
        double len = std::sqrt(dot3(nn, nn));
        for (double& x : nn) x /= len;
        V3 u = cross3(nn, std::abs(nn[0]) < .8 ? V3{1, 0, 0} : V3{0, 1, 0});
        len = std::sqrt(dot3(u, u));
        for (double& x : u) x /= len;
        const V3 w = cross3(nn, u);
        std::sort(cap.begin(), cap.end(), [&](V3 a, V3 b) {
            a = minus(a, c);
            b = minus(b, c);
            return std::atan2(dot3(a, w), dot3(a, u)) < std::atan2(dot3(b, w), dot3(b, u));
        });
        out.push_back(std::move(cap));
    }
    return out;
}

// Overlap of cell a with cell b: a's faces clipped by b's three slab pairs.
Solid overlap(const QCell& a, const QCell& b) {
    for (int i = 0; i < 3; ++i)
        if (a.hi[i] <= b.lo[i] || b.hi[i] <= a.lo[i]) return {};
    Solid s = a.faces();
    for (int i = 0; i < 3 && !s.empty(); ++i)
        for (double sign : {-1., 1.}) s = clip_half(s, b.dual[i], b.center, sign);
    return s;
}

struct QPoint {
    V3 x;
    double w;
};
// Degree-two, four-point tetrahedral rule ((5 + 3 sqrt 5) / 20 and (5 - sqrt 5) / 20
// barycentric weights) on the tetrahedra from the vertex mean to each face's fan.
std::vector<QPoint> quadrature(const Solid& s) {
    Loop all;
    for (const Loop& f : s) all.insert(all.end(), f.begin(), f.end());
    if (all.empty()) return {};
    const double big = (5.0 + 3.0 * std::sqrt(5.0)) / 20.0, small = (5.0 - std::sqrt(5.0)) / 20.0;
    const V3 anchor = centroid(all);
    std::vector<QPoint> pts;
    for (const Loop& f : s)
        for (size_t i = 1; i + 1 < f.size(); ++i) {
            const std::array<V3, 4> t{anchor, f[0], f[i], f[i + 1]};
            const double vol = std::abs(dot3(minus(t[1], t[0]), cross3(minus(t[2], t[0]), minus(t[3], t[0])))) / 6;
            if (vol <= 0) continue;
            for (int a = 0; a < 4; ++a) {
                V3 x{};
                for (int b = 0; b < 4; ++b)
                    for (int j = 0; j < 3; ++j) x[j] += (a == b ? big : small) * t[b][j];
                pts.push_back({x, vol / 4});
            }
        }
    return pts;
}

struct AmmMode {
    double omega, flux;
    V3 k, v;
    int branch, qpoint;
};

} // namespace

double amm_transmission(double z1, double z2, double c1, double c2) {
    if (!std::isfinite(z1) || !std::isfinite(z2) || !std::isfinite(c1) || !std::isfinite(c2) || z1 <= 0 || z2 <= 0 ||
        c1 < 0 || c2 < 0 || c1 > 1 || c2 > 1)
        throw std::invalid_argument("Invalid scalar acoustic impedance/angle.");
    const double a = z1 * c1, b = z2 * c2;
    return a + b > 0 ? std::clamp(4 * a * b / ((a + b) * (a + b)), 0.0, 1.0) : 0;
}

std::vector<AmmInterface> build_amm_tables(const std::vector<const ModeTables*>& mats,
                                           const std::vector<const RawMaterial*>& raws, int axis,
                                           const std::vector<double>& rho, double df, double dk) {
    const size_t nm = mats.size();
    if (axis < 0 || axis > 2 || nm < 2 || raws.size() != nm || rho.size() != nm || !std::isfinite(df) || df <= 0 ||
        !std::isfinite(dk) || dk <= 0)
        throw std::invalid_argument("AMM needs layered materials, mass densities and positive frequency/k-parallel bins.");
    const int ta = (axis + 1) % 3, tb = (axis + 2) % 3;

    // Modes, and the mirror partner of every mode that crosses the plane
    // (:44-180): the reciprocal-grid point of k with k_axis reversed, then the
    // branch of the same class (acoustic < 3 <= optical) with equal frequency
    // and mirrored velocity, the mode's own branch first on ties; optical modes
    // without one search their branch at equal frequency elsewhere.
    std::vector<std::vector<AmmMode>> modes(nm);
    std::vector<std::vector<int>> mirror(nm);
    std::vector<M3> qbasis(nm);
    for (size_t m = 0; m < nm; ++m) {
        if (!mats[m] || !raws[m] || !std::isfinite(rho[m]) || rho[m] <= 0)
            throw std::invalid_argument("AMM needs finite positive material densities.");
        const ModeTables& t = *mats[m];
        const RawMaterial& raw = *raws[m];
        const auto& mesh = raw.mesh;
        if (mesh[0] <= 0 || mesh[1] <= 0 || mesh[2] <= 0) throw std::runtime_error("AMM needs a complete uniform q mesh.");
        // fractional coordinates of a Cartesian k (columns of the reciprocal
        // matrix are the reciprocal vectors)
        const M3& rec = raw.reciprocal;
        const V3 r0{rec[0][0], rec[1][0], rec[2][0]}, r1{rec[0][1], rec[1][1], rec[2][1]},
            r2{rec[0][2], rec[1][2], rec[2][2]};
        const double det = dot3(r0, cross3(r1, r2));
        M3 inv{cross3(r1, r2), cross3(r2, r0), cross3(r0, r1)};
        for (auto& row : inv)
            for (double& x : row) x /= det;
        using Key = std::array<long long, 3>;
        const auto key_of = [&](const V3& k) {
            Key key{};
            for (int j = 0; j < 3; ++j) {
                const double x = dot3(inv[j], k) * 2 * mesh[j];
                const long long index = std::llround(x), period = 2LL * mesh[j];
                if (std::abs(x - index) > 1e-5)
                    throw std::runtime_error("Interface reflection does not preserve the material q grid.");
                key[j] = (index % period + period) % period;
            }
            return key;
        };
        const int n = t.n_active;
        mirror[m].assign(n, -1);
        std::map<Key, int> qpoint_at;
        for (int i = 0; i < n; ++i) {
            const int f = t.active_flat[i];
            const V3 v = t.velocity_flat[f];
            const V3 k = t.wavevector_flat[f];
            modes[m].push_back({t.omega_flat[f], std::abs(v[axis]) / t.energy_normalization, k, v, f % t.branches,
                                f / t.branches});
            qpoint_at[key_of(k)] = f / t.branches;
        }
        std::vector<int> optical_down; // optical modes moving toward -axis, by (frequency, index)
        for (int j = 0; j < n; ++j)
            if (modes[m][j].branch >= 3 && modes[m][j].v[axis] < -1e-7) optical_down.push_back(j);
        std::sort(optical_down.begin(), optical_down.end(), [&](int i, int j) {
            return modes[m][i].omega != modes[m][j].omega ? modes[m][i].omega < modes[m][j].omega : i < j;
        });
        const auto mirror_error = [&](const AmmMode& x, const AmmMode& y) {
            double dv = 0, scale = 0;
            for (int d = 0; d < 3; ++d) {
                dv += std::pow(x.v[d] - (d == axis ? -y.v[d] : y.v[d]), 2);
                scale += x.v[d] * x.v[d];
            }
            return std::array<double, 2>{dv, scale};
        };
        for (int i = 0; i < n; ++i) {
            const AmmMode& x = modes[m][i];
            if (x.v[axis] <= 1e-7) continue;
            V3 kr = x.k;
            kr[axis] *= -1;
            const auto it = qpoint_at.find(key_of(kr));
            if (it == qpoint_at.end()) throw std::runtime_error("AMM reflection has no reciprocal-grid partner.");
            int best = -1;
            double err = std::numeric_limits<double>::infinity();
            for (int br = 0; br < t.branches; ++br) {
                if ((x.branch < 3) != (br < 3)) continue;
                const int j = t.flat_to_active[it->second * t.branches + br];
                if (j < 0 || mirror[m][j] >= 0) continue;
                const AmmMode& y = modes[m][j];
                if (y.v[axis] >= -1e-7) continue;
                if (std::abs(x.omega - y.omega) > 1e-6 * std::max(x.omega, y.omega)) continue;
                auto [dv, scale] = mirror_error(x, y);
                if (dv > 1e-12 * std::max(1., scale)) continue;
                dv += br == x.branch ? 0 : 1e-24;
                if (dv < err) {
                    err = dv;
                    best = j;
                }
            }
            if (best < 0 && x.branch >= 3) {
                const double tol = 1e-6 * x.omega;
                auto first = std::lower_bound(optical_down.begin(), optical_down.end(), x.omega - tol,
                                              [&](int j, double w) { return modes[m][j].omega < w; });
                for (auto jt = first; jt != optical_down.end() && modes[m][*jt].omega <= x.omega + tol; ++jt) {
                    const AmmMode& y = modes[m][*jt];
                    if (mirror[m][*jt] >= 0 || y.branch != x.branch) continue;
                    const auto [dv, scale] = mirror_error(x, y);
                    if (dv <= 1e-12 * std::max(1., scale) && dv < err) {
                        err = dv;
                        best = *jt;
                    }
                }
            }
            if (best < 0)
                throw std::runtime_error("AMM needs mirror-symmetric frequency/velocity data for the interface axis (material " +
                                         std::to_string(m) + ", active " + std::to_string(i) + ", v " +
                                         std::to_string(x.v[0]) + "," + std::to_string(x.v[1]) + "," +
                                         std::to_string(x.v[2]) + ").");
            mirror[m][i] = best;
            mirror[m][best] = i;
        }
        for (int i = 0; i < n; ++i)
            if (std::abs(modes[m][i].v[axis]) >= 1e-7 && mirror[m][i] < 0)
                throw std::runtime_error("AMM reflection pairing is incomplete.");
    }

    // The q cell of every acoustic mode that crosses the plane, linearized in
    // (frequency, k_a, k_b) (:184-225).
    std::vector<std::vector<std::unique_ptr<QCell>>> cells(nm);
    for (size_t m = 0; m < nm; ++m) {
        const auto& mesh = raws[m]->mesh;
        if (static_cast<long long>(mesh[0]) * mesh[1] * mesh[2] != mats[m]->qpoints)
            throw std::runtime_error("Cell-overlap AMM requires a complete uniform q mesh.");
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) qbasis[m][i][j] = raws[m]->reciprocal[i][j] / mesh[j];
        cells[m].resize(modes[m].size());
        for (size_t i = 0; i < modes[m].size(); ++i) {
            const AmmMode& x = modes[m][i];
            if (x.branch >= 3 || std::abs(x.v[axis]) < 1e-7) continue;
            M3 e{};
            for (int j = 0; j < 3; ++j) {
                for (int d = 0; d < 3; ++d) e[0][j] += x.v[d] * qbasis[m][d][j] / kTwoPi;
                e[1][j] = qbasis[m][ta][j];
                e[2][j] = qbasis[m][tb][j];
            }
            cells[m][i] = std::make_unique<QCell>(V3{x.omega / kTwoPi, x.k[ta], x.k[tb]}, e);
        }
    }
    // Search bins of a cell (candidate search only; :245-266).
    const V3 bin{df, dk, dk};
    const auto bin_range = [&](const QCell& c) {
        std::array<std::array<int, 2>, 3> r{};
        long double count = 1;
        for (int j = 0; j < 3; ++j) {
            const double lo = std::floor(c.lo[j] / bin[j]), hi = std::floor(c.hi[j] / bin[j]);
            if (!std::isfinite(lo) || !std::isfinite(hi) || lo < std::numeric_limits<int>::min() ||
                hi >= std::numeric_limits<int>::max())
                throw std::runtime_error("AMM search grid is too fine; enlarge search bins (this does not change channel "
                                         "integrals).");
            r[j] = {static_cast<int>(lo), static_cast<int>(hi)};
            count *= hi - lo + 1;
        }
        if (count > 1000000)
            throw std::runtime_error("AMM search bins create too many candidates per cell; enlarge search bins.");
        return r;
    };
    const auto bin_key = [](int branch, int f, int x, int y) {
        return (static_cast<uint64_t>(static_cast<uint16_t>(branch)) << 48) ^
               (static_cast<uint64_t>(static_cast<uint16_t>(f)) << 32) ^
               (static_cast<uint64_t>(static_cast<uint16_t>(x)) << 16) ^ static_cast<uint64_t>(static_cast<uint16_t>(y));
    };

    std::vector<AmmInterface> out(nm - 1);
    for (size_t face = 0; face + 1 < nm; ++face) {
        AmmInterface& dst = out[face];
        std::array<std::vector<int>, 2> incoming;
        std::array<std::vector<std::vector<std::pair<int, double>>>, 2> edges;
        std::array<std::vector<double>, 2> load;
        for (int side = 0; side < 2; ++side) {
            const size_t m = face + side;
            const size_t n = modes[m].size();
            dst.reflected[side].assign(n, -1);
            edges[side].resize(n);
            load[side].assign(n, 0.0);
            for (size_t i = 0; i < n; ++i) {
                const AmmMode& x = modes[m][i];
                if (side == 0 ? x.v[axis] <= 1e-7 : x.v[axis] >= -1e-7) continue;
                dst.reflected[side][i] = mirror[m][i];
                if (mirror[m][i] < 0) throw std::runtime_error("Missing AMM reflection partner.");
                if (cells[m][i]) {
                    incoming[side].push_back(static_cast<int>(i));
                    dst.audit.incident_acoustic_flux[side] += x.flux;
                }
            }
        }
        // Identical banks on both sides: no interface in the acoustic subspace (:292-318).
        bool identical = modes[face].size() == modes[face + 1].size() &&
                         mats[face]->energy_normalization == mats[face + 1]->energy_normalization &&
                         rho[face] == rho[face + 1];
        for (size_t i = 0; identical && i < modes[face].size(); ++i) {
            const AmmMode &x = modes[face][i], &y = modes[face + 1][i];
            identical = x.omega == y.omega && x.k == y.k && x.v == y.v;
        }
        if (identical) {
            for (int side = 0; side < 2; ++side)
                for (int i : incoming[side]) {
                    edges[side][i].push_back({i, 1.0});
                    dst.audit.transmitted_flux[side] += modes[face + side][i].flux;
                }
        } else {
            // Candidates of each incident mode of side 0 among side 1's: same
            // branch, overlapping search bins (all of them when both lists are
            // short); visited in increasing index so that accumulation is stable.
            const bool direct = incoming[0].size() * incoming[1].size() <= 4096;
            std::unordered_map<uint64_t, std::vector<int>> by_bin;
            if (!direct)
                for (int j : incoming[1]) {
                    const auto r = bin_range(*cells[face + 1][j]);
                    for (int f = r[0][0]; f <= r[0][1]; ++f)
                        for (int x = r[1][0]; x <= r[1][1]; ++x)
                            for (int y = r[2][0]; y <= r[2][1]; ++y)
                                by_bin[bin_key(modes[face + 1][j].branch, f, x, y)].push_back(j);
                }
            struct Channel {
                int left, right;
                double g;
            };
            std::vector<Channel> channels;
            std::vector<int> seen(modes[face + 1].size(), -1);
            const auto impedance = [&](size_t m, int i, const V3& point) {
                const AmmMode& md = modes[m][i];
                const V3 y = cells[m][i]->local(point);
                V3 k = md.k;
                for (int d = 0; d < 3; ++d)
                    for (int j = 0; j < 3; ++j) k[d] += qbasis[m][d][j] * y[j];
                const double norm = std::sqrt(k[0] * k[0] + k[1] * k[1] + k[2] * k[2]);
                return std::array<double, 2>{norm > 0 ? rho[m] * point[0] * 2 * 3.14159265358979323846 / norm : 0,
                                             norm > 0 ? std::abs(k[axis]) / norm : 0};
            };
            for (int i : incoming[0]) {
                const QCell& c = *cells[face][i];
                std::vector<int> cand;
                if (direct) {
                    for (int j : incoming[1])
                        if (modes[face][i].branch == modes[face + 1][j].branch) cand.push_back(j);
                } else {
                    const auto r = bin_range(c);
                    for (int f = r[0][0]; f <= r[0][1]; ++f)
                        for (int x = r[1][0]; x <= r[1][1]; ++x)
                            for (int y = r[2][0]; y <= r[2][1]; ++y) {
                                const auto it = by_bin.find(bin_key(modes[face][i].branch, f, x, y));
                                if (it == by_bin.end()) continue;
                                for (int j : it->second)
                                    if (seen[j] != i) {
                                        seen[j] = i;
                                        cand.push_back(j);
                                    }
                            }
                }
                std::sort(cand.begin(), cand.end());
                for (int j : cand) {
                    const QCell& other = *cells[face + 1][j];
                    const auto pts = quadrature(overlap(c, other));
                    if (pts.empty()) continue;
                    double ov = 0, g = 0;
                    const double dens = std::min(modes[face][i].flux / c.volume, modes[face + 1][j].flux / other.volume);
                    for (const QPoint& q : pts) {
                        if (q.x[0] <= 0) continue;
                        const auto z1 = impedance(face, i, q.x), z2 = impedance(face + 1, j, q.x);
                        ov += q.w * dens;
                        if (z1[0] > 0 && z2[0] > 0)
                            g += q.w * dens * amm_transmission(z1[0], z2[0], std::min(1., z1[1]), std::min(1., z2[1]));
                    }
                    dst.audit.geometric_overlap_flux += ov;
                    if (g > 0) {
                        channels.push_back({i, j, g});
                        load[0][i] += ov;
                        load[1][j] += ov;
                    }
                }
            }
            // Both marginals capped by one factor per channel, keeping
            // reciprocity (:402-417).
            for (const Channel& ch : channels) {
                const double wl = modes[face][ch.left].flux, wr = modes[face + 1][ch.right].flux;
                const double factor = std::max({1., load[0][ch.left] / wl, load[1][ch.right] / wr});
                dst.audit.maximum_capacity_ratio = std::max(dst.audit.maximum_capacity_ratio, factor);
                const double g = ch.g / factor;
                dst.audit.capacity_removed_flux += ch.g - g;
                edges[0][ch.left].push_back({mirror[face + 1][ch.right], g / wl});
                edges[1][ch.right].push_back({mirror[face][ch.left], g / wr});
                dst.audit.transmitted_flux[0] += g;
                dst.audit.transmitted_flux[1] += g;
            }
        }
        for (int side = 0; side < 2; ++side) {
            for (int i : incoming[side]) {
                double sum = 0;
                for (const auto& e : edges[side][i]) sum += e.second;
                if (sum > 1 + 1e-10) throw std::runtime_error("AMM row exceeds unit probability.");
                if (edges[side][i].empty()) dst.audit.no_overlap_flux[side] += modes[face + side][i].flux;
            }
            dst.edge_begin[side].assign(1, 0);
            for (const auto& row : edges[side]) {
                for (const auto& e : row) {
                    dst.edge_mode[side].push_back(e.first);
                    dst.edge_prob[side].push_back(e.second);
                }
                dst.edge_begin[side].push_back(static_cast<int>(dst.edge_mode[side].size()));
            }
        }
    }
    return out;
}

} // namespace pmc
