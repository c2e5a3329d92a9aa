#include "boundary_tables.hpp"

#include <algorithm>
#include <cmath>
#include <iterator>
#include <limits>
#include <numeric>
#include <stdexcept>
#include <string>

namespace pmc {
namespace {
double dot(const Vec3& a, const Vec3& b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
double norm(const Vec3& a) { return std::sqrt(dot(a, a)); }
Vec3 sub(const Vec3& a, const Vec3& b) { return {a[0] - b[0], a[1] - b[1], a[2] - b[2]}; }
Vec3 mul(const Vec3& a, double s) { return {a[0] * s, a[1] * s, a[2] * s}; }
} // namespace

void build_alias(const std::vector<double>& w, std::vector<double>& prob, std::vector<int>& alias) {
    const int n = static_cast<int>(w.size());
    prob.assign(n, 1.0);
    alias.resize(n);
    std::iota(alias.begin(), alias.end(), 0);
    double sum = 0.0;
    for (double x : w) sum += std::max(0.0, x);
    if (!(sum > 0.0) || n == 0) return;
    std::vector<double> scaled(n);
    std::vector<int> small, large;
    for (int i = 0; i < n; ++i) {
        scaled[i] = std::max(0.0, w[i]) * n / sum;
        (scaled[i] < 1.0 ? small : large).push_back(i);
    }
    // Process in ascending index order for a reproducible table.
    std::reverse(small.begin(), small.end());
    std::reverse(large.begin(), large.end());
    while (!small.empty() && !large.empty()) {
        const int s = small.back();
        small.pop_back();
        const int l = large.back();
        large.pop_back();
        prob[s] = scaled[s];
        alias[s] = l;
        scaled[l] = (scaled[l] + scaled[s]) - 1.0;
        (scaled[l] < 1.0 ? small : large).push_back(l);
    }
    for (int i : large) prob[i] = 1.0;
    for (int i : small) prob[i] = 1.0; // rounding leftovers
}

std::vector<RoughTable> build_rough_tables(const Domain& d, const ModeTables& m, const std::vector<bool>& enabled,
                                           double equilibrium_temperature, std::vector<int>& facet_to_table) {
    // Follows RoughBoundarySampler::build (src/solver/RoughBoundarySampler.cpp:77-368) for one material.
    const int nf = static_cast<int>(d.facets.size());
    facet_to_table.assign(nf, -1);
    std::vector<RoughTable> tables;
    const int na = m.n_active;
    std::vector<Vec3> v(na);
    std::vector<double> vnorm(na), omega(na), knorm(na), domega(na);
    for (int a = 0; a < na; ++a) {
        const int f = m.active_flat[a];
        v[a] = m.velocity_flat[f];
        vnorm[a] = norm(v[a]);
        omega[a] = m.omega_flat[f];
        knorm[a] = m.knorm_flat[f];
        domega[a] = m.window_flat[f];
    }
    std::vector<int> shell(na), order(na);
    std::iota(order.begin(), order.end(), 0);
    std::sort(order.begin(), order.end(), [&](int a, int b) { return omega[a] < omega[b]; });
    int shells = 0;
    double shell_w = -1;
    for (int a : order) {
        if (shell_w < 0 || omega[a] - shell_w > 1e-8 * std::max(shell_w, 1e-12)) {
            ++shells;
            shell_w = omega[a];
        }
        shell[a] = shells - 1;
    }
    for (int facet = 0; facet < nf; ++facet) {
        if (!enabled[facet]) continue;
        RoughTable rd;
        rd.facet = facet;
        rd.strict_elastic = equilibrium_temperature > 0;
        rd.specularity.assign(na, 0.0);
        rd.spec_match.assign(na, -1);
        rd.diffuse_begin.assign(na, -1);
        rd.diffuse_end.assign(na, -1);
        const double eta = std::max(0.0, d.facets[facet].condition == 'R' ? d.facets[facet].value : 0.0);
        const Vec3 n_out = d.facets[facet].normal;
        const Vec3 n_in{-n_out[0], -n_out[1], -n_out[2]};
        std::vector<int> incoming, outgoing;
        std::vector<double> destruction(na, 0.0), creation(na, 0.0);
        for (int a = 0; a < na; ++a) {
            const double vn = dot(v[a], n_in);
            if (vn < 0.0) {
                incoming.push_back(a);
                destruction[a] = -vn;
            } else if (vn > 0.0) {
                outgoing.push_back(a);
                creation[a] = vn;
            }
            const double vmag = std::max(1e-12, vnorm[a]);
            const double incidence_cos = std::abs(vn) / vmag;
            const double x = rd.strict_elastic ? 2 * eta * std::abs(dot(m.wavevector_flat[m.active_flat[a]], n_in))
                                               : 2.0 * eta * incidence_cos * knorm[a];
            double p = std::exp(-(x * x));
            if (!std::isfinite(p)) p = 0.0;
            rd.specularity[a] = (vn < 0.0) ? std::clamp(p, 0.0, 1.0) : 0.0;
            if (rd.strict_elastic) {
                const double c = m.heat_capacity(equilibrium_temperature, m.active_flat[a]);
                destruction[a] *= c;
                creation[a] *= c;
            }
        }
        rd.outgoing_active = outgoing;
        rd.outgoing_sorted = outgoing;
        std::sort(rd.outgoing_sorted.begin(), rd.outgoing_sorted.end(),
                  [&omega](int a, int b) { return omega[a] < omega[b]; });
        std::vector<double> sorted_w(rd.outgoing_sorted.size());
        for (size_t k = 0; k < sorted_w.size(); ++k) sorted_w[k] = omega[rd.outgoing_sorted[k]];
        std::vector<int> shell_begin(shells, -1), shell_end(shells, -1);
        if (rd.strict_elastic) {
            std::vector<double> in_flux(shells, 0), out_flux(shells, 0);
            for (int a : incoming) in_flux[shell[a]] += destruction[a];
            for (int a : outgoing) out_flux[shell[a]] += creation[a];
            const double resolved = std::max(*std::max_element(in_flux.begin(), in_flux.end()),
                                             *std::max_element(out_flux.begin(), out_flux.end()));
            const double roundoff = 64 * std::numeric_limits<double>::epsilon() * resolved;
            for (int s = 0; s < shells; ++s)
                if (std::abs(in_flux[s] - out_flux[s]) > 1e-7 * std::max(in_flux[s], out_flux[s]) + roundoff)
                    throw std::invalid_argument("Callaway reflecting boundary lacks balanced equal-frequency "
                                                "incoming/outgoing flux. Check full-BZ material symmetry and velocities.");
            for (int pos = 0; pos < static_cast<int>(rd.outgoing_sorted.size()); ++pos) {
                const int s = shell[rd.outgoing_sorted[pos]];
                if (shell_begin[s] < 0) shell_begin[s] = pos;
                shell_end[s] = pos + 1;
            }
        }
        for (int a : incoming) {
            if (rd.outgoing_sorted.empty()) continue;
            const double tol = std::max(1e-12, domega[a]);
            const double w_lo = omega[a] - tol, w_hi = omega[a] + tol;
            const int ib = rd.strict_elastic ? shell_begin[shell[a]]
                                             : static_cast<int>(std::lower_bound(sorted_w.begin(), sorted_w.end(), w_lo) - sorted_w.begin());
            const int ie = rd.strict_elastic ? shell_end[shell[a]]
                                             : static_cast<int>(std::upper_bound(sorted_w.begin(), sorted_w.end(), w_hi) - sorted_w.begin());
            if (ie <= ib) continue;
            rd.diffuse_begin[a] = ib;
            rd.diffuse_end[a] = ie;
            const double vin_n = dot(v[a], n_in);
            const Vec3 vtry = sub(v[a], mul(n_in, 2.0 * vin_n));
            int best = -1;
            double best_metric = std::numeric_limits<double>::infinity();
            for (int pos = ib; pos < ie; ++pos) {
                const int ao = rd.outgoing_sorted[pos];
                const Vec3 dv = sub(vtry, v[ao]);
                const double refn = std::max({1e-12, norm(vtry), norm(v[ao])});
                const double metric = norm(dv) / refn;
                if (metric < best_metric) {
                    best_metric = metric;
                    best = ao;
                }
            }
            if (best >= 0) rd.spec_match[a] = best;
        }
        for (int a : incoming)
            if (rd.spec_match[a] < 0) rd.specularity[a] = 0.0;
        if (rd.strict_elastic) {
            std::vector<double> assigned(na, 0);
            for (int a : incoming)
                if (rd.spec_match[a] >= 0) assigned[rd.spec_match[a]] += destruction[a] * rd.specularity[a];
            for (int a : incoming) {
                const int ao = rd.spec_match[a];
                if (ao >= 0 && assigned[ao] > creation[ao]) rd.specularity[a] *= creation[ao] / assigned[ao];
            }
        }
        // Residual diffuse creation rates (include/solver/BoundaryFluxSampling.h:16-42).
        std::vector<double> residual = creation;
        for (int a = 0; a < na; ++a) {
            const int ao = rd.spec_match[a];
            if (ao < 0 || ao >= na) continue;
            const double flux = destruction[a], prob = rd.specularity[a];
            if (!std::isfinite(flux) || !std::isfinite(prob) || flux <= 0.0 || prob <= 0.0) continue;
            residual[ao] -= flux * std::clamp(prob, 0.0, 1.0);
        }
        for (double& r : residual)
            if (!std::isfinite(r) || r < 0.0) r = 0.0;
        double rate_sum = 0.0;
        for (int ao : outgoing) rate_sum += residual[ao];
        // Ordered nonnegative prefix (include/solver/BoundaryFluxSampling.h:44-62).
        rd.residual_prefix.assign(rd.outgoing_sorted.size() + 1, 0.0);
        for (size_t k = 0; k < rd.outgoing_sorted.size(); ++k) {
            const int idx = rd.outgoing_sorted[k];
            double w = (idx >= 0 && idx < na) ? residual[idx] : 0.0;
            if (!std::isfinite(w) || w < 0.0) w = 0.0;
            rd.residual_prefix[k + 1] = rd.residual_prefix[k] + w;
        }
        if (rate_sum > 0.0) {
            double cdf = 0.0;
            for (int ao : outgoing) {
                const double r = residual[ao];
                if (r <= 0.0) continue;
                cdf += r;
                rd.roulette_active.push_back(ao);
                rd.roulette_cdf.push_back(cdf);
            }
        }
        facet_to_table[facet] = static_cast<int>(tables.size());
        tables.push_back(std::move(rd));
    }
    return tables;
}

std::vector<ReservoirTable> build_reservoir_tables(const Domain& d, const ModeTables& m, double density,
                                                   double time_step, std::vector<int>& facet_to_reservoir) {
    // ReservoirSampler::configure (src/solver/ReservoirSampler.cpp:17-101) plus the
    // per-facet triangle pool of SurfaceMesh::sample_surface_points
    // (src/geometry/SurfaceSampling.cpp:20-73).
    const int nf = static_cast<int>(d.facets.size());
    facet_to_reservoir.assign(nf, -1);
    std::vector<ReservoirTable> out;
    for (int facet : d.reservoir_facets()) {
        ReservoirTable r;
        r.facet = facet;
        r.temperature = d.facets[facet].value;
        r.normal = d.facets[facet].normal;
        r.area = d.facets[facet].area;
        if (!std::isfinite(r.temperature) || r.temperature < 0 || !(r.area > 0))
            throw std::invalid_argument("Invalid reservoir temperature/area.");
        if (std::abs(norm(r.normal) - 1.0) > 1e-8) throw std::invalid_argument("Reservoir normal must have unit length.");
        const int na = m.n_active;
        r.cdf.assign(na, 0.0);
        std::vector<double> weight(na, 0.0);
        double expected = 0;
        if (density > 0.0 && time_step > 0.0) {
            const double density_area = density * r.area;
            const double thickness = static_cast<double>(na) / density_area;
            for (int a = 0; a < na; ++a) {
                const double parallel = -dot(r.normal, m.velocity_flat[m.active_flat[a]]);
                weight[a] = std::max(0.0, parallel * time_step / thickness);
                expected += weight[a];
                r.cdf[a] = expected;
            }
        }
        if (!std::isfinite(expected) || expected > static_cast<double>(std::numeric_limits<int>::max()))
            throw std::overflow_error("Reservoir initial emission count exceeds int capacity.");
        if (expected > 0)
            for (double& x : r.cdf) x /= expected;
        r.expected = expected;
        r.initial_count = static_cast<int>(std::llround(expected));
        build_alias(weight, r.alias_prob, r.alias_index);
        r.occupation.resize(na);
        for (int a = 0; a < na; ++a) r.occupation[a] = m.bose(r.temperature, m.active_flat[a]);
        // Area-weighted triangle choice as std::discrete_distribution builds it.
        const auto& faces = d.facets[facet].faces;
        std::vector<double> w;
        for (int f : faces) w.push_back(std::max(d.face_area[f], 0.0));
        const double sum = std::accumulate(w.begin(), w.end(), 0.0);
        if (!(sum > 0)) throw std::runtime_error("No faces available for surface sampling.");
        double acc = 0.0;
        for (size_t i = 0; i < w.size(); ++i) {
            acc += w[i] / sum;
            r.face_cdf.push_back(acc);
        }
        r.face_cdf.back() = 1.0;
        for (int f : faces)
            for (int k = 0; k < 3; ++k)
                for (int a = 0; a < 3; ++a) r.triangles.push_back(d.vertices[d.faces[f][k]][a]);
        facet_to_reservoir[facet] = static_cast<int>(out.size());
        out.push_back(std::move(r));
    }
    return out;
}

} // namespace pmc
