#include "material.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <limits>
#include <map>
#include <numeric>
#include <stdexcept>

namespace pmc {
namespace {
constexpr double kPi = 3.14159265358979323846;

std::array<double, 3> q_to_k(const std::array<std::array<double, 3>, 3>& r,
                             const std::array<double, 3>& q) {
    return {r[0][0] * q[0] + r[0][1] * q[1] + r[0][2] * q[2],
            r[1][0] * q[0] + r[1][1] * q[1] + r[1][2] * q[2],
            r[2][0] * q[0] + r[2][1] * q[1] + r[2][2] * q[2]};
}

double norm3(const std::array<double, 3>& v) {
    return std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
}

// Clamped piecewise-linear interpolation with the reference's tie and clamp rules
// (src/material/Thermodynamics.cpp:41-69).
double lerp(double x0, double x1, double y0, double y1, double x) {
    if (std::abs(x1 - x0) <= 1e-18) return y0;
    const double t = (x - x0) / (x1 - x0);
    return y0 * (1.0 - t) + y1 * t;
}

double interp_clamped(const std::vector<double>& xs, const std::vector<double>& ys, double x) {
    if (xs.empty() || ys.empty()) return 0.0;
    if (xs.size() == 1 || ys.size() == 1) return ys.front();
    if (x <= xs.front()) return ys.front();
    if (x >= xs.back()) return ys.back();
    const size_t hi = static_cast<size_t>(std::upper_bound(xs.begin(), xs.end(), x) - xs.begin());
    return lerp(xs[hi - 1], xs[hi], ys[hi - 1], ys[hi], x);
}
} // namespace

RawMaterial make_synthetic_material(const SyntheticSpec& s) {
    // Simple cubic lattice with the Si primitive-cell volume; Gamma-centred
    // s.mesh^3 grid folded to (-1/2, 1/2].  Acoustic branches follow a sine
    // dispersion whose slope at Gamma is the sound velocity and which flattens at
    // the zone corner; optical branches drop quadratically by optical_drop_thz.
    // Group velocities are isotropic, v = (dw/d|k|) k_hat, except that the
    // component normal to a zone face vanishes for q on that face (|q_a| = 1/2),
    // which keeps the mode set exactly mirror-symmetric.
    RawMaterial m;
    const int n = s.mesh;
    const double a = std::cbrt(s.cell_volume_a3);
    m.volume_a3 = a * a * a;
    const double b = 2.0 * kPi / a;
    m.reciprocal = {{{b, 0, 0}, {0, b, 0}, {0, 0, b}}};
    m.qpoints = n * n * n;
    m.branches = 6;
    m.mesh = {n, n, n};
    const int nm = m.qpoints * m.branches;
    m.frequency_thz.resize(nm);
    m.group_velocity.resize(nm);
    m.qpoint_fractions.resize(m.qpoints);
    const double k_max = std::sqrt(3.0) * 0.5 * b;
    const double acoustic_v[3] = {s.ta_velocity, s.ta_velocity, s.la_velocity};
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j)
            for (int l = 0; l < n; ++l) {
                const int q = (i * n + j) * n + l;
                auto fold = [n](int idx) {
                    return 2 * idx > n ? static_cast<double>(idx - n) / n : static_cast<double>(idx) / n;
                };
                const std::array<double, 3> f{fold(i), fold(j), fold(l)};
                m.qpoint_fractions[q] = f;
                const std::array<double, 3> k{b * f[0], b * f[1], b * f[2]};
                const double kn = norm3(k);
                std::array<double, 3> khat{0, 0, 0};
                if (kn > 0) khat = {k[0] / kn, k[1] / kn, k[2] / kn};
                for (int axis = 0; axis < 3; ++axis)
                    if (2 * std::abs(f[axis]) == 1.0) khat[axis] = 0.0;
                for (int br = 0; br < 6; ++br) {
                    double omega, slope;
                    if (br < 3) {
                        const double v0 = acoustic_v[br];
                        const double wmax = 2.0 * v0 * k_max / kPi;
                        omega = wmax * std::sin(0.5 * kPi * kn / k_max);
                        slope = v0 * std::cos(0.5 * kPi * kn / k_max);
                    } else {
                        const double drop = 2.0 * kPi * s.optical_drop_thz;
                        omega = 2.0 * kPi * s.optical_thz[br - 3] - drop * (kn / k_max) * (kn / k_max);
                        slope = -2.0 * drop * kn / (k_max * k_max);
                    }
                    const int fi = q * 6 + br;
                    m.frequency_thz[fi] = omega / (2.0 * kPi);
                    m.group_velocity[fi] = {slope * khat[0], slope * khat[1], slope * khat[2]};
                }
            }
    const int nt = static_cast<int>(std::floor(s.t_max / s.t_step + 0.5)) + 1;
    m.temperatures.resize(nt);
    for (int t = 0; t < nt; ++t) m.temperatures[t] = t * s.t_step;
    m.gamma.assign(static_cast<size_t>(nt) * nm, 0.0);
    if (s.callaway_split) {
        m.gamma_normal.assign(m.gamma.size(), 0.0);
        m.gamma_umklapp.assign(m.gamma.size(), 0.0);
        m.gamma_isotope.assign(nm, 0.0);
    }
    // Linewidth in THz from the inverse lifetime in 1/s: gamma = tau^-1 * 1e-12 / (4 pi).
    const double to_thz = 1e-12 / (4.0 * kPi);
    for (int fi = 0; fi < nm; ++fi) {
        const double w = 2.0 * kPi * m.frequency_thz[fi] * 1e12; // rad/s
        const double iso = s.rate_a * w * w * w * w * to_thz;
        if (s.callaway_split) m.gamma_isotope[fi] = iso;
        for (int t = 0; t < nt; ++t) {
            const double temp = m.temperatures[t];
            const double anh = temp > 0 ? s.rate_b * w * w * temp * std::exp(-s.rate_c / temp) * to_thz : 0.0;
            const size_t idx = static_cast<size_t>(t) * nm + fi;
            if (s.callaway_split) {
                m.gamma[idx] = anh;
                m.gamma_normal[idx] = 0.6 * anh;
                m.gamma_umklapp[idx] = anh - 0.6 * anh;
            } else {
                m.gamma[idx] = iso + anh;
            }
        }
    }
    return m;
}

void validate_raw_material(const RawMaterial& d) {
    // Same checks and messages as src/material/MaterialData.cpp:27-131.
    if (!std::isfinite(d.volume_a3) || d.volume_a3 <= 0)
        throw std::runtime_error("Unit-cell volume must be finite and positive");
    if (d.qpoints <= 0 || d.branches <= 0 || d.qpoints > std::numeric_limits<int>::max() / d.branches)
        throw std::runtime_error("Material mode dimensions must be positive and fit integer indices");
    const size_t nm = static_cast<size_t>(d.qpoints) * d.branches;
    if (d.frequency_thz.size() != nm) throw std::runtime_error("frequency shape mismatch");
    if (d.qpoint_fractions.size() != static_cast<size_t>(d.qpoints))
        throw std::runtime_error("qpoint shape mismatch");
    if (d.group_velocity.size() != nm) throw std::runtime_error("group_velocity shape mismatch");
    size_t full = 1;
    for (int c : d.mesh) {
        if (c <= 0) throw std::runtime_error("mesh entries must be positive integers");
        full *= static_cast<size_t>(c);
    }
    if (full * d.bz_refinement != static_cast<size_t>(d.qpoints)) // a completed table holds bz_refinement copies
        throw std::runtime_error("Full Brillouin-zone grid required: qpoint count differs from "
                                 "mesh product. Expand reduced data first.");
    for (double x : d.frequency_thz)
        if (!std::isfinite(x)) throw std::runtime_error("frequency contains non-finite values");
    for (const auto& v : d.group_velocity)
        for (double x : v)
            if (!std::isfinite(x)) throw std::runtime_error("group_velocity contains non-finite values");
    if (d.temperatures.empty() || d.temperatures.front() < 0 ||
        !std::is_sorted(d.temperatures.begin(), d.temperatures.end()) ||
        std::adjacent_find(d.temperatures.begin(), d.temperatures.end()) != d.temperatures.end())
        throw std::runtime_error("temperature must be nonempty, nonnegative and strictly increasing");
    if (d.gamma.size() != d.temperatures.size() * nm) throw std::runtime_error("gamma shape mismatch");
    for (double g : d.gamma) {
        if (!std::isfinite(g)) throw std::runtime_error("gamma contains non-finite values");
        if (g < 0) throw std::runtime_error("gamma must be nonnegative");
    }
    if (d.gamma_normal.empty() != d.gamma_umklapp.empty())
        throw std::runtime_error("gamma_N and gamma_U must both be supplied or both omitted");
    if (!d.gamma_isotope.empty() && d.gamma_isotope.size() != nm && d.gamma_isotope.size() != d.gamma.size())
        throw std::runtime_error("gamma_isotope shape mismatch");
}

double ModeTables::bose(double temperature, int flat) const {
    // src/material/Thermodynamics.cpp:71-82 (exp, not expm1; 1e-12 floor).
    const double w = omega_flat[flat];
    if (temperature <= 0.0 || w <= 0.0) return 0.0;
    const double x = kHbarEvPs * w / (kBoltzmannEvK * temperature);
    if (x > 700.0) return 0.0;
    const double ex = std::exp(x);
    return 1.0 / std::max(ex - 1.0, 1e-12);
}

double ModeTables::heat_capacity(double temperature, int flat) const {
    // src/material/Thermodynamics.cpp:11-27.
    const double hw = kHbarEvPs * omega_flat[flat];
    if (!(temperature > 0.0) || !(hw > 0.0)) return 0.0;
    const double x = hw / (kBoltzmannEvK * temperature);
    if (x < 1e-6) return kBoltzmannEvK * (1.0 - x * x / 12.0);
    if (x > 700.0) return 0.0;
    const double den = -std::expm1(-x);
    return kBoltzmannEvK * x * x * std::exp(-x) / (den * den);
}

double ModeTables::gamma_at(double temperature, int flat) const {
    // src/material/Thermodynamics.cpp:108-133.
    if (!std::isfinite(temperature) || temperature < 0.0)
        throw std::invalid_argument("Lifetime temperature must be finite and nonnegative");
    const size_t m = static_cast<size_t>(flat);
    const size_t nmode = static_cast<size_t>(n_flat);
    if (temperature <= temps.front()) return gamma_total[m];
    if (temperature >= temps.back()) return gamma_total[(temps.size() - 1) * nmode + m];
    const size_t hi = static_cast<size_t>(std::upper_bound(temps.begin(), temps.end(), temperature) - temps.begin());
    return lerp(temps[hi - 1], temps[hi], gamma_total[(hi - 1) * nmode + m], gamma_total[hi * nmode + m],
                temperature);
}

double ModeTables::relaxation_weight(double temperature, int flat, double dt) const {
    // src/material/Thermodynamics.cpp:95-106.
    const double g = gamma_at(temperature, flat);
    if (g == 0.0 || dt == 0.0) return 0.0;
    return -std::expm1(-(g * dt) * (4.0 * kPi));
}

double ModeTables::crystal_energy_density(double temperature) const {
    // src/material/Thermodynamics.cpp:151-163: active-list order, hbar*w*n.
    double e = 0.0;
    for (int a = 0; a < n_active; ++a) {
        const int f = active_flat[a];
        const double w = omega_flat[f];
        if (w <= 0.0) continue;
        e += kHbarEvPs * w * bose(temperature, f);
    }
    return e / energy_normalization + zero_point_density;
}

double ModeTables::energy_density_from_temperature(double t) const {
    return interp_clamped(lut_t, lut_e, t);
}

double ModeTables::temperature_from_energy_density(double e) const {
    return interp_clamped(lut_e, lut_t, e);
}

ModeTables build_mode_tables(const RawMaterial& raw, double lookup_dt, bool include_stationary) {
    if (!std::isfinite(lookup_dt) || lookup_dt <= 0.0)
        throw std::invalid_argument("Temperature lookup spacing must be finite and positive");
    validate_raw_material(raw);
    ModeTables t;
    t.include_stationary = include_stationary;
    t.qpoints = raw.qpoints;
    t.branches = raw.branches;
    t.n_flat = raw.qpoints * raw.branches;
    t.unit_cell_volume = raw.volume_a3;
    const int nm = t.n_flat;
    t.omega_flat.resize(nm);
    // src/PhononMaterial.cpp:87-90 evaluation order: max(0,f) * 2 * pi.
    for (int i = 0; i < nm; ++i) t.omega_flat[i] = std::max(0.0, raw.frequency_thz[i]) * 2.0 * M_PI;
    t.velocity_flat = raw.group_velocity;
    t.temps = raw.temperatures;
    t.gamma_total = raw.gamma;
    t.gamma_table = raw.gamma;
    t.gamma_extra.assign(raw.gamma.size(), 0.0);
    if (!raw.gamma_isotope.empty()) {
        for (size_t i = 0; i < t.gamma_total.size(); ++i) {
            const double extra = raw.gamma_isotope[raw.gamma_isotope.size() == static_cast<size_t>(nm) ? i % nm : i];
            t.gamma_extra[i] = 0.0 + extra;
            t.gamma_total[i] = raw.gamma[i] + t.gamma_extra[i];
        }
    }
    t.gamma_normal = raw.gamma_normal;
    t.gamma_umklapp = raw.gamma_umklapp;
    t.knorm_flat.assign(nm, 0.0);
    t.window_flat.assign(nm, 0.0);
    t.wavevector_flat.assign(nm, {0, 0, 0});
    const std::array<double, 3> mesh_q{1.0 / (2.0 * raw.mesh[0]), 1.0 / (2.0 * raw.mesh[1]),
                                       1.0 / (2.0 * raw.mesh[2])};
    const auto kg = q_to_k(raw.reciprocal, mesh_q);
    for (int q = 0; q < raw.qpoints; ++q) {
        const auto kv = q_to_k(raw.reciprocal, raw.qpoint_fractions[q]);
        const double kn = norm3(kv);
        for (int b = 0; b < raw.branches; ++b) {
            const int fi = q * raw.branches + b;
            const auto& v = t.velocity_flat[fi];
            t.knorm_flat[fi] = kn;
            t.wavevector_flat[fi] = kv;
            t.window_flat[fi] = std::sqrt((v[0] * kg[0]) * (v[0] * kg[0]) + (v[1] * kg[1]) * (v[1] * kg[1]) +
                                          (v[2] * kg[2]) * (v[2] * kg[2]));
        }
    }
    // Active modes: |v| > 1e-12 (or omega > 0 with stationary modes), by mesh
    // point, then branch, then the point's first-BZ copies, so that the copies
    // of a mode sit together (src/PhononMaterial.cpp:140-160; one copy without
    // completion).
    t.flat_to_active.assign(nm, -1);
    const int copies = static_cast<int>(raw.bz_refinement);
    for (int base = 0; base < raw.qpoints; base += copies)
        for (int b = 0; b < raw.branches; ++b)
            for (int copy = 0; copy < copies; ++copy) {
                const int fi = (base + copy) * raw.branches + b;
                const auto& v = t.velocity_flat[fi];
                if (include_stationary ? t.omega_flat[fi] > 0.0
                                       : std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]) > 1e-12) {
                    t.flat_to_active[fi] = static_cast<int>(t.active_flat.size());
                    t.active_flat.push_back(fi);
                }
            }
    t.n_active = static_cast<int>(t.active_flat.size());
    if (t.n_active == 0) throw std::runtime_error("No active modes found in material data.");
    t.energy_normalization = std::max(1.0, static_cast<double>(raw.qpoints) * raw.volume_a3);

    // Temperature lookup (src/material/Thermodynamics.cpp:178-224).
    double zp = 0.0;
    for (double w : t.omega_flat) zp += 0.5 * kHbarEvPs * std::max(0.0, w);
    t.zero_point_density = zp / t.energy_normalization;
    const double tmax = std::max(1000.0, *std::max_element(t.temps.begin(), t.temps.end()));
    const double dT = std::max(1e-6, lookup_dt);
    const double intervals = std::floor((tmax - 0.0) / dT + 0.5);
    if (!std::isfinite(intervals) || intervals > std::numeric_limits<int>::max() - 1)
        throw std::runtime_error("Temperature lookup table exceeds integer indexing capacity");
    const int n = std::max(2, static_cast<int>(intervals) + 1);
    t.lut_t.assign(n, 0.0);
    t.lut_e.assign(n, 0.0);
    std::vector<double> raw_e(n, 0.0);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) raw_e[i] = t.crystal_energy_density(0.0 + static_cast<double>(i) * dT);
    for (int i = 0; i < n; ++i) {
        t.lut_t[i] = 0.0 + static_cast<double>(i) * dT;
        double e = raw_e[i];
        if (i > 0) e = std::max(e, std::nextafter(t.lut_e[i - 1], std::numeric_limits<double>::infinity()));
        t.lut_e[i] = e;
    }
    return t;
}

namespace {
double det3(const std::array<std::array<double, 3>, 3>& b) {
    return b[0][0] * (b[1][1] * b[2][2] - b[1][2] * b[2][1]) - b[0][1] * (b[1][0] * b[2][2] - b[1][2] * b[2][0]) +
           b[0][2] * (b[1][0] * b[2][1] - b[1][1] * b[2][0]);
}

std::array<std::array<double, 3>, 3> inverse3(const std::array<std::array<double, 3>, 3>& b, double det) {
    std::array<std::array<double, 3>, 3> inv{};
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            inv[i][j] = (b[(j + 1) % 3][(i + 1) % 3] * b[(j + 2) % 3][(i + 2) % 3] -
                         b[(j + 1) % 3][(i + 2) % 3] * b[(j + 2) % 3][(i + 1) % 3]) /
                        det;
    return inv;
}

double length2(const std::array<std::array<double, 3>, 3>& b, const std::array<double, 3>& q) {
    double r = 0;
    for (int i = 0; i < 3; ++i) {
        double x = 0;
        for (int j = 0; j < 3; ++j) x += b[i][j] * q[j];
        r += x * x;
    }
    return r;
}
} // namespace

namespace {
// src/material/BzQuadrature.cpp:11-110: complete Gamma-centred mesh, no
// duplicates, first-BZ representatives; refinement = lcm of the numbers of
// equally short representatives.  reps (optional) receives each q point's
// representatives in the order of the search (i, j, k ascending).
unsigned bz_representatives(const RawMaterial& d, std::vector<std::vector<std::array<double, 3>>>* reps) {
    const auto& b = d.reciprocal;
    const double det = det3(b);
    if (!std::isfinite(det) || det == 0) throw std::invalid_argument("Callaway reciprocal lattice is singular.");
    const auto inv = inverse3(b, det);
    std::vector<std::array<int, 3>> keys;
    keys.reserve(d.qpoint_fractions.size());
    unsigned refinement = 1;
    for (const auto& q : d.qpoint_fractions) {
        std::array<int, 3> key{};
        for (int a = 0; a < 3; ++a) {
            const double x = (q[a] - std::floor(q[a])) * d.mesh[a];
            const long index = std::lround(x);
            if (std::abs(x - index) > 2e-6 * d.mesh[a])
                throw std::invalid_argument("Callaway requires a complete Gamma-centred full q mesh.");
            key[a] = static_cast<int>(index % d.mesh[a]);
        }
        keys.push_back(key);
        const double original = length2(b, q);
        int lo[3], hi[3];
        double count = 1;
        for (int a = 0; a < 3; ++a) {
            double nrm = 0;
            for (double x : inv[a]) nrm += x * x;
            const double extent = std::sqrt(original * nrm);
            if (!std::isfinite(extent) || extent > 100 || std::abs(q[a]) > 100)
                throw std::invalid_argument("Callaway first-BZ validation exceeds supported lattice conditioning.");
            lo[a] = static_cast<int>(std::ceil(q[a] - extent - 1e-10));
            hi[a] = static_cast<int>(std::floor(q[a] + extent + 1e-10));
            count *= hi[a] - lo[a] + 1;
        }
        if (count > 100000) throw std::invalid_argument("Callaway first-BZ validation work limit exceeded.");
        std::vector<double> lengths;
        std::vector<std::array<double, 3>> shifted;
        double shortest = original;
        for (int i = lo[0]; i <= hi[0]; ++i)
            for (int j = lo[1]; j <= hi[1]; ++j)
                for (int k = lo[2]; k <= hi[2]; ++k) {
                    const std::array<double, 3> p{q[0] - i, q[1] - j, q[2] - k};
                    const double r = length2(b, p);
                    shortest = std::min(shortest, r);
                    lengths.push_back(r);
                    shifted.push_back(p);
                }
        const double tolerance = 1e-10 * std::max(shortest, 1e-20);
        if (original > shortest + tolerance)
            throw std::invalid_argument("Callaway qpoint is outside the first Brillouin zone; supply first-BZ "
                                        "representatives.");
        std::vector<std::array<double, 3>> mine;
        for (size_t c = 0; c < lengths.size(); ++c)
            if (std::abs(lengths[c] - shortest) <= tolerance) mine.push_back(shifted[c]);
        const unsigned m = static_cast<unsigned>(mine.size());
        if (m == 0) throw std::runtime_error("Empty Callaway BZ representative set.");
        unsigned g = refinement, h = m;
        while (h) {
            const unsigned t = g % h;
            g = h;
            h = t;
        }
        const unsigned long long next = static_cast<unsigned long long>(refinement / g) * m;
        if (next > 96 || next * static_cast<unsigned long long>(d.qpoints) * d.branches >
                             static_cast<unsigned long long>(std::numeric_limits<int>::max()))
            throw std::length_error("Callaway equal-weight BZ refinement exceeds supported storage; use a smaller "
                                    "mesh.");
        refinement = static_cast<unsigned>(next);
        if (reps) reps->push_back(std::move(mine));
    }
    std::sort(keys.begin(), keys.end());
    if (std::adjacent_find(keys.begin(), keys.end()) != keys.end())
        throw std::invalid_argument("Callaway input q mesh contains duplicate points modulo reciprocal translations.");
    return refinement;
}
} // namespace

unsigned callaway_bz_refinement(const RawMaterial& d) { return bz_representatives(d, nullptr); }

RawMaterial complete_callaway_bz(const RawMaterial& d) {
    std::vector<std::vector<std::array<double, 3>>> reps;
    const unsigned refinement = bz_representatives(d, &reps);
    RawMaterial out = d;
    out.bz_refinement = refinement;
    if (refinement == 1) return out;
    // Source mode of every completed mode: q's representatives in turn, each
    // repeated refinement / |reps| times, every branch of q per copy.
    const size_t nb = static_cast<size_t>(d.branches), old_nm = static_cast<size_t>(d.qpoints) * nb;
    std::vector<size_t> source;
    source.reserve(old_nm * refinement);
    out.qpoint_fractions.clear();
    for (int q = 0; q < d.qpoints; ++q)
        for (const auto& r : reps[q])
            for (unsigned copy = 0; copy < refinement / reps[q].size(); ++copy) {
                out.qpoint_fractions.push_back(r);
                for (size_t br = 0; br < nb; ++br) source.push_back(static_cast<size_t>(q) * nb + br);
            }
    const auto expand = [&](const std::vector<double>& v) {
        std::vector<double> e;
        if (v.empty()) return e;
        e.reserve(v.size() * refinement);
        for (size_t t = 0; t < v.size() / old_nm; ++t)
            for (size_t i : source) e.push_back(v[t * old_nm + i]);
        return e;
    };
    out.frequency_thz = expand(d.frequency_thz);
    out.gamma = expand(d.gamma);
    out.gamma_normal = expand(d.gamma_normal);
    out.gamma_umklapp = expand(d.gamma_umklapp);
    out.gamma_isotope = expand(d.gamma_isotope);
    out.group_velocity.clear();
    for (size_t i : source) out.group_velocity.push_back(d.group_velocity[i]);
    out.qpoints = d.qpoints * static_cast<int>(refinement);
    return out;
}

CallawayTable callaway_data(const ModeTables& t, const RawMaterial& raw, double temperature) {
    // src/material/PhononMaterialCollisionData.cpp:32-165, with the mesh and
    // first-BZ checks of callaway_bz_refinement.
    if (t.gamma_normal.empty() || t.gamma_umklapp.empty())
        throw std::invalid_argument("Callaway requires gamma_N and gamma_U in kappa-fbz.hdf5; regenerate with phono3py "
                                    "--br --nu and expand consistently. No N/U split is inferred.");
    if (!std::isfinite(temperature) || temperature <= 0 || temperature < t.temps.front() ||
        temperature > t.temps.back())
        throw std::invalid_argument("Callaway reference temperature is outside the N/U data range.");
    for (size_t i = 0; i < t.gamma_table.size(); ++i) {
        const double sum = t.gamma_normal[i] + t.gamma_umklapp[i];
        if (!std::isfinite(sum) || std::abs(sum - t.gamma_table[i]) > 1e-7 * std::max(sum, t.gamma_table[i]) + 1e-14)
            throw std::invalid_argument("Callaway data inconsistent: gamma_N + gamma_U != gamma. Supply unmodified, "
                                        "same-run linewidths in the same units.");
    }
    // The completed mesh (src/material/PhononMaterialCollisionData.cpp:52-132):
    // every mesh point exactly bz_refinement times, every q point a shortest
    // representative within 1e-5.
    {
        std::map<std::array<int, 3>, unsigned> points;
        for (const auto& q : raw.qpoint_fractions) {
            std::array<int, 3> key{};
            for (int a = 0; a < 3; ++a) {
                const double x = (q[a] - std::floor(q[a])) * raw.mesh[a];
                const long index = std::lround(x);
                if (std::abs(x - index) > 2e-6 * raw.mesh[a])
                    throw std::invalid_argument("Callaway requires a complete Gamma-centred full q mesh.");
                key[a] = static_cast<int>(index % raw.mesh[a]);
            }
            if (++points[key] > raw.bz_refinement)
                throw std::invalid_argument("Callaway q mesh contains duplicate points modulo reciprocal translations.");
        }
        for (const auto& e : points)
            if (e.second != raw.bz_refinement)
                throw std::invalid_argument("Incomplete equal-weight Callaway BZ quadrature.");
        const auto& b = raw.reciprocal;
        const auto inv = inverse3(b, det3(b));
        for (const auto& q : raw.qpoint_fractions) {
            const double k2 = length2(b, q);
            int lo[3], hi[3];
            double count = 1;
            for (int a = 0; a < 3; ++a) {
                double nrm = 0;
                for (double x : inv[a]) nrm += x * x;
                const double extent = std::sqrt(k2 * nrm);
                if (!std::isfinite(extent) || extent > 100 || std::abs(q[a]) > 100)
                    throw std::invalid_argument("Callaway first-BZ validation exceeds supported lattice conditioning.");
                lo[a] = static_cast<int>(std::ceil(q[a] - extent - 1e-10));
                hi[a] = static_cast<int>(std::floor(q[a] + extent + 1e-10));
                count *= hi[a] - lo[a] + 1;
            }
            if (count > 100000) throw std::invalid_argument("Callaway first-BZ validation work limit exceeded.");
            for (int i = lo[0]; i <= hi[0]; ++i)
                for (int j = lo[1]; j <= hi[1]; ++j)
                    for (int l = lo[2]; l <= hi[2]; ++l)
                        if (length2(b, {q[0] - i, q[1] - j, q[2] - l}) < k2 - 1e-5 * std::max(k2, 1e-20))
                            throw std::invalid_argument("Callaway qpoint is outside the first Brillouin zone; supply "
                                                        "first-BZ representatives.");
        }
    }
    size_t hi = static_cast<size_t>(std::lower_bound(t.temps.begin(), t.temps.end(), temperature) - t.temps.begin());
    const size_t lo = hi > 0 ? hi - 1 : hi;
    const double fraction = hi == lo ? 0 : (temperature - t.temps[lo]) / (t.temps[hi] - t.temps[lo]);
    const size_t nm = static_cast<size_t>(t.n_flat);
    size_t positive = 0;
    for (double w : t.omega_flat)
        if (w > 0) ++positive;
    if (positive != t.active_flat.size())
        throw std::invalid_argument("Callaway requires all positive-frequency modes, including stationary modes.");
    CallawayTable c;
    const auto rate = [&](const std::vector<double>& table, size_t flat) {
        return 4 * std::acos(-1.0) * ((1 - fraction) * table[lo * nm + flat] + fraction * table[hi * nm + flat]);
    };
    for (int a = 0; a < t.n_active; ++a) {
        const size_t flat = static_cast<size_t>(t.active_flat[a]);
        c.omega.push_back(t.omega_flat[flat]);
        c.capacity.push_back(t.heat_capacity(temperature, static_cast<int>(flat)));
        c.wavevector.push_back(t.wavevector_flat[flat]);
        c.normal_rate.push_back(rate(t.gamma_normal, flat));
        c.resistive_rate.push_back(rate(t.gamma_umklapp, flat) + rate(t.gamma_extra, flat));
    }
    return c;
}

namespace {
constexpr char kMagic[8] = {'P', 'M', 'C', 'M', 'A', 'T', '0', '1'};

template <class T>
void put(std::ofstream& o, const T& v) { o.write(reinterpret_cast<const char*>(&v), sizeof(T)); }
template <class T>
void put_vec(std::ofstream& o, const std::vector<T>& v) {
    const uint64_t n = v.size();
    put(o, n);
    if (n) o.write(reinterpret_cast<const char*>(v.data()), static_cast<std::streamsize>(n * sizeof(T)));
}
template <class T>
void get(std::ifstream& i, T& v) {
    if (!i.read(reinterpret_cast<char*>(&v), sizeof(T))) throw std::runtime_error("Truncated material file");
}
template <class T>
void get_vec(std::ifstream& i, std::vector<T>& v) {
    uint64_t n = 0;
    get(i, n);
    if (n > (uint64_t(1) << 34)) throw std::runtime_error("Corrupt material file");
    v.resize(n);
    if (n && !i.read(reinterpret_cast<char*>(v.data()), static_cast<std::streamsize>(n * sizeof(T))))
        throw std::runtime_error("Truncated material file");
}
} // namespace

void save_raw_material(const RawMaterial& m, const std::string& path) {
    std::ofstream o(path, std::ios::binary | std::ios::trunc);
    if (!o) throw std::runtime_error("Cannot open material file for writing: " + path);
    o.write(kMagic, 8);
    put(o, m.volume_a3);
    put(o, m.reciprocal);
    put(o, m.qpoints);
    put(o, m.branches);
    put(o, m.mesh);
    put_vec(o, m.frequency_thz);
    put_vec(o, m.qpoint_fractions);
    put_vec(o, m.group_velocity);
    put_vec(o, m.temperatures);
    put_vec(o, m.gamma);
    put_vec(o, m.gamma_normal);
    put_vec(o, m.gamma_umklapp);
    put_vec(o, m.gamma_isotope);
    if (!o) throw std::runtime_error("Failed writing material file: " + path);
}

RawMaterial load_material(const std::string& spec, const std::string& input_directory) {
    if (spec.rfind("synthetic:", 0) == 0) {
        std::string rest = spec.substr(10);
        SyntheticSpec s;
        const auto colon = rest.find(':');
        std::string name = rest.substr(0, colon);
        if (colon != std::string::npos) {
            const std::string opt = rest.substr(colon + 1);
            if (opt != "callaway") throw std::runtime_error("Unknown synthetic material option: " + opt);
            s.callaway_split = true;
        }
        if (name.rfind("ge", 0) == 0) {
            // Ge-like: larger cell, slower sound, lower optical branches.
            s.cell_volume_a3 = 5.658 * 5.658 * 5.658 / 4.0;
            s.ta_velocity = 35.0;
            s.la_velocity = 53.0;
            s.optical_thz[0] = 9.0;
            s.optical_thz[1] = 8.6;
            s.optical_thz[2] = 8.2;
        } else if (name.rfind("si", 0) != 0) {
            throw std::runtime_error("Unknown synthetic material: " + name);
        }
        s.mesh = std::stoi(name.substr(2));
        if (s.mesh < 2 || s.mesh > 64) throw std::runtime_error("Synthetic mesh must be in [2, 64]");
        return make_synthetic_material(s);
    }
    namespace fs = std::filesystem;
    fs::path p(spec);
    if (p.is_relative()) p = fs::path(input_directory) / p;
    if (fs::is_directory(p)) p /= "material.pmcmat";
    std::ifstream in(p, std::ios::binary);
    if (!in) throw std::runtime_error("Material file does not exist: " + p.string());
    char magic[8];
    in.read(magic, 8);
    if (!in || std::memcmp(magic, kMagic, 8) != 0) throw std::runtime_error("Not a PhonoMC material table: " + p.string());
    RawMaterial m;
    get(in, m.volume_a3);
    get(in, m.reciprocal);
    get(in, m.qpoints);
    get(in, m.branches);
    get(in, m.mesh);
    get_vec(in, m.frequency_thz);
    get_vec(in, m.qpoint_fractions);
    get_vec(in, m.group_velocity);
    get_vec(in, m.temperatures);
    get_vec(in, m.gamma);
    get_vec(in, m.gamma_normal);
    get_vec(in, m.gamma_umklapp);
    get_vec(in, m.gamma_isotope);
    validate_raw_material(m);
    return m;
}

} // namespace pmc
