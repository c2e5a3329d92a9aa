#include "config.hpp"

#include <algorithm>
#include <cctype>
#include <cmath>
#include <filesystem>
#include <fstream>
#include <limits>
#include <map>
#include <set>
#include <sstream>
#include <stdexcept>

namespace pmc {
namespace {
std::string trim(const std::string& s) {
    size_t a = 0, b = s.size();
    while (a < b && std::isspace(static_cast<unsigned char>(s[a]))) ++a;
    while (b > a && std::isspace(static_cast<unsigned char>(s[b - 1]))) --b;
    return s.substr(a, b - a);
}

std::string lower(std::string s) {
    for (auto& c : s) c = static_cast<char>(std::tolower(static_cast<unsigned char>(c)));
    return s;
}

// Scans s once, tracking quotes and escapes; calls on(c, depth, in_quote) per char.
template <class F>
void scan(const std::string& s, F&& on) {
    bool in_quote = false, escape = false;
    char qc = 0;
    int depth = 0;
    for (size_t i = 0; i < s.size(); ++i) {
        const char c = s[i];
        if (escape) { escape = false; on(i, c, depth, true); continue; }
        if (in_quote && c == '\\') { escape = true; on(i, c, depth, true); continue; }
        if ((c == '"' || c == '\'') && (!in_quote || c == qc)) {
            in_quote = !in_quote;
            qc = in_quote ? c : 0;
            on(i, c, depth, true);
            continue;
        }
        if (!in_quote) {
            if (c == '[') ++depth;
            else if (c == ']') --depth;
        }
        if (!on(i, c, depth, in_quote)) return;
    }
}

std::string strip_comment(const std::string& line) {
    size_t cut = line.size();
    scan(line, [&](size_t i, char c, int, bool quoted) {
        if (!quoted && c == '#') { cut = i; return false; }
        return true;
    });
    return line.substr(0, cut);
}

int bracket_balance(const std::string& s) {
    int depth = 0;
    scan(s, [&](size_t, char, int d, bool) { depth = d; return true; });
    return depth;
}

bool quoted(const std::string& s) {
    return s.size() >= 2 && ((s.front() == '"' && s.back() == '"') || (s.front() == '\'' && s.back() == '\''));
}

std::string unquote(const std::string& s) {
    const std::string t = trim(s);
    return quoted(t) ? t.substr(1, t.size() - 2) : t;
}

double to_double(const std::string& s) {
    std::string t = trim(unquote(s));
    t.erase(std::remove(t.begin(), t.end(), '_'), t.end());
    if (t.empty()) throw std::runtime_error("Expected numeric value, got empty token.");
    size_t used = 0;
    double v;
    try {
        v = std::stod(t, &used);
    } catch (const std::invalid_argument&) {
        throw std::runtime_error("Expected a numeric value, got: " + s);
    } catch (const std::out_of_range&) {
        throw std::runtime_error("Numeric value out of range: " + s);
    }
    if (used != t.size() || !std::isfinite(v)) throw std::runtime_error("Expected a finite numeric value, got: " + s);
    return v;
}

int to_int(const std::string& s) {
    const double v = to_double(s);
    const double r = std::round(v);
    if (std::abs(v - r) > 1e-12 || r < std::numeric_limits<int>::min() || r > std::numeric_limits<int>::max())
        throw std::runtime_error("Expected an integer value, got: " + s);
    return static_cast<int>(r);
}

std::uint64_t to_u64(const std::string& s) {
    std::string t = trim(unquote(s));
    t.erase(std::remove(t.begin(), t.end(), '_'), t.end());
    const char* msg = "random_seed must be an unsigned 64-bit integer.";
    if (t.empty() || t.front() == '-') throw std::runtime_error(msg);
    size_t used = 0;
    unsigned long long v = 0;
    try {
        v = std::stoull(t, &used, 10);
    } catch (const std::exception&) {
        throw std::runtime_error(msg);
    }
    if (used != t.size()) throw std::runtime_error(msg);
    return v;
}

bool to_bool(const std::string& s) {
    const std::string t = lower(trim(s));
    if (t == "true") return true;
    if (t == "false") return false;
    throw std::runtime_error("Expected TOML boolean true or false, got: " + s);
}

std::vector<std::string> split_top(const std::string& s) {
    std::vector<std::string> out;
    std::string cur;
    scan(s, [&](size_t, char c, int depth, bool quoted_char) {
        if (!quoted_char && c == ',' && depth == 0) {
            out.push_back(trim(cur));
            cur.clear();
        } else {
            cur.push_back(c);
        }
        return true;
    });
    if (!trim(cur).empty()) out.push_back(trim(cur));
    return out;
}

std::vector<std::string> array_elements(const std::string& v) {
    const std::string t = trim(v);
    if (t.size() < 2 || t.front() != '[' || t.back() != ']') throw std::runtime_error("Expected TOML array value, got: " + t);
    const std::string inner = trim(t.substr(1, t.size() - 2));
    if (inner.empty()) return {};
    return split_top(inner);
}

std::vector<double> numbers(const std::string& v) {
    std::vector<double> out;
    for (const auto& e : array_elements(v)) out.push_back(to_double(e));
    return out;
}

std::vector<std::string> strings(const std::string& v) {
    std::vector<std::string> out;
    for (const auto& e : array_elements(v)) out.push_back(unquote(e));
    return out;
}

std::vector<std::array<double, 6>> regions(const std::string& v, const std::string& field) {
    std::vector<std::array<double, 6>> out;
    for (const auto& row : array_elements(v)) {
        const auto c = numbers(row);
        if (c.size() != 6)
            throw std::runtime_error(field + " row must be [xmin,ymin,zmin,xmax,ymax,zmax]. Point "
                                             "selectors [x,y,z] are no longer supported.");
        out.push_back({c[0], c[1], c[2], c[3], c[4], c[5]});
    }
    return out;
}

const std::set<std::string>& allowed_sections() {
    static const std::set<std::string> s{"sampling", "convergence", "geometry", "simulation", "boundary", "heat_source",
                                         "mode_excitation", "materials", "io", "scattering", "driving"};
    return s;
}

const std::set<std::string>& allowed_keys() {
    // Allow-list of src/config/ConfigParsing.cpp:403-480.
    static const std::set<std::string> s{
        "geometry.model", "geometry.sizes", "geometry.merge_coplanar_facets", "geometry.extruded_z",
        "simulation.particle_count", "simulation.time_step", "simulation.iterations",
        "simulation.convergence_write_interval", "simulation.random_seed", "simulation.compute_kappa",
        "simulation.write_cell_heat_flux", "simulation.profile_timers",
        "simulation.progress_temperature_summary_only", "simulation.temperature_lookup_dt",
        "simulation.background_temperature", "simulation.lifetime_temperature", "simulation.initial_temperature",
        "simulation.grid_xyz", "simulation.transport_axis", "driving.temperature_gradient",
        "sampling.resample_interval", "sampling.per_mode_sign", "driving.warm_start_ps", "convergence.enabled",
        "convergence.min_time_ps", "convergence.window_ps", "convergence.windows", "convergence.relative_tolerance",
        "convergence.absolute_tolerance", "scattering.model", "scattering.matrix_files",
        "scattering.callaway_formulation", "simulation.conservative_boundaries", "boundary.boundary_conditions",
        "boundary.boundary_values", "boundary.boundary_position", "boundary.periodic_pair", "heat_source.enabled",
        "heat_source.min", "heat_source.max", "heat_source.power_density", "heat_source.power_densities",
        "heat_source.total_power", "heat_source.spectrum", "heat_source.branches", "heat_source.branch_weights",
        "heat_source.frequency_min", "heat_source.frequency_max", "heat_source.frequency_center",
        "heat_source.frequency_sigma", "heat_source.profile", "heat_source.center", "heat_source.centers",
        "heat_source.sigma", "heat_source.half_width", "heat_source.time_profile", "heat_source.start_time",
        "heat_source.end_time", "heat_source.period", "heat_source.on_duration", "heat_source.duty_cycle",
        "heat_source.amplitude", "mode_excitation.enabled", "mode_excitation.frequency_min",
        "mode_excitation.frequency_max", "mode_excitation.occupation_multiplier", "materials.folders",
        "materials.interface_axis", "materials.interface_positions", "materials.interface_model",
        "materials.interface_frequency_bin_thz", "materials.interface_amm_fraction",
        "materials.interface_parallel_bin_inv_a", "materials.mass_densities_kg_m3", "io.material_folder",
        "io.output_folder"};
    return s;
}

std::map<std::string, std::string> read_assignments(std::istream& in) {
    std::map<std::string, std::string> kv;
    std::string section, line;
    while (std::getline(in, line)) {
        const std::string clean = trim(strip_comment(line));
        if (clean.empty()) continue;
        if (clean.front() == '[' && clean.back() == ']') {
            section = lower(trim(clean.substr(1, clean.size() - 2)));
            if (!allowed_sections().count(section)) throw std::runtime_error("Unknown TOML section: [" + section + "]");
            continue;
        }
        const auto eq = clean.find('=');
        if (eq == std::string::npos) throw std::runtime_error("Invalid TOML assignment: " + clean);
        const std::string key = lower(trim(clean.substr(0, eq)));
        std::string value = trim(clean.substr(eq + 1));
        if (value.empty()) throw std::runtime_error("Empty TOML value for key: " + key);
        if (section.empty())
            throw std::runtime_error("Top-level TOML keys are not supported; place '" + key +
                                     "' in its documented section.");
        if (value.front() == '[') {
            while (bracket_balance(value) > 0) {
                std::string next;
                if (!std::getline(in, next)) break;
                next = trim(strip_comment(next));
                if (!next.empty()) value += " " + next;
            }
            if (bracket_balance(value) != 0) throw std::runtime_error("Unclosed TOML array for key: " + section + "." + key);
        }
        const std::string full = section + "." + key;
        if (!kv.emplace(full, value).second) throw std::runtime_error("Duplicate TOML key: " + full);
    }
    for (const auto& [k, v] : kv)
        if (!allowed_keys().count(k)) throw std::runtime_error("Unknown TOML key: " + k);
    return kv;
}

SimConfig build(const std::map<std::string, std::string>& kv, const std::string& input_directory) {
    SimConfig c;
    c.input_directory = input_directory;
    auto get = [&](const char* k) -> const std::string* {
        auto it = kv.find(k);
        return it == kv.end() ? nullptr : &it->second;
    };
    if (auto v = get("geometry.model")) {
        c.model = unquote(*v);
        if (lower(c.model) == "cylinder")
            throw std::runtime_error("geometry.model='cylinder' is no longer supported; use 'box' or a mesh file.");
    }
    if (auto v = get("geometry.sizes")) {
        c.sizes = numbers(*v);
        for (double& x : c.sizes) x *= 10.0;
    }
    if (auto v = get("geometry.merge_coplanar_facets")) c.merge_coplanar_facets = to_bool(*v);
    if (auto v = get("geometry.extruded_z"); v && to_bool(*v))
        throw std::runtime_error("geometry.extruded_z prisms are not supported by this build.");
    if (auto v = get("simulation.particle_count")) c.particle_count = to_double(*v);
    if (auto v = get("simulation.time_step")) c.time_step = to_double(*v);
    if (auto v = get("simulation.iterations")) c.iterations = to_int(*v);
    if (auto v = get("simulation.convergence_write_interval")) c.convergence_write_interval = to_int(*v);
    if (auto v = get("simulation.random_seed")) c.random_seed = to_u64(*v);
    if (auto v = get("simulation.compute_kappa")) c.compute_kappa = to_bool(*v);
    if (auto v = get("simulation.transport_axis")) {
        const auto a = lower(unquote(trim(*v)));
        if (a != "x" && a != "y" && a != "z") throw std::runtime_error("simulation.transport_axis must be x, y, or z.");
        c.transport_axis = a == "x" ? 0 : a == "y" ? 1 : 2;
    }
    if (get("driving.temperature_gradient")) c.has_gradient_drive = true;
    if (auto v = get("sampling.resample_interval"); v && to_int(*v) != 0)
        throw std::runtime_error("Solver resampling is unsupported: sampling.resample_interval must be 0. "
                                 "Existing-carrier merging changes quadrature weights.");
    if (auto v = get("convergence.enabled")) c.has_convergence_stop = to_bool(*v);
    if (auto v = get("simulation.profile_timers")) c.profile_timers = to_bool(*v);
    if (auto v = get("simulation.progress_temperature_summary_only")) c.progress_temperature_summary_only = to_bool(*v);
    if (auto v = get("simulation.temperature_lookup_dt")) c.temperature_lookup_dt = to_double(*v);
    if (auto v = get("simulation.background_temperature")) {
        if (quoted(trim(*v)))
            throw std::runtime_error("background_temperature must be an unquoted finite non-negative number; "
                                     "the local background mode has been removed.");
        c.background_temperature = to_double(*v);
    }
    if (auto v = get("simulation.lifetime_temperature")) {
        const std::string raw = trim(*v);
        if (quoted(raw)) {
            if (lower(unquote(raw)) != "local")
                throw std::runtime_error("lifetime_temperature string value must be 'local'. Numeric fixed "
                                         "temperatures must not be quoted.");
            c.lifetime_local = true;
            c.lifetime_temperature = 300.0;
        } else {
            c.lifetime_local = false;
            c.lifetime_temperature = to_double(raw);
        }
    }
    if (auto v = get("scattering.model")) {
        const auto m = lower(unquote(trim(*v)));
        if (m == "rta") c.collision = CollisionModel::Rta;
        else if (m == "callaway") c.collision = CollisionModel::Callaway;
        else throw std::runtime_error("scattering.model must be 'rta' or 'callaway'.");
    }
    if (auto v = get("scattering.callaway_formulation")) {
        const auto f = lower(unquote(trim(*v)));
        if (f != "linearized" && f != "nonlinear") throw std::runtime_error("callaway_formulation must be linearized or nonlinear");
        c.callaway_nonlinear = f == "nonlinear";
    }
    if (auto v = get("simulation.write_cell_heat_flux")) c.write_cell_heat_flux = to_bool(*v);
    if (auto v = get("simulation.conservative_boundaries")) c.conservative_boundaries = to_bool(*v);
    if (auto v = get("scattering.matrix_files"); v && !strings(*v).empty())
        throw std::runtime_error("scattering.matrix_files is not supported by the public rta/callaway models.");
    if (auto v = get("simulation.initial_temperature")) {
        const std::string t = trim(*v);
        if (!t.empty() && t.front() == '[')
            throw std::runtime_error("simulation.initial_temperature must be a numeric scalar or 'linear', not an array.");
        if (lower(unquote(t)) == "linear") {
            c.initial_linear = true;
        } else {
            if (quoted(t)) throw std::runtime_error("Numeric initial_temperature values must not be quoted.");
            c.initial_linear = false;
            c.initial_uniform_temperature = to_double(t);
        }
    }
    if (auto v = get("simulation.grid_xyz")) {
        const auto g = numbers(*v);
        if (g.size() != 3) throw std::runtime_error("grid_xyz must contain 3 values [nx, ny, nz].");
        for (double x : g)
            if (std::abs(x - std::round(x)) > 1e-12 || x < 1 || x > std::numeric_limits<int>::max())
                throw std::runtime_error("grid_xyz values must be positive representable integers.");
        c.grid = {static_cast<int>(std::llround(g[0])), static_cast<int>(std::llround(g[1])),
                  static_cast<int>(std::llround(g[2]))};
    } else {
        throw std::runtime_error("Missing required key: simulation.grid_xyz.");
    }
    if (auto v = get("boundary.boundary_conditions")) {
        size_t i = 0;
        for (const auto& s : strings(*v)) {
            const auto t = lower(trim(unquote(s)));
            if (t != "t" && t != "p" && t != "r")
                throw std::runtime_error("Unsupported boundary condition at index " + std::to_string(i) + ": '" + s +
                                         "'. Only 'T', 'P', 'R' are supported.");
            c.boundary_conditions.push_back(static_cast<char>(std::toupper(t[0])));
            ++i;
        }
    }
    if (auto v = get("boundary.boundary_values")) c.boundary_values = numbers(*v);
    if (auto v = get("boundary.boundary_position")) c.boundary_boxes = regions(*v, "boundary_position");
    if (auto v = get("boundary.periodic_pair")) c.periodic_pairs = regions(*v, "periodic_pair");
    if (auto v = get("io.material_folder")) c.material_folder = unquote(*v);
    if (auto v = get("io.output_folder")) c.output_folder = unquote(*v);
    {
        // src/config/ConfigParsing.cpp:600-656.
        const bool other = get("materials.interface_axis") || get("materials.interface_positions") ||
                           get("materials.interface_model") || get("materials.interface_frequency_bin_thz") ||
                           get("materials.interface_amm_fraction") || get("materials.interface_parallel_bin_inv_a") ||
                           get("materials.mass_densities_kg_m3");
        const auto folders = get("materials.folders");
        if (!folders && other)
            throw std::runtime_error("materials.folders is required when any other [materials] key is present.");
        if (folders) {
            c.has_layered_materials = true;
            for (const auto& f : strings(*folders)) c.material_folders.push_back(unquote(trim(f)));
            if (c.material_folders.empty()) throw std::runtime_error("materials.folders must not be empty.");
        }
        if (auto v = get("materials.interface_axis")) {
            const auto a = lower(unquote(*v));
            if (a != "x" && a != "y" && a != "z") throw std::runtime_error("materials.interface_axis must be x, y, or z.");
            c.interface_axis = a == "x" ? 0 : a == "y" ? 1 : 2;
        }
        if (auto v = get("materials.interface_positions"))
            for (double x : numbers(*v)) c.interface_positions.push_back(10.0 * x);
        if (auto v = get("materials.interface_model")) {
            // parse_interface_model (src/config/ConfigParsing.cpp:157-169): long or short names
            const auto m = lower(unquote(*v));
            if (m == "dmm" || m == "diffuse_mismatch") c.interface_model = "dmm";
            else if (m == "amm" || m == "acoustic_mismatch") c.interface_model = "amm";
            else if (m == "mmm" || m == "mixed_mismatch") c.interface_model = "mmm";
            else throw std::runtime_error("materials.interface_model must be dmm, amm or mmm.");
        }
        if (auto v = get("materials.interface_frequency_bin_thz")) c.interface_bin_thz = to_double(*v);
        if (auto v = get("materials.interface_amm_fraction")) c.interface_amm_fraction = to_double(*v);
        if (auto v = get("materials.interface_parallel_bin_inv_a")) c.interface_parallel_bin = to_double(*v);
        if (auto v = get("materials.mass_densities_kg_m3")) c.mass_densities = numbers(*v);
        if (c.interface_model == "mmm" && !get("materials.interface_amm_fraction"))
            throw std::runtime_error("MMM requires explicit materials.interface_amm_fraction.");
        if (c.interface_model == "amm" && get("materials.interface_amm_fraction") && c.interface_amm_fraction != 1)
            throw std::runtime_error("AMM explicit fraction must equal 1; use MMM for intermediate weights.");
    }
    if (auto v = get("mode_excitation.enabled")) c.has_mode_excitation = to_bool(*v);
    if (auto v = get("mode_excitation.frequency_min")) c.excite_fmin = to_double(*v);
    if (auto v = get("mode_excitation.frequency_max")) c.excite_fmax = to_double(*v);
    if (auto v = get("mode_excitation.occupation_multiplier")) c.excite_multiplier = to_double(*v);

    // Heat source (src/config/ConfigParsing.cpp:658-766).
    bool enabled_set = false;
    if (auto v = get("heat_source.enabled")) { c.heat_source_enabled = to_bool(*v); enabled_set = true; }
    if (auto v = get("heat_source.min")) c.heat_min = numbers(*v);
    if (auto v = get("heat_source.max")) c.heat_max = numbers(*v);
    if (auto v = get("heat_source.total_power")) {
        if (get("heat_source.power_density") || get("heat_source.power_densities"))
            throw std::runtime_error("Specify either total_power or density inputs, not both.");
        c.heat_total_power = to_double(*v);
    }
    if (auto v = get("heat_source.spectrum")) {
        const auto s = lower(unquote(*v));
        if (s == "thermal") c.heat_spectrum = 0;
        else if (s == "weighted") c.heat_spectrum = 1;
        else if (s == "gaussian") c.heat_spectrum = 2;
        else throw std::runtime_error("heat_source.spectrum must be 'thermal', 'weighted', or 'gaussian'.");
    }
    if (auto v = get("heat_source.branches")) {
        for (double b : numbers(*v)) {
            if (!std::isfinite(b) || b < 0 || b != std::floor(b) || b > 2147483647.0)
                throw std::runtime_error("heat_source.branches must contain nonnegative integer indices.");
            c.heat_branches.push_back(static_cast<int>(b));
        }
    }
    if (auto v = get("heat_source.branch_weights")) c.heat_branch_weights = numbers(*v);
    if (auto v = get("heat_source.frequency_center")) c.heat_frequency_center = to_double(*v);
    if (auto v = get("heat_source.frequency_sigma")) c.heat_frequency_sigma = to_double(*v);
    if (auto v = get("heat_source.power_density")) c.heat_power_density = to_double(*v);
    if (auto v = get("heat_source.power_densities")) c.heat_power_densities = numbers(*v);
    if (auto v = get("heat_source.frequency_min")) c.heat_frequency_min = to_double(*v);
    if (auto v = get("heat_source.frequency_max")) c.heat_frequency_max = to_double(*v);
    if (auto v = get("heat_source.profile")) {
        const auto p = lower(trim(unquote(*v)));
        if (p == "uniform") c.heat_profile = HeatProfile::Uniform;
        else if (p == "gaussian") c.heat_profile = HeatProfile::Gaussian;
        else throw std::runtime_error("heat_source.profile must be either 'uniform' or 'gaussian'.");
    }
    if (auto v = get("heat_source.center")) c.heat_center = numbers(*v);
    if (auto v = get("heat_source.centers")) {
        for (const auto& row : array_elements(*v)) {
            const auto xyz = numbers(row);
            if (xyz.size() != 3) throw std::runtime_error("heat_source.centers row must be [x,y,z].");
            c.heat_centers.insert(c.heat_centers.end(), xyz.begin(), xyz.end());
        }
    }
    if (auto v = get("heat_source.sigma")) c.heat_sigma = numbers(*v);
    if (auto v = get("heat_source.half_width")) c.heat_half_width = numbers(*v);
    if (auto v = get("heat_source.time_profile")) {
        const auto p = lower(trim(unquote(*v)));
        if (p == "constant") c.heat_time_square = false;
        else if (p == "square") c.heat_time_square = true;
        else throw std::runtime_error("heat_source.time_profile must be either 'constant' or 'square'.");
    }
    if (auto v = get("heat_source.start_time")) c.heat_time_start = to_double(*v);
    if (auto v = get("heat_source.end_time")) c.heat_time_end = to_double(*v);
    if (auto v = get("heat_source.period")) c.heat_period = to_double(*v);
    if (auto v = get("heat_source.on_duration")) c.heat_on_duration = to_double(*v);
    if (auto v = get("heat_source.duty_cycle")) c.heat_duty_cycle = to_double(*v);
    if (auto v = get("heat_source.amplitude")) c.heat_amplitude = to_double(*v);
    if (!enabled_set) { // infer_heat_source_enabled (src/config/ConfigParsing.cpp:503-513)
        const bool has_power = c.heat_total_power.has_value() || std::abs(c.heat_power_density) > 0.0;
        if (has_power) {
            c.heat_source_enabled = c.heat_profile == HeatProfile::Gaussian
                                        ? ((c.heat_center.size() == 3 || c.heat_centers.size() >= 3) &&
                                           c.heat_sigma.size() == 3)
                                        : (c.heat_min.size() == 3 && c.heat_max.size() == 3);
        }
    }
    for (auto* v : {&c.heat_min, &c.heat_max, &c.heat_center, &c.heat_centers, &c.heat_sigma, &c.heat_half_width})
        for (double& x : *v) x *= 10.0;
    return c;
}
} // namespace

// validate_heat_source_config (src/config/ConfigValidation.cpp:196-344), same messages.
static void validate_heat_source(const SimConfig& c) {
    if (c.heat_total_power) {
        if (!std::isfinite(*c.heat_total_power) || !(*c.heat_total_power > 0.0))
            throw std::runtime_error("heat_source.total_power must be finite and positive (W).");
        if (c.heat_power_density != 0.0 || !c.heat_power_densities.empty())
            throw std::runtime_error("heat_source.total_power cannot be combined with power_density or power_densities.");
    } else if (!(c.heat_power_density > 0.0) || !std::isfinite(c.heat_power_density)) {
        throw std::runtime_error("enabled heat_source requires positive total_power or power_density.");
    }
    for (const auto* values : {&c.heat_min, &c.heat_max, &c.heat_center, &c.heat_centers, &c.heat_sigma,
                               &c.heat_half_width})
        for (double v : *values)
            if (!std::isfinite(v)) throw std::runtime_error("Heat-source spatial parameters must be finite.");
    std::vector<int> branches = c.heat_branches;
    std::sort(branches.begin(), branches.end());
    if ((!branches.empty() && branches.front() < 0) || std::adjacent_find(branches.begin(), branches.end()) != branches.end())
        throw std::runtime_error("heat_source.branches must contain unique nonnegative indices.");
    // Spectrum parameters (src/config/ConfigValidation.cpp:232-262).
    if (c.heat_spectrum == 2) {
        if (!c.heat_frequency_center || !c.heat_frequency_sigma || !std::isfinite(*c.heat_frequency_center) ||
            *c.heat_frequency_center < 0 || !std::isfinite(*c.heat_frequency_sigma) || *c.heat_frequency_sigma <= 0)
            throw std::runtime_error("Gaussian spectrum requires finite frequency_center >= 0 and frequency_sigma > 0, in THz.");
    } else if (c.heat_frequency_center || c.heat_frequency_sigma) {
        throw std::runtime_error("frequency_center/frequency_sigma require heat_source.spectrum='gaussian'.");
    }
    if (c.heat_spectrum != 1) {
        if (!c.heat_branch_weights.empty())
            throw std::runtime_error("branch_weights requires heat_source.spectrum='weighted'.");
    } else {
        bool positive = false;
        for (size_t b = 0; b < c.heat_branch_weights.size(); ++b) {
            const double w = c.heat_branch_weights[b];
            if (!std::isfinite(w) || w < 0.0)
                throw std::runtime_error("heat_source.branch_weights must be finite and nonnegative.");
            if (branches.empty() || std::binary_search(branches.begin(), branches.end(), static_cast<int>(b)))
                positive = positive || w > 0.0;
        }
        if (!positive) throw std::runtime_error("weighted heat_source needs positive weights on selected branches.");
    }
    for (double q : c.heat_power_densities)
        if (!(q > 0.0) || !std::isfinite(q))
            throw std::runtime_error("heat_source.power_densities values must be finite and positive.");
    if (!(c.heat_frequency_min < 0.0) && !(std::isfinite(c.heat_frequency_min) && c.heat_frequency_min >= 0.0))
        throw std::runtime_error("heat_source.frequency_min must be finite and non-negative if provided.");
    if (!(c.heat_frequency_max < 0.0) && !(std::isfinite(c.heat_frequency_max) && c.heat_frequency_max >= 0.0))
        throw std::runtime_error("heat_source.frequency_max must be finite and non-negative if provided.");
    if (c.heat_frequency_min >= 0.0 && c.heat_frequency_max >= 0.0 && c.heat_frequency_max < c.heat_frequency_min)
        throw std::runtime_error("heat_source.frequency_max must be no smaller than frequency_min.");
    if (c.heat_profile == HeatProfile::Uniform) {
        if (c.heat_min.size() != 3 || c.heat_max.size() != 3)
            throw std::runtime_error("uniform heat_source requires min and max arrays with 3 values.");
    } else {
        const bool single = c.heat_center.size() == 3, multi = !c.heat_centers.empty();
        if ((!single && !multi) || c.heat_sigma.size() != 3)
            throw std::runtime_error("gaussian heat_source requires center=[x,y,z] or centers=[[x,y,z],...] and "
                                     "sigma=[sx,sy,sz].");
        if (multi && c.heat_centers.size() % 3 != 0)
            throw std::runtime_error("heat_source.centers must contain one or more [x,y,z] rows.");
        const size_t nsrc = multi ? c.heat_centers.size() / 3 : 1;
        if (!c.heat_power_densities.empty() && c.heat_power_densities.size() != nsrc)
            throw std::runtime_error("heat_source.power_densities must match the number of heat_source.centers.");
        if (!c.heat_half_width.empty() && c.heat_half_width.size() != 3)
            throw std::runtime_error("heat_source.half_width must contain 3 values if provided.");
        if ((!c.heat_min.empty() || !c.heat_max.empty()) && (c.heat_min.size() != 3 || c.heat_max.size() != 3))
            throw std::runtime_error("gaussian heat_source min/max clipping box must be omitted or contain 3 values "
                                     "each.");
    }
    if (!std::isfinite(c.heat_amplitude) || c.heat_amplitude < 0.0)
        throw std::runtime_error("heat_source.amplitude must be finite and non-negative.");
    if (!std::isfinite(c.heat_on_duration)) throw std::runtime_error("heat_source.on_duration must be finite.");
    if (!std::isfinite(c.heat_time_start)) throw std::runtime_error("heat_source.start_time must be finite.");
    if (!std::isfinite(c.heat_time_end)) throw std::runtime_error("heat_source.end_time must be finite.");
    if (!std::isfinite(c.heat_duty_cycle) || c.heat_duty_cycle < 0.0 || c.heat_duty_cycle > 1.0)
        throw std::runtime_error("heat_source.duty_cycle must be in [0, 1].");
    if (c.heat_time_square) {
        if (!(c.heat_period > 0.0) || !std::isfinite(c.heat_period))
            throw std::runtime_error("square heat_source.time_profile requires a finite positive period.");
        if (c.heat_on_duration >= 0.0 && (!std::isfinite(c.heat_on_duration) || c.heat_on_duration > c.heat_period))
            throw std::runtime_error("heat_source.on_duration must be finite and no larger than period.");
    }
}

// validate_material_config (src/config/ConfigValidation.cpp:109-193), same messages and order.
static void validate_materials(const SimConfig& c) {
    if (!std::isfinite(c.interface_bin_thz) || c.interface_bin_thz <= 0)
        throw std::runtime_error("materials.interface_frequency_bin_thz must be finite and positive.");
    if (!std::isfinite(c.interface_amm_fraction) || c.interface_amm_fraction < 0 || c.interface_amm_fraction > 1)
        throw std::runtime_error("materials.interface_amm_fraction must be in [0,1].");
    if (!std::isfinite(c.interface_parallel_bin) || c.interface_parallel_bin <= 0)
        throw std::runtime_error("materials.interface_parallel_bin_inv_a must be finite and positive.");
    if (c.interface_model == "dmm" && c.interface_amm_fraction != 0)
        throw std::runtime_error("DMM cannot have nonzero AMM fraction.");
    if (c.interface_model == "amm" || (c.interface_model == "mmm" && c.interface_amm_fraction > 0)) {
        if (c.material_folders.size() < 2 || c.mass_densities.size() != c.material_folders.size())
            throw std::runtime_error("AMM/MMM requires one mass density per material layer.");
        for (double rho : c.mass_densities)
            if (!std::isfinite(rho) || rho <= 0)
                throw std::runtime_error("Material mass densities must be finite and positive.");
    }
    if (c.material_folders.empty()) return;
    if (c.material_folders.size() < 2)
        throw std::runtime_error("materials.folders must contain at least two folders; use io.material_folder for a "
                                 "single material.");
    if (c.interface_positions.size() + 1 != c.material_folders.size())
        throw std::runtime_error("materials.interface_positions must contain exactly folders.size()-1 values.");
    for (size_t k = 1; k < c.interface_positions.size(); ++k)
        if (!(c.interface_positions[k - 1] < c.interface_positions[k]))
            throw std::runtime_error("materials.interface_positions must be strictly increasing.");
    const double extent = c.sizes.size() == 3 ? c.sizes[static_cast<size_t>(c.interface_axis)] : 0.0;
    const int cells = c.interface_axis == 0 ? c.grid[0] : c.interface_axis == 1 ? c.grid[1] : c.grid[2];
    const double width = extent / static_cast<double>(cells);
    const double tol = std::max(1e-10, 1e-10 * extent);
    for (double x : c.interface_positions) {
        if (!std::isfinite(x)) throw std::runtime_error("materials.interface_positions must be finite.");
        if (c.model != "box") continue; // mesh bounds are checked after loading
        if (!(x > 0.0) || !(x < extent))
            throw std::runtime_error("materials.interface_positions must lie strictly inside the domain along "
                                     "interface_axis.");
        const double g = x / width;
        if (std::abs(g - std::round(g)) > tol / width)
            throw std::runtime_error("materials.interface_positions must coincide with grid-cell boundaries.");
    }
    for (const auto& f : c.material_folders)
        if (trim(f).empty()) throw std::runtime_error("materials.folders entries must not be empty.");
}

void validate_config(const SimConfig& c) {
    // Subset of src/config/ConfigValidation.cpp:457-502 and :475-501 with the same messages.
    if (!(c.temperature_lookup_dt > 0.0) || !std::isfinite(c.temperature_lookup_dt))
        throw std::runtime_error("temperature_lookup_dt must be a finite positive value.");
    if (c.callaway_nonlinear && c.collision != CollisionModel::Callaway)
        throw std::runtime_error("Nonlinear Callaway requires model=callaway and finite-domain thermal driving");
    if (c.collision == CollisionModel::Callaway && !(c.background_temperature > 0))
        throw std::runtime_error("Callaway requires a positive fixed background_temperature.");
    const size_t selectors = c.boundary_boxes.size();
    if (selectors == 0) {
        if (!c.boundary_conditions.empty() || !c.boundary_values.empty() || !c.periodic_pairs.empty())
            throw std::runtime_error("boundary.boundary_position is required when boundary conditions, values, or "
                                     "periodic pairs are provided.");
    } else {
        if (c.boundary_conditions.size() != selectors)
            throw std::runtime_error("boundary.boundary_conditions must contain exactly one entry per "
                                     "boundary_position region.");
        if (c.boundary_values.size() != selectors)
            throw std::runtime_error("boundary.boundary_values must contain exactly one entry per "
                                     "boundary_position region; use 0 for periodic boundaries.");
        size_t periodic = 0;
        for (size_t i = 0; i < selectors; ++i) {
            if (c.boundary_conditions[i] == 'T' && c.boundary_values[i] < 0.0)
                throw std::runtime_error("Thermal reservoir boundary values must be non-negative temperatures.");
            if (c.boundary_conditions[i] == 'R' && c.boundary_values[i] < 0.0)
                throw std::runtime_error("Rough boundary values must be non-negative roughness values.");
            if (c.boundary_conditions[i] == 'P') {
                ++periodic;
                if (c.boundary_values[i] != 0.0) throw std::runtime_error("Periodic boundary values must be 0.");
            }
        }
        if (c.periodic_pairs.size() != periodic || c.periodic_pairs.size() % 2 != 0)
            throw std::runtime_error("boundary.periodic_pair must contain every periodic boundary region exactly "
                                     "once, in pairs.");
    }
    if (c.heat_source_enabled) validate_heat_source(c);
    // validate_mode_excitation_config (src/config/ConfigValidation.cpp:81-107).
    if (c.has_mode_excitation) {
        if (!(c.excite_multiplier > 0.0) || !std::isfinite(c.excite_multiplier))
            throw std::runtime_error("mode_excitation.occupation_multiplier must be finite and positive.");
        if (!(c.excite_fmin < 0.0) && !(std::isfinite(c.excite_fmin) && c.excite_fmin >= 0.0))
            throw std::runtime_error("mode_excitation.frequency_min must be finite and non-negative if provided.");
        if (!(c.excite_fmax < 0.0) && !(std::isfinite(c.excite_fmax) && c.excite_fmax >= 0.0))
            throw std::runtime_error("mode_excitation.frequency_max must be finite and non-negative if provided.");
        if (c.excite_fmin >= 0.0 && c.excite_fmax >= 0.0 && c.excite_fmax < c.excite_fmin)
            throw std::runtime_error("mode_excitation.frequency_max must be no smaller than frequency_min.");
    }
    validate_materials(c);
    if (!std::isfinite(c.background_temperature) || c.background_temperature < 0.0)
        throw std::runtime_error("background_temperature must be a finite non-negative value.");
    if (!std::isfinite(c.lifetime_temperature) || c.lifetime_temperature < 0.0)
        throw std::runtime_error("lifetime_temperature must be a finite non-negative value.");
    if (!c.initial_linear && (!std::isfinite(c.initial_uniform_temperature) || c.initial_uniform_temperature < 0.0))
        throw std::runtime_error("initial_temperature must be a finite non-negative temperature.");
    if (c.convergence_write_interval <= 0)
        throw std::runtime_error("convergence_write_interval must be a positive integer.");
    const double count = std::round(c.particle_count);
    if (!std::isfinite(count) || count < 1 || count > std::numeric_limits<int>::max())
        throw std::runtime_error("particle_count must round to an integer in [1, INT_MAX].");
    if (!(c.time_step > 0.0) || !std::isfinite(c.time_step))
        throw std::runtime_error("time_step must be a finite positive value.");
    if (c.iterations <= 0) throw std::runtime_error("iterations must be a positive integer.");
    if (c.transport_axis < -1 || c.transport_axis > 2) throw std::runtime_error("Invalid transport_axis.");
    if (c.model == "box" &&
        (c.sizes.size() != 3 || std::any_of(c.sizes.begin(), c.sizes.end(), [](double v) {
             return !std::isfinite(v) || !(v > 0.0);
         })))
        throw std::runtime_error("geometry.sizes for box must contain three positive values.");
}

SimConfig parse_toml_text(const std::string& text, const std::string& input_directory) {
    std::istringstream in(text);
    auto c = build(read_assignments(in), input_directory);
    validate_config(c);
    return c;
}

SimConfig parse_toml_file(const std::string& path) {
    namespace fs = std::filesystem;
    std::string ext = fs::path(path).extension().string();
    if (lower(ext) != ".toml") throw std::runtime_error("Only sectioned TOML input files (*.toml) are supported.");
    std::ifstream in(path);
    if (!in) throw std::runtime_error("Failed to open TOML input file: " + path);
    auto c = build(read_assignments(in), fs::absolute(fs::path(path)).parent_path().string());
    validate_config(c);
    return c;
}

std::string create_indexed_output_folder(const std::string& output_folder) {
    // Same naming as src/config/OutputDirectory.cpp:9-51: <stem>_<next index>.
    if (output_folder.empty()) throw std::runtime_error("output_folder cannot be empty.");
    namespace fs = std::filesystem;
    fs::path base = fs::absolute(fs::path(output_folder)).lexically_normal();
    if (base.filename().empty()) base = base.parent_path();
    const fs::path parent = base.parent_path();
    const std::string stem = base.filename().string();
    int max_idx = -1;
    if (fs::exists(parent) && fs::is_directory(parent)) {
        for (const auto& e : fs::directory_iterator(parent)) {
            if (!e.is_directory()) continue;
            const auto name = e.path().filename().string();
            const auto prefix = stem + "_";
            if (name.rfind(prefix, 0) != 0) continue;
            const auto suffix = name.substr(prefix.size());
            if (suffix.empty() || !std::all_of(suffix.begin(), suffix.end(), [](unsigned char ch) { return std::isdigit(ch); }))
                continue;
            max_idx = std::max(max_idx, std::stoi(suffix));
        }
    }
    const fs::path out = parent / (stem + "_" + std::to_string(max_idx + 1));
    fs::create_directories(out);
    return out.string();
}

} // namespace pmc
