// pmc_import_material FOLDER OUT.pmcmat: a phono3py material folder (POSCAR and
// kappa-fbz.hdf5, kappa.hdf5 or one other *.hdf5) converted to the raw binary
// material table that load_material reads, so that hosts without the HDF5
// library (the GPU boxes) run real materials.  Follows the reference's reader
// (read_material_data and read_poscar_lattice, src/material/MaterialDataReader.cpp:84-302):
// the same file choice, datasets, shapes, unit-weight check and messages.
// Built only where the HDF5 C++ library is installed.
#include "material.hpp"

#include <H5Cpp.h>

#include <algorithm>
#include <cmath>
#include <filesystem>
#include <fstream>
#include <iostream>
#include <limits>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace {
namespace fs = std::filesystem;

std::vector<double> dataset(const H5::H5File& file, const std::string& name, std::vector<hsize_t>& dims) {
    if (H5Lexists(file.getId(), name.c_str(), H5P_DEFAULT) <= 0) throw std::runtime_error("Missing HDF5 dataset: " + name);
    H5::DataSet ds = file.openDataSet(name);
    H5::DataSpace sp = ds.getSpace();
    const int nd = sp.getSimpleExtentNdims();
    if (nd <= 0) throw std::runtime_error("Dataset " + name + " has invalid rank");
    dims.assign(static_cast<size_t>(nd), 0);
    sp.getSimpleExtentDims(dims.data());
    size_t n = 1;
    for (hsize_t d : dims) {
        if (d == 0 || d > std::numeric_limits<size_t>::max() / n)
            throw std::runtime_error("Dataset " + name + " is empty or too large");
        n *= static_cast<size_t>(d);
    }
    std::vector<double> out(n);
    ds.read(out.data(), H5::PredType::NATIVE_DOUBLE);
    return out;
}

bool has(const H5::H5File& file, const char* name) { return H5Lexists(file.getId(), name, H5P_DEFAULT) > 0; }

// Lattice vectors as rows (scaled), unit-cell volume |det A| and the angular
// reciprocal lattice 2 pi A^-1.
void read_poscar(const fs::path& path, pmc::RawMaterial& m) {
    if (!fs::exists(path)) throw std::runtime_error("POSCAR not found: " + path.string());
    std::ifstream in(path);
    if (!in) throw std::runtime_error("Failed to open POSCAR: " + path.string());
    std::vector<std::string> lines;
    for (std::string line; std::getline(in, line);) {
        const auto b = line.find_first_not_of(" \t\r\n");
        if (b != std::string::npos) lines.push_back(line.substr(b));
    }
    if (lines.size() < 5) throw std::runtime_error("POSCAR has insufficient lines.");
    const double scale = std::stod(lines[1]);
    if (!std::isfinite(scale)) throw std::runtime_error("Non-finite POSCAR scale.");
    double a[3][3];
    for (int i = 0; i < 3; ++i) {
        std::istringstream is(lines[2 + i]);
        if (!(is >> a[i][0] >> a[i][1] >> a[i][2])) throw std::runtime_error("Failed to parse POSCAR lattice vector.");
        for (double& x : a[i]) {
            x *= scale;
            if (!std::isfinite(x)) throw std::runtime_error("Non-finite POSCAR lattice.");
        }
    }
    const double det = a[0][0] * (a[1][1] * a[2][2] - a[1][2] * a[2][1]) -
                       a[0][1] * (a[1][0] * a[2][2] - a[1][2] * a[2][0]) +
                       a[0][2] * (a[1][0] * a[2][1] - a[1][1] * a[2][0]);
    m.volume_a3 = std::abs(det);
    if (!std::isfinite(m.volume_a3) || !(m.volume_a3 > 0.0))
        throw std::runtime_error("POSCAR unit-cell volume is non-positive.");
    if (std::abs(det) <= 1e-20) throw std::runtime_error("Failed to invert POSCAR lattice.");
    // adjugate over the determinant, times 2 pi
    const double c = 2.0 * M_PI / det;
    m.reciprocal[0] = {(a[1][1] * a[2][2] - a[1][2] * a[2][1]) * c, (a[0][2] * a[2][1] - a[0][1] * a[2][2]) * c,
                       (a[0][1] * a[1][2] - a[0][2] * a[1][1]) * c};
    m.reciprocal[1] = {(a[1][2] * a[2][0] - a[1][0] * a[2][2]) * c, (a[0][0] * a[2][2] - a[0][2] * a[2][0]) * c,
                       (a[0][2] * a[1][0] - a[0][0] * a[1][2]) * c};
    m.reciprocal[2] = {(a[1][0] * a[2][1] - a[1][1] * a[2][0]) * c, (a[0][1] * a[2][0] - a[0][0] * a[2][1]) * c,
                       (a[0][0] * a[1][1] - a[0][1] * a[1][0]) * c};
    for (const auto& row : m.reciprocal)
        for (double x : row)
            if (!std::isfinite(x)) throw std::runtime_error("Non-finite reciprocal lattice.");
}

fs::path hdf5_file(const fs::path& folder) {
    for (const char* name : {"kappa-fbz.hdf5", "kappa.hdf5"})
        if (fs::exists(folder / name)) return folder / name;
    std::vector<fs::path> found;
    for (const auto& e : fs::directory_iterator(folder))
        if (e.is_regular_file() && e.path().extension() == ".hdf5") found.push_back(e.path());
    if (found.size() > 1) throw std::runtime_error("Ambiguous HDF5 files: provide kappa-fbz.hdf5 or kappa.hdf5.");
    if (found.empty()) throw std::runtime_error("No HDF5 file found under material folder: " + folder.string());
    return found.front();
}

pmc::RawMaterial import_folder(const fs::path& folder) {
    pmc::RawMaterial m;
    try {
        read_poscar(folder / "POSCAR", m);
    } catch (const std::exception& e) {
        throw std::runtime_error(std::string("Strict POSCAR loading failed: ") + e.what());
    }
    H5::H5File file(hdf5_file(folder).string(), H5F_ACC_RDONLY);
    std::vector<hsize_t> d_omega, d_q, d_mesh, d_v, d_t, d_g;
    m.frequency_thz = dataset(file, "frequency", d_omega);
    if (d_omega.size() != 2) throw std::runtime_error("frequency rank mismatch");
    if (d_omega[0] > static_cast<hsize_t>(std::numeric_limits<int>::max()) / d_omega[1])
        throw std::runtime_error("frequency has too many modes for integer mode indices");
    m.qpoints = static_cast<int>(d_omega[0]);
    m.branches = static_cast<int>(d_omega[1]);
    const auto q = dataset(file, "qpoint", d_q);
    if (d_q.size() != 2 || d_q[0] != d_omega[0] || d_q[1] != 3) throw std::runtime_error("qpoint shape mismatch");
    m.qpoint_fractions.resize(m.qpoints);
    for (int k = 0; k < m.qpoints; ++k) m.qpoint_fractions[k] = {q[3 * k], q[3 * k + 1], q[3 * k + 2]};
    const auto mesh = dataset(file, "mesh", d_mesh);
    if (d_mesh.size() != 1 || d_mesh[0] != 3) throw std::runtime_error("mesh must have shape (3)");
    for (int a = 0; a < 3; ++a) {
        const double c = mesh[a];
        if (!std::isfinite(c) || c < 1 || c != std::floor(c) || c > std::numeric_limits<int>::max())
            throw std::runtime_error("mesh entries must be positive integers");
        m.mesh[a] = static_cast<int>(c);
    }
    if (has(file, "weight")) {
        std::vector<hsize_t> d_w;
        const auto w = dataset(file, "weight", d_w);
        if (d_w.size() != 1 || d_w[0] != d_omega[0] || std::any_of(w.begin(), w.end(), [](double x) { return x != 1.0; }))
            throw std::runtime_error("Full Brillouin-zone data requires one unit weight per qpoint");
    }
    const auto v = dataset(file, "group_velocity", d_v);
    if (d_v.size() != 3 || d_v[0] != d_omega[0] || d_v[1] != d_omega[1] || d_v[2] != 3)
        throw std::runtime_error("group_velocity shape mismatch");
    m.group_velocity.resize(v.size() / 3);
    for (size_t k = 0; k < m.group_velocity.size(); ++k) m.group_velocity[k] = {v[3 * k], v[3 * k + 1], v[3 * k + 2]};
    m.temperatures = dataset(file, "temperature", d_t);
    if (d_t.size() != 1) throw std::runtime_error("temperature rank mismatch");
    m.gamma = dataset(file, "gamma", d_g);
    if (d_g.size() != 3 || d_g[0] != m.temperatures.size() || d_g[1] != d_omega[0] || d_g[2] != d_omega[1])
        throw std::runtime_error("gamma shape mismatch");
    for (const auto& [name, dst] : {std::make_pair("gamma_N", &m.gamma_normal), std::make_pair("gamma_U", &m.gamma_umklapp)}) {
        if (!has(file, name)) continue;
        std::vector<hsize_t> dims;
        *dst = dataset(file, name, dims);
        if (dims != d_g) throw std::runtime_error(std::string(name) + " shape mismatch");
    }
    // Extra resistive linewidths: the raw table holds one temperature-independent channel.
    for (const char* name : {"gamma_isotope", "gamma_impurity", "gamma_defect"}) {
        if (!has(file, name)) continue;
        std::vector<hsize_t> dims;
        auto values = dataset(file, name, dims);
        if (dims != d_omega && dims != d_g)
            throw std::runtime_error(std::string(name) +
                                     " shape mismatch: expected (q, branch) or (temperature, q, branch)");
        if (dims != d_omega || !m.gamma_isotope.empty())
            throw std::runtime_error(std::string(name) +
                                     ": this table format keeps one (q, branch) resistive linewidth only");
        m.gamma_isotope = std::move(values);
    }
    pmc::validate_raw_material(m);
    return m;
}
} // namespace

int main(int argc, char** argv) {
    if (argc != 3) {
        std::cerr << "usage: pmc_import_material MATERIAL_FOLDER OUT.pmcmat\n";
        return 2;
    }
    try {
        const auto m = import_folder(argv[1]);
        pmc::save_raw_material(m, argv[2]);
        std::cout << "imported " << m.qpoints << " q points x " << m.branches << " branches, " << m.temperatures.size()
                  << " temperatures" << (m.gamma_normal.empty() ? "" : ", N/U split")
                  << (m.gamma_isotope.empty() ? "" : ", isotope linewidth") << " -> " << argv[2] << '\n';
    } catch (const H5::Exception& e) {
        std::cerr << "Fatal error: HDF5: " << e.getDetailMsg() << '\n';
        return 1;
    } catch (const std::exception& e) {
        std::cerr << "Fatal error: " << e.what() << '\n';
        return 1;
    }
    return 0;
}
