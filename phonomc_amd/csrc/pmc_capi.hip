// C ABI of the device step (include/pmc_capi.h): owns device buffers and
// sequences the step kernels on one HIP stream.
#include "pmc_capi.h"
#include "pmc_device.hpp"
#include "pmc_kernels.hpp"
#include "pmc_nlcallaway.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

using pmc_dev::DevState;
using pmc_dev::FacetDev;
using pmc_dev::CellRta;
using pmc_dev::ResDev;
using pmc_dev::RoughDev;
using pmc_dev::View;

namespace {
struct Carriers {
    float *px = nullptr, *py = nullptr, *pz = nullptr;
    int* mode = nullptr;
    double* occ = nullptr;
    int* cell = nullptr;
    unsigned* uid = nullptr;
};
} // namespace

struct pmc_ctx {
    pmc_setup setup{};
    hipStream_t stream = nullptr;
    std::vector<void*> allocations;
    Carriers buf[2];
    int cur = 0;
    View view{};
    bool has_material = false, has_geometry = false, has_tables = false, initialized = false;
    bool latched = false;
    pmc_batch* owner = nullptr; // batch that steps this context (pmc_batch_create .. pmc_batch_destroy)
    pmc_error error{};
    int64_t host_step = 0;
    double elapsed_ps = 0.0;   // accumulated like the reference's elapsed_time_
    pmc_heat_source hs{};      // scalars of the time profile (pointers not kept)
    int error_interval = 8;   // steps between reads of the device error word
    int n_reservoirs = 0;
    std::vector<std::pair<int, int>> rough_range; // (mode_begin, mode_count) of every uploaded reflection table
    double expected_injection = 0;
    bool timing = false;
    hipError_t hip_error = hipSuccess; // first failure of an unchecked-in-place HIP call
    static constexpr int kEvents = 64;
    hipEvent_t ev[kEvents] = {};
    void note(hipError_t e) {
        if (e != hipSuccess && hip_error == hipSuccess) hip_error = e;
    }
    double kernel_ms[PMC_KERNEL_CLASSES] = {};
    int64_t kernel_launches[PMC_KERNEL_CLASSES] = {};

    template <class T>
    T* alloc(size_t n) {
        void* p = nullptr;
        if (n == 0) n = 1;
        if (hipMalloc(&p, n * sizeof(T)) != hipSuccess) {
            note(hipErrorOutOfMemory);
            return nullptr;
        }
        allocations.push_back(p);
        return static_cast<T*>(p);
    }
    template <class T>
    T* upload(const T* host, size_t n) {
        T* d = alloc<T>(n);
        if (d && host && n) note(hipMemcpy(d, host, n * sizeof(T), hipMemcpyHostToDevice));
        return d;
    }
};

namespace {
int fail(pmc_ctx* ctx, int code, const std::string& msg) {
    if (ctx) {
        ctx->error.code = code;
        std::snprintf(ctx->error.message, sizeof(ctx->error.message), "%s", msg.c_str());
    }
    return code;
}

int check_hip(pmc_ctx* ctx, hipError_t e, const char* what) {
    if (e == hipSuccess && ctx) e = ctx->hip_error;
    if (e == hipSuccess) return PMC_OK;
    return fail(ctx, PMC_ERR_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

// View with the current carrier buffer as source and the other one as the
// destination of the cell sort.
View& view(pmc_ctx* ctx) {
    View& v = ctx->view;
    const Carriers& c = ctx->buf[ctx->cur];
    v.px = c.px;
    v.py = c.py;
    v.pz = c.pz;
    v.mode = c.mode;
    v.occ = c.occ;
    v.cell = c.cell;
    v.uid = c.uid;
    const Carriers& d = ctx->buf[1 - ctx->cur];
    v.o_px = d.px;
    v.o_py = d.py;
    v.o_pz = d.pz;
    v.o_mode = d.mode;
    v.o_occ = d.occ;
    v.o_cell = d.cell;
    v.o_uid = d.uid;
    return v;
}

// Reads the device error word and latches the context with the reference's message.
int collect_error(pmc_ctx* ctx) {
    DevState st{};
    hipError_t e = hipMemcpyAsync(&st, ctx->view.st, sizeof(DevState), hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) {
        ctx->latched = true;
        return check_hip(ctx, e, "step");
    }
    if (st.error_code == 0) return PMC_OK;
    ctx->latched = true;
    ctx->error.code = st.error_code;
    ctx->error.carrier = st.error_slot;
    std::memcpy(ctx->error.position, st.error_pos, sizeof(st.error_pos));
    ctx->error.facet = st.error_facet;
    char msg[256];
    const long long i = st.error_slot;
    const double x = st.error_pos[0], y = st.error_pos[1], z = st.error_pos[2];
    switch (st.error_code) {
    case PMC_ERR_EXCESS_EVENTS:
        std::snprintf(msg, sizeof msg,
                      "Particle %lld exceeded 64 transport events in one timestep at (%f, %f, %f). Reduce "
                      "simulation.time_step or inspect narrow STL features; no recovery was applied.",
                      i, x, y, z);
        break;
    case PMC_ERR_ESCAPED:
        std::snprintf(msg, sizeof msg,
                      "Particle %lld escaped the domain at (%f, %f, %f). Refusing to relocate or rethermalize it; "
                      "inspect the geometry and time step.",
                      i, x, y, z);
        break;
    case PMC_ERR_INVALID_STATE:
        if (st.error_facet == -3)
            std::snprintf(msg, sizeof msg, "Live carrier count %lld does not match the boundary ledger.", i);
        else if (st.error_facet == -4)
            std::snprintf(msg, sizeof msg, "Non-finite heat-source energy per carrier.");
        else if (st.error_facet == -5)
            std::snprintf(msg, sizeof msg, "Could not bracket the occupation-source pseudo-temperature.");
        else if (st.error_facet == -6)
            std::snprintf(msg, sizeof msg, "Heat source would produce an invalid occupation; cell budget retained.");
        else if (st.error_facet == -12)
            std::snprintf(msg, sizeof msg, "Source sample has no thermal capacity.");
        else if (st.error_facet == -9)
            std::snprintf(msg, sizeof msg, "Callaway exponential work limit exceeded; reduce time_step.");
        else if (st.error_facet == -10)
            std::snprintf(msg, sizeof msg, "Callaway exponential did not converge.");
        else if (st.error_facet == -20 - pmc_nl::kNoEnergyConstraint)
            std::snprintf(msg, sizeof msg, "Nonlinear Callaway has no representable energy constraint");
        else if (st.error_facet == -20 - pmc_nl::kSingularJacobian)
            std::snprintf(msg, sizeof msg, "Singular nonlinear Callaway moment Jacobian");
        else if (st.error_facet == -20 - pmc_nl::kLineSearch)
            std::snprintf(msg, sizeof msg, "Nonlinear Callaway moment line search failed");
        else if (st.error_facet == -20 - pmc_nl::kNoConvergence)
            std::snprintf(msg, sizeof msg, "Nonlinear Callaway moment solve did not converge");
        else if (st.error_facet == -20 - pmc_nl::kInvalidOccupation)
            std::snprintf(msg, sizeof msg, "Nonlinear Callaway occupation is invalid");
        else if (st.error_facet == -20 - pmc_nl::kTemperatureRange) {
            // the rate temperature: the fixed lifetime temperature or cell i's
            double t = ctx->view.lifetime_T;
            if (ctx->view.lifetime_local && i >= 0 && i < ctx->view.ncells) {
                pmc_dev::CellRta cr{};
                if (hipMemcpy(&cr, ctx->view.crta + i, sizeof cr, hipMemcpyDeviceToHost) == hipSuccess) t = cr.t0;
            }
            std::snprintf(msg, sizeof msg, "Nonlinear Callaway temperature %f K is outside the supplied N/U table", t);
        }
        else
            std::snprintf(msg, sizeof msg, "Invalid particle in grid energy statistics.");
        break;
    case PMC_ERR_CAPACITY:
        std::snprintf(msg, sizeof msg, "Carrier slot capacity exhausted while injecting into slot %lld.", i);
        break;
    case PMC_ERR_UNSUPPORTED:
        if (st.error_facet == -14)
            std::snprintf(msg, sizeof msg,
                          "Particle %lld was mirrored on a facet that is not axis-aligned (a rough wall kept its "
                          "incoming mode because the remapped occupation would be negative, or a periodic facet "
                          "without a partner); the mirrored direction is not a table velocity.",
                          i);
        else if (st.error_facet == -8)
            std::snprintf(msg, sizeof msg,
                          "Cell %lld holds more carriers than the Callaway kernel takes per cell (%d); use more "
                          "cells or fewer carriers.",
                          i, ctx->view.cw_threads * pmc_dev::kCwPer);
        else
            std::snprintf(msg, sizeof msg, "Rough facet without a reflection table reached by particle %lld.", i);
        break;
    default:
        if (st.error_facet == -7)
            std::snprintf(msg, sizeof msg,
                          "Full-matrix boundary mode mapping failed while advancing particles for particle %lld. An "
                          "inward discrete material mode and finite signed energy weight are required; no "
                          "velocity-only reflection was applied.",
                          i);
        else if (st.error_facet == -11)
            std::snprintf(msg, sizeof msg, "sample_volume_points_naive: insufficient accepted samples.");
        else if (st.error_facet == -13)
            std::snprintf(msg, sizeof msg, "AMM incoming state is tangent or unsupported.");
        else
            std::snprintf(msg, sizeof msg,
                          "Boundary intersection failed while advancing particles for particle %lld at (%f, %f, %f). "
                          "Check that the STL is closed and manifold.",
                          i, x, y, z);
    }
    std::snprintf(ctx->error.message, sizeof(ctx->error.message), "%s", msg);
    return st.error_code;
}

float elapsed(hipEvent_t a, hipEvent_t b) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, a, b) != hipSuccess) return 0.f;
    return ms;
}
} // namespace

namespace {
// Mesh upload: faces and star tetrahedra in BVH leaf order, node boxes padded
// by the mesh tolerance and rounded outward.
// Checks a pmc_bvh_node array over `items` and threads it: padded boxes, and
// for interior nodes the first node after the subtree in place of the right
// child.  Returns false on a malformed hierarchy.
bool threaded_nodes(const pmc_bvh_node* nodes, int n, int items, double pad, std::vector<pmc_dev::BvhDev>& out) {
    if (n <= 0 || !nodes) return false;
    out.assign(n, {});
    std::vector<int> end(n + 1, 0);
    for (int i = n - 1; i >= 0; --i) { // right children lie after their parent
        const pmc_bvh_node& nd = nodes[i];
        if (nd.count > 0) {
            if (nd.index < 0 || nd.index + nd.count > items) return false;
            end[i] = i + 1;
        } else {
            if (nd.count != 0 || i + 1 >= n || nd.index <= i + 1 || nd.index >= n) return false;
            if (end[i + 1] != nd.index) return false; // the left subtree ends where the right child starts
            end[i] = end[nd.index];
        }
        for (int a = 0; a < 3; ++a) {
            out[i].bmin[a] = std::nextafter(static_cast<float>(static_cast<double>(nd.bmin[a]) - pad), -INFINITY);
            out[i].bmax[a] = std::nextafter(static_cast<float>(static_cast<double>(nd.bmax[a]) + pad), INFINITY);
        }
        out[i].index = nd.count > 0 ? nd.index : end[i];
        out[i].count = nd.count;
    }
    if (end[0] != n) return false; // the root's subtree is the whole array
    std::vector<int> seen(items, 0); // every item in exactly one leaf
    for (int i = 0; i < n; ++i)
        for (int k = 0; k < nodes[i].count; ++k)
            if (seen[nodes[i].index + k]++) return false;
    for (int v : seen)
        if (v != 1) return false;
    return true;
}

int upload_mesh(pmc_ctx* ctx, const pmc_geometry* g) {
    View& v = ctx->view;
    const pmc_mesh& M = *g->mesh;
    if (M.n_faces <= 0 || M.n_face_nodes <= 0 || !M.face_nodes || !M.face_order)
        return fail(ctx, PMC_ERR_ARGUMENT, "mesh without faces or face hierarchy");
    const int nf = M.n_faces;
    std::vector<double4> plane(nf);
    std::vector<double> tri(9 * static_cast<size_t>(nf)), box(6 * static_cast<size_t>(nf));
    std::vector<int> facet(nf);
    for (int j = 0; j < nf; ++j) {
        const int f = M.face_order[j];
        if (f < 0 || f >= nf) return fail(ctx, PMC_ERR_ARGUMENT, "face order out of range");
        plane[j] = make_double4(M.face_normal[3 * f], M.face_normal[3 * f + 1], M.face_normal[3 * f + 2], M.face_k[f]);
        for (int k = 0; k < 3; ++k) {
            const int vi = M.faces[3 * f + k];
            if (vi < 0 || vi >= M.n_vertices) return fail(ctx, PMC_ERR_ARGUMENT, "face vertex out of range");
            for (int a = 0; a < 3; ++a) tri[9 * j + 3 * k + a] = M.vertices[3 * vi + a];
        }
        for (int k = 0; k < 6; ++k) box[6 * j + k] = M.face_bounds[6 * f + k];
        facet[j] = M.face_facet[f];
        if (facet[j] < 0 || facet[j] >= g->n_facets) return fail(ctx, PMC_ERR_ARGUMENT, "face without facet");
    }
    std::vector<pmc_dev::BvhDev> fn;
    if (!threaded_nodes(M.face_nodes, M.n_face_nodes, nf, v.mesh_tol, fn))
        return fail(ctx, PMC_ERR_ARGUMENT, "malformed face hierarchy");
    v.nfaces = nf;
    v.face_plane = ctx->upload(plane.data(), plane.size());
    v.face_tri = ctx->upload(tri.data(), tri.size());
    v.face_box = ctx->upload(box.data(), box.size());
    v.face_facet = ctx->upload(facet.data(), facet.size());
    v.face_nodes = ctx->upload(fn.data(), fn.size());
    v.n_face_nodes = M.n_face_nodes;
    v.field = nullptr;
    v.field_inside = nullptr;
    if (M.field) {
        const size_t nv = static_cast<size_t>(M.field_dims[0]) * M.field_dims[1] * M.field_dims[2];
        if (M.field_dims[0] <= 0 || M.field_dims[1] <= 0 || M.field_dims[2] <= 0 || nv > (size_t(1) << 30) ||
            !(M.field_h > 0.0) || !(M.field_unit > 0.0))
            return fail(ctx, PMC_ERR_ARGUMENT, "malformed distance field");
        v.field = ctx->upload(reinterpret_cast<const unsigned char*>(M.field), nv);
        v.field_inside =
            M.field_inside ? ctx->upload(reinterpret_cast<const unsigned char*>(M.field_inside), nv) : nullptr;
        for (int a = 0; a < 3; ++a) v.field_dims[a] = M.field_dims[a];
        // Rounded so the device bound never exceeds the host's: a shorter unit
        // and a longer voxel edge.
        v.field_inv_h = std::nextafter(static_cast<float>(1.0 / M.field_h), 0.0f);
        v.field_unit = std::nextafter(static_cast<float>(M.field_unit), 0.0f);
    }
    v.nsimplices = M.n_simplices;
    if (M.n_simplices > 0) {
        const int ns = M.n_simplices;
        std::vector<double> sv(12 * static_cast<size_t>(ns)), si(9 * static_cast<size_t>(ns)),
            sb(6 * static_cast<size_t>(ns)), cdf(ns);
        std::vector<int> pos(ns, -1);
        for (int j = 0; j < ns; ++j) {
            const int t = M.simplex_order[j];
            if (t < 0 || t >= ns || pos[t] >= 0) return fail(ctx, PMC_ERR_ARGUMENT, "simplex order is not a permutation");
            pos[t] = j;
            for (int k = 0; k < 12; ++k) sv[12 * j + k] = M.simplex_vertices[12 * t + k];
            for (int k = 0; k < 9; ++k) si[9 * j + k] = M.simplex_inverse[9 * t + k];
            for (int k = 0; k < 6; ++k) sb[6 * j + k] = M.simplex_bounds[6 * t + k];
        }
        double acc = 0.0;
        for (int t = 0; t < ns; ++t) cdf[t] = (acc += M.simplex_volume[t]);
        std::vector<pmc_dev::BvhDev> sn;
        // Tetrahedron boxes are tested with the 1e-10 barycentric tolerance only;
        // the padding covers it.
        if (!threaded_nodes(M.simplex_nodes, M.n_simplex_nodes, ns, 1e-6, sn))
            return fail(ctx, PMC_ERR_ARGUMENT, "malformed simplex hierarchy");
        v.simplex_v = ctx->upload(sv.data(), sv.size());
        v.simplex_inv = ctx->upload(si.data(), si.size());
        v.simplex_box = ctx->upload(sb.data(), sb.size());
        v.simplex_cdf = ctx->upload(cdf.data(), cdf.size());
        v.simplex_pos = ctx->upload(pos.data(), pos.size());
        v.simplex_nodes = ctx->upload(sn.data(), sn.size());
        v.n_simplex_nodes = M.n_simplex_nodes;
    }
    return PMC_OK;
}
} // namespace

extern "C" {

const char* pmc_status_string(int32_t code) {
    switch (code) {
    case PMC_OK: return "ok";
    case PMC_ERR_ARGUMENT: return "invalid argument";
    case PMC_ERR_DEVICE: return "device error";
    case PMC_ERR_EXCESS_EVENTS: return "excess transport events";
    case PMC_ERR_BOUNDARY: return "boundary failure";
    case PMC_ERR_ESCAPED: return "escaped carrier";
    case PMC_ERR_INVALID_STATE: return "invalid state";
    case PMC_ERR_CAPACITY: return "capacity exhausted";
    case PMC_ERR_LATCHED: return "context latched after failure";
    case PMC_ERR_UNSUPPORTED: return "unsupported";
    default: return "unknown";
    }
}

int pmc_create(const pmc_setup* setup, pmc_ctx** out) {
    if (!setup || !out) return PMC_ERR_ARGUMENT;
    *out = nullptr;
    if (!(setup->time_step_ps > 0) || !std::isfinite(setup->time_step_ps) || setup->capacity <= 0 ||
        setup->capacity > (1ll << 30) || !(setup->push_distance_a >= 0) ||
        (setup->temperature_inversion != 0 && setup->temperature_inversion != 1) ||
        (setup->conservative_walls != 0 && setup->conservative_walls != 1))
        return PMC_ERR_ARGUMENT;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0) return PMC_ERR_DEVICE;
    auto* ctx = new pmc_ctx();
    ctx->setup = *setup;
    if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
        delete ctx;
        return PMC_ERR_DEVICE;
    }
    for (auto& e : ctx->ev) ctx->note(hipEventCreate(&e));
    View& v = ctx->view;
    v.capacity = static_cast<int>(setup->capacity);
    v.dt = setup->time_step_ps;
    v.dtf = static_cast<float>(setup->time_step_ps);
    v.push = setup->push_distance_a;
    v.key0 = static_cast<uint32_t>(setup->seed & 0xffffffffu);
    v.key1 = static_cast<uint32_t>(setup->seed >> 32) ^ (setup->instance * 0x9E3779B9u);
    v.lifetime_local = setup->lifetime_local;
    v.lifetime_T = setup->lifetime_temperature;
    v.sampling = setup->mode_sampling;
    v.inversion = setup->temperature_inversion;
    v.series_on = 1;
    v.series_radius = 0.01;
    v.t_bg = setup->background_temperature;
    v.refill_fixed = setup->refill_policy == 1 ? 1 : 0;
    v.conservative = setup->conservative_walls != 0 ? 1 : 0;
    v.cw_threads = pmc_dev::kCwThreads;
    for (auto& b : ctx->buf) {
        const size_t n = static_cast<size_t>(setup->capacity);
        b.px = ctx->alloc<float>(n);
        b.py = ctx->alloc<float>(n);
        b.pz = ctx->alloc<float>(n);
        b.mode = ctx->alloc<int>(n);
        b.occ = ctx->alloc<double>(n);
        b.cell = ctx->alloc<int>(n);
        b.uid = ctx->alloc<unsigned>(n);
        if (!b.px || !b.py || !b.pz || !b.mode || !b.occ || !b.cell || !b.uid) {
            pmc_destroy(ctx);
            return PMC_ERR_DEVICE;
        }
    }
    v.st = ctx->alloc<DevState>(1);
    if (!v.st) {
        pmc_destroy(ctx);
        return PMC_ERR_DEVICE;
    }
    ctx->note(hipMemset(v.st, 0, sizeof(DevState)));
    *out = ctx;
    return PMC_OK;
}

int pmc_upload_material(pmc_ctx* ctx, const pmc_mode_table* t) {
    if (!ctx || !t || t->n_active <= 0 || t->n_flat <= 0 || t->n_temps <= 0 || t->lut_n < 2)
        return fail(ctx, PMC_ERR_ARGUMENT, "invalid mode table");
    View& v = ctx->view;
    v.na = t->n_active;
    v.nflat = t->n_flat;
    v.active_flat = ctx->upload(t->active_flat, t->n_active);
    v.omega = ctx->upload(t->omega, t->n_active);
    std::vector<float4> vel(t->n_active);
    for (int a = 0; a < t->n_active; ++a) {
        const double* p = t->velocity + 3 * a;
        vel[a] = make_float4(static_cast<float>(p[0]), static_cast<float>(p[1]), static_cast<float>(p[2]),
                             static_cast<float>(std::sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2])));
    }
    v.vel = ctx->upload(vel.data(), vel.size());
    v.veld = ctx->upload(t->velocity, 3 * static_cast<size_t>(t->n_active));
    v.neq = ctx->upload(t->neq_background, t->n_active);
    std::vector<double2> mtab(t->n_active);
    for (int a = 0; a < t->n_active; ++a) mtab[a] = make_double2(t->omega[a], t->neq_background[a]);
    v.mtab = ctx->upload(mtab.data(), mtab.size());
    v.ntemps = t->n_temps;
    v.temps = ctx->upload(t->temps, t->n_temps);
    // Rate table as row pairs in active-mode order: a carrier's two
    // interpolation rows come in one 16-byte load, and the few pairs the cell
    // temperatures use stay cache resident.
    std::vector<double2> gamma2(static_cast<size_t>(t->n_temps) * t->n_active);
    for (int a = 0; a < t->n_active; ++a) {
        const int f = t->active_flat[a];
        if (f < 0 || f >= t->n_flat) return fail(ctx, PMC_ERR_ARGUMENT, "active mode outside the flat table");
    }
    for (int r = 0; r < t->n_temps; ++r) {
        const int r1 = std::min(r + 1, t->n_temps - 1);
        for (int a = 0; a < t->n_active; ++a) {
            const int f = t->active_flat[a];
            gamma2[static_cast<size_t>(r) * t->n_active + a] =
                make_double2(t->gamma_total[static_cast<size_t>(r) * t->n_flat + f],
                             t->gamma_total[static_cast<size_t>(r1) * t->n_flat + f]);
        }
    }
    v.gamma2 = ctx->upload(gamma2.data(), gamma2.size());
    v.w = t->energy_weight;
    v.e_bg = t->background_energy_density;
    v.zp = t->zero_point_density;
    v.e_norm = t->energy_normalization;
    if (v.inversion == 1 && !(t->energy_normalization >= 1.0 && std::isfinite(t->energy_normalization)))
        return fail(ctx, PMC_ERR_ARGUMENT, "Newton temperature inversion needs the energy normalization (>= 1)");
    v.lut_n = t->lut_n;
    v.lut_t = ctx->upload(t->lut_t, t->lut_n);
    v.lut_e = ctx->upload(t->lut_e, t->lut_n);
    if (!v.active_flat || !v.omega || !v.vel || !v.veld || !v.neq || !v.mtab || !v.temps || !v.gamma2 || !v.lut_t ||
        !v.lut_e)
        return fail(ctx, PMC_ERR_DEVICE, "material upload failed");
    ctx->has_material = true;
    return PMC_OK;
}

int pmc_upload_geometry(pmc_ctx* ctx, const pmc_geometry* g) {
    if (!ctx || !g || g->n_facets <= 0 || g->n_facets > pmc_dev::kMaxFacetsDev || g->n_cells <= 0)
        return fail(ctx, PMC_ERR_ARGUMENT, "invalid geometry");
    if (g->kind != 0 && !(g->kind == 1 && g->mesh))
        return fail(ctx, PMC_ERR_UNSUPPORTED, "geometry must be the analytic box or a triangle mesh");
    View& v = ctx->view;
    v.geom_kind = g->kind;
    // FP32 positions carry about 2^-24 of the largest coordinate; eight of those
    // quanta are taken as on the surface (box walls or mesh faces).
    double reach = 0.0;
    for (int a = 0; a < 3; ++a) reach = std::max({reach, std::abs(g->bmin[a]), std::abs(g->bmax[a])});
    v.mesh_tol = std::max(1e-7, 8.0 * std::ldexp(reach, -24));
    for (int a = 0; a < 3; ++a) {
        v.bmin[a] = g->bmin[a];
        v.bmax[a] = g->bmax[a];
        v.fbmin[a] = static_cast<float>(g->bmin[a]);
        v.fbmax[a] = static_cast<float>(g->bmax[a]);
        v.grid[a] = g->grid[a];
        v.inv_cell[a] = g->inv_cell[a];
    }
    for (int k = 0; k < 6; ++k) {
        if (g->kind == 1) {
            v.plane_facet[k] = -1;
            continue;
        }
        if (g->box_plane_facet[k] < 0 || g->box_plane_facet[k] >= g->n_facets)
            return fail(ctx, PMC_ERR_ARGUMENT, "box plane without facet");
        v.plane_facet[k] = g->box_plane_facet[k];
    }
    std::vector<FacetDev> f(g->n_facets);
    for (int i = 0; i < g->n_facets; ++i) {
        for (int a = 0; a < 3; ++a) {
            f[i].n[a] = g->facets[i].normal[a];
            f[i].shift[a] = g->facets[i].shift[a];
        }
        f[i].cond = g->facets[i].condition;
        f[i].rough = g->facets[i].rough_table;
        f[i].res = g->facets[i].reservoir;
        f[i].pair = g->facets[i].periodic_pair;
        // Carriers keep a table mode plus per-axis velocity signs, so a mirror
        // reflection (a periodic facet without a partner, outside conservative
        // transport) needs an axis-aligned facet.  A facet with a reflection
        // table sends carriers into table modes that point into the body, on
        // any orientation.
        f[i].axis = -1;
        for (int a = 0; a < 3; ++a) {
            const double o1 = f[i].n[(a + 1) % 3], o2 = f[i].n[(a + 2) % 3];
            if (std::abs(std::abs(f[i].n[a]) - 1.0) <= 1e-9 && std::abs(o1) <= 1e-9 && std::abs(o2) <= 1e-9)
                f[i].axis = a;
        }
        const bool mirrors = f[i].cond == 'R' || (f[i].cond == 'P' && f[i].pair < 0);
        if (mirrors && f[i].axis < 0 && f[i].rough < 0) {
            char msg[200];
            std::snprintf(msg, sizeof msg,
                          "facet %d reflects carriers by a mirror but is not axis-aligned; the device takes "
                          "mirror reflections on axis-aligned facets only",
                          i);
            return fail(ctx, PMC_ERR_UNSUPPORTED, msg);
        }
    }
    if (g->kind == 1) {
        const int rc = upload_mesh(ctx, g);
        if (rc) return rc;
    }
    v.nfacets = g->n_facets;
    v.facets = ctx->upload(f.data(), f.size());
    const size_t nvox = static_cast<size_t>(g->grid[0]) * g->grid[1] * g->grid[2];
    v.voxel = ctx->upload(g->voxel_to_cell, nvox);
    v.ncells = g->n_cells;
    v.cell_vol = ctx->upload(g->cell_volume, g->n_cells);
    const size_t nc = static_cast<size_t>(g->n_cells);
    v.cell_e = ctx->alloc<double>(nc);
    v.cell_tally = ctx->alloc<double>(pmc_dev::kTallyStride * nc);
    // Default event split: more event blocks at small carrier counts on a mesh,
    // where the face searches are latency-bound (pmc_autotune can revise it).
    v.event_split = g->kind == 1 ? pmc_launch::event_split(static_cast<int>(v.capacity)) : 1;
    v.partial_rows = pmc_launch::partial_rows(v);
    v.partials = ctx->alloc<double>(static_cast<size_t>(v.partial_rows) * pmc_dev::kPartialCols);
    v.pcols = pmc_dev::kPAbsorbedRes;
    v.cell_T = ctx->alloc<double>(nc);
    v.cell_tx = ctx->alloc<double>(nc);
    v.cell_flux = ctx->alloc<double>(3 * nc);
    v.cell_count = ctx->alloc<int>(nc);
    v.cell_start = ctx->alloc<int>(nc + 1);
    v.cell_cursor = ctx->alloc<int>(nc);
    v.crta = ctx->alloc<CellRta>(nc);
    v.chunk_start = ctx->alloc<int>(nc + 1);
    v.max_chunks = static_cast<int>((v.capacity + pmc_dev::kChunk - 1) / pmc_dev::kChunk + nc);
    v.chunk_cell = ctx->alloc<int>(static_cast<size_t>(v.max_chunks));
    v.chunk_acc = ctx->alloc<double>(static_cast<size_t>(v.max_chunks) * pmc_dev::kAccCols);
    v.chunk_out = ctx->alloc<double>(2 * static_cast<size_t>(v.max_chunks));
    v.aw = ctx->alloc<double>(static_cast<size_t>(v.capacity));
    const size_t blocks = static_cast<size_t>(pmc_launch::wave_blocks(v.capacity));
    v.events = ctx->alloc<int>(blocks * pmc_dev::kSlotsPerBlock);
    v.block_events = ctx->alloc<int>(blocks);
    if (!v.facets || !v.voxel || !v.cell_vol || !v.cell_e || !v.cell_T || !v.cell_tx || !v.cell_tally || !v.partials ||
        !v.cell_flux || !v.cell_count || !v.cell_start || !v.cell_cursor || !v.crta || !v.events || !v.block_events ||
        !v.chunk_start || !v.chunk_cell || !v.chunk_acc || !v.chunk_out || !v.aw)
        return fail(ctx, PMC_ERR_DEVICE, "geometry upload failed");
    ctx->note(hipMemset(v.chunk_start, 0, (nc + 1) * sizeof(int)));
    ctx->note(hipMemset(v.cell_start, 0, (nc + 1) * sizeof(int)));
    ctx->note(hipMemset(v.cell_count, 0, nc * sizeof(int)));
    ctx->note(hipMemset(v.block_events, 0, blocks * sizeof(int)));
    ctx->note(hipMemset(v.cell_tally, 0, pmc_dev::kTallyStride * nc * sizeof(double)));
    ctx->note(hipMemset(v.partials, 0, static_cast<size_t>(v.partial_rows) * pmc_dev::kPartialCols * sizeof(double)));
    ctx->note(hipMemset(v.crta, 0, nc * sizeof(CellRta)));
    ctx->has_geometry = true;
    return PMC_OK;
}

int pmc_upload_boundary_tables(pmc_ctx* ctx, int32_t n_rough, const pmc_rough_table* rough, int32_t n_res,
                               const pmc_reservoir* res) {
    if (!ctx || !ctx->has_material || n_rough < 0 || n_res < 0 || n_res > pmc_dev::kMaxReservoirs)
        return fail(ctx, PMC_ERR_ARGUMENT, "invalid boundary tables");
    View& v = ctx->view;
    const int na = v.na;
    std::vector<RoughDev> rd(std::max(1, n_rough));
    for (int r = 0; r < n_rough; ++r) {
        const auto& t = rough[r];
        rd[r].spec = ctx->upload(t.specularity, na);
        rd[r].match = ctx->upload(t.spec_match, na);
        rd[r].dbeg = ctx->upload(t.diffuse_begin, na);
        rd[r].dend = ctx->upload(t.diffuse_end, na);
        rd[r].n_out = t.n_outgoing;
        rd[r].out_active = ctx->upload(t.outgoing_active, t.n_outgoing);
        rd[r].out_sorted = ctx->upload(t.outgoing_sorted, t.n_outgoing);
        rd[r].prefix = ctx->upload(t.residual_prefix, t.n_outgoing + 1);
        rd[r].n_roul = t.n_roulette;
        rd[r].roul_active = ctx->upload(t.roulette_active, t.n_roulette);
        rd[r].roul_cdf = ctx->upload(t.roulette_cdf, t.n_roulette);
        rd[r].strict = t.strict_elastic;
        if (t.mode_count < 0 || t.mode_begin < 0 || t.mode_begin + t.mode_count > na)
            return fail(ctx, PMC_ERR_ARGUMENT, "reflection table mode range outside the active modes");
        rd[r].mode_lo = t.mode_count ? t.mode_begin : 0;
        rd[r].mode_n = t.mode_count ? t.mode_count : na;
    }
    ctx->rough_range.clear();
    for (int r = 0; r < n_rough; ++r) ctx->rough_range.push_back({rough[r].mode_begin, rough[r].mode_count});
    v.rough = ctx->upload(rd.data(), rd.size());
    std::vector<ResDev> dd(std::max(1, n_res));
    double expected = 0;
    std::vector<int> initial(pmc_dev::kMaxReservoirs, 0);
    for (int r = 0; r < n_res; ++r) {
        const auto& t = res[r];
        dd[r].facet = t.facet;
        dd[r].temperature = t.temperature;
        for (int a = 0; a < 3; ++a) dd[r].n[a] = t.normal[a];
        dd[r].initial = t.initial_count;
        dd[r].expected = t.expected_count;
        dd[r].cdf = ctx->upload(t.cdf, na);
        dd[r].aprob = ctx->upload(t.alias_prob, na);
        dd[r].aidx = ctx->upload(t.alias_index, na);
        dd[r].occ = ctx->upload(t.occupation, na);
        dd[r].nfaces = t.n_faces;
        dd[r].face_cdf = ctx->upload(t.face_cdf, t.n_faces);
        dd[r].tri = ctx->upload(t.triangles, 9 * static_cast<size_t>(t.n_faces));
        expected += t.expected_count;
        initial[r] = t.initial_count;
    }
    v.res = ctx->upload(dd.data(), dd.size());
    v.nres = n_res;
    v.pcols = pmc_dev::kPAbsorbedRes + n_res;
    ctx->n_reservoirs = n_res;
    ctx->expected_injection = expected;
    // BoundaryLedger::configure: the first step refills the rounded expected counts.
    ctx->note(hipMemcpy(reinterpret_cast<char*>(v.st) + offsetof(DevState, absorbed_now), initial.data(),
              sizeof(int) * pmc_dev::kMaxReservoirs, hipMemcpyHostToDevice));
    ctx->has_tables = true;
    return PMC_OK;
}

int pmc_upload_heat_source(pmc_ctx* ctx, const pmc_heat_source* h) {
    if (!ctx || !h || !ctx->has_material || !ctx->has_geometry || h->n_cells != ctx->view.ncells || !h->rate ||
        !h->mode_eligible)
        return fail(ctx, PMC_ERR_ARGUMENT, "invalid heat source");
    if (h->spectrum < 0 || h->spectrum > 2 || (h->spectrum != 0 && !h->mode_weight))
        return fail(ctx, PMC_ERR_ARGUMENT, "heat-source spectrum must be 0, 1 or 2 with mode weights for 1 and 2");
    View& v = ctx->view;
    if (h->spectrum != 0)
        for (int a = 0; a < v.na; ++a)
            if (!std::isfinite(h->mode_weight[a]) || (h->spectrum == 1 && h->mode_weight[a] < 0.0))
                return fail(ctx, PMC_ERR_ARGUMENT, "invalid heat-source mode weight");
    const size_t nc = static_cast<size_t>(v.ncells);
    v.hs_spectrum = h->spectrum;
    v.hs_w = h->spectrum != 0 ? ctx->upload(h->mode_weight, static_cast<size_t>(v.na)) : nullptr;
    v.hs_rate = ctx->upload(h->rate, nc);
    v.hs_elig = ctx->upload(h->mode_eligible, static_cast<size_t>(v.na));
    v.hs_pending = ctx->alloc<double>(nc);
    v.hs_cell = ctx->alloc<double>(pmc_dev::kSrcCellCols * nc);
    v.hs_acc = ctx->alloc<double>(static_cast<size_t>(v.max_chunks) * pmc_dev::kSrcCols);
    v.chunk_src = ctx->alloc<double>(static_cast<size_t>(v.max_chunks));
    if (!v.hs_rate || !v.hs_elig || !v.hs_pending || !v.hs_cell || !v.hs_acc || !v.chunk_src)
        return fail(ctx, PMC_ERR_DEVICE, "heat source upload failed");
    ctx->note(hipMemset(v.hs_pending, 0, nc * sizeof(double)));
    ctx->note(hipMemset(v.hs_cell, 0, pmc_dev::kSrcCellCols * nc * sizeof(double)));
    ctx->hs = *h;
    ctx->hs.rate = nullptr;
    ctx->hs.mode_eligible = nullptr;
    ctx->hs.mode_weight = nullptr;
    v.hs_square = h->square;
    v.hs_start = h->start_ps;
    v.hs_end = h->end_ps;
    v.hs_period = h->period_ps;
    v.hs_on_duration = h->on_duration_ps;
    v.hs_duty = h->duty_cycle;
    v.hs_amplitude = h->amplitude;
    v.hs_on = 1;
    return check_hip(ctx, hipGetLastError(), "heat source upload");
}

static int finish_population(pmc_ctx* ctx, int64_t count) {
    View& v = view(ctx);
    DevState st{};
    ctx->note(hipMemcpy(&st, v.st, sizeof(DevState), hipMemcpyDeviceToHost));
    st.n_slots = static_cast<int>(count);
    st.n_live = static_cast<int>(count);
    st.n_slots_start = static_cast<int>(count);
    st.next_uid = static_cast<unsigned>(count);
    ctx->note(hipMemcpy(v.st, &st, sizeof(DevState), hipMemcpyHostToDevice));
    // Cell sort, then the initial refresh (src/solver/Observables.cpp:12-21).
    pmc_launch::count_all(v, ctx->stream);
    pmc_launch::refresh(v, ctx->stream, 0);
    pmc_launch::sort(v, ctx->stream);
    ctx->cur = 1 - ctx->cur;
    View& w = view(ctx);
    pmc_launch::refresh(w, ctx->stream, 3);
    if (w.inversion == 1) pmc_launch::invert(w, ctx->stream, 1);
    int rc = check_hip(ctx, hipGetLastError(), "initial refresh");
    if (rc) return rc;
    ctx->initialized = true;
    return collect_error(ctx);
}

int pmc_upload_callaway(pmc_ctx* ctx, const pmc_callaway_table* t) {
    if (!ctx) return PMC_ERR_ARGUMENT;
    if (!t || !t->capacity || !t->wavevector || !t->normal_rate || !t->resistive_rate)
        return fail(ctx, PMC_ERR_ARGUMENT, "Callaway table arrays missing");
    if (!ctx->has_material || !ctx->has_geometry)
        return fail(ctx, PMC_ERR_ARGUMENT, "material and geometry must be uploaded before the Callaway table");
    if (ctx->view.layers > 1)
        return fail(ctx, PMC_ERR_ARGUMENT, "upload the Callaway table before the layered-material tables");
    if (ctx->view.inversion == 1)
        return fail(ctx, PMC_ERR_UNSUPPORTED,
                    "Newton temperature inversion applies to RTA transport; linearized Callaway uses "
                    "T_bg + (e - E_bg) / C_bg");
    if (ctx->initialized)
        return fail(ctx, PMC_ERR_ARGUMENT, "the Callaway table must be uploaded before the population is created");
    if (ctx->view.nl_callaway) return fail(ctx, PMC_ERR_ARGUMENT, "the context already runs nonlinear Callaway");
    if (!(t->heat_capacity_density > 0.0) || !std::isfinite(t->heat_capacity_density) ||
        !std::isfinite(t->background_energy_density))
        return fail(ctx, PMC_ERR_ARGUMENT, "Linearized temperature requires positive heat capacity.");
    View& v = ctx->view;
    const int na = v.na;
    std::vector<double> omega(na);
    ctx->note(hipMemcpy(omega.data(), v.omega, na * sizeof(double), hipMemcpyDeviceToHost));
    int rc = check_hip(ctx, hipGetLastError(), "Callaway table");
    if (rc) return rc;
    // Packed per-mode constants, 64 B per mode: (g_N, g_R, B0, B1), (B2, B3, hw, 0)
    // with B = sqrt(C) (1, q / w) (CallawayOperator::CallawayOperator,
    // src/material/CallawayOperator.cpp:62-95, and its checks).
    std::vector<double> pack(8 * static_cast<size_t>(na), 0.0);
    for (int a = 0; a < na; ++a) {
        const double w = omega[a], cap = t->capacity[a], gn = t->normal_rate[a], gr = t->resistive_rate[a];
        if (!std::isfinite(w) || w <= 0 || !std::isfinite(cap) || cap <= 0 || !std::isfinite(gn) || gn < 0 ||
            !std::isfinite(gr) || gr < 0)
            return fail(ctx, PMC_ERR_ARGUMENT,
                        "Callaway needs positive frequencies/capacities and finite nonnegative rates.");
        const double b0 = std::sqrt(cap);
        double* p = pack.data() + 8 * static_cast<size_t>(a);
        p[0] = gn;
        p[1] = gr;
        p[2] = b0;
        for (int k = 0; k < 3; ++k) {
            const double q = t->wavevector[3 * a + k];
            if (!std::isfinite(q)) return fail(ctx, PMC_ERR_ARGUMENT, "Nonfinite Callaway wavevector.");
            p[3 + k] = b0 * q / w;
        }
        p[6] = 6.582119569e-4 * w;
    }
    v.cw_cap = ctx->upload(t->capacity, na);
    v.cw_q = ctx->upload(t->wavevector, 3 * static_cast<size_t>(na));
    v.cw_gn = ctx->upload(t->normal_rate, na);
    v.cw_gr = ctx->upload(t->resistive_rate, na);
    v.cw_pack = reinterpret_cast<const double4*>(ctx->upload(pack.data(), pack.size()));
    v.cw_dir = ctx->alloc<double>(4 * static_cast<size_t>(v.capacity));
    v.cw_cell = ctx->alloc<double>(pmc_dev::kCwRec * static_cast<size_t>(v.ncells));
    v.cw_big = ctx->alloc<int>(v.ncells);
    if (!v.cw_dir || !v.cw_cell || !v.cw_big) return fail(ctx, PMC_ERR_DEVICE, "Callaway direction buffer");
    v.cw_cbg = t->heat_capacity_density;
    v.e_bg = t->background_energy_density;
    v.callaway = 1;
    v.conservative = 1;
    v.linearized = 1;
    v.cw_threads = pmc_dev::kCwThreads;
    return check_hip(ctx, hipGetLastError(), "Callaway table");
}

int pmc_upload_callaway_nonlinear(pmc_ctx* ctx, const pmc_callaway_rates* r) {
    if (!ctx) return PMC_ERR_ARGUMENT;
    if (!r || r->n_temps <= 0 || !r->temps || !r->gamma_normal || !r->gamma_umklapp || !r->gamma_extra ||
        !r->wavevector)
        return fail(ctx, PMC_ERR_ARGUMENT, "nonlinear Callaway rate tables missing");
    if (!ctx->has_material || !ctx->has_geometry)
        return fail(ctx, PMC_ERR_ARGUMENT, "material and geometry must be uploaded before the Callaway tables");
    if (ctx->initialized)
        return fail(ctx, PMC_ERR_ARGUMENT, "the Callaway tables must be uploaded before the population is created");
    if (ctx->view.callaway) return fail(ctx, PMC_ERR_ARGUMENT, "the context already runs linearized Callaway");
    if (ctx->view.inversion == 1)
        return fail(ctx, PMC_ERR_UNSUPPORTED, "Newton temperature inversion applies to RTA transport");
    if (ctx->view.layers > 1)
        return fail(ctx, PMC_ERR_ARGUMENT, "upload the nonlinear Callaway tables before the layered-material tables");
    View& v = ctx->view;
    const size_t na = static_cast<size_t>(v.na), nt = static_cast<size_t>(r->n_temps);
    for (size_t k = 0; k < nt; ++k)
        if (!std::isfinite(r->temps[k]) || (k && !(r->temps[k] > r->temps[k - 1])))
            return fail(ctx, PMC_ERR_ARGUMENT, "rate temperatures must be finite and increasing");
    for (size_t k = 0; k < nt * na; ++k)
        if (!std::isfinite(r->gamma_normal[k]) || r->gamma_normal[k] < 0 || !std::isfinite(r->gamma_umklapp[k]) ||
            r->gamma_umklapp[k] < 0 || !std::isfinite(r->gamma_extra[k]) || r->gamma_extra[k] < 0)
            return fail(ctx, PMC_ERR_ARGUMENT, "Invalid nonlinear Callaway mode");
    for (size_t k = 0; k < 3 * na; ++k)
        if (!std::isfinite(r->wavevector[k])) return fail(ctx, PMC_ERR_ARGUMENT, "Invalid nonlinear Callaway wavevector");
    v.nl_nt = r->n_temps;
    v.nl_temps = ctx->upload(r->temps, nt);
    v.nl_gn = ctx->upload(r->gamma_normal, nt * na);
    v.nl_gu = ctx->upload(r->gamma_umklapp, nt * na);
    v.nl_ge = ctx->upload(r->gamma_extra, nt * na);
    v.nl_q = ctx->upload(r->wavevector, 3 * na);
    v.nl_scratch = ctx->alloc<double>(7 * static_cast<size_t>(v.capacity));
    if (!v.nl_temps || !v.nl_gn || !v.nl_gu || !v.nl_ge || !v.nl_q || !v.nl_scratch)
        return fail(ctx, PMC_ERR_DEVICE, "nonlinear Callaway tables");
    v.nl_callaway = 1;
    v.conservative = 1;
    v.linearized = 0;
    return check_hip(ctx, hipGetLastError(), "nonlinear Callaway tables");
}

int pmc_initialize_carriers(pmc_ctx* ctx, const pmc_init* init) {
    if (!ctx || !init || !ctx->has_material || !ctx->has_geometry || !ctx->has_tables)
        return fail(ctx, PMC_ERR_ARGUMENT, "tables must be uploaded before initialization");
    if (init->count <= 0 || init->count > ctx->setup.capacity) return fail(ctx, PMC_ERR_ARGUMENT, "invalid carrier count");
    if (init->temperature_mode == 2 && !init->cell_temperature)
        return fail(ctx, PMC_ERR_ARGUMENT, "per-cell initial temperature field missing");
    if (init->excite && (!(init->excite_multiplier > 0.0) || !std::isfinite(init->excite_multiplier)))
        return fail(ctx, PMC_ERR_ARGUMENT, "mode_excitation.occupation_multiplier must be finite and positive.");
    View& v = view(ctx);
    v.ex_on = init->excite ? 1 : 0;
    v.ex_fmin = init->excite_fmin_thz;
    v.ex_fmax = init->excite_fmax_thz;
    v.ex_mult = init->excite_multiplier;
    double* field = nullptr;
    if (init->temperature_mode == 2) field = ctx->upload(init->cell_temperature, v.ncells);
    pmc_launch::init(v, ctx->stream, static_cast<int>(init->count), init->temperature_mode, init->uniform_temperature,
                     init->cold_point, init->hot_point, init->cold_temperature, init->hot_temperature, field);
    int rc = check_hip(ctx, hipGetLastError(), "init");
    if (rc) return rc;
    return finish_population(ctx, init->count);
}

int pmc_seed_carriers(pmc_ctx* ctx, int64_t n, const float* position, const int32_t* mode, const double* occupation) {
    if (!ctx || n <= 0 || n > ctx->setup.capacity || !position || !mode || !occupation || !ctx->has_geometry ||
        !ctx->has_material || !ctx->has_tables)
        return fail(ctx, PMC_ERR_ARGUMENT, "invalid carrier seed");
    View& v = view(ctx);
    std::vector<float> x(n), y(n), z(n);
    std::vector<int> cell(n);
    std::vector<unsigned> uid(n);
    for (int64_t i = 0; i < n; ++i) {
        x[i] = position[3 * i];
        y[i] = position[3 * i + 1];
        z[i] = position[3 * i + 2];
        if ((mode[i] & pmc_dev::kModeMask) >= v.na || mode[i] < 0) return fail(ctx, PMC_ERR_ARGUMENT, "mode out of range");
        uid[i] = static_cast<unsigned>(i);
        // Cell index as the device computes it (FP64 on the FP32 position).
        int idx[3];
        const double p[3] = {x[i], y[i], z[i]};
        for (int a = 0; a < 3; ++a) {
            const int nn = v.grid[a];
            if (nn <= 1 || !(v.bmax[a] - v.bmin[a] > 0.0)) idx[a] = 0;
            else {
                int k = static_cast<int>(std::floor((p[a] - v.bmin[a]) * v.inv_cell[a]));
                idx[a] = std::min(std::max(k, 0), nn - 1);
            }
        }
        cell[i] = (idx[0] * v.grid[1] + idx[1]) * v.grid[2] + idx[2];
    }
    std::vector<int> vox(static_cast<size_t>(v.grid[0]) * v.grid[1] * v.grid[2]);
    ctx->note(hipMemcpy(vox.data(), v.voxel, vox.size() * sizeof(int), hipMemcpyDeviceToHost));
    for (auto& c : cell) c = vox[c];
    ctx->note(hipMemcpy(v.px, x.data(), n * sizeof(float), hipMemcpyHostToDevice));
    ctx->note(hipMemcpy(v.py, y.data(), n * sizeof(float), hipMemcpyHostToDevice));
    ctx->note(hipMemcpy(v.pz, z.data(), n * sizeof(float), hipMemcpyHostToDevice));
    ctx->note(hipMemcpy(v.mode, mode, n * sizeof(int), hipMemcpyHostToDevice));
    ctx->note(hipMemcpy(v.occ, occupation, n * sizeof(double), hipMemcpyHostToDevice));
    ctx->note(hipMemcpy(v.cell, cell.data(), n * sizeof(int), hipMemcpyHostToDevice));
    ctx->note(hipMemcpy(v.uid, uid.data(), n * sizeof(unsigned), hipMemcpyHostToDevice));
    return finish_population(ctx, n);
}

} // extern "C"

namespace {
// Launches one kernel (or kernel group) of class c, bracketed by events when timing.
struct StepTimer {
    pmc_ctx* ctx;
    int used = 0;
    int cls[pmc_ctx::kEvents / 2];
    template <class F>
    void run(int c, F&& launch) {
        if (!ctx->timing || used >= pmc_ctx::kEvents / 2) {
            launch();
            return;
        }
        ctx->note(hipEventRecord(ctx->ev[2 * used], ctx->stream));
        launch();
        ctx->note(hipEventRecord(ctx->ev[2 * used + 1], ctx->stream));
        cls[used++] = c;
    }
    void collect(bool whole_step = true) {
        if (!used) return;
        ctx->note(hipEventSynchronize(ctx->ev[2 * used - 1]));
        for (int k = 0; k < used; ++k) {
            ctx->kernel_ms[cls[k]] += elapsed(ctx->ev[2 * k], ctx->ev[2 * k + 1]);
            ctx->kernel_launches[cls[k]] += 1;
        }
        if (!whole_step) return;
        ctx->kernel_ms[PMC_KERNEL_STEP] += elapsed(ctx->ev[0], ctx->ev[2 * used - 1]);
        ctx->kernel_launches[PMC_KERNEL_STEP] += 1;
    }
};
} // namespace

extern "C" {

int pmc_step(pmc_ctx* ctx, int32_t nsteps) {
    if (!ctx) return PMC_ERR_ARGUMENT;
    if (ctx->latched) return PMC_ERR_LATCHED; // the first failure's message stays in ctx->error
    if (ctx->owner) return fail(ctx, PMC_ERR_ARGUMENT, "context belongs to a batch: step it with pmc_batch_step");
    if (!ctx->initialized) return fail(ctx, PMC_ERR_ARGUMENT, "carriers not initialized");
    if (nsteps < 0) return fail(ctx, PMC_ERR_ARGUMENT, "negative step count");
    for (int s = 0; s < nsteps; ++s) {
        View& v = view(ctx);
        hipStream_t st = ctx->stream;
        StepTimer T{ctx};
        T.run(PMC_KERNEL_BEGIN, [&] {
            pmc_launch::begin(v, st);
            if (ctx->host_step % 100 == 0) pmc_launch::containment(v, st);
        });
        if (v.hs_on)
            T.run(PMC_KERNEL_SOURCE, [&] { pmc_launch::source(v, st, 0); });
        T.run(PMC_KERNEL_TRANSPORT, [&] { pmc_launch::transport(v, st); });
        T.run(PMC_KERNEL_EVENTS, [&] { pmc_launch::events(v, st); });
        T.run(PMC_KERNEL_INJECT, [&] { pmc_launch::inject(v, st); });
        T.run(PMC_KERNEL_REFRESH, [&] { pmc_launch::refresh(v, st, 1); });
        T.run(PMC_KERNEL_SORT, [&] { pmc_launch::sort(v, st); });
        ctx->cur = 1 - ctx->cur;
        View& w = view(ctx);
        T.run(PMC_KERNEL_REFRESH, [&] {
            if (w.inversion == 1) pmc_launch::invert(w, st, 0);
            pmc_launch::cell_prep(w, st);
        });
        if (w.callaway) {
            T.run(PMC_KERNEL_CALLAWAY, [&] { pmc_launch::callaway(w, st); });
        } else if (w.nl_callaway) {
            T.run(PMC_KERNEL_CALLAWAY, [&] { pmc_launch::nlcallaway(w, st); });
        } else {
            T.run(PMC_KERNEL_MOMENTS, [&] { pmc_launch::moments(w, st); });
            T.run(PMC_KERNEL_NEWTON, [&] { pmc_launch::newton(w, st); });
            T.run(PMC_KERNEL_RELAX, [&] { pmc_launch::relax(w, st); });
        }
        if (w.hs_on) T.run(PMC_KERNEL_SOURCE, [&] { pmc_launch::source(w, st, 1); });
        T.run(PMC_KERNEL_REFRESH, [&] {
            pmc_launch::refresh(w, st, 2);
            if (w.inversion == 1) pmc_launch::invert(w, st, 1);
        });
        int rc = check_hip(ctx, hipGetLastError(), "step launch");
        if (rc) {
            ctx->latched = true;
            return rc;
        }
        T.collect();
        ctx->host_step++;
        ctx->elapsed_ps += v.dt;
        if (ctx->host_step % ctx->error_interval == 0 && s != nsteps - 1) {
            rc = collect_error(ctx);
            if (rc) return rc;
        }
    }
    return collect_error(ctx);
}

int pmc_read_cells(pmc_ctx* ctx, double* temperature, double* energy_density) {
    if (!ctx || !ctx->has_geometry) return PMC_ERR_ARGUMENT;
    const View& v = ctx->view;
    ctx->note(hipStreamSynchronize(ctx->stream));
    if (temperature) ctx->note(hipMemcpy(temperature, v.cell_T, v.ncells * sizeof(double), hipMemcpyDeviceToHost));
    if (energy_density) ctx->note(hipMemcpy(energy_density, v.cell_e, v.ncells * sizeof(double), hipMemcpyDeviceToHost));
    return check_hip(ctx, hipGetLastError(), "read cells");
}

int pmc_read_heat_flux(pmc_ctx* ctx, double* flux) {
    if (!ctx || !ctx->initialized || !flux) return PMC_ERR_ARGUMENT;
    View& v = view(ctx);
    StepTimer T{ctx};
    T.run(PMC_KERNEL_OUTPUT, [&] { pmc_launch::flux(v, ctx->stream); });
    T.collect(false);
    ctx->note(hipStreamSynchronize(ctx->stream));
    ctx->note(hipMemcpy(flux, v.cell_flux, 3 * v.ncells * sizeof(double), hipMemcpyDeviceToHost));
    return check_hip(ctx, hipGetLastError(), "heat flux");
}

int pmc_read_source_pending(pmc_ctx* ctx, double* pending) {
    if (!ctx || !pending || !ctx->has_geometry) return PMC_ERR_ARGUMENT;
    const View& v = ctx->view;
    ctx->note(hipStreamSynchronize(ctx->stream));
    if (v.hs_on) ctx->note(hipMemcpy(pending, v.hs_pending, v.ncells * sizeof(double), hipMemcpyDeviceToHost));
    else std::fill(pending, pending + v.ncells, 0.0);
    return check_hip(ctx, hipGetLastError(), "read source pending");
}

int pmc_read_ledger(pmc_ctx* ctx, pmc_ledger* out) {
    if (!ctx || !out) return PMC_ERR_ARGUMENT;
    DevState st{};
    ctx->note(hipStreamSynchronize(ctx->stream));
    ctx->note(hipMemcpy(&st, ctx->view.st, sizeof(DevState), hipMemcpyDeviceToHost));
    out->step = st.step;
    out->live_carriers = st.n_live;
    out->slots = st.n_slots;
    out->step_absorbed = st.step_absorbed;
    out->step_injected = st.step_injected;
    out->step_absorbed_energy_ev = st.step_absorbed_e;
    out->step_injected_energy_ev = st.step_injected_e;
    out->step_lifetime_residual_ev = st.step_residual_e;
    out->step_balance_residual_ev = st.step_balance_e;
    out->total_absorbed = st.total_absorbed;
    out->total_injected = st.total_injected;
    out->total_absorbed_energy_ev = st.total_absorbed_e;
    out->total_injected_energy_ev = st.total_injected_e;
    out->total_lifetime_residual_ev = st.total_residual_e;
    out->total_balance_residual_ev = st.total_balance_e;
    out->thermal_energy_ev = st.thermal_e;
    out->carrier_steps = st.carrier_steps;
    out->rough_events = st.rough_events;
    out->interface_events = st.iface_events;
    out->interface_transmitted = st.iface_transmitted;
    out->rough_specular = st.rough_specular;
    out->rough_diffuse = st.rough_diffuse;
    out->newton_passes = st.newton_passes;
    out->newton_evals = st.newton_evals;
    out->step_source_energy_ev = st.step_source_e;
    out->total_source_energy_ev = st.total_source_e;
    out->source_updates = st.source_updates;
    out->source_exact_passes = st.source_passes;
    out->step_source_updates = st.step_source_updates;
    out->source_time_ps = st.source_time;
    out->source_pending_ev = 0.0;
    if (ctx->view.hs_on) { // PrescribedHeatSource::pending_energy_ev: the cells summed in order
        std::vector<double> pending(ctx->view.ncells);
        ctx->note(hipMemcpy(pending.data(), ctx->view.hs_pending, pending.size() * sizeof(double),
                            hipMemcpyDeviceToHost));
        for (double e : pending) out->source_pending_ev += e;
    }
    return check_hip(ctx, hipGetLastError(), "read ledger");
}

int pmc_read_carriers(pmc_ctx* ctx, int64_t* n, float* position, int32_t* mode, double* occupation, int32_t* cell) {
    if (!ctx || !n) return PMC_ERR_ARGUMENT;
    View& v = view(ctx);
    DevState st{};
    ctx->note(hipStreamSynchronize(ctx->stream));
    ctx->note(hipMemcpy(&st, v.st, sizeof(DevState), hipMemcpyDeviceToHost));
    const int64_t count = st.n_slots;
    // *n is the caller's capacity (slots) whenever it asks for data.
    if ((position || mode || occupation || cell) && (*n < 0 || count > *n)) {
        *n = count;
        return fail(ctx, PMC_ERR_ARGUMENT, "carrier buffers hold fewer slots than are in use");
    }
    if (position) {
        std::vector<float> x(count), y(count), z(count);
        ctx->note(hipMemcpy(x.data(), v.px, count * sizeof(float), hipMemcpyDeviceToHost));
        ctx->note(hipMemcpy(y.data(), v.py, count * sizeof(float), hipMemcpyDeviceToHost));
        ctx->note(hipMemcpy(z.data(), v.pz, count * sizeof(float), hipMemcpyDeviceToHost));
        for (int64_t i = 0; i < count; ++i) {
            position[3 * i] = x[i];
            position[3 * i + 1] = y[i];
            position[3 * i + 2] = z[i];
        }
    }
    if (mode) ctx->note(hipMemcpy(mode, v.mode, count * sizeof(int), hipMemcpyDeviceToHost));
    if (occupation) ctx->note(hipMemcpy(occupation, v.occ, count * sizeof(double), hipMemcpyDeviceToHost));
    if (cell) ctx->note(hipMemcpy(cell, v.cell, count * sizeof(int), hipMemcpyDeviceToHost));
    *n = count;
    return check_hip(ctx, hipGetLastError(), "read carriers");
}

int pmc_read_uids(pmc_ctx* ctx, int64_t n, uint32_t* uid) {
    if (!ctx || !uid) return PMC_ERR_ARGUMENT;
    View& v = view(ctx);
    ctx->note(hipStreamSynchronize(ctx->stream));
    DevState st{};
    ctx->note(hipMemcpy(&st, v.st, sizeof(DevState), hipMemcpyDeviceToHost));
    if (n < 0 || n > st.n_slots) return fail(ctx, PMC_ERR_ARGUMENT, "uid count outside the slots in use");
    ctx->note(hipMemcpy(uid, v.uid, n * sizeof(unsigned), hipMemcpyDeviceToHost));
    return check_hip(ctx, hipGetLastError(), "read uids");
}

int pmc_last_error(pmc_ctx* ctx, pmc_error* error) {
    if (!ctx || !error) return PMC_ERR_ARGUMENT;
    *error = ctx->error;
    return PMC_OK;
}

int pmc_set_tuning(pmc_ctx* ctx, const pmc_tuning* t) {
    if (!ctx || !t) return PMC_ERR_ARGUMENT;
    if (ctx->owner) return fail(ctx, PMC_ERR_ARGUMENT, "context belongs to a batch: tuning is fixed until the batch is destroyed");
    if (!(t->series_radius >= 0.0) || !(t->series_radius <= 0.5))
        return fail(ctx, PMC_ERR_ARGUMENT, "series_radius must lie in [0, 0.5]");
    ctx->view.series_on = t->series_on ? 1 : 0;
    ctx->view.series_radius = t->series_radius;
    return PMC_OK;
}

int pmc_get_tuning(pmc_ctx* ctx, pmc_tuning* t) {
    if (!ctx || !t) return PMC_ERR_ARGUMENT;
    *t = pmc_tuning{};
    t->series_on = ctx->view.series_on;
    t->series_radius = ctx->view.series_radius;
    return PMC_OK;
}

int pmc_kernel_times(pmc_ctx* ctx, double* total_ms, int64_t* launches, int32_t n) {
    if (!ctx) return PMC_ERR_ARGUMENT;
    for (int i = 0; i < n && i < PMC_KERNEL_CLASSES; ++i) {
        if (total_ms) total_ms[i] = ctx->kernel_ms[i];
        if (launches) launches[i] = ctx->kernel_launches[i];
    }
    return PMC_OK;
}

int pmc_reset_kernel_times(pmc_ctx* ctx) {
    if (!ctx) return PMC_ERR_ARGUMENT;
    for (int i = 0; i < PMC_KERNEL_CLASSES; ++i) {
        ctx->kernel_ms[i] = 0;
        ctx->kernel_launches[i] = 0;
    }
    return PMC_OK;
}

int pmc_set_timing(pmc_ctx* ctx, int32_t enabled) {
    if (!ctx) return PMC_ERR_ARGUMENT;
    ctx->timing = enabled != 0;
    return PMC_OK;
}

void pmc_destroy(pmc_ctx* ctx) {
    if (!ctx) return;
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    for (void* p : ctx->allocations) (void)hipFree(p);
    for (auto& e : ctx->ev)
        if (e) (void)hipEventDestroy(e);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

} // extern "C"

// ---------------------------------------------------------------- layered materials
extern "C" int pmc_upload_layers(pmc_ctx* ctx, const pmc_layers* L) {
    if (!ctx || !L) return PMC_ERR_ARGUMENT;
    View& v = ctx->view;
    if (!ctx->has_material || !ctx->has_geometry) return fail(ctx, PMC_ERR_ARGUMENT, "layers need material and geometry");
    if (ctx->initialized) return fail(ctx, PMC_ERR_ARGUMENT, "layers must be uploaded before the population is created");
    if (v.linearized) {
        if (!L->heat_capacity_density)
            return fail(ctx, PMC_ERR_ARGUMENT, "linearized layers need each material's heat-capacity density");
        for (int k = 0; k < L->count && k < pmc_dev::kMaxLayers; ++k)
            if (!(L->heat_capacity_density[k] > 0.0) || !std::isfinite(L->heat_capacity_density[k]))
                return fail(ctx, PMC_ERR_ARGUMENT, "linearized layers need positive heat-capacity densities");
    }
    if (v.inversion != 0)
        return fail(ctx, PMC_ERR_UNSUPPORTED, "Newton temperature inversion covers one material; layers use their tables");
    if (v.geom_kind != 0) return fail(ctx, PMC_ERR_UNSUPPORTED, "layered materials need the box geometry");
    const int n = L->count;
    if (n < 2 || n > pmc_dev::kMaxLayers || L->axis < 0 || L->axis > 2 || !L->planes || !L->mode_begin ||
        !L->cell_material || !L->background_energy_density || !L->lut_begin || !L->spatial_volume || !L->bank_begin ||
        !L->bank_mode || !L->bank_omega || !L->bank_prefix || !(L->band_width > 0.0))
        return fail(ctx, PMC_ERR_ARGUMENT, "invalid layered-material tables");
    if (L->mode_begin[0] != 0 || L->mode_begin[n] != v.na || L->lut_begin[0] != 0 || L->lut_begin[n] != v.lut_n ||
        L->bank_begin[0] != 0)
        return fail(ctx, PMC_ERR_ARGUMENT, "layer mode or table ranges do not cover the tables");
    double vmin = L->spatial_volume[0];
    for (int k = 0; k < n; ++k) {
        if (L->mode_begin[k + 1] <= L->mode_begin[k] || L->lut_begin[k + 1] < L->lut_begin[k] + 2 ||
            !(L->spatial_volume[k] > 0.0))
            return fail(ctx, PMC_ERR_ARGUMENT, "empty material in the layered tables");
        vmin = std::min(vmin, L->spatial_volume[k]);
    }
    for (int k = 0; k + 1 < n; ++k)
        if (!(L->planes[k] > v.bmin[L->axis] && L->planes[k] < v.bmax[L->axis]) ||
            (k && !(L->planes[k] > L->planes[k - 1])))
            return fail(ctx, PMC_ERR_ARGUMENT, "material interfaces must be increasing and inside the box");
    for (int b = 0; b < 2 * n; ++b)
        if (L->bank_begin[b + 1] < L->bank_begin[b]) return fail(ctx, PMC_ERR_ARGUMENT, "malformed interface banks");
    const int nb = L->bank_begin[2 * n];
    for (int j = 0; j < nb; ++j)
        if (L->bank_mode[j] < 0 || L->bank_mode[j] >= v.na)
            return fail(ctx, PMC_ERR_ARGUMENT, "interface bank mode out of range");
    std::vector<int> cm(L->cell_material, L->cell_material + v.ncells);
    for (int c : cm)
        if (c < 0 || c >= n) return fail(ctx, PMC_ERR_ARGUMENT, "cell material out of range");
    // Band start positions of every bank: the two lower_bound searches of a DMM
    // band draw become two loads.  Same products j * band as the device's
    // floor(w / band) * band, so the same band boundaries.  Narrower bands than
    // the index holds (over 2^20 bands) keep the two searches on the device.
    double wmax = 0.0;
    for (int i = 0; i < nb; ++i) wmax = std::max(wmax, L->bank_omega[i]);
    const double bands = std::floor(wmax / L->band_width) + 1.0;
    const int nbands = bands <= static_cast<double>(1 << 20) ? static_cast<int>(bands) : 0;
    std::vector<int32_t> starts(static_cast<size_t>(2 * n) * (nbands + 1));
    for (int b = 0; b < 2 * n && nbands; ++b) {
        const double* om = L->bank_omega + L->bank_begin[b];
        const int cnt = L->bank_begin[b + 1] - L->bank_begin[b];
        for (int j = 0; j <= nbands; ++j)
            starts[static_cast<size_t>(b) * (nbands + 1) + j] =
                static_cast<int32_t>(std::lower_bound(om, om + cnt, static_cast<double>(j) * L->band_width) - om);
    }
    // Reflection tables by material: each must cover exactly its material's modes.
    std::vector<int32_t> fr;
    if (L->facet_rough) {
        fr.assign(L->facet_rough, L->facet_rough + static_cast<size_t>(n) * v.nfacets);
        for (int k = 0; k < n; ++k)
            for (int f = 0; f < v.nfacets; ++f) {
                const int t = fr[static_cast<size_t>(k) * v.nfacets + f];
                if (t < -1 || t >= static_cast<int>(ctx->rough_range.size()))
                    return fail(ctx, PMC_ERR_ARGUMENT, "layer reflection table out of range");
                if (t >= 0 && (ctx->rough_range[t].first != L->mode_begin[k] ||
                               ctx->rough_range[t].second != L->mode_begin[k + 1] - L->mode_begin[k]))
                    return fail(ctx, PMC_ERR_ARGUMENT, "layer reflection table does not cover its material's modes");
            }
    }
    v.layer_axis = L->axis;
    for (int k = 0; k + 1 < n; ++k) v.layer_plane[k] = static_cast<float>(L->planes[k]);
    for (int k = 0; k <= n; ++k) {
        v.mode_begin[k] = L->mode_begin[k];
        v.lut_begin[k] = L->lut_begin[k];
    }
    for (int k = 0; k < n; ++k) {
        v.e_bg_mat[k] = L->background_energy_density[k];
        v.cbg_mat[k] = v.linearized ? L->heat_capacity_density[k] : 0.0;
        v.layer_accept[k] = vmin / L->spatial_volume[k];
    }
    for (int b = 0; b <= 2 * n; ++b) v.bank_begin[b] = L->bank_begin[b];
    v.band = L->band_width;
    v.cell_mat = ctx->upload(cm.data(), cm.size());
    v.layer_rough = fr.empty() ? nullptr : ctx->upload(fr.data(), fr.size());
    v.bank_mode = ctx->upload(L->bank_mode, static_cast<size_t>(nb));
    v.bank_omega = ctx->upload(L->bank_omega, static_cast<size_t>(nb));
    v.bank_prefix = ctx->upload(L->bank_prefix, static_cast<size_t>(nb) + 2 * n);
    // AMM rows (all-or-nothing: validated before anything is uploaded).
    const double fa = L->amm_fraction;
    if (!(fa >= 0.0 && fa <= 1.0)) return fail(ctx, PMC_ERR_ARGUMENT, "AMM fraction must be in [0,1].");
    int n_rows = 0, n_edges = 0;
    if (fa > 0.0) {
        if (!L->amm_block || !L->amm_reflected || !L->amm_edge_begin)
            return fail(ctx, PMC_ERR_ARGUMENT, "AMM tables missing");
        if (L->amm_block[0] != 0) return fail(ctx, PMC_ERR_ARGUMENT, "malformed AMM blocks");
        for (int f = 0; f + 1 < n; ++f)
            for (int sd = 0; sd < 2; ++sd) {
                const int b = 2 * f + sd, src = f + sd;
                if (L->amm_block[b + 1] - L->amm_block[b] != L->mode_begin[src + 1] - L->mode_begin[src] + 1)
                    return fail(ctx, PMC_ERR_ARGUMENT, "AMM block sizes do not match the materials");
            }
        n_rows = L->amm_block[2 * (n - 1)];
        n_edges = L->amm_edge_begin[n_rows - 1];
        if (n_edges > 0 && (!L->amm_edge_mode || !L->amm_edge_prob))
            return fail(ctx, PMC_ERR_ARGUMENT, "AMM edge arrays missing");
        for (int r = 0; r < n_rows; ++r) {
            if (L->amm_reflected[r] < -1 || L->amm_reflected[r] >= v.na || L->amm_edge_begin[r] < 0 ||
                L->amm_edge_begin[r] > n_edges || (r && L->amm_edge_begin[r] < L->amm_edge_begin[r - 1]))
                return fail(ctx, PMC_ERR_ARGUMENT, "AMM rows out of range");
        }
        for (int e = 0; e < n_edges; ++e)
            if (L->amm_edge_mode[e] < 0 || L->amm_edge_mode[e] >= v.na || !std::isfinite(L->amm_edge_prob[e]) ||
                L->amm_edge_prob[e] < 0.0)
                return fail(ctx, PMC_ERR_ARGUMENT, "AMM edges out of range");
    }
    v.band_start = nbands ? ctx->upload(starts.data(), starts.size()) : nullptr;
    v.band_count = nbands;
    v.amm_fraction = fa;
    if (fa > 0.0) {
        for (int b = 0; b <= 2 * (n - 1); ++b) v.amm_block[b] = L->amm_block[b];
        v.amm_refl = ctx->upload(L->amm_reflected, static_cast<size_t>(n_rows));
        v.amm_ebeg = ctx->upload(L->amm_edge_begin, static_cast<size_t>(n_rows));
        v.amm_emode = n_edges ? ctx->upload(L->amm_edge_mode, static_cast<size_t>(n_edges)) : nullptr;
        v.amm_eprob = n_edges ? ctx->upload(L->amm_edge_prob, static_cast<size_t>(n_edges)) : nullptr;
    }
    v.layers = n;
    return check_hip(ctx, hipGetLastError(), "layers");
}

// ---------------------------------------------------------------- launch configuration
extern "C" {

int pmc_get_launch_config(pmc_ctx* ctx, pmc_launch_config* out) {
    if (!ctx || !out) return PMC_ERR_ARGUMENT;
    out->event_split = ctx->view.event_split;
    out->callaway_threads = ctx->view.cw_threads;
    return PMC_OK;
}

int pmc_set_launch_config(pmc_ctx* ctx, const pmc_launch_config* c) {
    if (!ctx || !c) return PMC_ERR_ARGUMENT;
    if (ctx->owner) return fail(ctx, PMC_ERR_ARGUMENT, "context belongs to a batch: its launch split is fixed");
    if (!ctx->has_geometry) return fail(ctx, PMC_ERR_ARGUMENT, "geometry must be uploaded before the launch configuration");
    const int s = c->event_split, t = c->callaway_threads;
    if (!(s == 1 || s == 2 || s == 4 || s == 8) || s > pmc_dev::kMaxEventSplit)
        return fail(ctx, PMC_ERR_ARGUMENT, "event_split must be 1, 2, 4 or 8");
    if (!(t == 256 || t == 384 || t == pmc_dev::kCwThreads))
        return fail(ctx, PMC_ERR_ARGUMENT, "callaway_threads must be 256, 384 or 1024");
    ctx->view.event_split = s;
    ctx->view.cw_threads = t;
    return PMC_OK;
}

int pmc_autotune(pmc_ctx* ctx, int32_t steps, pmc_launch_config* chosen, double* report_ms, int32_t report_len) {
    if (!ctx || steps <= 0 || report_len < 0) return PMC_ERR_ARGUMENT;
    if (ctx->latched) return PMC_ERR_LATCHED;
    if (ctx->owner) return fail(ctx, PMC_ERR_ARGUMENT, "context belongs to a batch: tune it before pmc_batch_create");
    if (!ctx->initialized) return fail(ctx, PMC_ERR_ARGUMENT, "carriers not initialized");
    if (report_ms)
        for (int k = 0; k < report_len; ++k) report_ms[k] = 0.0;
    const bool timing = ctx->timing;
    ctx->timing = true;
    auto trial = [&](int cls, double& ms) {
        pmc_reset_kernel_times(ctx);
        const int rc = pmc_step(ctx, steps);
        ms = ctx->kernel_launches[cls] ? ctx->kernel_ms[cls] / static_cast<double>(ctx->kernel_launches[cls]) : 0.0;
        return rc;
    };
    int rc = PMC_OK;
    // Event blocks per transport block.
    const int splits[4] = {1, 2, 4, 8};
    int best_split = ctx->view.event_split;
    double best = -1.0;
    for (int k = 0; k < 4 && rc == PMC_OK; ++k) {
        ctx->view.event_split = splits[k];
        double ms = 0.0;
        rc = trial(PMC_KERNEL_EVENTS, ms);
        if (k < report_len && report_ms) report_ms[k] = ms;
        if (rc == PMC_OK && (best < 0.0 || ms < best)) {
            best = ms;
            best_split = splits[k];
        }
    }
    ctx->view.event_split = best_split;
    // Callaway workgroup size, among those the fullest cell fits with a margin.
    if (rc == PMC_OK && ctx->view.callaway) {
        const View& v = view(ctx);
        std::vector<int> start(static_cast<size_t>(v.ncells) + 1);
        ctx->note(hipStreamSynchronize(ctx->stream));
        ctx->note(hipMemcpy(start.data(), v.cell_start, start.size() * sizeof(int), hipMemcpyDeviceToHost));
        int fullest = 0;
        for (int c = 0; c < v.ncells; ++c) fullest = std::max(fullest, start[c + 1] - start[c]);
        const int sizes[3] = {256, 384, pmc_dev::kCwThreads};
        int best_t = ctx->view.cw_threads;
        best = -1.0;
        for (int k = 0; k < 3 && rc == PMC_OK; ++k) {
            if (sizes[k] != pmc_dev::kCwThreads && 3 * fullest > 2 * sizes[k] * pmc_dev::kCwPer)
                continue; // the fullest cell needs 1.5x headroom
            ctx->view.cw_threads = sizes[k];
            double ms = 0.0;
            rc = trial(PMC_KERNEL_CALLAWAY, ms);
            if (4 + k < report_len && report_ms) report_ms[4 + k] = ms;
            if (rc == PMC_OK && (best < 0.0 || ms < best)) {
                best = ms;
                best_t = sizes[k];
            }
        }
        ctx->view.cw_threads = best_t;
    }
    pmc_reset_kernel_times(ctx);
    ctx->timing = timing;
    if (chosen) pmc_get_launch_config(ctx, chosen);
    return rc;
}

} // extern "C"

// ---------------------------------------------------------------- batches
struct pmc_batch {
    std::vector<pmc_ctx*> ctx;
    hipStream_t stream = nullptr;
    View* views[2] = {};                    // device copies of the instances' views, by carrier buffer
    // Per kernel kind, the launches of one batched step: the device prefix
    // offsets of the instances' block counts, shifted to the launch's first
    // block, and the launch's block count.  A kind of more than kMaxBlocks
    // blocks in all is split (a launch's work-items must stay below 2^32).
    struct Piece {
        int* offs;
        int blocks;
    };
    static constexpr int kMaxBlocks = 1 << 21;
    std::vector<Piece> pieces[pmc_launch::kBatchKinds];
    int* errors = nullptr;                  // device: error word per instance
    bool mesh = false, callaway = false, source = false;
    bool use_graph = true;
    hipGraphExec_t exec[2] = {};
    int failed = -1;
    std::vector<void*> allocations;
};

namespace {
template <class T>
T* batch_upload(pmc_batch* b, const T* host, size_t n) {
    void* p = nullptr;
    if (hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T)) != hipSuccess) return nullptr;
    b->allocations.push_back(p);
    if (hipMemcpy(p, host, n * sizeof(T), hipMemcpyHostToDevice) != hipSuccess) return nullptr;
    return static_cast<T*>(p);
}

// The step's kernels after the optional containment check, for the instances'
// carrier buffer `cur` (the sort writes the other one).
void batch_kind(pmc_batch* b, int kind, int buf) {
    const int n = static_cast<int>(b->ctx.size());
    for (const pmc_batch::Piece& p : b->pieces[kind])
        pmc_launch::batch_launch(kind, b->mesh, b->views[buf], p.offs, n, p.blocks, b->stream);
}

void batch_sequence(pmc_batch* b, int cur) {
    using namespace pmc_launch;
    auto run = [&](int kind, int buf) { batch_kind(b, kind, buf); };
    run(kBatchBegin, cur);
    if (b->source) { // first half-step deposit (src/MonteCarloSolver.cpp:366-369)
        run(kBatchSourceSums0, cur);
        run(kBatchSourceSolve0, cur);
        run(kBatchSourceIncrements, cur);
        run(kBatchSourceApply0, cur);
    }
    run(kBatchTransport, cur);
    run(kBatchEvents, cur);
    run(kBatchInject, cur);
    run(kBatchRefresh1, cur);
    run(kBatchSort, cur);
    run(kBatchCellPrep, 1 - cur);
    if (b->callaway) {
        run(kBatchCallawayWave, 1 - cur);
        run(kBatchCallaway, 1 - cur);
    } else {
        run(kBatchMoments, 1 - cur);
        run(kBatchNewton, 1 - cur);
        run(kBatchRelax, 1 - cur);
    }
    if (b->source) { // second half-step deposit (:377-379)
        run(kBatchSourceSums1, 1 - cur);
        run(kBatchSourceSolve1, 1 - cur);
        run(kBatchSourceIncrements, 1 - cur);
        run(kBatchSourceApply1, 1 - cur);
    }
    run(kBatchRefresh2, 1 - cur);
}

// Reads every instance's error word; latches the first failed context.
int batch_check(pmc_batch* b) {
    const int n = static_cast<int>(b->ctx.size());
    std::vector<int> codes(n, 0);
    pmc_launch::batch_errors(b->views[0], n, b->errors, b->stream);
    hipError_t e = hipMemcpyAsync(codes.data(), b->errors, n * sizeof(int), hipMemcpyDeviceToHost, b->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
    if (e != hipSuccess) {
        b->failed = 0;
        for (pmc_ctx* c : b->ctx) c->latched = true;
        return fail(b->ctx[0], PMC_ERR_DEVICE, std::string("batch: ") + hipGetErrorString(e));
    }
    for (int i = 0; i < n; ++i)
        if (codes[i]) {
            b->failed = i;
            return collect_error(b->ctx[i]);
        }
    return PMC_OK;
}
} // namespace

extern "C" {

int pmc_batch_create(pmc_ctx* const* contexts, int32_t n, pmc_batch** out) {
    if (!contexts || n <= 0 || !out) return PMC_ERR_ARGUMENT;
    *out = nullptr;
    pmc_ctx* c0 = contexts[0];
    if (!c0) return PMC_ERR_ARGUMENT;
    for (int i = 0; i < n; ++i) {
        pmc_ctx* c = contexts[i];
        if (!c) return PMC_ERR_ARGUMENT;
        if (c->latched) return PMC_ERR_LATCHED;
        if (!c->initialized) return fail(c, PMC_ERR_ARGUMENT, "batch: carriers not initialized");
        if (c->view.layers > 1) return fail(c, PMC_ERR_UNSUPPORTED, "batch: layered instances are not batched");
        if (c->view.inversion != 0)
            return fail(c, PMC_ERR_UNSUPPORTED, "batch: instances with Newton temperature inversion are not batched");
        if (c->view.nl_callaway)
            return fail(c, PMC_ERR_UNSUPPORTED, "batch: nonlinear Callaway instances are not batched");
        if (c->owner) return fail(c, PMC_ERR_ARGUMENT, "batch: a context already belongs to a batch");
        if (c->view.geom_kind != c0->view.geom_kind || c->view.callaway != c0->view.callaway)
            return fail(c, PMC_ERR_ARGUMENT, "batch: instances differ in geometry kind or collision model");
        if (c->cur != c0->cur || c->host_step != c0->host_step)
            return fail(c, PMC_ERR_ARGUMENT, "batch: instances are at different steps");
        for (int j = 0; j < i; ++j)
            if (contexts[j] == c) return fail(c, PMC_ERR_ARGUMENT, "batch: a context appears twice");
        if (hipStreamSynchronize(c->stream) != hipSuccess) return fail(c, PMC_ERR_DEVICE, "batch: stream sync");
    }
    auto* b = new pmc_batch();
    b->ctx.assign(contexts, contexts + n);
    b->mesh = c0->view.geom_kind == 1;
    b->callaway = c0->view.callaway != 0;
    for (pmc_ctx* c : b->ctx) b->source = b->source || c->view.hs_on;
    auto bail = [&](const char* what) {
        pmc_batch_destroy(b);
        return fail(c0, PMC_ERR_DEVICE, std::string("batch: ") + what);
    };
    if (hipStreamCreate(&b->stream) != hipSuccess) return bail("stream");
    for (int cur = 0; cur < 2; ++cur) {
        std::vector<View> host(n);
        for (int i = 0; i < n; ++i) {
            const int keep = b->ctx[i]->cur;
            b->ctx[i]->cur = cur;
            host[i] = view(b->ctx[i]);
            b->ctx[i]->cur = keep;
            view(b->ctx[i]);
        }
        if (!(b->views[cur] = batch_upload(b, host.data(), host.size()))) return bail("views");
    }
    for (int k = 0; k < pmc_launch::kBatchKinds; ++k) {
        std::vector<int> offs(n + 1, 0);
        // Each instance gets exactly its single-launch grid.  The kernels that map
        // blocks to XCDs have counts that are multiples of 8, so their ranges
        // start at multiples of 8 and keep the mapping.
        long long total = 0;
        for (int i = 0; i < n; ++i) {
            total += pmc_launch::batch_blocks(b->ctx[i]->view, k);
            if (total > INT_MAX / 2) return bail("too many workgroups");
            offs[i + 1] = static_cast<int>(total);
        }
        for (int base = 0; base < offs[n]; base += pmc_batch::kMaxBlocks) {
            std::vector<int> shifted(n + 1);
            for (int i = 0; i <= n; ++i) shifted[i] = offs[i] - base;
            int* d = batch_upload(b, shifted.data(), shifted.size());
            if (!d) return bail("offsets");
            b->pieces[k].push_back({d, std::min(pmc_batch::kMaxBlocks, offs[n] - base)});
        }
    }
    std::vector<int> zero(n, 0);
    if (!(b->errors = batch_upload(b, zero.data(), zero.size()))) return bail("error words");
    for (pmc_ctx* c : b->ctx) c->owner = b;
    *out = b;
    return PMC_OK;
}

int pmc_batch_set_graph(pmc_batch* b, int32_t enabled) {
    if (!b) return PMC_ERR_ARGUMENT;
    b->use_graph = enabled != 0;
    return PMC_OK;
}

int pmc_batch_failed(const pmc_batch* b) { return b ? b->failed : -1; }

int pmc_batch_step(pmc_batch* b, int32_t nsteps) {
    if (!b) return PMC_ERR_ARGUMENT;
    if (b->failed >= 0) return PMC_ERR_LATCHED;
    if (nsteps < 0) return fail(b->ctx[0], PMC_ERR_ARGUMENT, "negative step count");
    const int n = static_cast<int>(b->ctx.size());
    for (pmc_ctx* c : b->ctx)
        if (c->cur != b->ctx[0]->cur || c->host_step != b->ctx[0]->host_step || c->owner != b)
            return fail(c, PMC_ERR_INVALID_STATE, "batch: an instance left the batch's step");
    for (int s = 0; s < nsteps; ++s) {
        pmc_ctx* c0 = b->ctx[0];
        const int cur = c0->cur;
        if (c0->host_step % 100 == 0) batch_kind(b, pmc_launch::kBatchContainment, cur);
        if (b->use_graph) {
            if (!b->exec[cur]) { // capture this buffer parity's step once
                hipGraph_t g = nullptr;
                hipError_t e = hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal);
                if (e == hipSuccess) {
                    batch_sequence(b, cur);
                    e = hipStreamEndCapture(b->stream, &g);
                }
                if (e == hipSuccess) e = hipGraphInstantiate(&b->exec[cur], g, nullptr, nullptr, 0);
                if (g) (void)hipGraphDestroy(g);
                if (e != hipSuccess) {
                    b->failed = 0;
                    for (pmc_ctx* c : b->ctx) c->latched = true;
                    return fail(c0, PMC_ERR_DEVICE, std::string("batch graph: ") + hipGetErrorString(e));
                }
            }
            if (hipGraphLaunch(b->exec[cur], b->stream) != hipSuccess) {
                b->failed = 0;
                for (pmc_ctx* c : b->ctx) c->latched = true;
                return fail(c0, PMC_ERR_DEVICE, "batch graph launch failed");
            }
        } else {
            batch_sequence(b, cur);
        }
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) {
            b->failed = 0;
            for (pmc_ctx* c : b->ctx) c->latched = true;
            return fail(c0, PMC_ERR_DEVICE, std::string("batch launch: ") + hipGetErrorString(e));
        }
        for (pmc_ctx* c : b->ctx) {
            c->cur = 1 - c->cur;
            view(c);
            c->host_step++;
            c->elapsed_ps += c->view.dt;
        }
        if (c0->host_step % c0->error_interval == 0 && s != nsteps - 1) {
            const int rc = batch_check(b);
            if (rc) return rc;
        }
    }
    return batch_check(b);
}

void pmc_batch_destroy(pmc_batch* b) {
    if (!b) return;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    for (pmc_ctx* c : b->ctx)
        if (c && c->owner == b) c->owner = nullptr;
    for (auto& x : b->exec)
        if (x) (void)hipGraphExecDestroy(x);
    for (void* p : b->allocations) (void)hipFree(p);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
}

} // extern "C"
