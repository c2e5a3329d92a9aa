// CDNA4 (gfx950) kernels of the deviational Monte Carlo time step.
//
// One lane per carrier slot.  Carriers are a structure of arrays in HBM
// (FP32 position, packed mode word, FP64 occupation, cell id, Philox id) and are
// kept roughly cell-sorted so that wave-segmented reductions issue about one
// FP64 atomic per wave per tallied quantity.
#include "pmc_device.hpp"
#include "pmc_kernels.hpp"
#include "pmc_nlcallaway.hpp"

#include <hip/hip_runtime.h>
#include <math.h>

namespace pmc_dev {

// ---------------------------------------------------------------- random numbers
// Philox4x32-10 (Salmon et al., SC'11), counter = (id, step, draw, stream).
__device__ __forceinline__ uint4 philox(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0,
                                        uint32_t k1) {
    const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        const uint32_t hi0 = __umulhi(M0, c0), lo0 = M0 * c0;
        const uint32_t hi1 = __umulhi(M1, c2), lo1 = M1 * c2;
        const uint32_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
        c0 = n0;
        c1 = lo1;
        c2 = n2;
        c3 = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    return make_uint4(c0, c1, c2, c3);
}

// 53-bit uniform in [0, 1) from two words (lo, hi).
__device__ __forceinline__ double u53(uint32_t lo, uint32_t hi) {
    const uint64_t x = ((static_cast<uint64_t>(hi) << 32) | lo) >> 11;
    return static_cast<double>(x) * 0x1.0p-53;
}

// Uniforms of one transport event, drawn on demand in the order the reference
// consumes them (specular test, diffuse window, roulette).
struct EventUniforms {
    uint32_t uid, step, ev, k0, k1;
    int used;
    double u[4];
    __device__ double next() {
        if ((used & 1) == 0) {
            const uint4 w = philox(uid, step, 2 * ev + (used >> 1), kStreamTransport, k0, k1);
            u[0] = u53(w.x, w.y);
            u[1] = u53(w.z, w.w);
        }
        const double r = u[used & 1];
        ++used;
        return r;
    }
};

// ---------------------------------------------------------------- physics helpers
__device__ __forceinline__ double bose(double t, double w) {
    // src/material/Thermodynamics.cpp:71-82.
    if (t <= 0.0 || w <= 0.0) return 0.0;
    const double x = kHbar * w / (kBoltz * t);
    if (x > 700.0) return 0.0;
    const double ex = exp(x);
    return 1.0 / fmax(ex - 1.0, 1e-12);
}

__device__ __forceinline__ int grid_index(const View& V, double x, double y, double z) {
    // src/SimulationDomain.cpp:92-126 in FP64 on the FP32 position.
    const double p[3] = {x, y, z};
    int idx[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const int n = V.grid[a];
        if (n <= 1 || !(V.bmax[a] - V.bmin[a] > 0.0)) {
            idx[a] = 0;
        } else {
            int k = static_cast<int>(floor((p[a] - V.bmin[a]) * V.inv_cell[a]));
            idx[a] = k < 0 ? 0 : (k >= n ? n - 1 : k);
        }
    }
    return V.voxel[(idx[0] * V.grid[1] + idx[1]) * V.grid[2] + idx[2]];
}

__device__ __forceinline__ float4 carrier_velocity(const View& V, int mw) {
    float4 v = V.vel[mw & kModeMask];
    if (mw & (1 << kFlipShift)) v.x = -v.x;
    if (mw & (2 << kFlipShift)) v.y = -v.y;
    if (mw & (4 << kFlipShift)) v.z = -v.z;
    return v;
}

// ---------------------------------------------------------------- triangle mesh
// Squared distance from p to the triangle T (9 doubles): the Voronoi-region
// closest point of Ericson, Real-Time Collision Detection 5.1.5, the method of
// src/SurfaceMesh.cpp:111-159.
__device__ double tri_dist2(const double p[3], const double* T) {
    const double a[3] = {T[0], T[1], T[2]};
    double ab[3], ac[3], ap[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        ab[k] = T[3 + k] - a[k];
        ac[k] = T[6 + k] - a[k];
        ap[k] = p[k] - a[k];
    }
    auto d3 = [](const double* x, const double* y) { return x[0] * y[0] + x[1] * y[1] + x[2] * y[2]; };
    const double d1 = d3(ab, ap), d2 = d3(ac, ap);
    double v = 0.0, w = 0.0; // closest point a + v ab + w ac
    const double bp[3] = {p[0] - T[3], p[1] - T[4], p[2] - T[5]};
    const double cp[3] = {p[0] - T[6], p[1] - T[7], p[2] - T[8]};
    const double d3b = d3(ab, bp), d4 = d3(ac, bp), d5 = d3(ab, cp), d6 = d3(ac, cp);
    const double vc = d1 * d4 - d3b * d2, vb = d5 * d2 - d1 * d6, va = d3b * d6 - d5 * d4;
    if (d1 <= 0.0 && d2 <= 0.0) {
    } else if (d3b >= 0.0 && d4 <= d3b) {
        v = 1.0;
    } else if (vc <= 0.0 && d1 >= 0.0 && d3b <= 0.0) {
        v = d1 / (d1 - d3b);
    } else if (d6 >= 0.0 && d5 <= d6) {
        w = 1.0;
    } else if (vb <= 0.0 && d2 >= 0.0 && d6 <= 0.0) {
        w = d2 / (d2 - d6);
    } else if (va <= 0.0 && (d4 - d3b) >= 0.0 && (d5 - d6) >= 0.0) {
        w = (d4 - d3b) / ((d4 - d3b) + (d5 - d6));
        v = 1.0 - w;
    } else {
        const double den = 1.0 / (va + vb + vc);
        v = vb * den;
        w = vc * den;
    }
    double r2 = 0.0;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const double d = ap[k] - v * ab[k] - w * ac[k];
        r2 += d * d;
    }
    return r2;
}

// Slab test of the ray p + v t, t in [0, tmax], against a node box, in FP32:
// node boxes are padded by mesh_tol, eight FP32 quanta of the coordinates, which
// covers the rounding of this test.
__device__ __forceinline__ bool ray_box(const BvhDev& N, const float p[3], const float v[3], const float inv[3],
                                        float tmax) {
    float t0 = 0.0f, t1 = tmax;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        if (v[a] == 0.0f) {
            if (p[a] < N.bmin[a] || p[a] > N.bmax[a]) return false;
            continue;
        }
        float ta = (N.bmin[a] - p[a]) * inv[a], tb = (N.bmax[a] - p[a]) * inv[a];
        if (ta > tb) {
            const float x = ta;
            ta = tb;
            tb = x;
        }
        t0 = fmaxf(t0, ta);
        t1 = fminf(t1, tb);
    }
    return t0 <= t1;
}

__device__ __forceinline__ bool point_box(const BvhDev& N, const double p[3]) {
    return p[0] >= N.bmin[0] && p[0] <= N.bmax[0] && p[1] >= N.bmin[1] && p[1] <= N.bmax[1] && p[2] >= N.bmin[2] &&
           p[2] <= N.bmax[2];
}

// Nearest face the carrier leaves through along p + v t, 0 <= t <= tmax: the
// restatement of SurfaceMesh::trace_boundary_intersection
// (src/SurfaceMesh.cpp:1164-1270) for FP32 positions, limited to one flight.
// Only faces with v.n > 1e-10 count (from inside a closed surface the first
// crossing is always one of them); a point up to mesh_tol beyond such a face's
// plane is on it (t = 0); a crossing counts when it lies within mesh_tol of the
// triangle; an exact tie goes to the lower face.  Returns the face (BVH order),
// or -1 when no face lies within the flight.
__device__ int mesh_trace(const View& V, const double p[3], const double v[3], double tmax, double& t_hit) {
    const float pf[3] = {static_cast<float>(p[0]), static_cast<float>(p[1]), static_cast<float>(p[2])};
    const float vf[3] = {static_cast<float>(v[0]), static_cast<float>(v[1]), static_cast<float>(v[2])};
    const float inv[3] = {1.0f / vf[0], 1.0f / vf[1], 1.0f / vf[2]};
    const double tol = V.mesh_tol, tol2 = tol * tol;
    int node = 0, bf = -1;
    double best = tmax;
    while (node < V.n_face_nodes) { // threaded: a missed box skips its subtree
        const BvhDev N = V.face_nodes[node];
        // best rounded up, so the FP32 bound never cuts a box the FP64 one keeps
        if (!ray_box(N, pf, vf, inv, static_cast<float>(best) * 1.000001f + 1e-30f)) {
            node = N.count ? node + 1 : N.index;
            continue;
        }
        if (N.count) {
            for (int f = N.index; f < N.index + N.count; ++f) {
                const double4 P = V.face_plane[f];
                const double den = v[0] * P.x + v[1] * P.y + v[2] * P.z;
                if (!(den > 1e-10)) continue;
                const double s = p[0] * P.x + p[1] * P.y + p[2] * P.z + P.w;
                double t = -s / den;
                if (t < 0.0) {
                    if (s > tol) continue;
                    t = 0.0;
                }
                if (t > best || (t == best && bf >= 0 && f > bf)) continue;
                const double q[3] = {p[0] + v[0] * t, p[1] + v[1] * t, p[2] + v[2] * t};
                const double* B = V.face_box + 6 * f;
                if (q[0] < B[0] - tol || q[1] < B[1] - tol || q[2] < B[2] - tol || q[0] > B[3] + tol ||
                    q[1] > B[4] + tol || q[2] > B[5] + tol)
                    continue;
                if (tri_dist2(q, V.face_tri + 9 * f) > tol2) continue;
                best = t;
                bf = f;
            }
        }
        ++node;
    }
    t_hit = best;
    return bf;
}

// Distance-field voxel of p (FP32 lookup, clamped to the grid).
__device__ __forceinline__ size_t field_voxel(const View& V, const float p[3]) {
    int k[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const int i = static_cast<int>((p[a] - V.fbmin[a]) * V.field_inv_h);
        k[a] = i < 0 ? 0 : (i >= V.field_dims[a] ? V.field_dims[a] - 1 : i);
    }
    return (static_cast<size_t>(k[0]) * V.field_dims[1] + k[1]) * V.field_dims[2] + k[2];
}

// Lower bound of the distance from p to every face, from the distance field
// (0 without one).  Callers keep a margin of 2 mesh_tol for the FP32 lookup.
__device__ __forceinline__ float mesh_clearance(const View& V, const float p[3]) {
    if (!V.field) return 0.0f;
    return V.field[field_voxel(V, p)] * V.field_unit;
}

// Whether a flight of `t` ps at FP32 speed `speed` from p stays clear of every
// face by the distance field (then no face search is needed).
__device__ __forceinline__ bool mesh_clear(const View& V, const float p[3], float speed, float t) {
    const float reach = speed * t * 1.00001f + static_cast<float>(2.0 * V.mesh_tol);
    return reach < mesh_clearance(V, p);
}

// Odd number of distinct +x crossings (SurfaceMesh::contains_point_ray_cast,
// src/SurfaceMesh.cpp:1100-1123: Moller-Trumbore with 1e-10 tolerances,
// crossings closer than 1e-9 counted once).
__device__ bool mesh_ray_cast(const View& V, const double p[3]) {
    constexpr int kMaxHits = 32;
    double hits[kMaxHits];
    int nh = 0;
    const float pf[3] = {static_cast<float>(p[0]), static_cast<float>(p[1]), static_cast<float>(p[2])};
    const float vf[3] = {1.0f, 0.0f, 0.0f}, inv[3] = {1.0f, INFINITY, INFINITY};
    int node = 0;
    while (node < V.n_face_nodes) {
        const BvhDev N = V.face_nodes[node];
        if (!ray_box(N, pf, vf, inv, INFINITY)) {
            node = N.count ? node + 1 : N.index;
            continue;
        }
        for (int f = N.index; f < N.index + N.count; ++f) { // no iterations on an interior node
            const double* T = V.face_tri + 9 * f;
            const double e1[3] = {T[3] - T[0], T[4] - T[1], T[5] - T[2]};
            const double e2[3] = {T[6] - T[0], T[7] - T[1], T[8] - T[2]};
            const double pv[3] = {0.0, -e2[2], e2[1]}; // d x e2 with d = +x
            const double det = e1[1] * pv[1] + e1[2] * pv[2];
            if (fabs(det) < 1e-10) continue;
            const double id = 1.0 / det;
            const double sv[3] = {p[0] - T[0], p[1] - T[1], p[2] - T[2]};
            const double u = (sv[1] * pv[1] + sv[2] * pv[2]) * id;
            if (u < -1e-10 || u > 1.0 + 1e-10) continue;
            const double qv[3] = {sv[1] * e1[2] - sv[2] * e1[1], sv[2] * e1[0] - sv[0] * e1[2],
                                  sv[0] * e1[1] - sv[1] * e1[0]};
            const double w = qv[0] * id;
            if (w < -1e-10 || u + w > 1.0 + 1e-10) continue;
            const double t = (e2[0] * qv[0] + e2[1] * qv[1] + e2[2] * qv[2]) * id;
            if (!(t > 1e-10)) continue;
            if (nh == kMaxHits) return true; // a degenerate stack of crossings: not an escape
            int k = nh++;
            while (k > 0 && hits[k - 1] > t) {
                hits[k] = hits[k - 1];
                --k;
            }
            hits[k] = t;
        }
        ++node;
    }
    if (nh == 0) return false;
    int unique = 1;
    for (int k = 1; k < nh; ++k)
        if (fabs(hits[k] - hits[k - 1]) > 1e-9) ++unique;
    return unique & 1;
}

// Membership of an FP32 position: inside a star tetrahedron (1e-10 barycentric
// tolerance, src/SurfaceMesh.cpp:1126-1151), else the ray cast, else within
// mesh_tol of a face (a carrier that FP32 rounding put just outside the surface
// it sits on).
__device__ bool mesh_contains(const View& V, const double p[3]) {
    if (V.nsimplices > 0) {
        const double tol = 1e-10;
        int node = 0;
        while (node < V.n_simplex_nodes) {
            const BvhDev N = V.simplex_nodes[node];
            if (!point_box(N, p)) {
                node = N.count ? node + 1 : N.index;
                continue;
            }
            for (int j = N.index; j < N.index + N.count; ++j) {
                const double* B = V.simplex_box + 6 * j;
                if (p[0] < B[0] - tol || p[1] < B[1] - tol || p[2] < B[2] - tol || p[0] > B[3] + tol ||
                    p[1] > B[4] + tol || p[2] > B[5] + tol)
                    continue;
                const double* o = V.simplex_v + 12 * j;
                const double* A = V.simplex_inv + 9 * j;
                const double q[3] = {p[0] - o[0], p[1] - o[1], p[2] - o[2]};
                const double l1 = A[0] * q[0] + A[1] * q[1] + A[2] * q[2];
                const double l2 = A[3] * q[0] + A[4] * q[1] + A[5] * q[2];
                const double l3 = A[6] * q[0] + A[7] * q[1] + A[8] * q[2];
                const double l0 = 1.0 - l1 - l2 - l3;
                if (l0 >= -tol && l1 >= -tol && l2 >= -tol && l3 >= -tol && l0 <= 1.0 + tol && l1 <= 1.0 + tol &&
                    l2 <= 1.0 + tol && l3 <= 1.0 + tol)
                    return true;
            }
            ++node;
        }
    }
    if (mesh_ray_cast(V, p)) return true;
    const double tol2 = V.mesh_tol * V.mesh_tol;
    int node = 0;
    while (node < V.n_face_nodes) {
        const BvhDev N = V.face_nodes[node];
        if (!point_box(N, p)) {
            node = N.count ? node + 1 : N.index;
            continue;
        }
        for (int f = N.index; f < N.index + N.count; ++f)
            if (tri_dist2(p, V.face_tri + 9 * f) <= tol2) return true;
        ++node;
    }
    return false;
}

// Material of a global active mode (layered domains).
__device__ __forceinline__ int mode_material(const View& V, int m) {
    int k = 0;
    while (k + 1 < V.layers && m >= V.mode_begin[k + 1]) ++k;
    return k;
}

// InterfaceModeSampler::sample_band (src/solver/InterfaceModeSampler.cpp:72-109):
// in bank (material, direction), the modes whose frequency lies in the band
// [floor(w / bw) bw, (floor(w / bw) + 1) bw), drawn by their cumulative flux with
// uniform u.  Returns the global mode or -1; flux = the band's total flux.  The
// band's bounds come from the band index built at upload (the reference's two
// lower_bound searches, done once per band).
__device__ int sample_band(const View& V, int material, int direction, double w, double u, double& flux) {
    const int b = 2 * material + direction;
    const int b0 = V.bank_begin[b];
    const double* pre = V.bank_prefix + b0 + b; // bank size + 1 entries
    const double bin = floor(w / V.band);
    int begin, end;
    if (V.band_start) {
        if (!(bin < static_cast<double>(V.band_count))) { // above every frequency of every bank
            flux = 0.0;
            return -1;
        }
        // [lower_bound(bin * band), lower_bound((bin + 1) * band)) from the band index.
        const int* bs = V.band_start + b * (V.band_count + 1) + static_cast<int>(bin);
        begin = bs[0];
        end = bs[1];
    } else { // bands too narrow for the index: the two searches
        const double* om = V.bank_omega + b0;
        const int cnt = V.bank_begin[b + 1] - b0;
        const double lo_w = bin * V.band, hi_w = (bin + 1.0) * V.band;
        int l = 0, h = cnt;
        while (l < h) {
            const int mid = (l + h) >> 1;
            if (om[mid] < lo_w) l = mid + 1;
            else h = mid;
        }
        begin = l;
        h = cnt;
        while (l < h) {
            const int mid = (l + h) >> 1;
            if (om[mid] < hi_w) l = mid + 1;
            else h = mid;
        }
        end = l;
    }
    flux = end > begin ? pre[end] - pre[begin] : 0.0;
    if (!(flux > 0.0)) return -1;
    const double target = pre[begin] + u * flux;
    int l = begin + 1, h = end + 1; // upper_bound over pre[begin+1 .. end]
    while (l < h) {
        const int mid = (l + h) >> 1;
        if (pre[mid] <= target) l = mid + 1;
        else h = mid;
    }
    int pos = l < begin + 1 ? begin + 1 : (l > end ? end : l);
    return V.bank_mode[b0 + pos - 1];
}

__device__ __forceinline__ double relax_weight(const View& V, const CellRta& cr, int m) {
    // PhononMaterial::mode_relaxation_weight (src/material/Thermodynamics.cpp:95-106)
    // with mode_gamma's interpolation (:108-133); row and weight are per cell.
    // Row-pair table: (gamma[glo], gamma[glo + 1]) of a mode in one load.
    const double2 y = V.gamma2[static_cast<size_t>(cr.glo) * V.na + m];
    double g = y.x;
    if (cr.ghi != cr.glo) g = y.x * (1.0 - cr.gt) + y.y * cr.gt;
    if (g == 0.0 || V.dt == 0.0) return 0.0;
    return -expm1(-(g * V.dt) * (4.0 * kPi));
}

// Coefficients of P_k(n) = sum_j c[k][j] n^j with P_0 = n and
// P_{k+1} = n (n + 1) dP_k/dn, so that d^k n_BE / d beta^k = (-hw/kB)^k P_k(n_BE).
struct BoseSeries {
    double c[kSeriesOrder + 1][kSeriesOrder + 2];
};
constexpr BoseSeries make_bose_series() {
    BoseSeries s{};
    s.c[0][1] = 1.0;
    for (int k = 0; k < kSeriesOrder; ++k)
        for (int j = 1; j <= k + 1; ++j) {
            const double d = j * s.c[k][j]; // coefficient of n^(j-1) in dP_k/dn
            s.c[k + 1][j] += d;             // times n
            s.c[k + 1][j + 1] += d;         // times n^2
        }
    return s;
}
constexpr BoseSeries kBoseSeries = make_bose_series();

// ---------------------------------------------------------------- wave reductions
__device__ __forceinline__ double wave_sum(double x) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
    return x;
}

// Adds v[k] of every active lane to base[key*K + k], one atomic per distinct key
// present in the wave.  All lanes of the wave must call it.
template <int K, int STRIDE = K>
__device__ __forceinline__ void seg_atomic_add(bool active, int key, const double (&v)[K], double* base) {
    const int lane = __lane_id();
    uint64_t pending = __ballot(active);
    while (pending) {
        const int leader = __ffsll(static_cast<long long>(pending)) - 1;
        const int lkey = __shfl(key, leader, 64);
        const bool mine = ((pending >> lane) & 1ull) && key == lkey;
        const uint64_t mask = __ballot(mine);
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const double s = wave_sum(mine ? v[k] : 0.0);
            if (lane == leader) unsafeAtomicAdd(&base[static_cast<size_t>(lkey) * STRIDE + k], s);
        }
        pending &= ~mask;
    }
}

__device__ __forceinline__ void seg_count(bool active, int key, int* base) {
    const int lane = __lane_id();
    uint64_t pending = __ballot(active);
    while (pending) {
        const int leader = __ffsll(static_cast<long long>(pending)) - 1;
        const int lkey = __shfl(key, leader, 64);
        const bool mine = ((pending >> lane) & 1ull) && key == lkey;
        const uint64_t mask = __ballot(mine);
        if (lane == leader) atomicAdd(&base[lkey], __popcll(mask));
        pending &= ~mask;
    }
}

__device__ void raise_error(const View& V, int code, long long slot, const float* p, int facet) {
    if (atomicCAS(&V.st->error_code, 0, code) == 0) {
        V.st->error_slot = slot;
        if (p) {
            V.st->error_pos[0] = p[0];
            V.st->error_pos[1] = p[1];
            V.st->error_pos[2] = p[2];
        }
        V.st->error_facet = facet;
    }
}

// advance() result for a discrete-mode reflection that found no inward mode
// (conservative transport); reported as PMC_ERR_BOUNDARY with facet -7.
constexpr int kErrModeMapping = 64;
// advance() result for a carrier that meets an AMM interface in a direction its
// mode has no row for (a mirrored carrier); PMC_ERR_BOUNDARY with facet -13.
constexpr int kErrAmmRow = 65;
// advance() result for a mirror reflection on a facet that is not axis-aligned:
// the mirrored velocity is no table velocity with sign flips, so the carrier
// state cannot hold it; PMC_ERR_UNSUPPORTED with facet -14.
constexpr int kErrObliqueMirror = 66;

__device__ __forceinline__ void raise_advance_error(const View& V, int rc, long long slot, const float* p) {
    if (rc == -kErrModeMapping) raise_error(V, PMC_ERR_BOUNDARY, slot, p, -7);
    else if (rc == -kErrAmmRow) raise_error(V, PMC_ERR_BOUNDARY, slot, p, -13);
    else if (rc == -kErrObliqueMirror) raise_error(V, PMC_ERR_UNSUPPORTED, slot, p, -14);
    else raise_error(V, -rc, slot, p, -1);
}

// A block's place in its instance's launch: G.b of G.n blocks.  A
// single-instance launch is the hardware grid; a batched launch covers many
// instances, each with a range of blocks that starts at a multiple of 8.
struct Grid {
    int b, n;
};

__device__ __forceinline__ Grid launch_grid() { return {static_cast<int>(blockIdx.x), static_cast<int>(gridDim.x)}; }

// Batched launch: the instance owning this block, from the prefix offsets of
// the instances' block counts (an instance with no blocks owns none).
__device__ __forceinline__ int batch_instance(const int* __restrict__ offs, int n, Grid& G) {
    const int b = static_cast<int>(blockIdx.x);
    int lo = 0, hi = n - 1; // last instance whose range starts at or before b
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (offs[mid] <= b) lo = mid;
        else hi = mid - 1;
    }
    G = {b - offs[lo], offs[lo + 1] - offs[lo]};
    return lo;
}

// Logical block index grouping consecutive carrier chunks on one XCD (blocks are
// dispatched round-robin over the 8 XCDs of an MI355X).
__device__ __forceinline__ int xcd_block(const Grid G) {
    if ((G.n & 7) != 0) return G.b;
    return (G.b & 7) * (G.n >> 3) + (G.b >> 3);
}

__device__ __forceinline__ int carrier_slot(const Grid G) { return xcd_block(G) * blockDim.x + threadIdx.x; }

// ---------------------------------------------------------------- boundary events
// RoughBoundarySampler::sample (src/solver/RoughBoundarySampler.cpp:459-533)
// with sample_diffuse (:373-456) and flux_weighted_window_index
// (include/solver/BoundaryFluxSampling.h:64-88).  Returns the outgoing active mode
// and occupation; kind = 1 specular, 2 diffuse, 0 unchanged.
// signed_w: deviational occupations may go negative (the signed_weights
// argument of RoughBoundarySampler::sample, linearized transport).
__device__ void rough_sample(const View& V, const RoughDev& R, int in_ai, double in_occ, EventUniforms& U,
                             int& out_ai, double& out_occ, int& kind, bool signed_w) {
    out_ai = in_ai;
    out_occ = in_occ;
    kind = 0;
    double p_spec = R.spec[in_ai];
    p_spec = p_spec < 0.0 ? 0.0 : (p_spec > 1.0 ? 1.0 : p_spec);
    const double w_in = V.omega[in_ai];
    const double neq_in = V.neq[in_ai];
    if (U.next() <= p_spec) {
        const int o = R.match[in_ai];
        if (o >= 0 && o < V.na) {
            const double w_out = V.omega[o];
            const double mapped = V.neq[o] + w_in * (in_occ - neq_in) / fmax(1e-30, w_out);
            if (isfinite(mapped) && (signed_w || mapped >= 0.0)) {
                out_ai = o;
                out_occ = mapped;
            }
            kind = 1;
            return;
        }
    }
    int o = -1;
    const int begin = R.dbeg[in_ai], end = R.dend[in_ai];
    bool done = false;
    if (begin >= 0 && end > begin && end <= R.n_out) {
        const double u = U.next();
        const double fb = R.prefix[begin], fe = R.prefix[end];
        const double wf = fe - fb;
        if (wf > 0.0 && isfinite(wf)) {
            const double uc = u < 0.0 ? 0.0 : (u > 0x1.fffffffffffffp-1 ? 0x1.fffffffffffffp-1 : u);
            const double target = fb + uc * wf;
            // upper_bound over prefix[begin+1 .. end]
            int lo = begin + 1, hi = end + 1;
            while (lo < hi) {
                const int mid = (lo + hi) >> 1;
                if (R.prefix[mid] <= target) lo = mid + 1;
                else hi = mid;
            }
            int pos = lo - 1;
            pos = pos < begin ? begin : (pos > end - 1 ? end - 1 : pos);
            o = R.out_sorted[pos];
            done = true;
        }
    }
    if (!done) {
        if (R.strict) {
            return; // no cross-frequency fallback: caller mirrors the incoming mode
        }
        if (R.n_roul > 0 && R.roul_cdf[R.n_roul - 1] > 0.0) {
            const double r = U.next() * R.roul_cdf[R.n_roul - 1];
            int lo = 0, hi = R.n_roul;
            while (lo < hi) { // lower_bound
                const int mid = (lo + hi) >> 1;
                if (R.roul_cdf[mid] < r) lo = mid + 1;
                else hi = mid;
            }
            if (lo >= R.n_roul) lo = R.n_roul - 1;
            o = R.roul_active[lo];
        } else if (R.n_out > 0) {
            int k = static_cast<int>(U.next() * R.n_out);
            o = R.out_active[k < R.n_out ? k : R.n_out - 1];
        } else {
            int k = static_cast<int>(U.next() * R.mode_n);
            o = R.mode_lo + (k < R.mode_n ? k : R.mode_n - 1);
        }
    }
    kind = 2;
    const double w_out = V.omega[o];
    const double mapped = V.neq[o] + w_in * (in_occ - neq_in) / fmax(1e-30, w_out);
    if (isfinite(mapped) && (signed_w || mapped >= 0.0)) {
        out_ai = o;
        out_occ = mapped;
    }
}

struct Tally {
    double absorbed_e;
    int absorbed_res;  // reservoir index or -1
    bool absorbed;
    int rough_kind;    // last rough event kind seen (for counters)
    int rough_events, rough_spec, rough_diff;
    int iface_events, iface_transmitted;
};

// Free flight for `remaining` ps with up to 64 wall events
// (TransportOperator::advance_particle, src/solver/TransportOperator.cpp:405-466,
// against the analytic box of SimulationDomain::trace_boundary_intersection,
// src/SimulationDomain.cpp:342-450, or the triangle mesh (mesh_trace), and
// interact_boundary, src/solver/BoundaryInteraction.cpp:53-162).  FP32
// positions: a carrier on a wall always moves away from it (box planes are
// chosen by velocity sign, mesh faces must be left through), so no push
// distance is needed; positions are clamped into the bounding box.
// Returns 0 alive, 1 absorbed, or a negative error.
template <int kGeom>
__device__ int advance(const View& V, float p[3], int& mw, double& occ, unsigned uid, float remaining, Tally& T,
                       bool& mode_changed, bool& occ_changed) {
    const uint32_t step = static_cast<uint32_t>(V.st->step);
    EventUniforms U;
    for (int ev = 0; ev < 64; ++ev) {
        if (!(remaining > 1e-14f)) return 0;
        const float4 v = carrier_velocity(V, mw);
        const float vv[3] = {v.x, v.y, v.z};
        int facet, ha;
        if constexpr ((kGeom == kGeomMesh)) {
            // Triangle mesh: FP64 trace from the FP32 position.
            const double* vt = V.veld + 3 * static_cast<size_t>(mw & kModeMask);
            const double vd[3] = {(mw & (1 << kFlipShift)) ? -vt[0] : vt[0], (mw & (2 << kFlipShift)) ? -vt[1] : vt[1],
                                  (mw & (4 << kFlipShift)) ? -vt[2] : vt[2]};
            if (v.w == 0.0f) return 0; // stationary mode
            const double pd[3] = {p[0], p[1], p[2]};
            double th = INFINITY;
            const int f = mesh_clear(V, p, v.w, remaining) ? -1
                                                           : mesh_trace(V, pd, vd, static_cast<double>(remaining), th);
            if (f < 0) { // no face within the flight
#pragma unroll
                for (int a = 0; a < 3; ++a) {
                    float x = fmaf(vv[a], remaining, p[a]);
                    p[a] = fminf(fmaxf(x, V.fbmin[a]), V.fbmax[a]);
                }
                return 0;
            }
            remaining -= static_cast<float>(th);
#pragma unroll
            for (int a = 0; a < 3; ++a) p[a] = static_cast<float>(fma(vd[a], th, pd[a]));
            facet = V.face_facet[f];
            ha = V.facets[facet].axis;
        } else {
            float tbest = INFINITY;
            int plane = -1;
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                if (vv[a] != 0.0f) {
                    const int side = vv[a] > 0.0f ? 1 : 0;
                    const float target = side ? V.fbmax[a] : V.fbmin[a];
                    float t = (target - p[a]) / vv[a];
                    t = t > 0.0f ? t : 0.0f;
                    if (t < tbest) {
                        tbest = t;
                        plane = 2 * a + side;
                    }
                }
            }
            // Layered materials: the next interface plane along the axis
            // (LayeredMaterialLayout::next_crossing, src/solver/LayeredMaterialLayout.cpp:33-60);
            // the outer wall wins exact ties.
            if (kGeom == kGeomLayered) {
                const int ax = V.layer_axis, mat = mode_material(V, mw & kModeMask);
                const float va = vv[ax];
                int dest = -1;
                float at = 0.0f;
                if (va > 0.0f && mat + 1 < V.layers) {
                    at = V.layer_plane[mat];
                    dest = mat + 1;
                } else if (va < 0.0f && mat > 0) {
                    at = V.layer_plane[mat - 1];
                    dest = mat - 1;
                }
                const float ti = dest >= 0 ? (at - p[ax]) / va : INFINITY;
                if (ti > 1e-14f && ti < tbest && !(ti > remaining)) {
                    remaining -= ti;
#pragma unroll
                    for (int a = 0; a < 3; ++a) {
                        float x = fmaf(vv[a], ti, p[a]);
                        p[a] = fminf(fmaxf(x, V.fbmin[a]), V.fbmax[a]);
                    }
                    p[ax] = at;
                    if (ev == 0 || U.uid != uid) {
                        U.uid = uid;
                        U.step = step;
                        U.k0 = V.key0;
                        U.k1 = V.key1;
                    }
                    U.ev = static_cast<uint32_t>(ev);
                    U.used = 0;
                    const int m = mw & kModeMask;
                    const double w = V.omega[m];
                    const int inc = va > 0.0f ? 1 : 0;
                    // InterfaceInteraction::interact (src/solver/InterfaceInteraction.cpp:46-153):
                    // mixed mismatch draws the channel first; acoustic mismatch walks the
                    // mode's row of transmitted edges with one draw (AcousticInterfaceSampler::sample,
                    // src/solver/AcousticInterfaceSampler.cpp:457-471); diffuse mismatch
                    // draws a reflection band in this material, a transmission band in
                    // the next and the choice.
                    const bool mixed = V.amm_fraction > 0.0 && V.amm_fraction < 1.0;
                    const double um = mixed ? U.next() : 1.0;
                    const bool amm = V.amm_fraction >= 1.0 || (mixed && um < V.amm_fraction);
                    bool transmit = false;
                    int sel = -1;
                    if (amm) {
                        const double vax = V.veld[3 * static_cast<size_t>(m) + ax] * ((mw >> (kFlipShift + ax)) & 1 ? -1.0 : 1.0);
                        if (fabs(vax) < 1e-7) { // tangent: mirror along the axis
                            T.iface_events++;
                            mw ^= (1 << ax) << kFlipShift;
                            mode_changed = true;
                            p[ax] = va > 0.0f ? nextafterf(at, -INFINITY) : nextafterf(at, INFINITY);
                            continue;
                        }
                        const int face = dest > mat ? mat : dest, side = dest > mat ? 0 : 1;
                        const int r = V.amm_block[2 * face + side] + (m - V.mode_begin[mat]);
                        sel = V.amm_refl[r];
                        if (sel < 0) return -kErrAmmRow;
                        double x = U.next();
                        for (int e = V.amm_ebeg[r], e1 = V.amm_ebeg[r + 1]; e < e1; ++e) {
                            x -= V.amm_eprob[e];
                            if (x < 0.0) {
                                sel = V.amm_emode[e];
                                transmit = true;
                                break;
                            }
                        }
                    } else {
                        const double u1 = U.next(), u2 = U.next(), u3 = U.next();
                        double fr, ft;
                        const int mr = sample_band(V, mat, 1 - inc, w, u1, fr);
                        const int mt = sample_band(V, dest, inc, w, u2, ft);
                        const double tr = ft > 0.0 ? ft / fmax(1e-300, fr + ft) : 0.0;
                        transmit = mt >= 0 && u3 < tr;
                        sel = transmit ? mt : mr;
                    }
                    T.iface_events++;
                    double nocc = 0.0;
                    // Energy-preserving remap with the (common) energy weight, :111-122.
                    if (sel >= 0) nocc = V.neq[sel] + (V.w * w * (occ - V.neq[m])) / fmax(1e-300, V.w * V.omega[sel]);
                    // Linearized transport (:100-105, :124-133): no outgoing mode, or one that
                    // does not leave in the expected direction along the axis, fails; signed
                    // occupations are kept.
                    if (V.linearized) {
                        if (sel < 0) return -kErrModeMapping;
                        const double vs = V.veld[3 * static_cast<size_t>(sel) + ax];
                        const bool out_up = transmit ? inc == 1 : inc == 0;
                        if (!isfinite(nocc) || !isfinite(vs) || (out_up ? !(vs > 0.0) : !(vs < 0.0)))
                            return -kErrModeMapping;
                    }
                    if (sel < 0 || !isfinite(nocc) || (!V.linearized && nocc < 0.0)) {
                        mw ^= (1 << ax) << kFlipShift; // mirror along the axis, same mode
                        mode_changed = true;
                        p[ax] = va > 0.0f ? nextafterf(at, -INFINITY) : nextafterf(at, INFINITY);
                        continue;
                    }
                    if (transmit) T.iface_transmitted++;
                    mw = sel;
                    occ = nocc;
                    mode_changed = true;
                    occ_changed = true;
                    // Place the carrier strictly on its new material's side.
                    const bool up = transmit ? inc == 1 : inc == 0;
                    p[ax] = up ? nextafterf(at, INFINITY) : nextafterf(at, -INFINITY);
                    continue;
                }
            }
            if (plane < 0 || tbest > remaining) {
#pragma unroll
                for (int a = 0; a < 3; ++a) {
                    float x = fmaf(vv[a], remaining, p[a]);
                    p[a] = fminf(fmaxf(x, V.fbmin[a]), V.fbmax[a]);
                }
                return 0;
            }
            remaining -= tbest;
            ha = plane >> 1;
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                float x = fmaf(vv[a], tbest, p[a]);
                p[a] = fminf(fmaxf(x, V.fbmin[a]), V.fbmax[a]);
            }
            p[ha] = (plane & 1) ? V.fbmax[ha] : V.fbmin[ha];
            facet = V.plane_facet[plane];
        }
        const FacetDev& F = V.facets[facet];
        const int m = mw & kModeMask;
        if (F.cond == 'T') {
            T.absorbed = true;
            T.absorbed_res = F.res;
            T.absorbed_e = (kHbar * V.omega[m] * (occ - V.neq[m])) * V.w;
            return 1;
        }
        if (F.cond == 'P' && F.pair >= 0) {
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                float x = p[a] + static_cast<float>(F.shift[a]);
                p[a] = fminf(fmaxf(x, V.fbmin[a]), V.fbmax[a]);
            }
            continue;
        }
        // Conservative transport reflects through the table on every reflecting
        // facet; otherwise only rough facets have one.  Layered materials keep one
        // table per (material, facet).
        int rtab = F.rough;
        if (kGeom == kGeomLayered && V.layer_rough) rtab = V.layer_rough[mode_material(V, m) * V.nfacets + facet];
        if (rtab >= 0 && (F.cond == 'R' || V.conservative)) {
            if (ev == 0 || U.uid != uid) {
                U.uid = uid;
                U.step = step;
                U.k0 = V.key0;
                U.k1 = V.key1;
            }
            U.ev = static_cast<uint32_t>(ev);
            U.used = 0;
            int out_ai, kind;
            double out_occ;
            rough_sample(V, V.rough[rtab], m, occ, U, out_ai, out_occ, kind, V.conservative != 0);
            T.rough_events++;
            if (kind == 1) T.rough_spec++;
            else if (kind == 2) T.rough_diff++;
            // New velocity is the table velocity of the outgoing mode, mirrored if it
            // points out of the facet; under conservative transport it must point in
            // (src/solver/BoundaryInteraction.cpp:90-113).  The sampler returns modes
            // that point in, except when it keeps the incoming mode because a remapped
            // occupation would be negative; on a facet that is not axis-aligned that
            // mirror is no table direction (ha < 0).
            int nw = out_ai;
            const double* vo = V.veld + 3 * static_cast<size_t>(out_ai);
            const double vn = vo[0] * F.n[0] + vo[1] * F.n[1] + vo[2] * F.n[2];
            if (V.conservative) {
                const double wo = V.omega[out_ai];
                if (!isfinite(out_occ) || !(wo > 0.0) || !(vn < 0.0)) return -kErrModeMapping;
            } else if (vn > 0.0) {
                if (ha < 0) return -kErrObliqueMirror;
                nw |= (1 << ha) << kFlipShift;
            }
            if (nw != mw) mode_changed = true;
            if (out_occ != occ) occ_changed = true;
            mw = nw;
            occ = out_occ;
            continue;
        }
        if (F.cond == 'R') return -PMC_ERR_UNSUPPORTED; // table-less rough facet
        if (V.conservative) return -kErrModeMapping; // reflecting facet without a table
        // Specular mirror (periodic facet without a partner).
        if (ha < 0) return -kErrObliqueMirror;
        mw ^= (1 << ha) << kFlipShift;
        mode_changed = true;
    }
    return remaining > 1e-14f ? -PMC_ERR_EXCESS_EVENTS : 0;
}

// ---------------------------------------------------------------- carrier passes
// Every carrier pass is wave-persistent: a wave walks kRowsPerWave rows of 64
// consecutive (cell-sorted) slots, keeps lane-local running sums per cell and
// flushes them with one segmented reduction + atomic per distinct cell only
// when some lane's cell changes.  Rare scalar events (absorption, injection,
// rough reflections) go to LDS and leave as one partial row per block.

__device__ __forceinline__ int wave_global(const Grid G) {
    return xcd_block(G) * (blockDim.x >> 6) + (threadIdx.x >> 6);
}

template <int K, int STRIDE>
struct SegAcc {
    int key;
    bool has;
    double v[K];
    __device__ void reset() {
        has = false;
        key = 0;
#pragma unroll
        for (int k = 0; k < K; ++k) v[k] = 0.0;
    }
    __device__ void flush(double* base) {
        seg_atomic_add<K, STRIDE>(has, key, v, base);
        reset();
    }
    // All lanes of the wave must call add() together.
    __device__ void add(bool active, int k, const double (&x)[K], double* base) {
        if (__ballot(active && has && k != key)) flush(base);
        if (active) {
            key = k;
            has = true;
#pragma unroll
            for (int j = 0; j < K; ++j) v[j] += x[j];
        }
    }
};

struct BlockPartials {
    double* s; // shared, kPartialCols
    __device__ void zero() {
        for (int c = threadIdx.x; c < kPartialCols; c += blockDim.x) s[c] = 0.0;
        __syncthreads();
    }
    __device__ void add(int col, double x) { atomicAdd(&s[col], x); }
    __device__ void write(const View& V, int row) {
        __syncthreads();
        for (int c = threadIdx.x; c < V.pcols; c += blockDim.x) V.partials[row * V.pcols + c] = s[c];
    }
};

__device__ __forceinline__ void record_events(BlockPartials& P, const Tally& T) {
    if (T.absorbed) {
        P.add(kPAbsorbedE, T.absorbed_e);
        P.add(kPAbsorbedN, 1.0);
        if (T.absorbed_res >= 0) P.add(kPAbsorbedRes + T.absorbed_res, 1.0);
    }
    if (T.rough_events) {
        P.add(kPRoughEvents, T.rough_events);
        if (T.rough_spec) P.add(kPRoughSpecular, T.rough_spec);
        if (T.rough_diff) P.add(kPRoughDiffuse, T.rough_diff);
    }
    if (T.iface_events) {
        P.add(kPIfaceEvents, T.iface_events);
        if (T.iface_transmitted) P.add(kPIfaceTransmitted, T.iface_transmitted);
    }
}

// Claims destination slots in cell order: the lanes of one wave that share a
// key take consecutive slots with one atomic on that key's cursor.  All lanes
// of the wave must call it; returns -1 for inactive lanes.
__device__ __forceinline__ int seg_reserve(bool active, int key, int* cursor) {
    const int lane = __lane_id();
    uint64_t pending = __ballot(active);
    int dst = -1;
    while (pending) {
        const int leader = __ffsll(static_cast<long long>(pending)) - 1;
        const int lkey = __shfl(key, leader, 64);
        const bool mine = ((pending >> lane) & 1ull) && key == lkey;
        const uint64_t mask = __ballot(mine);
        int base = 0;
        if (lane == leader) base = atomicAdd(&cursor[lkey], __popcll(mask));
        base = __shfl(base, leader, 64);
        if (mine) dst = base + __popcll(mask & ((1ull << lane) - 1ull));
        pending &= ~mask;
    }
    return dst;
}

// ---------------------------------------------------------------- per-block cell tables
// Open-addressing table of the cells one block touches (LDS).  Carriers of a
// cell-sorted block fall into a handful of cells, so counts and slot
// reservations reach global memory once per (block, cell) instead of once per
// carrier or row.  A cell that finds no entry within kHashProbes probes falls
// back to a direct global atomic.
constexpr int kHashSlots = 1024;
constexpr int kHashProbes = 32;

struct BlockHash {
    int* key;
    int* cnt;
    int* base;
    __device__ void init() {
        for (int t = threadIdx.x; t < kHashSlots; t += blockDim.x) {
            key[t] = -1;
            cnt[t] = 0;
        }
        __syncthreads();
    }
    __device__ int slot(int cell) {
        unsigned s = (static_cast<unsigned>(cell) * 2654435761u) >> 22;
        for (int p = 0; p < kHashProbes; ++p) {
            const int k = key[s];
            if (k == cell) return static_cast<int>(s);
            if (k == -1) {
                const int old = atomicCAS(&key[s], -1, cell);
                if (old == -1 || old == cell) return static_cast<int>(s);
            }
            s = (s + 1) & (kHashSlots - 1);
        }
        return -1;
    }
    // Adds the wave's active lanes to their cells' counts (all lanes call).
    __device__ void count(bool active, int cell, int* global_count) {
        int s = -1;
        if (active) {
            s = slot(cell);
            if (s < 0) {
                atomicAdd(&global_count[cell], 1);
                active = false;
            }
        }
        const int lane = __lane_id();
        uint64_t pending = __ballot(active);
        while (pending) {
            const int leader = __ffsll(static_cast<long long>(pending)) - 1;
            const int ls = __shfl(s, leader, 64);
            const uint64_t mask = __ballot(((pending >> lane) & 1ull) && s == ls);
            if (lane == leader) atomicAdd(&cnt[ls], __popcll(mask));
            pending &= ~mask;
        }
    }
    // Adds the block's counts to the global per-cell counts.
    __device__ void flush_counts(int* global_count) {
        __syncthreads();
        for (int t = threadIdx.x; t < kHashSlots; t += blockDim.x)
            if (key[t] >= 0 && cnt[t] > 0) atomicAdd(&global_count[key[t]], cnt[t]);
    }
};

#define PMC_BLOCK_HASH(H)                                                                                              \
    __shared__ int H##_key[kHashSlots];                                                                                \
    __shared__ int H##_cnt[kHashSlots];                                                                                \
    __shared__ int H##_base[kHashSlots];                                                                               \
    BlockHash H{H##_key, H##_cnt, H##_base};                                                                           \
    H.init()

// Emission counts of the step: the previous step's absorptions per reservoir
// (one-to-one refill), or floor(expected) plus one with probability
// frac(expected) (ReservoirSampler::sample_flux_emission_count,
// src/solver/ReservoirSampler.cpp:117-125; Philox stream 4, counter (reservoir, step)).
__device__ __forceinline__ void begin_body(const View& V, const Grid G) {
    DevState* st = V.st;
    const int t = threadIdx.x;
    if (t < kMaxReservoirs) {
        int count = 0;
        if (t < V.nres) {
            if (V.refill_fixed) {
                const double expected = V.res[t].expected;
                count = static_cast<int>(floor(expected));
                const double fraction = expected - count;
                if (fraction > 0.0) {
                    const uint4 w = philox(static_cast<uint32_t>(t), static_cast<uint32_t>(st->step), 0, kStreamFlux,
                                           V.key0, V.key1);
                    if (u53(w.x, w.y) < fraction) ++count;
                }
            } else {
                count = st->absorbed_now[t];
            }
        }
        st->emit_count[t] = count;
    }
    if (t == 0) {
        st->n_slots_start = st->n_slots;
        st->carrier_steps += st->n_live;
        st->step_source_e = 0.0;
        st->step_source_updates = 0;
    }
    __syncthreads();
    if (t == 0) {
        int total = 0;
        for (int r = 0; r < V.nres; ++r) total += st->emit_count[r];
        st->emit_total = total;
    }
}

__global__ void k_begin(View V) { begin_body(V, launch_grid()); }

__global__ void kb_begin(const View* __restrict__ views, const int* __restrict__ offs, int n) {
    Grid G;
    const int i = batch_instance(offs, n, G);
    begin_body(views[i], G);
}

// Free flight of every carrier that reaches no wall within the step: the first
// iteration of advance() when it ends without an event, bit for bit.  Carriers
// that do reach a wall are listed per block for k_events and left untouched.
// Counts carriers per new cell for the cell sort.
template <int kGeom>
__device__ __forceinline__ void transport_body(const View& V, const Grid G) {
    __shared__ int ev_n;
    if (threadIdx.x == 0) ev_n = 0;
    PMC_BLOCK_HASH(H);
    const int n = V.st->n_slots_start;
    const int blk = xcd_block(G);
    const int base = (blk * kWavesPerBlock + (threadIdx.x >> 6)) * 64 * kRowsPerWave;
    const int lane = __lane_id();
    int* queue = V.events + static_cast<size_t>(blk) * kSlotsPerBlock;
    const float dt = V.dtf;
    const bool moves = dt > 1e-14f;
    for (int r0 = 0; r0 < kRowsPerWave; r0 += kUnroll) {
        if (base + r0 * 64 >= n) break;
        int mw[kUnroll];
        float x[kUnroll], y[kUnroll], z[kUnroll];
#pragma unroll
        for (int j = 0; j < kUnroll; ++j) {
            const int i = base + (r0 + j) * 64 + lane;
            mw[j] = i < n ? V.mode[i] : -1;
        }
#pragma unroll
        for (int j = 0; j < kUnroll; ++j) { // independent of the mode load: all slots below n hold data
            const int i = base + (r0 + j) * 64 + lane;
            x[j] = y[j] = z[j] = 0.0f;
            if (i < n) {
                x[j] = V.px[i];
                y[j] = V.py[i];
                z[j] = V.pz[i];
            }
        }
#pragma unroll
        for (int j = 0; j < kUnroll; ++j) {
            const int i = base + (r0 + j) * 64 + lane;
            const bool alive = mw[j] >= 0;
            bool event = false;
            int cell = 0;
            if (alive) {
                float p[3] = {x[j], y[j], z[j]};
                if (moves) {
                    const float4 v = carrier_velocity(V, mw[j]);
                    const float vv[3] = {v.x, v.y, v.z};
                    if constexpr ((kGeom == kGeomMesh)) {
                        // Carriers the distance field cannot clear go to k_events,
                        // which searches the faces with whole waves of such carriers.
                        event = v.w != 0.0f && !mesh_clear(V, p, v.w, dt);
                    } else {
                        float tbest = INFINITY;
                        int plane = -1;
#pragma unroll
                        for (int a = 0; a < 3; ++a) {
                            if (vv[a] != 0.0f) {
                                const int side = vv[a] > 0.0f ? 1 : 0;
                                const float target = side ? V.fbmax[a] : V.fbmin[a];
                                float t = (target - p[a]) / vv[a];
                                t = t > 0.0f ? t : 0.0f;
                                if (t < tbest) {
                                    tbest = t;
                                    plane = 2 * a + side;
                                }
                            }
                        }
                        event = plane >= 0 && !(tbest > dt);
                        if (kGeom == kGeomLayered && !event) { // a material interface within the step
                            const int ax = V.layer_axis, mat = mode_material(V, mw[j] & kModeMask);
                            const float va = vv[ax];
                            float ti = INFINITY;
                            if (va > 0.0f && mat + 1 < V.layers) ti = (V.layer_plane[mat] - p[ax]) / va;
                            else if (va < 0.0f && mat > 0) ti = (V.layer_plane[mat - 1] - p[ax]) / va;
                            event = ti > 1e-14f && !(ti > dt);
                        }
                    }
                    if (!event) {
#pragma unroll
                        for (int a = 0; a < 3; ++a) {
                            const float q = fmaf(vv[a], dt, p[a]);
                            p[a] = fminf(fmaxf(q, V.fbmin[a]), V.fbmax[a]);
                        }
                    }
                }
                if (!event) {
                    V.px[i] = p[0];
                    V.py[i] = p[1];
                    V.pz[i] = p[2];
                    cell = grid_index(V, p[0], p[1], p[2]);
                    V.cell[i] = cell;
                }
            }
            const uint64_t em = __ballot(event);
            if (em) {
                const int leader = __ffsll(static_cast<long long>(em)) - 1;
                int b = 0;
                if (lane == leader) b = atomicAdd(&ev_n, __popcll(em));
                b = __shfl(b, leader, 64);
                if (event) queue[b + __popcll(em & ((1ull << lane) - 1ull))] = i;
            }
            H.count(alive && !event, cell, V.cell_count);
        }
    }
    H.flush_counts(V.cell_count);
    if (threadIdx.x == 0) V.block_events[blk] = ev_n;
}

template <int kGeom>
__global__ void __launch_bounds__(256) k_transport(View V) { transport_body<kGeom>(V, launch_grid()); }

template <int kGeom>
__global__ void __launch_bounds__(256) kb_transport(const View* __restrict__ views, const int* __restrict__ offs, int n) {
    Grid G;
    const int i = batch_instance(offs, n, G);
    transport_body<kGeom>(views[i], G);
}

// Carriers with wall events this step: the whole event loop of
// TransportOperator::advance_particle (src/solver/TransportOperator.cpp:405-466)
// for the slots k_transport listed.  One block per transport block.
template <int kGeom>
__device__ __forceinline__ void events_body(const View& V, const Grid G) {
    // event_split blocks share one transport block's list (more waves for
    // latency-bound event work when there are few transport blocks).
    const int kSplit = V.event_split;
    __shared__ double sp[kPartialCols];
    BlockPartials P{sp};
    P.zero();
    PMC_BLOCK_HASH(H);
    const int row = xcd_block(G);
    const int blk = row / kSplit, part = row % kSplit;
    const int cnt = V.block_events[blk];
    const int* queue = V.events + static_cast<size_t>(blk) * kSlotsPerBlock;
    for (int k0 = part * blockDim.x; k0 < cnt; k0 += kSplit * blockDim.x) {
        const int k = k0 + threadIdx.x;
        Tally T{};
        T.absorbed_res = -1;
        bool alive = false;
        int cell = 0;
        if (k < cnt) {
            const int i = queue[k];
            int mw = V.mode[i];
            float p[3] = {V.px[i], V.py[i], V.pz[i]};
            double occ = V.occ[i];
            bool mc = false, oc = false;
            const int rc = advance<kGeom>(V, p, mw, occ, V.uid[i], V.dtf, T, mc, oc);
            if (rc < 0) {
                raise_advance_error(V, rc, i, p);
                V.mode[i] = -1; // keeps the cell counts equal to the live slots
            } else if (rc == 1) {
                V.mode[i] = -1;
            } else {
                alive = true;
                V.px[i] = p[0];
                V.py[i] = p[1];
                V.pz[i] = p[2];
                if (mc) V.mode[i] = mw;
                if (oc) V.occ[i] = occ;
                cell = grid_index(V, p[0], p[1], p[2]);
                V.cell[i] = cell;
                if (!isfinite(occ)) raise_error(V, PMC_ERR_INVALID_STATE, i, p, -1);
            }
            record_events(P, T);
        }
        H.count(alive, cell, V.cell_count);
    }
    H.flush_counts(V.cell_count);
    P.write(V, row);
}

template <int kGeom>
__global__ void __launch_bounds__(256) k_events(View V) { events_body<kGeom>(V, launch_grid()); }

template <int kGeom>
__global__ void __launch_bounds__(256) kb_events(const View* __restrict__ views, const int* __restrict__ offs, int n) {
    Grid G;
    const int i = batch_instance(offs, n, G);
    events_body<kGeom>(views[i], G);
}

// Reservoir injection (ReservoirInjector::inject, src/solver/ReservoirInjector.cpp:62-132)
// followed by the remaining flight of each injected carrier
// (TransportOperator::advance_injected, src/solver/TransportOperator.cpp:134-164).
template <int kGeom>
__device__ __forceinline__ void inject_body(const View& V, const Grid G) {
    const int row0 = event_rows_for(V); // partial rows after the event blocks

    __shared__ double sp[kPartialCols];
    BlockPartials P{sp};
    P.zero();
    PMC_BLOCK_HASH(H);
    DevState* st = V.st;
    const int total = st->emit_total;
    const int base_slot = st->n_slots_start;
    const uint32_t step = static_cast<uint32_t>(st->step);
    for (int k0 = G.b * blockDim.x; k0 < total; k0 += G.n * blockDim.x) {
        const int k = k0 + threadIdx.x;
        const bool active = k < total;
        Tally T{};
        T.absorbed_res = -1;
        bool alive = false;
        int cell = 0;
        const int slot = base_slot + k;
        if (active && slot >= V.capacity) {
            raise_error(V, PMC_ERR_CAPACITY, slot, nullptr, -1);
        } else if (active) {
            int r = 0, accn = 0;
            while (r < V.nres - 1 && k >= accn + st->emit_count[r]) accn += st->emit_count[r++];
            const ResDev& R = V.res[r];
            const uint4 w0 = philox(static_cast<uint32_t>(k), step, 0, kStreamInject, V.key0, V.key1);
            const uint4 w1 = philox(static_cast<uint32_t>(k), step, 1, kStreamInject, V.key0, V.key1);
            const uint4 w2 = philox(static_cast<uint32_t>(k), step, 2, kStreamInject, V.key0, V.key1);
            const double u_mode = u53(w0.x, w0.y), u_time = u53(w0.z, w0.w);
            const double u_face = u53(w1.x, w1.y), r1 = u53(w1.z, w1.w), r2 = u53(w2.x, w2.y);
            int m;
            if (V.sampling == 1) { // inverse CDF, std::lower_bound
                int lo = 0, hi = V.na;
                while (lo < hi) {
                    const int mid = (lo + hi) >> 1;
                    if (R.cdf[mid] < u_mode) lo = mid + 1;
                    else hi = mid;
                }
                m = lo < V.na ? lo : V.na - 1;
            } else { // Walker alias
                const double x = u_mode * V.na;
                int j = static_cast<int>(x);
                j = j < V.na ? j : V.na - 1;
                m = (x - j) < R.aprob[j] ? j : R.aidx[j];
            }
            const double t_emit = V.dt * u_time;
            int f = 0, fh = R.nfaces - 1; // std::discrete_distribution: lower_bound on the area CDF
            while (f < fh) {
                const int mid = (f + fh) >> 1;
                if (R.face_cdf[mid] < u_face) f = mid + 1;
                else fh = mid;
            }
            const double* tri = R.tri + 9 * f;
            const double sr1 = sqrt(r1);
            const double b0 = 1.0 - sr1, b1 = sr1 * (1.0 - r2), b2 = sr1 * r2;
            double pt[3];
#pragma unroll
            for (int a = 0; a < 3; ++a) pt[a] = b0 * tri[a] + b1 * tri[3 + a] + b2 * tri[6 + a];
            const double* vd = V.veld + 3 * static_cast<size_t>(m);
            double vx = vd[0], vy = vd[1], vz = vd[2];
            int mw = m;
            if (vx * R.n[0] + vy * R.n[1] + vz * R.n[2] > 0.0) {
                vx = -vx;
                vy = -vy;
                vz = -vz;
                mw |= 7 << kFlipShift;
            }
            const double speed = fmax(1e-12, sqrt(vx * vx + vy * vy + vz * vz));
            const double s = V.push / speed;
            float p[3] = {static_cast<float>(pt[0] + vx * s), static_cast<float>(pt[1] + vy * s),
                          static_cast<float>(pt[2] + vz * s)};
#pragma unroll
            for (int a = 0; a < 3; ++a) p[a] = fminf(fmaxf(p[a], V.fbmin[a]), V.fbmax[a]);
            double occ = R.occ[m];
            P.add(kPInjectedE, (kHbar * V.omega[m] * (occ - V.neq[m])) * V.w);
            P.add(kPInjectedN, 1.0);
            const unsigned uid = st->next_uid + static_cast<unsigned>(k);
            const double remain = V.dt - t_emit;
            int rc = 0;
            bool mc = false, oc = false;
            if (remain > 1e-14) rc = advance<kGeom>(V, p, mw, occ, uid, static_cast<float>(remain), T, mc, oc);
            V.uid[slot] = uid;
            if (rc < 0) {
                raise_advance_error(V, rc, slot, p);
                V.mode[slot] = -1;
            } else if (rc == 1) {
                V.mode[slot] = -1;
            } else {
                alive = true;
                cell = grid_index(V, p[0], p[1], p[2]);
                V.px[slot] = p[0];
                V.py[slot] = p[1];
                V.pz[slot] = p[2];
                V.mode[slot] = mw;
                V.occ[slot] = occ;
                V.cell[slot] = cell;
            }
            record_events(P, T);
        }
        H.count(alive, cell, V.cell_count);
    }
    H.flush_counts(V.cell_count);
    P.write(V, row0 + G.b);
}

template <int kGeom>
__global__ void __launch_bounds__(256) k_inject(View V) {
    // The View is this kernel's only argument, at offset 0 of the kernarg
    // segment.  Reading it there keeps the compiler from copying all of it to
    // per-lane scratch first, which it does here when handed the parameter.
#if defined(__HIP_DEVICE_COMPILE__)
    (void)V;
    const View& KV = *(const View*)__builtin_amdgcn_kernarg_segment_ptr();
#else
    const View& KV = V;
#endif
    inject_body<kGeom>(KV, launch_grid());
}

template <int kGeom>
__global__ void __launch_bounds__(256) kb_inject(const View* __restrict__ views, const int* __restrict__ offs, int n) {
    Grid G;
    const int i = batch_instance(offs, n, G);
    inject_body<kGeom>(views[i], G);
}

__device__ double lut_inverse(const View& V, double e, int base, int n) {
    // PhononMaterial::temperature_from_energy_density: clamped linear interpolation
    // over rows base .. base+n-1 (one material's table).
    const double* te = V.lut_e + base;
    const double* tt = V.lut_t + base;
    if (e <= te[0]) return tt[0];
    if (e >= te[n - 1]) return tt[n - 1];
    int lo = 0, hi = n; // upper_bound
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (te[mid] <= e) lo = mid + 1;
        else hi = mid;
    }
    const int h = lo, l = lo - 1;
    const double x0 = te[l], x1 = te[h], y0 = tt[l], y1 = tt[h];
    if (fabs(x1 - x0) <= 1e-18) return y0;
    const double t = (e - x0) / (x1 - x0);
    return y0 * (1.0 - t) + y1 * t;
}

// Background energy density of cell c (its material's, in a layered domain).
__device__ __forceinline__ double cell_background(const View& V, int c) {
    return V.layers > 1 ? V.e_bg_mat[V.cell_mat[c]] : V.e_bg;
}

// Temperature of cell c at energy density e: its material's lookup table, or
// under linearized transport T_bg + (e - E_bg) / C_bg (src/solver/GridStatistics.cpp:193-225).
__device__ __forceinline__ double cell_temperature(const View& V, double e, int c) {
    if (V.linearized) {
        if (V.layers > 1) {
            const int k = V.cell_mat[c];
            return V.t_bg + (e - V.e_bg_mat[k]) / V.cbg_mat[k];
        }
        return V.t_bg + (e - V.e_bg) / V.cw_cbg;
    }
    if (V.layers > 1) {
        const int m = V.cell_mat[c];
        return lut_inverse(V, e, V.lut_begin[m], V.lut_begin[m + 1] - V.lut_begin[m]);
    }
    return lut_inverse(V, e, 0, V.lut_n);
}

__device__ void gamma_row(const View& V, double temp, int& lo, int& hi, double& t) {
    const int n = V.ntemps;
    if (temp <= V.temps[0]) { lo = hi = 0; t = 0; return; }
    if (temp >= V.temps[n - 1]) { lo = hi = n - 1; t = 0; return; }
    int a = 0, b = n;
    while (a < b) {
        const int mid = (a + b) >> 1;
        if (V.temps[mid] <= temp) a = mid + 1;
        else b = mid;
    }
    hi = a;
    lo = a - 1;
    const double x0 = V.temps[lo], x1 = V.temps[hi];
    if (fabs(x1 - x0) <= 1e-18) { hi = lo; t = 0; return; }
    t = (temp - x0) / (x1 - x0);
}

__device__ double block_sum(double x, double* red) {
    red[threadIdx.x] = x;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    const double r = red[0];
    __syncthreads();
    return r;
}

// Ledger rows, carrier counts -> cell offsets, and cell energies -> temperatures
// (GridStatistics::update_energy_density / refresh_temperatures,
// src/solver/GridStatistics.cpp:95-235; bucket offsets of
// CellParticleIndex::ensure, src/solver/CellParticleIndex.cpp:14-44).
//   phase 0: initial counts -> offsets
//   phase 1: transport/injection ledger rows (BoundaryLedger::finish_transport,
//            src/solver/BoundaryLedger.cpp:86-115) + counts -> offsets
//   phase 2: cell energies after relaxation -> temperatures, energy ledger closure
//   phase 3: initial cell energies -> temperatures and thermal energy
__device__ __forceinline__ void refresh_body(const View& V, const Grid G, int phase) {
    DevState* st = V.st;
    __shared__ double red[1024];
    __shared__ int scan[1024];
    __shared__ int scan2[1024];
    if (phase == 1) {
        // Every column of the partial rows written this step in one pass: each
        // thread sums whole rows, then every column is reduced across the block.
        const int ncols = V.pcols; // rows are packed to the columns in use
        const int rows = event_rows_for(V) + kInjectBlocks; // rows written this step
        double acc[kPartialCols];
#pragma unroll
        for (int c = 0; c < kPartialCols; ++c) acc[c] = 0.0;
        for (int r = threadIdx.x; r < rows; r += blockDim.x) {
            const double* row = V.partials + static_cast<size_t>(r) * ncols;
#pragma unroll
            for (int c = 0; c < kPartialCols; ++c)
                if (c < ncols) acc[c] += row[c];
        }
        static_assert(16 * kPartialCols <= 1024, "per-wave column sums must fit the reduction buffer");
        const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, waves = blockDim.x >> 6;
#pragma unroll
        for (int c = 0; c < kPartialCols; ++c) {
            const double w = wave_sum(acc[c]);
            if (lane == 0) red[wave * kPartialCols + c] = w;
        }
        __syncthreads();
        if (static_cast<int>(threadIdx.x) < ncols) {
            const int c = threadIdx.x;
            double s = 0.0;
            for (int w = 0; w < waves; ++w) s += red[w * kPartialCols + c];
            switch (c) {
            case kPAbsorbedE: st->step_absorbed_e = s; break;
            case kPInjectedE: st->step_injected_e = s; break;
            case kPAbsorbedN: st->step_absorbed = static_cast<long long>(s); break;
            case kPInjectedN: st->step_injected = static_cast<long long>(s); break;
            case kPRoughEvents: st->rough_events += static_cast<long long>(s); break;
            case kPRoughSpecular: st->rough_specular += static_cast<long long>(s); break;
            case kPRoughDiffuse: st->rough_diffuse += static_cast<long long>(s); break;
            case kPIfaceEvents: st->iface_events += static_cast<long long>(s); break;
            case kPIfaceTransmitted: st->iface_transmitted += static_cast<long long>(s); break;
            default:
                if (c >= kPAbsorbedRes && c < kPAbsorbedRes + kMaxReservoirs)
                    st->absorbed_now[c - kPAbsorbedRes] = static_cast<int>(s);
            }
        }
        __syncthreads();
    }
    const int nc = V.ncells;
    const int per = (nc + static_cast<int>(blockDim.x) - 1) / static_cast<int>(blockDim.x);
    const int c0 = min(nc, static_cast<int>(threadIdx.x) * per), c1 = min(nc, c0 + per);
    if (phase <= 1) {
        int mine = 0, mine_chunks = 0;
        for (int c = c0; c < c1; ++c) {
            mine += V.cell_count[c];
            mine_chunks += (V.cell_count[c] + kChunk - 1) / kChunk;
        }
        scan[threadIdx.x] = mine;
        scan2[threadIdx.x] = mine_chunks;
        __syncthreads();
        for (int off = 1; off < static_cast<int>(blockDim.x); off <<= 1) {
            const bool take = static_cast<int>(threadIdx.x) >= off;
            const int add = take ? scan[threadIdx.x - off] : 0;
            const int add2 = take ? scan2[threadIdx.x - off] : 0;
            __syncthreads();
            scan[threadIdx.x] += add;
            scan2[threadIdx.x] += add2;
            __syncthreads();
        }
        int offset = scan[threadIdx.x] - mine;
        int coffset = scan2[threadIdx.x] - mine_chunks;
        const int total = scan[blockDim.x - 1];
        const int total_chunks = scan2[blockDim.x - 1];
        for (int c = c0; c < c1; ++c) {
            const int count = V.cell_count[c];
            V.cell_start[c] = offset;
            V.cell_cursor[c] = offset;
            V.chunk_start[c] = coffset;
            const int nch = (count + kChunk - 1) / kChunk;
            for (int k = 0; k < nch; ++k) V.chunk_cell[coffset + k] = c;
            offset += count;
            coffset += nch;
            V.cell_count[c] = 0;
            V.cell_tally[static_cast<size_t>(kTallyStride) * c + kTallyEnergy] = 0.0;
        }
        if (threadIdx.x == 0) {
            V.cell_start[nc] = total;
            V.chunk_start[nc] = total_chunks;
            if (phase == 1) {
                const int expect =
                    st->n_live - static_cast<int>(st->step_absorbed) + static_cast<int>(st->step_injected);
                if (expect != total) raise_error(V, PMC_ERR_INVALID_STATE, total, nullptr, -3);
                const int grown = st->n_slots_start + st->emit_total;
                st->n_slots = grown < V.capacity ? grown : V.capacity;
                st->next_uid += static_cast<unsigned>(st->emit_total);
            }
            st->n_live = total;
        }
        return;
    }
    double local = 0.0, residual = 0.0;
    // Chunk partials of the relaxation.  With at most four cells per wave, one
    // wave per cell with lanes striding the cell's chunks, so a few cells of many
    // chunks do not serialize on one thread (the sums wait in cell_e for the
    // per-cell pass); with more cells, each thread sums its own cells' chunks.
    const int waves = static_cast<int>(blockDim.x) >> 6;
    const bool wave_cells = phase == 2 && nc <= 4 * waves;
    if (wave_cells) {
        const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
        for (int c = wave; c < nc; c += waves) {
            double es = 0.0, rs = 0.0;
            for (int k = V.chunk_start[c] + lane; k < V.chunk_start[c + 1]; k += 64) {
                es += V.chunk_out[2 * static_cast<size_t>(k)];
                rs += V.chunk_out[2 * static_cast<size_t>(k) + 1];
                if (V.hs_on) es += V.chunk_src[k];
            }
            es = wave_sum(es);
            rs = wave_sum(rs);
            if (lane == 0) {
                V.cell_e[c] = es;
                residual += rs;
            }
        }
        __syncthreads();
    }
    for (int c = c0; c < c1; ++c) {
        double esum = V.cell_tally[static_cast<size_t>(kTallyStride) * c + kTallyEnergy];
        if (wave_cells) {
            esum = V.cell_e[c];
        } else if (phase == 2) { // in chunk order
            esum = 0.0;
            for (int k = V.chunk_start[c]; k < V.chunk_start[c + 1]; ++k) {
                esum += V.chunk_out[2 * static_cast<size_t>(k)];
                residual += V.chunk_out[2 * static_cast<size_t>(k) + 1];
                if (V.hs_on) esum += V.chunk_src[k];
            }
        }
        const double e = esum / V.cell_vol[c] + cell_background(V, c);
        V.cell_e[c] = e;
        V.cell_T[c] = cell_temperature(V, e, c);
        local += (e - V.zp) * V.cell_vol[c];
    }
    const double energy = block_sum(local, red);
    const double res = block_sum(residual, red);
    if (threadIdx.x == 0) {
        if (phase == 3) st->thermal_e = energy;
        if (phase == 2) {
            // EnergyLedger::finish_step (src/solver/EnergyLedger.cpp:58-90) and
            // BoundaryLedger::commit_step (src/solver/BoundaryLedger.cpp:117-141).
            st->step_residual_e = res;
            const double balance =
                energy - st->thermal_e - st->step_source_e - 0.0 - st->step_injected_e + st->step_absorbed_e - res;
            st->step_balance_e = balance;
            st->total_absorbed += st->step_absorbed;
            st->total_injected += st->step_injected;
            st->total_absorbed_e += st->step_absorbed_e;
            st->total_injected_e += st->step_injected_e;
            st->total_residual_e += res;
            st->total_source_e += st->step_source_e;
            st->total_balance_e += balance;
            st->thermal_e = energy;
            st->step += 1;
            st->elapsed += V.dt;
        }
        st->n_slots = st->n_live;
    }
}

__global__ void __launch_bounds__(1024) k_refresh(View V, int phase) { refresh_body(V, launch_grid(), phase); }

__global__ void __launch_bounds__(1024) kb_refresh(const View* __restrict__ views, const int* __restrict__ offs, int n, int phase) {
    Grid G;
    const int i = batch_instance(offs, n, G);
    refresh_body(views[i], G, phase);
}

// Counting-sort scatter of the live carriers into cell order (destination
// buffer), dropping absorbed slots: ParticleStorage::compact_alive
// (src/solver/ParticleStorage.cpp:75-105) and the cell buckets of
// CellParticleIndex in one pass.  Each block ranks its carriers per cell in
// LDS, reserves one destination range per (block, cell) from the global cell
// cursor, then streams the carriers out.
__device__ __forceinline__ void sort_body(const View& V, const Grid G) {
    __shared__ double esum[kHashSlots];
    for (int t = threadIdx.x; t < kHashSlots; t += blockDim.x) esum[t] = 0.0;
    PMC_BLOCK_HASH(H);
    const int n = V.st->n_slots;
    const int base = (xcd_block(G) * kWavesPerBlock + (threadIdx.x >> 6)) * 64 * kSortRows;
    const int lane = __lane_id();
    int mw[kSortRows], cl[kSortRows];
    double oc[kSortRows];
    float px[kSortRows], py[kSortRows], pz[kSortRows];
    unsigned id[kSortRows];
#pragma unroll
    for (int r = 0; r < kSortRows; ++r) {
        const int i = base + r * 64 + lane;
        mw[r] = -1;
        cl[r] = 0;
        oc[r] = 0.0;
        px[r] = py[r] = pz[r] = 0.0f;
        id[r] = 0;
        if (i < n) { // all slots below n hold data; dead ones are skipped below
            mw[r] = V.mode[i];
            cl[r] = V.cell[i];
            oc[r] = V.occ[i];
            px[r] = V.px[i];
            py[r] = V.py[i];
            pz[r] = V.pz[i];
            id[r] = V.uid[i];
        }
    }
    int tok[kSortRows]; // (hash slot << 16 | rank), -2 - destination, or -1 (dead)
    int run_slot = -1;  // lane-local running energy sum of one hash slot
    double run_e = 0.0;
#pragma unroll
    for (int r = 0; r < kSortRows; ++r) {
        bool active = mw[r] >= 0;
        const int cell = cl[r];
        double e = 0.0;
        if (active) {
            // Deviational energy for the refresh after transport (update_energy_density).
            const double2 mt = V.mtab[mw[r] & kModeMask];
            e = (kHbar * mt.x * (oc[r] - mt.y)) * V.w;
        }
        int s = -1;
        tok[r] = -1;
        if (active) {
            s = H.slot(cell);
            if (s < 0) {
                tok[r] = -2 - atomicAdd(&V.cell_cursor[cell], 1);
                atomicAdd(&V.cell_tally[static_cast<size_t>(kTallyStride) * cell + kTallyEnergy], e);
                active = false;
            } else if (s != run_slot) {
                if (run_slot >= 0) atomicAdd(&esum[run_slot], run_e);
                run_slot = s;
                run_e = 0.0;
            }
            if (active) run_e += e;
        }
        uint64_t pending = __ballot(active);
        while (pending) {
            const int leader = __ffsll(static_cast<long long>(pending)) - 1;
            const int ls = __shfl(s, leader, 64);
            const bool mine = ((pending >> lane) & 1ull) && s == ls;
            const uint64_t mask = __ballot(mine);
            int b = 0;
            if (lane == leader) b = atomicAdd(&H.cnt[ls], __popcll(mask));
            b = __shfl(b, leader, 64);
            if (mine) tok[r] = (ls << 16) | (b + __popcll(mask & ((1ull << lane) - 1ull)));
            pending &= ~mask;
        }
    }
    if (run_slot >= 0) atomicAdd(&esum[run_slot], run_e);
    __syncthreads();
    for (int t = threadIdx.x; t < kHashSlots; t += blockDim.x)
        if (H.key[t] >= 0 && H.cnt[t] > 0) {
            H.base[t] = atomicAdd(&V.cell_cursor[H.key[t]], H.cnt[t]);
            atomicAdd(&V.cell_tally[static_cast<size_t>(kTallyStride) * H.key[t] + kTallyEnergy], esum[t]);
        }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kSortRows; ++r) {
        if (tok[r] == -1) continue;
        const int dst = tok[r] >= 0 ? H.base[tok[r] >> 16] + (tok[r] & 0xffff) : -2 - tok[r];
        V.o_px[dst] = px[r];
        V.o_py[dst] = py[r];
        V.o_pz[dst] = pz[r];
        V.o_mode[dst] = mw[r];
        V.o_occ[dst] = oc[r];
        V.o_cell[dst] = cl[r];
        V.o_uid[dst] = id[r];
    }
}

__global__ void __launch_bounds__(256) k_sort(View V) { sort_body(V, launch_grid()); }

__global__ void __launch_bounds__(256) kb_sort(const View* __restrict__ views, const int* __restrict__ offs, int n) {
    Grid G;
    const int i = batch_instance(offs, n, G);
    sort_body(views[i], G);
}

// Sums v[k] over the block of W waves; every thread receives the same totals
// (wave sums combined in wave order).
template <int K, int W>
__device__ __forceinline__ void block_reduce(double (&v)[K], double* lds) {
    const int wave = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < K; ++k) v[k] = wave_sum(v[k]);
    if (__lane_id() == 0) {
#pragma unroll
        for (int k = 0; k < K; ++k) lds[k * W + wave] = v[k];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < K; ++k) {
        double s = 0.0;
#pragma unroll
        for (int w = 0; w < W; ++w) s += lds[k * W + w];
        v[k] = s;
    }
    __syncthreads();
}

// Adds one carrier's terms of the pseudo-temperature sums (the evaluate lambda
// of src/solver/RtaCollisionOperator.cpp:91-116 and its f(0), f(high) and
// energy-scale sums, :118-146), expanded around the cell temperature T0:
// n_BE(T) = n0 + sum_k (-x0)^k P_k(n0)/k! u^k with u = T0/T - 1, x0 = hw/(kB T0),
// P_k in q = n0(n0+1), r = 2 n0 + 1.  Carriers the reference skips (hw <= 0 or
// a <= 0) add nothing.  Returns the relaxation weight a.
__device__ __forceinline__ double rta_accumulate(const View& V, const CellRta& cr, int m, double n,
                                                 double (&acc)[kAccCols]) {
    const double a = relax_weight(V, cr, m);
    const double w = fmax(0.0, V.omega[m]);
    const double hw = kHbar * w;
    if (!(hw > 0.0) || !(a > 0.0)) return a;
    const double g = hw * a;
    const double t0 = fmax(0.0, cr.t0);
    const double n0 = bose(t0, w);
    acc[kAccD] += g * (n - n0);
    if (t0 > 0.0) {
        const double hk = hw * (1.0 / kBoltz); // hbar omega / kB
        const double x0 = hk * cr.inv_t0;
        if (x0 >= 1e-6) {
            static_assert(kSeriesOrder == 6, "closed forms below cover six terms");
            const double q = n0 * (n0 + 1.0), r = 2.0 * n0 + 1.0;
            const double P[6] = {q,
                                 q * r,
                                 q * (6.0 * q + 1.0),
                                 q * r * (12.0 * q + 1.0),
                                 q * ((120.0 * q + 30.0) * q + 1.0),
                                 q * r * ((360.0 * q + 60.0) * q + 1.0)};
            constexpr double kInvFact[6] = {1.0, 1.0 / 2.0, 1.0 / 6.0, 1.0 / 24.0, 1.0 / 120.0, 1.0 / 720.0};
            double c = g;
#pragma unroll
            for (int k = 1; k <= kSeriesOrder; ++k) {
                c *= -x0; // g (-x0)^k
                acc[kAccM1 + k - 1] += (c * kInvFact[k - 1]) * P[k - 1];
            }
        } else {
            acc[kAccBad] += 1.0;
        }
    }
    acc[kAccS0] += g * n;
    acc[kAccEscale] += g * fmax(1.0, n);
    const double inv_xh = cr.high / (hw * (1.0 / kBoltz)); // 1/x at T = high
    const double lb = inv_xh < 1e9 ? fmax(0.0, inv_xh - 0.5) : 0.0;
    acc[kAccUHigh] += g * (n - lb);
    return a;
}

// Per-cell RTA constants after the transport refresh: cell temperature from the
// deviational energy tallied by the sort (GridStatistics::refresh_temperatures)
// and the rate-table row of the lifetime temperature.
__device__ __forceinline__ CellRta cell_rta(const View& V, int c) {
    const double e = V.cell_tally[static_cast<size_t>(kTallyStride) * c + kTallyEnergy] / V.cell_vol[c] +
                     cell_background(V, c);
    const double T = V.inversion == 1 ? V.cell_tx[c] : cell_temperature(V, e, c);
    CellRta cr{};
    cr.t0 = T;
    cr.high = fmax(1000.0, 2.0 * fmax(1.0, T));
    cr.tr = T;
    cr.inv_t0 = T > 0.0 ? 1.0 / T : 0.0;
    gamma_row(V, V.lifetime_local ? T : V.lifetime_T, cr.glo, cr.ghi, cr.gt);
    return cr;
}

// Per-cell constants of the collision step, once per cell after the sort.
__device__ __forceinline__ void cell_prep_body(const View& V, const Grid G) {
    const int c = G.b * blockDim.x + threadIdx.x;
    if (c < V.ncells) V.crta[c] = cell_rta(V, c);
    if (c == 0 && V.callaway) V.st->cw_big = 0; // the Callaway kernels that follow list large cells anew
}

__global__ void k_cell_prep(View V) { cell_prep_body(V, launch_grid()); }

__global__ void kb_cell_prep(const View* __restrict__ views, const int* __restrict__ offs, int n) {
    Grid G;
    const int i = batch_instance(offs, n, G);
    cell_prep_body(views[i], G);
}

// Per-cell Newton inversion of the exact crystal energy (temperature_inversion 1):
// solves sum_a hw_a n_BE(T, w_a) / e_norm + zp = e over the active modes, the sum
// PhononMaterial::crystal_energy_density evaluates (src/material/Thermodynamics.cpp:151-163)
// and from which the reference tabulates its lookup table (:178-224).  One block
// per cell, started at the table's temperature, with dn/dT = n (n + 1) x / T.
// source 0: the post-transport energy of the tally (into cell_tx, read by the
// per-cell collision constants); source 1: the refreshed energy (cell_e, into cell_T).
__global__ void __launch_bounds__(256) k_invert(View V, int source) {
    __shared__ double red[2 * 4];
    const int c = blockIdx.x;
    if (c >= V.ncells) return;
    const double e = source == 0 ? V.cell_tally[static_cast<size_t>(kTallyStride) * c + kTallyEnergy] / V.cell_vol[c] +
                                       V.e_bg
                                 : V.cell_e[c];
    double T = lut_inverse(V, e, 0, V.lut_n);
    const double target = (e - V.zp) * V.e_norm; // sum hw n_BE(T) the cell holds
    if (target > 0.0 && T > 0.0) {
        for (int it = 0; it < 32; ++it) {
            double v[2] = {0.0, 0.0};
            const double inv_t = 1.0 / T;
            for (int a = threadIdx.x; a < V.na; a += blockDim.x) {
                const double w = V.omega[a];
                if (!(w > 0.0)) continue;
                const double x = kHbar * w / (kBoltz * T);
                if (x > 700.0) continue;
                const double n = 1.0 / fmax(exp(x) - 1.0, 1e-12);
                const double hw = kHbar * w;
                v[0] += hw * n;
                v[1] += hw * n * (n + 1.0) * x * inv_t;
            }
            block_reduce<2, 4>(v, red);
            const double f = v[0] - target;
            if (!(v[1] > 0.0)) break;
            double next = T - f / v[1];
            if (!(next > 0.0)) next = 0.5 * T;
            const double step = next - T;
            T = next;
            if (fabs(step) <= 1e-14 * T) break;
        }
    } else {
        T = 0.0;
    }
    if (threadIdx.x == 0) {
        if (source == 0) V.cell_tx[c] = T;
        else V.cell_T[c] = T;
    }
}

// Pseudo-temperature sums of one chunk at the cell temperature (terms of
// rta_accumulate), written as the chunk's partial row.
__device__ __forceinline__ void moments_body(const View& V, const Grid G) {
    __shared__ double lds[kAccCols * kChunkWaves];
    const int b = G.b;
    if (b >= V.chunk_start[V.ncells]) return;
    const int c = V.chunk_cell[b];
    const int begin = V.cell_start[c] + (b - V.chunk_start[c]) * kChunk;
    const int end = min(begin + kChunk, V.cell_start[c + 1]);
    const CellRta cr = V.crta[c];
    double acc[kAccCols];
#pragma unroll
    for (int k = 0; k < kAccCols; ++k) acc[k] = 0.0;
    constexpr int U = kChunk / kChunkThreads; // every carrier of the chunk in one round of loads
    for (int i0 = begin + static_cast<int>(threadIdx.x); i0 < end; i0 += kChunkThreads * U) {
        int m[U];
        double n[U];
#pragma unroll
        for (int j = 0; j < U; ++j) {
            const int i = i0 + j * kChunkThreads;
            m[j] = -1;
            n[j] = 0.0;
            if (i < end) {
                m[j] = V.mode[i] & kModeMask;
                n[j] = V.occ[i];
            }
        }
#pragma unroll
        for (int j = 0; j < U; ++j)
            if (m[j] >= 0) V.aw[i0 + j * kChunkThreads] = rta_accumulate(V, cr, m[j], n[j], acc);
    }
    block_reduce<kAccCols, kChunkWaves>(acc, lds);
    if (threadIdx.x == 0) {
        double* out = V.chunk_acc + static_cast<size_t>(b) * kAccCols;
#pragma unroll
        for (int k = 0; k < kAccCols; ++k) out[k] = acc[k];
    }
}

__global__ void __launch_bounds__(kChunkThreads) k_moments(View V) { moments_body(V, launch_grid()); }

__global__ void __launch_bounds__(kChunkThreads) kb_moments(const View* __restrict__ views, const int* __restrict__ offs, int n) {
    Grid G;
    const int i = batch_instance(offs, n, G);
    moments_body(views[i], G);
}

// Exact pseudo-temperature sums of one cell over its carriers (the evaluate
// lambda of src/solver/RtaCollisionOperator.cpp:91-116), one wave; the result
// is identical in every lane.
__device__ void exact_sums(const View& V, const CellRta& cr, int begin, int end, double t, double& value,
                           double& slope) {
    const double T = fmax(0.0, t);
    double f = 0.0, s = 0.0;
    for (int i = begin + __lane_id(); i < end; i += 64) {
        const int m = V.mode[i] & kModeMask;
        const double w = fmax(0.0, V.mtab[m].x);
        const double hw = kHbar * w;
        const double a = V.aw[i];
        if (!(hw > 0.0) || !(a > 0.0)) continue;
        const double neq = bose(T, w);
        f += hw * a * (V.occ[i] - neq);
        if (T > 0.0) s -= hw * a * (hw * neq * (neq + 1.0) / (kBoltz * T * T));
    }
    value = wave_sum(f);
    slope = wave_sum(s);
}

// Pseudo-temperature sums from the per-cell series at temperature t.  Returns
// false (caller sums exactly) outside |u| <= radius or when the geometric tail
// estimate of the truncated terms exceeds 1e-3 of the convergence tolerance.
__device__ __forceinline__ bool series_sums(const double (&acc)[kAccCols], double t0, double t, double tol,
                                            double radius, double& value, double& slope) {
    if (!(t > 0.0)) return false;
    const double u = t0 / t - 1.0;
    if (!(fabs(u) <= radius)) return false;
    double sum = 0.0, dsum = 0.0, up = 1.0, last = 0.0, prev = 0.0;
#pragma unroll
    for (int k = 1; k <= kSeriesOrder; ++k) {
        const double mk = acc[kAccM1 + k - 1];
        dsum += k * mk * up;
        const double term = mk * up * u;
        sum += term;
        prev = last;
        last = term;
        up *= u;
    }
    if (last != 0.0) {
        const double rho = fabs(last) / fabs(prev);
        if (!(rho <= 0.5)) return false;
        if (!(fabs(last) * rho / (1.0 - rho) <= 1e-3 * tol)) return false;
    }
    value = acc[kAccD] - sum;
    slope = (t0 / (t * t)) * dsum;
    return true;
}

// Per-cell bracket + safeguarded Newton solve of sum hw a (n - n_BE(T_R)) = 0
// (src/solver/RtaCollisionOperator.cpp:118-165), one wave per cell with every
// lane running the same iteration.  f(0) and the energy scale are exact chunk
// sums; f(high) is summed exactly only when its upper bound is not negative;
// f(T) and f'(T) come from the series where trusted, otherwise from the cell's
// carriers.
__device__ __forceinline__ void newton_body(const View& V, const Grid G) {
    const int c = G.b * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (c >= V.ncells) return;
    const int begin = V.cell_start[c], end = V.cell_start[c + 1];
    if (end <= begin) return;
    const CellRta cr = V.crta[c];
    const int k0 = V.chunk_start[c], k1 = V.chunk_start[c + 1];
    double acc[kAccCols];
#pragma unroll
    for (int k = 0; k < kAccCols; ++k) {
        double x = 0.0;
        for (int ch = k0 + __lane_id(); ch < k1; ch += 64) x += V.chunk_acc[static_cast<size_t>(ch) * kAccCols + k];
        acc[k] = wave_sum(x);
    }
    const double T = cr.t0;
    const double s0 = acc[kAccS0], escale = acc[kAccEscale];
    long long exact = 0, evals = 0;
    double low = 0.0, high = cr.high;
    bool known_high = false;
    double f_high = 0.0, d = 0.0;
    if (s0 > 0.0 && !(acc[kAccUHigh] < -1e-12 * escale)) {
        exact_sums(V, cr, begin, end, high, f_high, d);
        ++exact;
        for (int x = 0; f_high > 0.0 && x < 20; ++x) {
            high *= 2.0;
            exact_sums(V, cr, begin, end, high, f_high, d);
            ++exact;
        }
        known_high = true;
    }
    double t = T < low ? low : (high < T ? high : T);
    if (s0 <= 0.0) {
        t = low;
    } else if (known_high && f_high >= 0.0) {
        t = high;
    } else {
        const double tol = 1e-13 * fmax(1.0, escale);
        const bool series = V.series_on && acc[kAccBad] == 0.0 && cr.t0 > 0.0;
        for (int iter = 0; iter < 40; ++iter) {
            double value = 0.0, slope = 0.0;
            if (!(series && series_sums(acc, cr.t0, t, tol, V.series_radius, value, slope))) {
                exact_sums(V, cr, begin, end, t, value, slope);
                ++exact;
            }
            ++evals;
            if (fabs(value) <= tol) break;
            if (value > 0.0) low = t;
            else high = t;
            double cand = 0.5 * (low + high);
            if (slope < 0.0 && isfinite(slope)) {
                const double nt = t - value / slope;
                if (nt > low && nt < high && isfinite(nt)) cand = nt;
            }
            t = cand;
        }
    }
    if (__lane_id() == 0) {
        V.crta[c].tr = t;
        if (exact)
            atomicAdd(reinterpret_cast<unsigned long long*>(&V.st->newton_passes),
                      static_cast<unsigned long long>(exact));
        if (evals)
            atomicAdd(reinterpret_cast<unsigned long long*>(&V.st->newton_evals),
                      static_cast<unsigned long long>(evals));
    }
}

__global__ void __launch_bounds__(256) k_newton(View V) { newton_body(V, launch_grid()); }

__global__ void __launch_bounds__(256) kb_newton(const View* __restrict__ views, const int* __restrict__ offs, int n) {
    Grid G;
    const int i = batch_instance(offs, n, G);
    newton_body(views[i], G);
}

// Relaxation n <- max(0, n + a (n_BE(T_R) - n)) and residual
// (src/solver/RtaCollisionOperator.cpp:167-183) of one chunk, with the
// deviational energy after relaxation for the following refresh.
__device__ __forceinline__ void relax_body(const View& V, const Grid G) {
    __shared__ double lds[2 * kChunkWaves];
    const int b = G.b;
    if (b >= V.chunk_start[V.ncells]) return;
    const int c = V.chunk_cell[b];
    const int begin = V.cell_start[c] + (b - V.chunk_start[c]) * kChunk;
    const int end = min(begin + kChunk, V.cell_start[c + 1]);
    const CellRta cr = V.crta[c];
    double r[2] = {0.0, 0.0};
    for (int i0 = begin + static_cast<int>(threadIdx.x); i0 < end; i0 += kChunkThreads * kUnroll) {
        int m[kUnroll];
        double n[kUnroll], aj[kUnroll];
#pragma unroll
        for (int j = 0; j < kUnroll; ++j) {
            const int i = i0 + j * kChunkThreads;
            m[j] = -1;
            n[j] = 0.0;
            aj[j] = 0.0;
            if (i < end) {
                m[j] = V.mode[i] & kModeMask;
                n[j] = V.occ[i];
                aj[j] = V.aw[i];
            }
        }
#pragma unroll
        for (int j = 0; j < kUnroll; ++j) {
            if (m[j] < 0) continue;
            const int i = i0 + j * kChunkThreads;
            const double2 mt = V.mtab[m[j]];
            const double w = fmax(0.0, mt.x);
            const double hw = kHbar * w;
            const double a = aj[j];
            const double old = n[j];
            const double neq = bose(cr.tr, w);
            const double nn = old + a * (neq - old);
            const double occ = fmax(0.0, nn);
            if (occ != old) V.occ[i] = occ;
            r[0] += (kHbar * mt.x * (occ - mt.y)) * V.w;
            r[1] += (hw * nn - hw * old) * V.w;
            if (!isfinite(occ)) raise_error(V, PMC_ERR_INVALID_STATE, i, nullptr, -1);
        }
    }
    block_reduce<2, kChunkWaves>(r, lds);
    if (threadIdx.x == 0) {
        V.chunk_out[2 * static_cast<size_t>(b)] = r[0];
        V.chunk_out[2 * static_cast<size_t>(b) + 1] = r[1];
    }
}

__global__ void __launch_bounds__(kChunkThreads) k_relax(View V) { relax_body(V, launch_grid()); }

__global__ void __launch_bounds__(kChunkThreads) kb_relax(const View* __restrict__ views, const int* __restrict__ offs, int n) {
    Grid G;
    const int i = batch_instance(offs, n, G);
    relax_body(views[i], G);
}

// ---------------------------------------------------------------- Callaway
// Linearized Callaway collisions of one cell: CarrierCollisionOperator::apply
// (src/solver/CarrierCollisionOperator.cpp:38-142) with CallawayOperator
// (src/material/CallawayOperator.cpp:62-201); the cell's carriers are the
// quadrature.  In x_i = e_i / sqrt(C_i), e_i = hw_i (n_i - n_eq,i), the generator
// is L x = -(g_N + g_R) x + U_N U_N^T x + U_R U_R^T x, where U_N spans
// sqrt(g_N) B_k over k = 0..3 (energy and the three crystal-momentum components,
// B_i = sqrt(C_i) (1, q_i / w_i)) and U_R spans sqrt(g_R) B_0.  Writing the
// orthonormal vectors as combinations of the scaled basis, the gain terms become
// g_N,i B_i . (P_N m_N) + g_R,i B0_i m_R / G_R with the moments
// m_N = sum_i g_N,i B_i x_i, m_R = sum_i g_R,i B0_i x_i, the Gram matrix
// G_N = sum_i g_N,i B_i B_i^T and P_N = sum_k c_k c_k^T from modified
// Gram-Schmidt (two passes) of the normalized basis in G_N's inner product: the
// reference's orthogonalization of the n-vectors carried out in the 4-dim
// coefficient space, while every direction keeps a squared length of at least
// 1e-6; otherwise on explicit per-carrier vectors, skipping a direction below
// the reference's 1e-12 (callaway_setup_body).  exp(dt L) is the reference's
// Taylor scheme: ceil(dt max rate / 4) substeps, terms until the geometric tail
// bound is at most 2e-14/steps of |x|.  Every Taylor term is one pass over the
// registers with one block sum.
// A double from the lane that DPP control `kCtrl` names (both 32-bit halves).
template <int kCtrl>
__device__ __forceinline__ double dpp_double(double v) {
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), kCtrl, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), kCtrl, 0xF, 0xF, false);
    return __hiloint2double(hi, lo);
}
__device__ __forceinline__ double read_lane(double v, int lane) {
    return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), lane),
                            __builtin_amdgcn_readlane(__double2loint(v), lane));
}
constexpr int kDppXor1 = 0xB1, kDppXor2 = 0x4E;          // quad_perm [1,0,3,2], [2,3,0,1]
constexpr int kDppHalfMirror = 0x141, kDppMirror = 0x140; // within 8 and 16 lanes

// Sum over the wavefront as one wave-uniform value: in each row of 16 lanes,
// partner exchanges (xor 1, xor 2, half-row mirror, row mirror) leave every
// lane with the same row sum (each stage adds two equal partials in either
// order), then the four row sums are read out of lanes 0, 16, 32 and 48.
template <int K>
__device__ __forceinline__ void wave_sum(double (&v)[K]) {
#pragma unroll
    for (int k = 0; k < K; ++k) v[k] += dpp_double<kDppXor1>(v[k]);
#pragma unroll
    for (int k = 0; k < K; ++k) v[k] += dpp_double<kDppXor2>(v[k]);
#pragma unroll
    for (int k = 0; k < K; ++k) v[k] += dpp_double<kDppHalfMirror>(v[k]);
#pragma unroll
    for (int k = 0; k < K; ++k) v[k] += dpp_double<kDppMirror>(v[k]);
#pragma unroll
    for (int k = 0; k < K; ++k)
        v[k] = ((read_lane(v[k], 0) + read_lane(v[k], 16)) + read_lane(v[k], 32)) + read_lane(v[k], 48);
}
__device__ __forceinline__ double wave_max(double a) {
    a = fmax(a, dpp_double<kDppXor1>(a));
    a = fmax(a, dpp_double<kDppXor2>(a));
    a = fmax(a, dpp_double<kDppHalfMirror>(a));
    a = fmax(a, dpp_double<kDppMirror>(a));
    return fmax(fmax(read_lane(a, 0), read_lane(a, 16)), fmax(read_lane(a, 32), read_lane(a, 48)));
}

// Block sums of K values with one barrier: DPP wave sums, lane 0 of each wave
// stores its K sums in one of two LDS buffers (alternated by `parity`, so a
// wave never overwrites a buffer another wave may still read), and every
// thread adds the W rows in wave order, which leaves the same totals in every
// thread.  Consecutive calls on one lds must alternate parity; buffer p starts
// at lds + p * stride (stride >= K W for every K used), so lds holds 2 stride
// doubles.
template <int K, int W>
__device__ __forceinline__ void block_sum_sync1(double (&v)[K], double* lds, int parity, int stride) {
    const int wave = threadIdx.x >> 6;
    wave_sum<K>(v);
    double* buf = lds + parity * stride;
    if (__lane_id() == 0) {
#pragma unroll
        for (int k = 0; k < K; ++k) buf[wave * K + k] = v[k];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < K; ++k) {
        double s = buf[k];
        for (int w = 1; w < W; ++w) s += buf[w * K + k];
        v[k] = s;
    }
}

// Block sums of K values for many waves: wave sums to LDS, K threads add the
// W wave rows (in wave order), every thread reads the K totals.  lds holds
// K * W + K doubles.
template <int K, int W>
__device__ __forceinline__ void block_reduce_wide(double (&v)[K], double* lds) {
    const int wave = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < K; ++k) v[k] = wave_sum(v[k]);
    if (__lane_id() == 0) {
#pragma unroll
        for (int k = 0; k < K; ++k) lds[k * W + wave] = v[k];
    }
    __syncthreads();
    if (static_cast<int>(threadIdx.x) < K) {
        double s = 0.0;
        for (int w = 0; w < W; ++w) s += lds[threadIdx.x * W + w];
        lds[K * W + threadIdx.x] = s;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < K; ++k) v[k] = lds[K * W + k];
    __syncthreads();
}

template <int W>
__device__ __forceinline__ void block_max2(double& a, double& b, double* lds) {
    const int wave = threadIdx.x >> 6;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        a = fmax(a, __shfl_xor(a, off, 64));
        b = fmax(b, __shfl_xor(b, off, 64));
    }
    if (__lane_id() == 0) {
        lds[wave] = a;
        lds[W + wave] = b;
    }
    __syncthreads();
    a = 0.0;
    b = 0.0;
    for (int w = 0; w < W; ++w) {
        a = fmax(a, lds[w]);
        b = fmax(b, lds[W + w]);
    }
    __syncthreads();
}

// Inner product <a, b> = a^T G b of basis coefficient vectors.
__device__ __forceinline__ double cw_inner(const double (&G)[4][4], const double (&a)[4], const double (&b)[4]) {
    double s = 0.0;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        double gb = 0.0;
#pragma unroll
        for (int j = 0; j < 4; ++j) gb += G[i][j] * b[j];
        s += a[i] * gb;
    }
    return s;
}

// Linearized Callaway collisions of a large cell (more than kCwWaveMax
// carriers, listed by the one-wave kernels below) in two kernels with one
// workgroup of T threads per cell (up to kCwPer carriers per thread).
// k_callaway_setup sums the cell's Gram matrices and moments and builds the
// normal-gain projector into a per-cell record; k_callaway runs the Taylor
// steps with the carriers' x in registers, the record read as uniform data.
//
// Record columns (kCwRec doubles per cell).
constexpr int kCwP = 0;           // 16: projector in basis-moment space (row-major)
constexpr int kCwGrInv = 16;      // 1 / sum g_R B0^2 (resistive gain)
constexpr int kCwRmax = 17, kCwXmax = 18;
constexpr int kCwMoments = 19;    // 4 normal moments of x (basis or explicit directions)
constexpr int kCwMomR = 23, kCwXX = 24;
constexpr int kCwExplicit = 25;   // 1: gain weights from the explicit directions in cw_dir
constexpr int kCwEvolve = 26;     // 1: the cell evolves (nonzero x, rates and dt)

template <int T>
__device__ __forceinline__ void callaway_setup_body(const View& V, const int c) {
    constexpr int kW = T / 64;
    __shared__ double lds[18 * kW + 18];
    const int begin = V.cell_start[c], end = V.cell_start[c + 1];
    const int n = end - begin;
    const int k0 = V.chunk_start[c], k1 = V.chunk_start[c + 1];
    double* rec = V.cw_cell + static_cast<size_t>(kCwRec) * c;
    for (int k = k0 + static_cast<int>(threadIdx.x); k < k1; k += T) {
        V.chunk_out[2 * static_cast<size_t>(k)] = 0.0;
        V.chunk_out[2 * static_cast<size_t>(k) + 1] = 0.0;
    }
    if (n <= 0 || n > T * kCwPer) {
        if (threadIdx.x == 0) {
            rec[kCwEvolve] = 0.0;
            if (n > 0) raise_error(V, PMC_ERR_UNSUPPORTED, c, nullptr, -8);
        }
        return;
    }
    const int t = threadIdx.x;
    // x = e / sqrt(C); Gram matrices, largest rate and |x|, moments of x.
    double acc[17];
#pragma unroll
    for (int k = 0; k < 17; ++k) acc[k] = 0.0;
    double rmax = 0.0, xmax = 0.0;
    for (int j = t; j < n; j += T) {
        const int m = V.mode[begin + j] & kModeMask;
        const double4 p0 = V.cw_pack[2 * m], p1 = V.cw_pack[2 * m + 1];
        const double e = p1.z * (V.occ[begin + j] - V.mtab[m].y);
        if (!isfinite(e)) raise_error(V, PMC_ERR_INVALID_STATE, begin + j, nullptr, -1);
        const double xi = e / p0.z;
        const double b[4] = {p0.z, p0.w, p1.x, p1.y};
        int q = 0;
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int d = a; d < 4; ++d) acc[q++] += p0.x * b[a] * b[d];
        acc[10] += p0.y * b[0] * b[0];
#pragma unroll
        for (int a = 0; a < 4; ++a) acc[11 + a] += p0.x * b[a] * xi;
        acc[15] += p0.y * b[0] * xi;
        acc[16] += xi * xi;
        rmax = fmax(rmax, p0.x + p0.y);
        xmax = fmax(xmax, fabs(xi));
    }
    block_reduce_wide<17, kW>(acc, lds);
    block_max2<kW>(rmax, xmax, lds);
    // Normal-gain directions (CallawayOperator.cpp:23-57): modified Gram-Schmidt,
    // two passes, of the basis sqrt(g_N) B, each normalized first; a direction
    // whose length after projection is below 1e-12 is dropped.  Every thread
    // runs it in coefficient space with the Gram matrix G.  That is exact enough
    // when every direction keeps a length of at least 1e-3 (squared 1e-6); a
    // cell with a shorter one (dependent or nearly dependent carrier momenta, as
    // in cells of a few carriers) redoes it on explicit per-carrier vectors,
    // like the reference, kept in cw_dir for the Taylor steps.
    double Gm[4][4], Q[4][4];
    bool keep[4];
    bool explicit_dirs = false;
    {
        int q = 0;
        for (int a = 0; a < 4; ++a)
            for (int d = a; d < 4; ++d) {
                Gm[a][d] = acc[q];
                Gm[d][a] = acc[q];
                ++q;
            }
    }
    for (int k = 0; k < 4; ++k) {
        keep[k] = false;
        for (int i = 0; i < 4; ++i) Q[k][i] = 0.0;
        if (!(Gm[k][k] > 0.0)) continue; // inactive mechanism or absent momentum component
        double cf[4] = {0.0, 0.0, 0.0, 0.0};
        cf[k] = 1.0 / sqrt(Gm[k][k]);
        for (int pass = 0; pass < 2; ++pass)
            for (int j = 0; j < k; ++j) {
                if (!keep[j]) continue;
                const double proj = cw_inner(Gm, Q[j], cf);
                for (int i = 0; i < 4; ++i) cf[i] -= proj * Q[j][i];
            }
        const double len2 = cw_inner(Gm, cf, cf);
        if (!(len2 >= 1e-6)) {
            explicit_dirs = true;
            break;
        }
        const double inv = 1.0 / sqrt(len2);
        for (int i = 0; i < 4; ++i) Q[k][i] = cf[i] * inv;
        keep[k] = true;
    }
    double mom[4] = {acc[11], acc[12], acc[13], acc[14]};
    int nkeep = 0;
    if (explicit_dirs) {
        // Column nkeep of the cell's cw_dir rows is the candidate; kept
        // directions stay in columns 0 .. nkeep-1, unused columns end at zero.
        for (int k = 0; k < 4; ++k) {
            if (!(Gm[k][k] > 0.0)) continue;
            const double inv0 = 1.0 / sqrt(Gm[k][k]);
            for (int j = t; j < n; j += T) {
                const int m = V.mode[begin + j] & kModeMask;
                const double4 p0 = V.cw_pack[2 * m], p1 = V.cw_pack[2 * m + 1];
                const double b = k == 0 ? p0.z : k == 1 ? p0.w : k == 2 ? p1.x : p1.y;
                V.cw_dir[4 * static_cast<size_t>(begin + j) + nkeep] = sqrt(p0.x) * b * inv0;
            }
            for (int pass = 0; pass < 2; ++pass)
                for (int i = 0; i < nkeep; ++i) {
                    double proj[1] = {0.0};
                    for (int j = t; j < n; j += T) {
                        const double* row = V.cw_dir + 4 * static_cast<size_t>(begin + j);
                        proj[0] += row[i] * row[nkeep];
                    }
                    block_reduce_wide<1, kW>(proj, lds);
                    for (int j = t; j < n; j += T) {
                        double* row = V.cw_dir + 4 * static_cast<size_t>(begin + j);
                        row[nkeep] -= proj[0] * row[i];
                    }
                }
            double len2[1] = {0.0};
            for (int j = t; j < n; j += T) {
                const double v = V.cw_dir[4 * static_cast<size_t>(begin + j) + nkeep];
                len2[0] += v * v;
            }
            block_reduce_wide<1, kW>(len2, lds);
            if (!(len2[0] >= 1e-24)) continue; // rank-deficient cell geometry
            const double inv = 1.0 / sqrt(len2[0]);
            for (int j = t; j < n; j += T) V.cw_dir[4 * static_cast<size_t>(begin + j) + nkeep] *= inv;
            ++nkeep;
        }
        // Moments of x on the explicit directions.
        for (int i = 0; i < 4; ++i) mom[i] = 0.0;
        for (int j = t; j < n; j += T) {
            double* row = V.cw_dir + 4 * static_cast<size_t>(begin + j);
            for (int i = nkeep; i < 4; ++i) row[i] = 0.0;
            const int m = V.mode[begin + j] & kModeMask;
            const double4 p0 = V.cw_pack[2 * m], p1 = V.cw_pack[2 * m + 1];
            const double xi = p1.z * (V.occ[begin + j] - V.mtab[m].y) / p0.z;
            const double sg = sqrt(p0.x);
            for (int i = 0; i < 4; ++i) mom[i] += sg * row[i] * xi;
        }
        block_reduce_wide<4, kW>(mom, lds);
    }
    if (t == 0) {
        for (int a = 0; a < 4; ++a)
            for (int d = 0; d < 4; ++d) {
                double v = 0.0;
                if (explicit_dirs) {
                    v = (a == d && a < nkeep) ? 1.0 : 0.0;
                } else {
                    for (int k = 0; k < 4; ++k)
                        if (keep[k]) v += Q[k][a] * Q[k][d];
                }
                rec[kCwP + 4 * a + d] = v;
            }
        rec[kCwGrInv] = acc[10] > 0.0 ? 1.0 / acc[10] : 0.0;
        rec[kCwRmax] = rmax;
        rec[kCwXmax] = xmax;
        for (int a = 0; a < 4; ++a) rec[kCwMoments + a] = mom[a];
        rec[kCwMomR] = acc[15];
        rec[kCwXX] = acc[16];
        rec[kCwExplicit] = explicit_dirs ? 1.0 : 0.0;
        rec[kCwEvolve] = (xmax > 0.0 && V.dt > 0.0 && rmax > 0.0) ? 1.0 : 0.0;
    }
}

template <int T>
__device__ __forceinline__ void callaway_body(const View& V, const int c) {
    constexpr int kW = T / 64;
    __shared__ double lds[18 * kW + 18];
    const int begin = V.cell_start[c], end = V.cell_start[c + 1];
    const int n = end - begin;
    if (n <= 0 || n > T * kCwPer) return;
    const int k0 = V.chunk_start[c];
    const double* rec = V.cw_cell + static_cast<size_t>(kCwRec) * c;
    const int t = threadIdx.x;
    const bool evolve = rec[kCwEvolve] != 0.0;
    const bool explicit_dirs = rec[kCwExplicit] != 0.0;
    int par = 0; // LDS buffer of the next block sum
    int md[kCwPer];
    double x[kCwPer], s[kCwPer];
    const double xmax = rec[kCwXmax];
    const double inv = evolve ? 1.0 / xmax : 0.0;
#pragma unroll
    for (int r = 0; r < kCwPer; ++r) {
        const int j = t + r * T;
        md[r] = -1;
        x[r] = 0.0;
        if (j < n) {
            const int m = V.mode[begin + j] & kModeMask;
            md[r] = m;
            x[r] = V.cw_pack[2 * m + 1].z * (V.occ[begin + j] - V.mtab[m].y) / V.cw_pack[2 * m].z * inv;
        }
    }
    if (evolve) {
        const double dt = V.dt, rmax = rec[kCwRmax];
        const double steps_d = fmax(1.0, ceil(dt * rmax / 4.0));
        if (!(steps_d <= 100000.0) || steps_d * n > 200000000.0) {
            if (t == 0) raise_error(V, PMC_ERR_INVALID_STATE, c, nullptr, -9);
            return;
        }
        const int steps = static_cast<int>(steps_d);
        const double h = dt / steps, scaled_norm = h * rmax, tol = 2e-14 / steps;
        const double gr_inv = rec[kCwGrInv];
        double mN[4], mR = rec[kCwMomR] * inv, xx = rec[kCwXX] * inv * inv;
#pragma unroll
        for (int a = 0; a < 4; ++a) mN[a] = rec[kCwMoments + a] * inv;
        for (int st = 0; st < steps; ++st) {
            const double reference = sqrt(xx);
#pragma unroll
            for (int r = 0; r < kCwPer; ++r) s[r] = x[r];
            double sN[4] = {mN[0], mN[1], mN[2], mN[3]}, sR = mR; // moments of the running sum
            bool converged = false;
            double ss = xx;
            for (int k = 1; k <= 40; ++k) {
                double a4[4];
#pragma unroll
                for (int i = 0; i < 4; ++i)
                    a4[i] = rec[kCwP + 4 * i] * mN[0] + rec[kCwP + 4 * i + 1] * mN[1] + rec[kCwP + 4 * i + 2] * mN[2] +
                            rec[kCwP + 4 * i + 3] * mN[3];
                const double rr = mR * gr_inv, hk = h / k;
                double red[7] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
#pragma unroll
                for (int r = 0; r < kCwPer; ++r) {
                    if (md[r] < 0) continue;
                    const double4 p0 = V.cw_pack[2 * md[r]];
                    const double2 p1 = *reinterpret_cast<const double2*>(V.cw_pack + 2 * md[r] + 1);
                    // Gain weights of the carrier: g_N B (basis moments through P),
                    // or sqrt(g_N) times its explicit directions.
                    double w0, w1, w2, w3;
                    if (explicit_dirs) {
                        const double sg = sqrt(p0.x);
                        const double* row = V.cw_dir + 4 * static_cast<size_t>(begin + t + r * T);
                        w0 = sg * row[0];
                        w1 = sg * row[1];
                        w2 = sg * row[2];
                        w3 = sg * row[3];
                    } else {
                        w0 = p0.x * p0.z;
                        w1 = p0.x * p0.w;
                        w2 = p0.x * p1.x;
                        w3 = p0.x * p1.y;
                    }
                    const double act = -(p0.x + p0.y) * x[r] + (w0 * a4[0] + w1 * a4[1] + w2 * a4[2] + w3 * a4[3]) +
                                       p0.y * p0.z * rr;
                    const double tn = act * hk;
                    x[r] = tn;
                    s[r] += tn;
                    red[0] += w0 * tn;
                    red[1] += w1 * tn;
                    red[2] += w2 * tn;
                    red[3] += w3 * tn;
                    red[4] += p0.y * p0.z * tn;
                    red[5] += tn * tn;
                    red[6] += s[r] * s[r];
                }
                block_sum_sync1<7, kW>(red, lds, par, 7 * kW);
                par ^= 1;
#pragma unroll
                for (int a = 0; a < 4; ++a) {
                    mN[a] = red[a];
                    sN[a] += red[a];
                }
                mR = red[4];
                sR += red[4];
                ss = red[6];
                if (scaled_norm < k + 2) {
                    const double tail = sqrt(red[5]) * scaled_norm / (k + 1) / (1.0 - scaled_norm / (k + 2));
                    if (tail <= tol * reference) {
                        converged = true;
                        break;
                    }
                }
            }
            if (!converged) {
                if (t == 0) raise_error(V, PMC_ERR_INVALID_STATE, c, nullptr, -10);
                return;
            }
#pragma unroll
            for (int r = 0; r < kCwPer; ++r) x[r] = s[r];
#pragma unroll
            for (int a = 0; a < 4; ++a) mN[a] = sN[a];
            mR = sR;
            xx = ss;
        }
    }
    // Occupations from the evolved energies: n = n_eq + (x |x|max sqrt(C)) / hw.
    double out[2] = {0.0, 0.0};
#pragma unroll
    for (int r = 0; r < kCwPer; ++r) {
        if (md[r] < 0) continue;
        const int i = begin + t + r * T;
        const int m = md[r];
        const double b0 = V.cw_pack[2 * m].z, hw = V.cw_pack[2 * m + 1].z;
        const double2 mt = V.mtab[m];
        const double old = V.occ[i];
        const double evolved = evolve ? x[r] * xmax * b0 : hw * (old - mt.y);
        const double next = mt.y + evolved / hw;
        if (!isfinite(next)) raise_error(V, PMC_ERR_INVALID_STATE, i, nullptr, -1);
        V.occ[i] = next;
        out[0] += (kHbar * mt.x * (next - mt.y)) * V.w;
        out[1] += (hw * next - hw * old) * V.w;
    }
    block_sum_sync1<2, kW>(out, lds, par, 7 * kW);
    if (t == 0) {
        V.chunk_out[2 * static_cast<size_t>(k0)] = out[0];
        V.chunk_out[2 * static_cast<size_t>(k0) + 1] = out[1];
    }
    __syncthreads(); // the workgroup's next cell reuses the buffers from parity 0
}

// The workgroups of a context take the listed large cells in turn.
template <int T>
__device__ __forceinline__ void callaway_big_setup(const View& V, const Grid G) {
    const int count = min(V.st->cw_big, V.ncells);
    for (int i = G.b; i < count; i += G.n) callaway_setup_body<T>(V, V.cw_big[i]);
}
template <int T>
__device__ __forceinline__ void callaway_big(const View& V, const Grid G) {
    const int count = min(V.st->cw_big, V.ncells);
    for (int i = G.b; i < count; i += G.n) callaway_body<T>(V, V.cw_big[i]);
}

template <int T>
__global__ void __launch_bounds__(T) k_callaway_setup(View V) { callaway_big_setup<T>(V, launch_grid()); }
template <int T>
__global__ void __launch_bounds__(T) k_callaway(View V) { callaway_big<T>(V, launch_grid()); }

__global__ void __launch_bounds__(kCwThreads) kb_callaway_setup(const View* __restrict__ views,
                                                                const int* __restrict__ offs, int n) {
    Grid G;
    const int i = batch_instance(offs, n, G);
    callaway_big_setup<kCwThreads>(views[i], G);
}

__global__ void __launch_bounds__(kCwThreads) kb_callaway(const View* __restrict__ views, const int* __restrict__ offs, int n) {
    Grid G;
    const int i = batch_instance(offs, n, G);
    callaway_big<kCwThreads>(views[i], G);
}

// ---- one wavefront per cell
// The whole collision of a cell of at most 64 P carriers (the setup and the
// Taylor steps of the large-cell kernels, same arithmetic per carrier) on one
// wavefront: lane l holds carriers l, l + 64, ... with their gain weights.
// Cells of lo < n <= 64 P carriers are this class's (lo < 0 takes empty cells
// too); with `list`, a larger cell is appended to the large-cell list.
template <int P>
__device__ __forceinline__ void callaway_wave_body(const View& V, const int c, const int lo, const bool list) {
    const int lane = __lane_id();
    const int begin = V.cell_start[c], end = V.cell_start[c + 1];
    const int n = end - begin;
    if (n > 64 * P) {
        if (list && lane == 0) {
            const int slot = atomicAdd(&V.st->cw_big, 1);
            if (slot < V.ncells) V.cw_big[slot] = c;
        }
        return;
    }
    if (n <= lo) return;
    const int k0 = V.chunk_start[c], k1 = V.chunk_start[c + 1];
    for (int k = k0 + lane; k < k1; k += 64) {
        V.chunk_out[2 * static_cast<size_t>(k)] = 0.0;
        V.chunk_out[2 * static_cast<size_t>(k) + 1] = 0.0;
    }
    if (n <= 0) return;
    // Setup: x = e / sqrt(C), Gram matrices, largest rate and |x|, moments.
    int md[P];
    double x[P], dg[P], cr[P], w[P][4];
    double acc[17];
#pragma unroll
    for (int k = 0; k < 17; ++k) acc[k] = 0.0;
    double rmax = 0.0, xmax = 0.0;
#pragma unroll
    for (int r = 0; r < P; ++r) {
        const int j = lane + 64 * r;
        md[r] = -1;
        x[r] = dg[r] = cr[r] = 0.0;
        w[r][0] = w[r][1] = w[r][2] = w[r][3] = 0.0;
        if (j >= n) continue;
        const int m = V.mode[begin + j] & kModeMask;
        md[r] = m;
        const double4 p0 = V.cw_pack[2 * m], p1 = V.cw_pack[2 * m + 1];
        const double e = p1.z * (V.occ[begin + j] - V.mtab[m].y);
        if (!isfinite(e)) raise_error(V, PMC_ERR_INVALID_STATE, begin + j, nullptr, -1);
        const double xi = e / p0.z;
        x[r] = xi;
        dg[r] = p0.x + p0.y;
        cr[r] = p0.y * p0.z;
        const double b[4] = {p0.z, p0.w, p1.x, p1.y};
        int q = 0;
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int d = a; d < 4; ++d) acc[q++] += p0.x * b[a] * b[d];
        acc[10] += p0.y * b[0] * b[0];
#pragma unroll
        for (int a = 0; a < 4; ++a) {
            w[r][a] = p0.x * b[a];
            acc[11 + a] += w[r][a] * xi;
        }
        acc[15] += p0.y * b[0] * xi;
        acc[16] += xi * xi;
        rmax = fmax(rmax, p0.x + p0.y);
        xmax = fmax(xmax, fabs(xi));
    }
    wave_sum<17>(acc);
    rmax = wave_max(rmax);
    xmax = wave_max(xmax);
    // Normal-gain directions, as in callaway_setup_body: Gram space when every
    // direction keeps a length of at least 1e-3, else explicit vectors.
    double Gm[4][4], Q[4][4];
    bool keep[4];
    bool explicit_dirs = false;
    {
        int q = 0;
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int d = a; d < 4; ++d) {
                Gm[a][d] = acc[q];
                Gm[d][a] = acc[q];
                ++q;
            }
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        keep[k] = false;
#pragma unroll
        for (int i = 0; i < 4; ++i) Q[k][i] = 0.0;
    }
    for (int k = 0; k < 4; ++k) {
        if (!(Gm[k][k] > 0.0)) continue;
        double cf[4] = {0.0, 0.0, 0.0, 0.0};
        cf[k] = 1.0 / sqrt(Gm[k][k]);
        for (int pass = 0; pass < 2; ++pass)
            for (int j = 0; j < k; ++j) {
                if (!keep[j]) continue;
                const double proj = cw_inner(Gm, Q[j], cf);
                for (int i = 0; i < 4; ++i) cf[i] -= proj * Q[j][i];
            }
        const double len2 = cw_inner(Gm, cf, cf);
        if (!(len2 >= 1e-6)) {
            explicit_dirs = true;
            break;
        }
        const double inv = 1.0 / sqrt(len2);
        for (int i = 0; i < 4; ++i) Q[k][i] = cf[i] * inv;
        keep[k] = true;
    }
    double mom[4] = {acc[11], acc[12], acc[13], acc[14]};
    double Pm[4][4];
    if (explicit_dirs) {
        // Directions in w (column q of carrier r), built column by column.
        int nkeep = 0;
#pragma unroll
        for (int r = 0; r < P; ++r) w[r][0] = w[r][1] = w[r][2] = w[r][3] = 0.0;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            if (!(Gm[k][k] > 0.0)) continue;
            const double inv0 = 1.0 / sqrt(Gm[k][k]);
            double cand[P];
#pragma unroll
            for (int r = 0; r < P; ++r) {
                cand[r] = 0.0;
                if (md[r] < 0) continue;
                const double4 p0 = V.cw_pack[2 * md[r]], p1 = V.cw_pack[2 * md[r] + 1];
                const double b = k == 0 ? p0.z : k == 1 ? p0.w : k == 2 ? p1.x : p1.y;
                cand[r] = sqrt(p0.x) * b * inv0;
            }
            for (int pass = 0; pass < 2; ++pass)
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    if (i >= nkeep) break;
                    double proj[1] = {0.0};
#pragma unroll
                    for (int r = 0; r < P; ++r) proj[0] += w[r][i] * cand[r];
                    wave_sum<1>(proj);
#pragma unroll
                    for (int r = 0; r < P; ++r) cand[r] -= proj[0] * w[r][i];
                }
            double len2[1] = {0.0};
#pragma unroll
            for (int r = 0; r < P; ++r) len2[0] += cand[r] * cand[r];
            wave_sum<1>(len2);
            if (!(len2[0] >= 1e-24)) continue; // rank-deficient cell geometry
            const double inv = 1.0 / sqrt(len2[0]);
#pragma unroll
            for (int r = 0; r < P; ++r) {
                const double v = cand[r] * inv;
#pragma unroll
                for (int q = 0; q < 4; ++q)
                    if (q == nkeep) w[r][q] = v;
            }
            ++nkeep;
        }
        // Gain weights sqrt(g_N) times the directions; moments of x on them.
#pragma unroll
        for (int i = 0; i < 4; ++i) mom[i] = 0.0;
#pragma unroll
        for (int r = 0; r < P; ++r) {
            if (md[r] < 0) continue;
            const double sg = sqrt(V.cw_pack[2 * md[r]].x);
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                w[r][i] = sg * w[r][i];
                mom[i] += w[r][i] * x[r];
            }
        }
        wave_sum<4>(mom);
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int d = 0; d < 4; ++d) Pm[a][d] = (a == d && a < nkeep) ? 1.0 : 0.0;
    } else {
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int d = 0; d < 4; ++d) {
                double v = 0.0;
#pragma unroll
                for (int k = 0; k < 4; ++k)
                    if (keep[k]) v += Q[k][a] * Q[k][d];
                Pm[a][d] = v;
            }
    }
    const double gr_inv = acc[10] > 0.0 ? 1.0 / acc[10] : 0.0;
    const bool evolve = xmax > 0.0 && V.dt > 0.0 && rmax > 0.0;
    if (evolve) {
        const double inv = 1.0 / xmax;
#pragma unroll
        for (int r = 0; r < P; ++r) x[r] = x[r] * inv;
        const double dt = V.dt;
        const double steps_d = fmax(1.0, ceil(dt * rmax / 4.0));
        if (!(steps_d <= 100000.0) || steps_d * n > 200000000.0) {
            if (lane == 0) raise_error(V, PMC_ERR_INVALID_STATE, c, nullptr, -9);
            return;
        }
        const int steps = static_cast<int>(steps_d);
        const double h = dt / steps, scaled_norm = h * rmax, tol = 2e-14 / steps;
        double mN[4], mR = acc[15] * inv, xx = acc[16] * inv * inv;
#pragma unroll
        for (int a = 0; a < 4; ++a) mN[a] = mom[a] * inv;
        double sacc[P];
        for (int st = 0; st < steps; ++st) {
            const double reference = sqrt(xx);
#pragma unroll
            for (int r = 0; r < P; ++r) sacc[r] = x[r];
            double sN[4] = {mN[0], mN[1], mN[2], mN[3]}, sR = mR;
            bool converged = false;
            double ss = xx;
            for (int k = 1; k <= 40; ++k) {
                double a4[4];
#pragma unroll
                for (int i = 0; i < 4; ++i)
                    a4[i] = Pm[i][0] * mN[0] + Pm[i][1] * mN[1] + Pm[i][2] * mN[2] + Pm[i][3] * mN[3];
                const double rr = mR * gr_inv, hk = h / k;
                double red[7] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
#pragma unroll
                for (int r = 0; r < P; ++r) {
                    if (md[r] < 0) continue;
                    const double act = -dg[r] * x[r] +
                                       (w[r][0] * a4[0] + w[r][1] * a4[1] + w[r][2] * a4[2] + w[r][3] * a4[3]) +
                                       cr[r] * rr;
                    const double tn = act * hk;
                    x[r] = tn;
                    sacc[r] += tn;
                    red[0] += w[r][0] * tn;
                    red[1] += w[r][1] * tn;
                    red[2] += w[r][2] * tn;
                    red[3] += w[r][3] * tn;
                    red[4] += cr[r] * tn;
                    red[5] += tn * tn;
                    red[6] += sacc[r] * sacc[r];
                }
                wave_sum<7>(red);
#pragma unroll
                for (int a = 0; a < 4; ++a) {
                    mN[a] = red[a];
                    sN[a] += red[a];
                }
                mR = red[4];
                sR += red[4];
                ss = red[6];
                if (scaled_norm < k + 2) {
                    const double tail = sqrt(red[5]) * scaled_norm / (k + 1) / (1.0 - scaled_norm / (k + 2));
                    if (tail <= tol * reference) {
                        converged = true;
                        break;
                    }
                }
            }
            if (!converged) {
                if (lane == 0) raise_error(V, PMC_ERR_INVALID_STATE, c, nullptr, -10);
                return;
            }
#pragma unroll
            for (int r = 0; r < P; ++r) x[r] = sacc[r];
#pragma unroll
            for (int a = 0; a < 4; ++a) mN[a] = sN[a];
            mR = sR;
            xx = ss;
        }
    }
    // Occupations from the evolved energies: n = n_eq + (x |x|max sqrt(C)) / hw.
    double out[2] = {0.0, 0.0};
#pragma unroll
    for (int r = 0; r < P; ++r) {
        if (md[r] < 0) continue;
        const int i = begin + lane + 64 * r;
        const int m = md[r];
        const double b0 = V.cw_pack[2 * m].z, hw = V.cw_pack[2 * m + 1].z;
        const double2 mt = V.mtab[m];
        const double old = V.occ[i];
        const double evolved = evolve ? x[r] * xmax * b0 : hw * (old - mt.y);
        const double next = mt.y + evolved / hw;
        if (!isfinite(next)) raise_error(V, PMC_ERR_INVALID_STATE, i, nullptr, -1);
        V.occ[i] = next;
        out[0] += (kHbar * mt.x * (next - mt.y)) * V.w;
        out[1] += (hw * next - hw * old) * V.w;
    }
    wave_sum<2>(out);
    if (lane == 0) {
        V.chunk_out[2 * static_cast<size_t>(k0)] = out[0];
        V.chunk_out[2 * static_cast<size_t>(k0) + 1] = out[1];
    }
}

// Four cells per 256-thread workgroup, one per wavefront.  Class 0: cells of
// up to kCwTinyMax carriers (empty ones included); class 1: up to kCwWaveMax,
// listing the larger cells.
template <int kClass>
__device__ __forceinline__ void callaway_wave_cells(const View& V, const Grid G) {
    const int c = G.b * 4 + static_cast<int>(threadIdx.x >> 6);
    if (c >= V.ncells) return;
    if (kClass == 0) callaway_wave_body<kCwTinyPer>(V, c, -1, false);
    else callaway_wave_body<kCwWavePer>(V, c, kCwTinyMax, true);
}
template <int kClass>
__global__ void __launch_bounds__(256) k_callaway_wave(View V) { callaway_wave_cells<kClass>(V, launch_grid()); }
template <int kClass>
__global__ void __launch_bounds__(256) kb_callaway_wave(const View* __restrict__ views, const int* __restrict__ offs,
                                                         int n) {
    Grid G;
    const int i = batch_instance(offs, n, G);
    callaway_wave_cells<kClass>(views[i], G);
}

// ---------------------------------------------------------------- nonlinear Callaway
// NonlinearCallawayCollision::apply (src/solver/NonlinearCallawayCollision.cpp:287-344):
// one wavefront per cell walks the cell's sorted slots in every pass of
// pmc_nl::cell (the N(dt/2) R(dt) N(dt/2) moment-matched BGK substeps), with
// wave sums by cross-lane DPP exchanges and per-slot values in the scratch
// arrays.  Rates at the post-transport cell temperature (lifetime_local) or at
// the fixed lifetime temperature.  The cell's energy after the collision and its
// lifetime residual go to its first chunk, as the other collision kernels write
// them for the energy refresh.
struct WaveMapReduce {
    template <int K, class F>
    __device__ __forceinline__ void map_reduce(int begin, int end, double (&acc)[K], F&& f) {
#pragma unroll
        for (int k = 0; k < K; ++k) acc[k] = 0.0;
        for (int i = begin + static_cast<int>(__lane_id()); i < end; i += 64) f(i, acc);
        wave_sum<K>(acc);
    }
};

__device__ __forceinline__ void nlcallaway_cell(const View& V, const int c) {
    const int lane = __lane_id();
    const int begin = V.cell_start[c], end = V.cell_start[c + 1];
    const int k0 = V.chunk_start[c], k1 = V.chunk_start[c + 1];
    for (int k = k0 + lane; k < k1; k += 64) {
        V.chunk_out[2 * static_cast<size_t>(k)] = 0.0;
        V.chunk_out[2 * static_cast<size_t>(k) + 1] = 0.0;
    }
    if (end <= begin) return;
    const int64_t cap = V.capacity;
    pmc_nl::Data D;
    D.mode = V.mode;
    D.mode_mask = kModeMask;
    D.occ = V.occ;
    D.omega = &V.mtab[0].x;
    D.neq_bg = &V.mtab[0].y;
    D.omega_stride = 2;
    D.q = V.nl_q;
    D.n_active = V.na;
    D.n_temps = V.nl_nt;
    D.temps = V.nl_temps;
    D.gn = V.nl_gn;
    D.gu = V.nl_gu;
    D.ge = V.nl_ge;
    D.a = V.nl_scratch;
    D.metric = V.nl_scratch + cap;
    D.n0 = V.nl_scratch + 2 * cap;
    D.b = V.nl_scratch + 3 * cap;
    D.bstride = cap;
    D.energy_weight = V.w;
    const double T = V.crta[c].t0;
    const double rate_T = V.lifetime_local ? T : V.lifetime_T;
    double out[2];
    WaveMapReduce x;
    const int rc = pmc_nl::cell(x, D, begin, end, T, rate_T, V.dt, out);
    if (rc) {
        // facet -20 - rc names the reason; the slot field carries the cell
        if (lane == 0) raise_error(V, PMC_ERR_INVALID_STATE, c, nullptr, -20 - rc);
        return;
    }
    if (lane == 0) {
        V.chunk_out[2 * static_cast<size_t>(k0)] = out[0];
        V.chunk_out[2 * static_cast<size_t>(k0) + 1] = out[1];
    }
}

// Four cells per 256-thread workgroup, one per wavefront.
__global__ void __launch_bounds__(256) k_nlcallaway(View V) {
    const int c = static_cast<int>(blockIdx.x) * 4 + static_cast<int>(threadIdx.x >> 6);
    if (c >= V.ncells) return;
    nlcallaway_cell(V, c);
}

// ---------------------------------------------------------------- heat source
// Prescribed lattice heat source, thermal spectrum: the deposit of
// PhononSourceOperator::apply (src/solver/PhononSourceOperator.cpp:120-341) as
// four kernels over the cell-sorted chunks.  half = 0 runs before transport
// (base temperature: the previous refresh), half = 1 after relaxation (base:
// the refresh after transport).

// Integrated drive time of the half step (MonteCarloSolver::
// local_heat_source_integrated_time_factor, src/solver/HeatSourceDriving.cpp:14-46)
// from the device's elapsed time: half 0 covers [t, t + dt/2), half 1
// [t + dt/2, t + dt), with t summed step by step as the reference sums it.
__device__ double source_factor(const View& V, int half) {
    const double t0 = V.st->elapsed, mid = t0 + 0.5 * V.dt, t1 = t0 + V.dt;
    const double a = half == 0 ? t0 : mid, z = half == 0 ? mid : t1;
    const double b = fmax(a, V.hs_start);
    const double e = V.hs_end >= 0.0 ? fmin(z, V.hs_end) : z;
    if (e <= b) return 0.0;
    if (!V.hs_square) return V.hs_amplitude * (e - b);
    const double on = V.hs_on_duration < 0.0 ? V.hs_duty * V.hs_period : V.hs_on_duration;
    if (on <= 0.0) return 0.0;
    auto on_time = [&](double rel) {
        if (!(rel > 0.0)) return 0.0;
        const double cycles = floor(rel / V.hs_period);
        return cycles * on + fmin(rel - cycles * V.hs_period, on);
    };
    return V.hs_amplitude * fmax(0.0, on_time(e - V.hs_start) - on_time(b - V.hs_start));
}

__device__ __forceinline__ double source_base_T(const View& V, int c, int half) {
    const double t = half == 0 ? V.cell_T[c] : V.crta[c].t0;
    return (isfinite(t) && t >= 0.0) ? t : V.t_bg;
}

// Eligible carriers (the eligible lambda, :146-157): hw > 1e-14 eV and a mode
// the spectrum filter keeps.
__device__ __forceinline__ bool source_eligible(const View& V, int m, double& w, double& hw) {
    w = V.mtab[m].x;
    hw = kHbar * w;
    return isfinite(hw) && hw > 1e-14 && V.hs_elig[m] != 0;
}

__device__ __forceinline__ int chunk_range(const View& V, int b, int& begin, int& end) {
    const int c = V.chunk_cell[b];
    begin = V.cell_start[c] + (b - V.chunk_start[c]) * kChunk;
    end = min(begin + kChunk, V.cell_start[c + 1]);
    return c;
}

// Largest value over the block of W waves (every thread receives it); -inf
// when every value is.
template <int W>
__device__ __forceinline__ double block_max_of(double a, double* lds) {
    const int wave = threadIdx.x >> 6;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) a = fmax(a, __shfl_xor(a, off, 64));
    if (__lane_id() == 0) lds[wave] = a;
    __syncthreads();
    a = -INFINITY;
#pragma unroll
    for (int w = 0; w < W; ++w) a = fmax(a, lds[w]);
    __syncthreads();
    return a;
}

// Per chunk: Taylor terms of sum hw n_BE(T) about the base temperature T_b of
// the eligible carriers of cells with pending energy.
__device__ __forceinline__ void source_sums_body(const View& V, const Grid G, int half) {
    __shared__ double lds[kSrcCols * kChunkWaves];
    const int b = G.b;
    if (b >= V.chunk_start[V.ncells]) return;
    const double factor = source_factor(V, half);
    int begin, end;
    const int c = chunk_range(V, b, begin, end);
    const double pend = V.hs_pending[c] + V.hs_rate[c] * factor;
    double acc[kSrcCols];
#pragma unroll
    for (int k = 0; k < kSrcCols; ++k) acc[k] = 0.0;
    if (pend > 0.0 && V.hs_spectrum == 2) {
        // Gaussian spectrum (PhononSourceOperator::apply, src/solver/PhononSourceOperator.cpp:186-206):
        // the chunk's largest log weight M and sum exp(l - M), combined per cell
        // about the cell's largest (a lane keeps a running maximum and rescales).
        double mx = -INFINITY, sx = 0.0;
        for (int i = begin + static_cast<int>(threadIdx.x); i < end; i += kChunkThreads) {
            const int m = V.mode[i] & kModeMask;
            double w, hw;
            if (!source_eligible(V, m, w, hw)) continue;
            acc[kSrcCount] += 1.0;
            const double l = V.hs_w[m];
            if (l > mx) {
                sx = sx * exp(mx - l) + 1.0;
                mx = l;
            } else {
                sx += exp(l - mx);
            }
        }
        const double M = block_max_of<kChunkWaves>(mx, lds);
        acc[kSrcM1] = sx > 0.0 ? sx * exp(mx - M) : 0.0;
        block_reduce<kSrcCols, kChunkWaves>(acc, lds);
        acc[kSrcM1 + 1] = M;
    } else if (pend > 0.0 && (V.hs_spectrum == 1 || V.linearized)) {
        // Weighted spectrum: the branch weights of the eligible carriers (:207-217);
        // thermal spectrum under linearized transport: their heat capacities at
        // T_bg (:221-236).  Either way one scale per cell.
        for (int i = begin + static_cast<int>(threadIdx.x); i < end; i += kChunkThreads) {
            const int m = V.mode[i] & kModeMask;
            double w, hw;
            if (!source_eligible(V, m, w, hw)) continue;
            acc[kSrcCount] += 1.0;
            acc[kSrcM1] += V.hs_spectrum == 1 ? V.hs_w[m] : V.cw_cap[m];
        }
        block_reduce<kSrcCols, kChunkWaves>(acc, lds);
    } else if (pend > 0.0) {
        const double tb = source_base_T(V, c, half);
        const double inv = tb > 0.0 ? 1.0 / tb : 0.0;
        for (int i = begin + static_cast<int>(threadIdx.x); i < end; i += kChunkThreads) {
            const int m = V.mode[i] & kModeMask;
            double w, hw;
            if (!source_eligible(V, m, w, hw)) continue;
            acc[kSrcCount] += 1.0;
            const double x0 = hw * (1.0 / kBoltz) * inv;
            if (!(tb > 0.0) || !(x0 >= 1e-6)) {
                acc[kSrcBad] += 1.0;
                continue;
            }
            const double n0 = bose(tb, w);
            const double q = n0 * (n0 + 1.0), r = 2.0 * n0 + 1.0;
            const double P[6] = {q,
                                 q * r,
                                 q * (6.0 * q + 1.0),
                                 q * r * (12.0 * q + 1.0),
                                 q * ((120.0 * q + 30.0) * q + 1.0),
                                 q * r * ((360.0 * q + 60.0) * q + 1.0)};
            constexpr double kInvFact[6] = {1.0, 1.0 / 2.0, 1.0 / 6.0, 1.0 / 24.0, 1.0 / 120.0, 1.0 / 720.0};
            double cf = hw;
#pragma unroll
            for (int k = 1; k <= kSeriesOrder; ++k) {
                cf *= -x0;
                acc[kSrcM1 + k - 1] += (cf * kInvFact[k - 1]) * P[k - 1];
            }
        }
        block_reduce<kSrcCols, kChunkWaves>(acc, lds);
    }
    if (threadIdx.x == 0) {
        double* out = V.hs_acc + static_cast<size_t>(b) * kSrcCols;
#pragma unroll
        for (int k = 0; k < kSrcCols; ++k) out[k] = acc[k];
    }
}

// Exact deposit energy sum_i hw (n_BE(T) - n_BE(T_b)) and its slope over the
// cell's eligible carriers (thermal_increment_energy, :240-257); one wave.
__device__ void source_exact(const View& V, int begin, int end, double tb, double t, double& energy,
                             double& slope) {
    double e = 0.0, s = 0.0;
    for (int i = begin + __lane_id(); i < end; i += 64) {
        const int m = V.mode[i] & kModeMask;
        double w, hw;
        if (!source_eligible(V, m, w, hw)) continue;
        const double h = kHbar * fmax(0.0, w);
        const double n0 = bose(tb, w), n = bose(t, w);
        e += h * (n - n0);
        if (t > 0.0) s += h * (h * n * (n + 1.0) / (kBoltz * t * t));
    }
    energy = wave_sum(e);
    slope = wave_sum(s);
}

// Per cell (one wave): accrue the pending energy and solve for the source
// pseudo-temperature T_s with sum hw (n_BE(T_s) - n_BE(T_b)) = pending / w to
// 1e-12 * max(1e-6, target).  Newton from T_b on the series when it is trusted,
// else the reference's bracket + safeguarded Newton on exact sums (:259-293).
// Cells without eligible carriers keep their budget (:162-164).
__device__ __forceinline__ void source_solve_body(const View& V, const Grid G, int half) {
    const int c = G.b * (blockDim.x >> 6) + (threadIdx.x >> 6);
    const double factor = source_factor(V, half);
    if (c >= V.ncells) return;
    const bool lead = __lane_id() == 0;
    // Integrated source time: every cell accrues rate * factor (the prescribed
    // energy is the total rate times this sum).
    if (c == 0 && lead) V.st->source_time += factor;
    double* cellv = V.hs_cell + static_cast<size_t>(kSrcCellCols) * c;
    const double pend = V.hs_pending[c] + V.hs_rate[c] * factor;
    if (!(pend > 0.0)) {
        if (lead) {
            V.hs_pending[c] = pend;
            cellv[3] = 0.0;
        }
        return;
    }
    const int k0 = V.chunk_start[c], k1 = V.chunk_start[c + 1];
    double acc[kSrcCols];
#pragma unroll
    for (int k = 0; k < kSrcCols; ++k) {
        double x = 0.0;
        for (int ch = k0 + __lane_id(); ch < k1; ch += 64) x += V.hs_acc[static_cast<size_t>(ch) * kSrcCols + k];
        acc[k] = wave_sum(x);
    }
    if (!(acc[kSrcCount] > 0.0)) { // no eligible carrier: keep the budget locally
        if (lead) {
            V.hs_pending[c] = pend;
            cellv[3] = 0.0;
        }
        return;
    }
    const int begin = V.cell_start[c], end = V.cell_start[c + 1];
    const double tb = source_base_T(V, c, half);
    const double target = pend / V.w;
    if (!isfinite(target)) {
        raise_error(V, PMC_ERR_INVALID_STATE, c, nullptr, -4);
        return;
    }
    if (V.hs_spectrum == 2) { // increments target * exp(l_i - P) / S / hw_i about the cell's largest P
        double mloc = -INFINITY;
        for (int ch = k0 + __lane_id(); ch < k1; ch += 64) {
            const double* row = V.hs_acc + static_cast<size_t>(ch) * kSrcCols;
            if (row[kSrcCount] > 0.0) mloc = fmax(mloc, row[kSrcM1 + 1]);
        }
        const double P = wave_max(mloc);
        double sloc = 0.0;
        for (int ch = k0 + __lane_id(); ch < k1; ch += 64) {
            const double* row = V.hs_acc + static_cast<size_t>(ch) * kSrcCols;
            if (row[kSrcCount] > 0.0) sloc += row[kSrcM1] * exp(row[kSrcM1 + 1] - P);
        }
        const double S = wave_sum(sloc);
        if (lead) {
            cellv[0] = P;
            cellv[1] = target / S;
            cellv[2] = target;
            cellv[3] = 1.0;
            V.hs_pending[c] = 0.0;
            atomicAdd(&V.st->step_source_e, pend);
        }
        return;
    }
    if (V.hs_spectrum == 1 || V.linearized) { // increments target * w_i / sum w / hw_i: one scale per cell
        if (!(acc[kSrcM1] > 0.0)) {
            raise_error(V, PMC_ERR_INVALID_STATE, c, nullptr, -12);
            return;
        }
        if (lead) {
            cellv[0] = tb;
            cellv[1] = target / acc[kSrcM1];
            cellv[2] = target;
            cellv[3] = 1.0;
            V.hs_pending[c] = 0.0;
            atomicAdd(&V.st->step_source_e, pend);
        }
        return;
    }
    const double tol = 1e-12 * fmax(1e-6, target);
    long long exact = 0;
    bool done = false;
    double t = tb;
    if (V.series_on && acc[kSrcBad] == 0.0 && tb > 0.0) {
        // Newton on d = T_s - T_b; u = T_b / T - 1 = -d / (T_b + d) keeps full
        // precision for the millikelvin increments of a half-step deposit.
        double d = 0.0;
        for (int it = 0; it < 20; ++it) {
            const double tt = tb + d;
            const double u = -d / tt;
            if (!(fabs(u) <= V.series_radius)) break;
            double sum = 0.0, dsum = 0.0, up = 1.0, last = 0.0, prev = 0.0;
#pragma unroll
            for (int k = 1; k <= kSeriesOrder; ++k) {
                const double mk = acc[kSrcM1 + k - 1];
                dsum += k * mk * up;
                const double term = mk * up * u;
                sum += term;
                prev = last;
                last = term;
                up *= u;
            }
            if (last != 0.0) {
                const double rho = fabs(last) / fabs(prev);
                if (!(rho <= 0.5) || !(fabs(last) * rho / (1.0 - rho) <= 1e-3 * tol)) break;
            }
            // d/dT of sum_k M_k u^k is -(T_b / T^2) sum_k k M_k u^(k-1).
            const double value = sum - target;
            const double slope = -(tb / (tt * tt)) * dsum;
            if (fabs(value) <= tol) {
                t = tt;
                done = true;
                break;
            }
            if (!(slope > 0.0) || !isfinite(slope)) break;
            d -= value / slope;
            if (!(d >= 0.0) || !isfinite(d)) break;
        }
    }
    if (!done) {
        double low = fmax(0.0, tb);
        double high = fmax(low + 1.0, 1.05 * low + 1.0);
        double he, d;
        source_exact(V, begin, end, tb, high, he, d);
        ++exact;
        for (int x = 0; he < target && x < 60; ++x) {
            high = 2.0 * high + 1.0;
            source_exact(V, begin, end, tb, high, he, d);
            ++exact;
        }
        if (!(he >= target) || !isfinite(he)) {
            raise_error(V, PMC_ERR_INVALID_STATE, c, nullptr, -5);
            return;
        }
        t = 0.5 * (low + high);
        for (int iter = 0; iter < 60; ++iter) {
            double e, slope;
            source_exact(V, begin, end, tb, t, e, slope);
            ++exact;
            const double residual = e - target;
            if (fabs(residual) <= tol) break;
            if (residual > 0.0) high = t;
            else low = t;
            double cand = 0.5 * (low + high);
            if (slope > 0.0 && isfinite(slope)) {
                const double nt = t - residual / slope;
                if (nt > low && nt < high && isfinite(nt)) cand = nt;
            }
            t = cand;
        }
    }
    if (lead) {
        cellv[0] = tb;
        cellv[1] = t;
        cellv[2] = target;
        cellv[3] = 1.0;
        V.hs_pending[c] = 0.0;
        atomicAdd(&V.st->step_source_e, pend);
        if (exact)
            atomicAdd(reinterpret_cast<unsigned long long*>(&V.st->source_passes),
                      static_cast<unsigned long long>(exact));
    }
}

// Per chunk: increments n_BE(T_s) - n_BE(T_b) of the eligible carriers, or
// scale * C_i / hw_i under linearized transport (kept in
// the scratch array), their applied energy and the largest-energy carrier
// (:296-311; ties go to the smallest carrier id).
__device__ __forceinline__ void source_increments_body(const View& V, const Grid G) {
    __shared__ double lds[3 * kChunkWaves];
    const int b = G.b;
    if (b >= V.chunk_start[V.ncells]) return;
    int begin, end;
    const int c = chunk_range(V, b, begin, end);
    const double* cellv = V.hs_cell + static_cast<size_t>(kSrcCellCols) * c;
    double* out = V.hs_acc + static_cast<size_t>(b) * kSrcCols;
    if (cellv[3] == 0.0) return;
    const double tb = cellv[0], ts = cellv[1];
    double applied = 0.0, best = -1.0;
    unsigned best_uid = 0xffffffffu;
    int best_slot = -1;
    for (int i = begin + static_cast<int>(threadIdx.x); i < end; i += kChunkThreads) {
        const int m = V.mode[i] & kModeMask;
        double w, hw;
        double inc = 0.0;
        if (source_eligible(V, m, w, hw)) {
            if (V.hs_spectrum == 2) inc = ts * exp(V.hs_w[m] - tb) / hw; // tb holds the cell's largest log weight
            else if (V.hs_spectrum == 1) inc = ts * V.hs_w[m] / hw;
            else inc = V.linearized ? ts * V.cw_cap[m] / hw : bose(ts, w) - bose(tb, w);
            const double e = hw * inc;
            applied += e;
            const unsigned id = V.uid[i];
            if (e > best || (e == best && id < best_uid)) {
                best = e;
                best_uid = id;
                best_slot = i;
            }
        }
        V.aw[i] = inc;
    }
    // Wave reduction of the applied sum and of the (energy, id) maximum.
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        applied += __shfl_xor(applied, off, 64);
        const double ob = __shfl_xor(best, off, 64);
        const unsigned ou = __shfl_xor(best_uid, off, 64);
        const int os = __shfl_xor(best_slot, off, 64);
        if (ob > best || (ob == best && ou < best_uid)) {
            best = ob;
            best_uid = ou;
            best_slot = os;
        }
    }
    const int wave = threadIdx.x >> 6;
    if (__lane_id() == 0) {
        lds[wave] = applied;
        lds[kChunkWaves + wave] = best;
        lds[2 * kChunkWaves + wave] = __longlong_as_double((static_cast<long long>(best_uid) << 32) |
                                                           static_cast<unsigned>(best_slot));
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double a = 0.0, bb = -1.0;
        unsigned bu = 0xffffffffu;
        int bs = -1;
        for (int k = 0; k < kChunkWaves; ++k) {
            a += lds[k];
            const long long packed = __double_as_longlong(lds[2 * kChunkWaves + k]);
            const unsigned u = static_cast<unsigned>(packed >> 32);
            const int s = static_cast<int>(packed & 0xffffffffll);
            if (lds[kChunkWaves + k] > bb || (lds[kChunkWaves + k] == bb && u < bu)) {
                bb = lds[kChunkWaves + k];
                bu = u;
                bs = s;
            }
        }
        out[kSrcApplied] = a;
        out[kSrcMaxE] = bb;
        out[kSrcMaxUid] = static_cast<double>(bu);
        out[kSrcMaxSlot] = static_cast<double>(bs);
    }
}

// Per chunk: roundoff correction on the cell's largest-energy carrier so the
// deposit equals the target exactly (:312-318), validation (:319-326) and the
// occupation update; records the energy change for the following refresh.
__device__ __forceinline__ void source_apply_body(const View& V, const Grid G, int half) {
    __shared__ double lds[2 * kChunkWaves];
    const int b = G.b;
    if (b >= V.chunk_start[V.ncells]) return;
    int begin, end;
    const int c = chunk_range(V, b, begin, end);
    const double* cellv = V.hs_cell + static_cast<size_t>(kSrcCellCols) * c;
    if (cellv[3] == 0.0) {
        if (half == 1 && threadIdx.x == 0) V.chunk_src[b] = 0.0;
        return;
    }
    // Cell totals from the chunk rows, in chunk order.
    double applied = 0.0, best = -1.0;
    unsigned best_uid = 0xffffffffu;
    int best_slot = -1;
    for (int ch = V.chunk_start[c]; ch < V.chunk_start[c + 1]; ++ch) {
        const double* row = V.hs_acc + static_cast<size_t>(ch) * kSrcCols;
        applied += row[kSrcApplied];
        const unsigned u = static_cast<unsigned>(row[kSrcMaxUid]);
        if (row[kSrcMaxE] > best || (row[kSrcMaxE] == best && u < best_uid)) {
            best = row[kSrcMaxE];
            best_uid = u;
            best_slot = static_cast<int>(row[kSrcMaxSlot]);
        }
    }
    const double target = cellv[2];
    double correction = 0.0;
    if (best_slot >= begin && best_slot < end) {
        const double hw = kHbar * V.mtab[V.mode[best_slot] & kModeMask].x;
        correction = (target - applied) / hw;
    }
    double r[2] = {0.0, 0.0}; // energy change (tally units), updated carriers
    for (int i = begin + static_cast<int>(threadIdx.x); i < end; i += kChunkThreads) {
        const int m = V.mode[i] & kModeMask;
        double w, hw;
        if (!source_eligible(V, m, w, hw)) continue;
        const double inc = V.aw[i] + (i == best_slot ? correction : 0.0);
        const double next = V.occ[i] + inc;
        if (!isfinite(inc) || !isfinite(next) || next < 0.0) {
            raise_error(V, PMC_ERR_INVALID_STATE, c, nullptr, -6);
            continue;
        }
        V.occ[i] = next;
        r[0] += (kHbar * w * inc) * V.w;
        r[1] += 1.0;
    }
    block_reduce<2, kChunkWaves>(r, lds);
    if (threadIdx.x == 0) {
        if (half == 1) V.chunk_src[b] = r[0];
        atomicAdd(reinterpret_cast<unsigned long long*>(&V.st->source_updates),
                  static_cast<unsigned long long>(r[1]));
        atomicAdd(reinterpret_cast<unsigned long long*>(&V.st->step_source_updates),
                  static_cast<unsigned long long>(r[1]));
    }
}

__global__ void __launch_bounds__(kChunkThreads) k_source_sums(View V, int half) {
    source_sums_body(V, launch_grid(), half);
}
__global__ void __launch_bounds__(256) k_source_solve(View V, int half) { source_solve_body(V, launch_grid(), half); }
__global__ void __launch_bounds__(kChunkThreads) k_source_increments(View V) {
    source_increments_body(V, launch_grid());
}
__global__ void __launch_bounds__(kChunkThreads) k_source_apply(View V, int half) {
    source_apply_body(V, launch_grid(), half);
}
__global__ void __launch_bounds__(kChunkThreads) kb_source_sums(const View* __restrict__ views,
                                                                const int* __restrict__ offs, int n, int half) {
    Grid G;
    const int i = batch_instance(offs, n, G);
    source_sums_body(views[i], G, half);
}
__global__ void __launch_bounds__(256) kb_source_solve(const View* __restrict__ views, const int* __restrict__ offs,
                                                       int n, int half) {
    Grid G;
    const int i = batch_instance(offs, n, G);
    source_solve_body(views[i], G, half);
}
__global__ void __launch_bounds__(kChunkThreads) kb_source_increments(const View* __restrict__ views,
                                                                      const int* __restrict__ offs, int n) {
    Grid G;
    const int i = batch_instance(offs, n, G);
    source_increments_body(views[i], G);
}
__global__ void __launch_bounds__(kChunkThreads) kb_source_apply(const View* __restrict__ views,
                                                                 const int* __restrict__ offs, int n, int half) {
    Grid G;
    const int i = batch_instance(offs, n, G);
    source_apply_body(views[i], G, half);
}

__global__ void k_zero_tally(View V) {
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < kTallyStride * V.ncells; c += gridDim.x * blockDim.x)
        V.cell_tally[c] = 0.0;
}

__global__ void k_zero_flux(View V) {
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < 3 * V.ncells; c += gridDim.x * blockDim.x) V.cell_flux[c] = 0.0;
}

// Heat-flux tally (GridStatistics::update_heat_flux, src/solver/GridStatistics.cpp:269-347).
__global__ void __launch_bounds__(256) k_flux(View V) {
    const int n = V.st->n_slots;
    const int base = wave_global(launch_grid()) * 64 * kRowsPerWave;
    const int lane = __lane_id();
    SegAcc<3, 3> acc;
    acc.reset();
    for (int r = 0; r < kRowsPerWave; ++r) {
        const int row = base + r * 64;
        if (row >= n) break;
        const int i = row + lane;
        const int mw = i < n ? V.mode[i] : -1;
        const bool alive = mw >= 0;
        int cell = 0;
        double q[3] = {0, 0, 0};
        if (alive) {
            cell = V.cell[i];
            const int m = mw & kModeMask;
            const double e = (kHbar * V.omega[m] * (V.occ[i] - V.neq[m])) * V.w;
            const double* vd = V.veld + 3 * static_cast<size_t>(m);
            q[0] = ((mw >> kFlipShift) & 1 ? -vd[0] : vd[0]) * e;
            q[1] = ((mw >> kFlipShift) & 2 ? -vd[1] : vd[1]) * e;
            q[2] = ((mw >> kFlipShift) & 4 ? -vd[2] : vd[2]) * e;
        }
        acc.add(alive, cell, q, V.cell_flux);
    }
    acc.flush(V.cell_flux);
}

__global__ void k_flux_finish(View V) {
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < V.ncells; c += gridDim.x * blockDim.x)
        for (int a = 0; a < 3; ++a) V.cell_flux[3 * c + a] = V.cell_flux[3 * c + a] * (1.0 / V.cell_vol[c]) * kFluxUnit;
}

// Domain membership check (TransportOperator::validate_containment,
// src/solver/TransportOperator.cpp:364-402).
__device__ __forceinline__ void containment_body(const View& V, const Grid G) {
    const int i = G.b * blockDim.x + threadIdx.x;
    const int n = V.st->n_slots;
    if (i >= n || V.mode[i] < 0) return;
    const float p[3] = {V.px[i], V.py[i], V.pz[i]};
    double scale = 1.0;
    for (int a = 0; a < 3; ++a) scale = fmax(scale, fabs(V.bmax[a] - V.bmin[a]));
    // The reference's 1e-10 of the extent, widened to the FP32 position quantum:
    // carriers are clamped to the FP32-rounded walls, which can lie that far out.
    const double tol = fmax(1e-10 * scale, V.mesh_tol);
    for (int a = 0; a < 3; ++a)
        if (!isfinite(p[a]) || p[a] < V.bmin[a] - tol || p[a] > V.bmax[a] + tol) {
            raise_error(V, PMC_ERR_ESCAPED, i, p, -1);
            return;
        }
    if (V.geom_kind == 1) {
        // A voxel clear of the surface answers for all its points; the full test
        // runs near the surface only.
        if (V.field_inside) {
            const size_t vox = field_voxel(V, p);
            if (V.field[vox] * V.field_unit > static_cast<float>(2.0 * V.mesh_tol)) {
                if (!V.field_inside[vox]) raise_error(V, PMC_ERR_ESCAPED, i, p, -1);
                return;
            }
        }
        const double pd[3] = {p[0], p[1], p[2]};
        if (!mesh_contains(V, pd)) raise_error(V, PMC_ERR_ESCAPED, i, p, -1);
    }
}

__global__ void __launch_bounds__(256) k_containment(View V) { containment_body(V, launch_grid()); }

__global__ void __launch_bounds__(256) kb_containment(const View* __restrict__ views, const int* __restrict__ offs, int n) {
    Grid G;
    const int i = batch_instance(offs, n, G);
    containment_body(views[i], G);
}

// ---------------------------------------------------------------- initialization
// Initial population (src/solver/Initialization.cpp:48-113 and :181-265):
// uniform positions in the box, uniform active modes, Bose occupations at the
// initial temperature of the carrier's position.
// Uniform point in a mesh (SurfaceMesh::sample_volume_points,
// src/geometry/SurfaceSampling.cpp:113-155): a star tetrahedron by volume, then
// Dirichlet(1,1,1,1) weights from normalized exponentials, on the carrier's
// Philox words a, b and a third one.  Without a star decomposition, rejection in
// the bounding box against the ray cast (:81-106), up to 256 trials.
__device__ bool mesh_volume_point(const View& V, uint32_t id, const uint4& a, const uint4& b, double p[3]) {
    if (V.nsimplices > 0) {
        const uint4 c = philox(id, 0, 2, kStreamInit, V.key0, V.key1);
        const double target = u53(a.x, a.y) * V.simplex_cdf[V.nsimplices - 1];
        int lo = 0, hi = V.nsimplices - 1; // first simplex whose cumulative volume exceeds target
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (V.simplex_cdf[mid] > target) hi = mid;
            else lo = mid + 1;
        }
        const double u[4] = {u53(a.z, a.w), u53(b.x, b.y), u53(c.x, c.y), u53(c.z, c.w)};
        double w[4], sum = 0.0;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            w[k] = -log(fmax(u[k], 1e-15));
            sum += w[k];
        }
        const double* v = V.simplex_v + 12 * V.simplex_pos[lo];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            double x = 0.0;
#pragma unroll
            for (int k = 0; k < 4; ++k) x += (w[k] / sum) * v[3 * k + d];
            p[d] = x;
        }
        return true;
    }
    for (uint32_t trial = 0; trial < 256; ++trial) {
        const uint4 ta = philox(id, 1 + trial, 0, kStreamInit, V.key0, V.key1);
        const uint4 tb = philox(id, 1 + trial, 1, kStreamInit, V.key0, V.key1);
        const double u[3] = {u53(ta.x, ta.y), u53(ta.z, ta.w), u53(tb.x, tb.y)};
        for (int d = 0; d < 3; ++d) p[d] = V.bmin[d] + (V.bmax[d] - V.bmin[d]) * u[d];
        if (mesh_ray_cast(V, p)) return true;
    }
    return false;
}

__global__ void __launch_bounds__(256) k_init(View V, int count, int tmode, double t0, double c0, double c1, double c2,
                                              double h0, double h1, double h2, double tcold, double thot,
                                              const double* field) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const uint4 a = philox(static_cast<uint32_t>(i), 0, 0, kStreamInit, V.key0, V.key1);
    const uint4 b = philox(static_cast<uint32_t>(i), 0, 1, kStreamInit, V.key0, V.key1);
    const double u[4] = {u53(a.x, a.y), u53(a.z, a.w), u53(b.x, b.y), u53(b.z, b.w)};
    double p[3];
    for (int k = 0; k < 3; ++k) p[k] = V.bmin[k] + (V.bmax[k] - V.bmin[k]) * u[k];
    if (V.geom_kind == 1 && !mesh_volume_point(V, static_cast<uint32_t>(i), a, b, p)) {
        raise_error(V, PMC_ERR_BOUNDARY, i, nullptr, -11);
        V.mode[i] = -1;
        V.uid[i] = static_cast<unsigned>(i);
        V.cell[i] = 0;
        return;
    }
    int layer = 0;
    if (V.layers > 1) {
        // Layered materials: carriers denser where the phase-space density is,
        // by rejection with the ratio of spatial volumes (src/solver/Initialization.cpp:49-66),
        // up to 64 trials on Philox words (carrier, trial + 1, 0 and 1).
        bool accepted = false;
        for (uint32_t trial = 0; trial < 64 && !accepted; ++trial) {
            const uint4 ta = philox(static_cast<uint32_t>(i), 1 + trial, 0, kStreamInit, V.key0, V.key1);
            const uint4 tb = philox(static_cast<uint32_t>(i), 1 + trial, 1, kStreamInit, V.key0, V.key1);
            const double q[3] = {u53(ta.x, ta.y), u53(ta.z, ta.w), u53(tb.x, tb.y)};
            for (int k = 0; k < 3; ++k) p[k] = V.bmin[k] + (V.bmax[k] - V.bmin[k]) * q[k];
            const float pa = fminf(fmaxf(static_cast<float>(p[V.layer_axis]), V.fbmin[V.layer_axis]),
                                   V.fbmax[V.layer_axis]);
            layer = 0;
            while (layer + 1 < V.layers && pa >= V.layer_plane[layer]) ++layer;
            accepted = u53(tb.z, tb.w) < V.layer_accept[layer];
        }
        if (!accepted) {
            raise_error(V, PMC_ERR_BOUNDARY, i, nullptr, -11);
            V.mode[i] = -1;
            V.uid[i] = static_cast<unsigned>(i);
            V.cell[i] = 0;
            return;
        }
    }
    float fp[3];
    for (int k = 0; k < 3; ++k) {
        fp[k] = static_cast<float>(p[k]);
        fp[k] = fminf(fmaxf(fp[k], V.fbmin[k]), V.fbmax[k]);
    }
    int m;
    if (V.layers > 1) { // uniform over the carrier's material (src/solver/Initialization.cpp:140-144)
        const int b0 = V.mode_begin[layer], nm = V.mode_begin[layer + 1] - b0;
        const int k = static_cast<int>(u[3] * nm);
        m = b0 + (k < nm ? k : nm - 1);
    } else {
        m = static_cast<int>(u[3] * V.na);
        m = m < V.na ? m : V.na - 1;
    }
    double T = t0;
    const int cell = grid_index(V, fp[0], fp[1], fp[2]);
    if (tmode == 1) {
        const double ax[3] = {h0 - c0, h1 - c1, h2 - c2};
        const double rel[3] = {p[0] - c0, p[1] - c1, p[2] - c2};
        const double len2 = ax[0] * ax[0] + ax[1] * ax[1] + ax[2] * ax[2];
        double s = (rel[0] * ax[0] + rel[1] * ax[1] + rel[2] * ax[2]) / len2;
        s = s < 0.0 ? 0.0 : (s > 1.0 ? 1.0 : s);
        T = tcold + s * (thot - tcold);
    } else if (tmode == 2) {
        T = field[cell];
    }
    V.px[i] = fp[0];
    V.py[i] = fp[1];
    V.pz[i] = fp[2];
    V.mode[i] = m;
    // Linearized transport: n_eq(T_bg) + C(T_bg) (T - T_bg) / hw
    // (PhononMaterial::linearized_bose_occupation, src/material/Thermodynamics.cpp:27-37).
    if (V.linearized) {
        const double hw = kHbar * V.omega[m];
        V.occ[i] = hw > 0.0 ? V.neq[m] + V.cw_cap[m] * (T - V.t_bg) / hw : 0.0;
    } else {
        V.occ[i] = bose(T, V.omega[m]);
    }
    if (V.ex_on) { // MonteCarloSolver::apply_initial_mode_excitation (src/solver/Initialization.cpp:146-178)
        const double f = V.omega[m] / 6.283185307179586476925286766559;
        if (isfinite(f) && !(V.ex_fmin >= 0.0 && f < V.ex_fmin) && !(V.ex_fmax >= 0.0 && f > V.ex_fmax))
            V.occ[i] *= V.ex_mult;
    }
    V.cell[i] = cell;
    V.uid[i] = static_cast<unsigned>(i);
}

// Live carriers per cell of an initial population (for the first cell sort).
__global__ void __launch_bounds__(256) k_count_all(View V) {
    PMC_BLOCK_HASH(H);
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int n = V.st->n_slots;
    const int mw = i < n ? V.mode[i] : -1;
    H.count(mw >= 0, mw >= 0 ? V.cell[i] : 0, V.cell_count);
    H.flush_counts(V.cell_count);
}

} // namespace pmc_dev

// Error word of every instance of a batch, for one read by the host.
__global__ void kb_errors(const pmc_dev::View* __restrict__ views, int n, int* out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = views[i].st->error_code;
}

// ---------------------------------------------------------------- launchers
namespace pmc_launch {
using namespace pmc_dev;

int wave_blocks(int capacity) { return wave_blocks_for(capacity); }

int event_split(int capacity) {
    const int b = wave_blocks(capacity);
    return std::max(1, std::min(8, 2048 / b));
}

int event_rows(const View& v) { return event_rows_for(v); }

// Allocated rows: enough for the largest event split a context may select.
int partial_rows(const View& v) { return wave_blocks(v.capacity) * kMaxEventSplit + kInjectBlocks; }

static int slot_blocks(int capacity) { return (capacity + 255) / 256; }
static int cell_blocks(int ncells) { return (ncells + 255) / 256; }
static const dim3 kWaveBlock(64 * kWavesPerBlock);

void begin(const View& v, hipStream_t s) { hipLaunchKernelGGL(k_begin, dim3(1), dim3(64), 0, s, v); }
void transport(const View& v, hipStream_t s) {
    if (v.geom_kind == 1) hipLaunchKernelGGL(k_transport<kGeomMesh>, dim3(wave_blocks(v.capacity)), kWaveBlock, 0, s, v);
    else if (v.layers > 1) hipLaunchKernelGGL(k_transport<kGeomLayered>, dim3(wave_blocks(v.capacity)), kWaveBlock, 0, s, v);
    else hipLaunchKernelGGL(k_transport<kGeomBox>, dim3(wave_blocks(v.capacity)), kWaveBlock, 0, s, v);
}
void events(const View& v, hipStream_t s) {
    if (v.geom_kind == 1) hipLaunchKernelGGL(k_events<kGeomMesh>, dim3(event_rows(v)), kWaveBlock, 0, s, v);
    else if (v.layers > 1) hipLaunchKernelGGL(k_events<kGeomLayered>, dim3(event_rows(v)), kWaveBlock, 0, s, v);
    else hipLaunchKernelGGL(k_events<kGeomBox>, dim3(event_rows(v)), kWaveBlock, 0, s, v);
}
void inject(const View& v, hipStream_t s) {
    if (v.geom_kind == 1) hipLaunchKernelGGL(k_inject<kGeomMesh>, dim3(kInjectBlocks), dim3(256), 0, s, v);
    else if (v.layers > 1) hipLaunchKernelGGL(k_inject<kGeomLayered>, dim3(kInjectBlocks), dim3(256), 0, s, v);
    else hipLaunchKernelGGL(k_inject<kGeomBox>, dim3(kInjectBlocks), dim3(256), 0, s, v);
}
void refresh(const View& v, hipStream_t s, int phase) { hipLaunchKernelGGL(k_refresh, dim3(1), dim3(1024), 0, s, v, phase); }
void sort(const View& v, hipStream_t s) {
    const int b = (v.capacity + 64 * kSortRows * kWavesPerBlock - 1) / (64 * kSortRows * kWavesPerBlock);
    hipLaunchKernelGGL(k_sort, dim3((b + 7) & ~7), kWaveBlock, 0, s, v);
}
void invert(const View& v, hipStream_t s, int source) {
    hipLaunchKernelGGL(k_invert, dim3(v.ncells), dim3(256), 0, s, v, source);
}
void cell_prep(const View& v, hipStream_t s) {
    hipLaunchKernelGGL(k_cell_prep, dim3(cell_blocks(v.ncells)), dim3(256), 0, s, v);
}
void moments(const View& v, hipStream_t s) {
    hipLaunchKernelGGL(k_moments, dim3(v.max_chunks), dim3(kChunkThreads), 0, s, v);
}
void newton(const View& v, hipStream_t s) {
    hipLaunchKernelGGL(k_newton, dim3((v.ncells + 3) / 4), dim3(256), 0, s, v);
}
void relax(const View& v, hipStream_t s) {
    hipLaunchKernelGGL(k_relax, dim3(v.max_chunks), dim3(kChunkThreads), 0, s, v);
}
void callaway(const View& v, hipStream_t s) {
    const dim3 wg((v.ncells + 3) / 4), big(std::min(v.ncells, kCwBigBlocks));
    hipLaunchKernelGGL(k_callaway_wave<0>, wg, dim3(256), 0, s, v);
    hipLaunchKernelGGL(k_callaway_wave<1>, wg, dim3(256), 0, s, v);
    switch (v.cw_threads) {
    case 256:
        hipLaunchKernelGGL(k_callaway_setup<256>, big, dim3(256), 0, s, v);
        hipLaunchKernelGGL(k_callaway<256>, big, dim3(256), 0, s, v);
        break;
    case 384:
        hipLaunchKernelGGL(k_callaway_setup<384>, big, dim3(384), 0, s, v);
        hipLaunchKernelGGL(k_callaway<384>, big, dim3(384), 0, s, v);
        break;
    default:
        hipLaunchKernelGGL(k_callaway_setup<kCwThreads>, big, dim3(kCwThreads), 0, s, v);
        hipLaunchKernelGGL(k_callaway<kCwThreads>, big, dim3(kCwThreads), 0, s, v);
        break;
    }
}
void nlcallaway(const View& v, hipStream_t s) {
    hipLaunchKernelGGL(k_nlcallaway, dim3((v.ncells + 3) / 4), dim3(256), 0, s, v);
}
void source(const View& v, hipStream_t s, int half) {
    hipLaunchKernelGGL(k_source_sums, dim3(v.max_chunks), dim3(kChunkThreads), 0, s, v, half);
    hipLaunchKernelGGL(k_source_solve, dim3((v.ncells + 3) / 4), dim3(256), 0, s, v, half);
    hipLaunchKernelGGL(k_source_increments, dim3(v.max_chunks), dim3(kChunkThreads), 0, s, v);
    hipLaunchKernelGGL(k_source_apply, dim3(v.max_chunks), dim3(kChunkThreads), 0, s, v, half);
}
void clear_cells(const View& v, hipStream_t s) {
    hipLaunchKernelGGL(k_zero_tally, dim3(cell_blocks(kTallyStride * v.ncells)), dim3(256), 0, s, v);
}
void flux(const View& v, hipStream_t s) {
    hipLaunchKernelGGL(k_zero_flux, dim3(cell_blocks(3 * v.ncells)), dim3(256), 0, s, v);
    hipLaunchKernelGGL(k_flux, dim3(wave_blocks(v.capacity)), kWaveBlock, 0, s, v);
    hipLaunchKernelGGL(k_flux_finish, dim3(cell_blocks(v.ncells)), dim3(256), 0, s, v);
}
// ---- batched launches: one grid over many instances (per-instance block ranges)
int batch_blocks(const View& v, int kind) {
    switch (kind) {
    case kBatchBegin:
    case kBatchRefresh1:
    case kBatchRefresh2: return 1;
    case kBatchContainment: return slot_blocks(v.capacity);
    case kBatchTransport: return wave_blocks(v.capacity);
    case kBatchEvents: return event_rows(v);
    case kBatchInject: return kInjectBlocks;
    case kBatchSort: {
        const int b = (v.capacity + 64 * kSortRows * kWavesPerBlock - 1) / (64 * kSortRows * kWavesPerBlock);
        return (b + 7) & ~7;
    }
    case kBatchCellPrep: return cell_blocks(v.ncells);
    case kBatchMoments:
    case kBatchRelax: return v.max_chunks;
    case kBatchNewton: return (v.ncells + 3) / 4;
    case kBatchCallawayWave: return (v.ncells + 3) / 4;
    // Workgroups walking the large-cell list: as many as cells when the
    // population can fill cells beyond one wavefront's reach, else a few.
    case kBatchCallaway:
        return static_cast<long long>(v.capacity) > static_cast<long long>(kCwTinyMax) * v.ncells
                   ? std::min(v.ncells, 256)
                   : std::min(v.ncells, 4);
    case kBatchSourceSums0:
    case kBatchSourceSums1:
    case kBatchSourceIncrements:
    case kBatchSourceApply0:
    case kBatchSourceApply1: return v.hs_on ? v.max_chunks : 0; // instances without a source get no blocks
    case kBatchSourceSolve0:
    case kBatchSourceSolve1: return v.hs_on ? (v.ncells + 3) / 4 : 0;
    default: return 0;
    }
}

void batch_launch(int kind, bool mesh, const View* views, const int* offs, int n, int total, hipStream_t s) {
    if (total <= 0) return;
    const dim3 g(total);
    switch (kind) {
    case kBatchBegin: hipLaunchKernelGGL(kb_begin, g, dim3(64), 0, s, views, offs, n); break;
    case kBatchContainment: hipLaunchKernelGGL(kb_containment, g, dim3(256), 0, s, views, offs, n); break;
    case kBatchTransport:
        if (mesh) hipLaunchKernelGGL(kb_transport<kGeomMesh>, g, kWaveBlock, 0, s, views, offs, n);
        else hipLaunchKernelGGL(kb_transport<kGeomBox>, g, kWaveBlock, 0, s, views, offs, n);
        break;
    case kBatchEvents:
        if (mesh) hipLaunchKernelGGL(kb_events<kGeomMesh>, g, kWaveBlock, 0, s, views, offs, n);
        else hipLaunchKernelGGL(kb_events<kGeomBox>, g, kWaveBlock, 0, s, views, offs, n);
        break;
    case kBatchInject:
        if (mesh) hipLaunchKernelGGL(kb_inject<kGeomMesh>, g, dim3(256), 0, s, views, offs, n);
        else hipLaunchKernelGGL(kb_inject<kGeomBox>, g, dim3(256), 0, s, views, offs, n);
        break;
    case kBatchRefresh1: hipLaunchKernelGGL(kb_refresh, g, dim3(1024), 0, s, views, offs, n, 1); break;
    case kBatchRefresh2: hipLaunchKernelGGL(kb_refresh, g, dim3(1024), 0, s, views, offs, n, 2); break;
    case kBatchSort: hipLaunchKernelGGL(kb_sort, g, kWaveBlock, 0, s, views, offs, n); break;
    case kBatchCellPrep: hipLaunchKernelGGL(kb_cell_prep, g, dim3(256), 0, s, views, offs, n); break;
    case kBatchMoments: hipLaunchKernelGGL(kb_moments, g, dim3(kChunkThreads), 0, s, views, offs, n); break;
    case kBatchNewton: hipLaunchKernelGGL(kb_newton, g, dim3(256), 0, s, views, offs, n); break;
    case kBatchRelax: hipLaunchKernelGGL(kb_relax, g, dim3(kChunkThreads), 0, s, views, offs, n); break;
    case kBatchCallawayWave:
        hipLaunchKernelGGL(kb_callaway_wave<0>, g, dim3(256), 0, s, views, offs, n);
        hipLaunchKernelGGL(kb_callaway_wave<1>, g, dim3(256), 0, s, views, offs, n);
        break;
    case kBatchCallaway:
        hipLaunchKernelGGL(kb_callaway_setup, g, dim3(kCwThreads), 0, s, views, offs, n);
        hipLaunchKernelGGL(kb_callaway, g, dim3(kCwThreads), 0, s, views, offs, n);
        break;
    case kBatchSourceSums0:
    case kBatchSourceSums1:
        hipLaunchKernelGGL(kb_source_sums, g, dim3(kChunkThreads), 0, s, views, offs, n, kind == kBatchSourceSums1);
        break;
    case kBatchSourceSolve0:
    case kBatchSourceSolve1:
        hipLaunchKernelGGL(kb_source_solve, g, dim3(256), 0, s, views, offs, n, kind == kBatchSourceSolve1);
        break;
    case kBatchSourceIncrements:
        hipLaunchKernelGGL(kb_source_increments, g, dim3(kChunkThreads), 0, s, views, offs, n);
        break;
    case kBatchSourceApply0:
    case kBatchSourceApply1:
        hipLaunchKernelGGL(kb_source_apply, g, dim3(kChunkThreads), 0, s, views, offs, n, kind == kBatchSourceApply1);
        break;
    default: break;
    }
}

void batch_errors(const View* views, int n, int* out, hipStream_t s) {
    hipLaunchKernelGGL(kb_errors, dim3((n + 255) / 256), dim3(256), 0, s, views, n, out);
}

void containment(const View& v, hipStream_t s) {
    hipLaunchKernelGGL(k_containment, dim3(slot_blocks(v.capacity)), dim3(256), 0, s, v);
}
void init(const View& v, hipStream_t s, int count, int tmode, double t0, const double* cold, const double* hot,
          double tcold, double thot, const double* field) {
    hipLaunchKernelGGL(k_init, dim3((count + 255) / 256), dim3(256), 0, s, v, count, tmode, t0, cold[0], cold[1],
                       cold[2], hot[0], hot[1], hot[2], tcold, thot, field);
}
void count_all(const View& v, hipStream_t s) {
    hipLaunchKernelGGL(k_count_all, dim3(slot_blocks(v.capacity)), dim3(256), 0, s, v);
}
} // namespace pmc_launch

