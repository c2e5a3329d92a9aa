// Nonlinear Callaway collisions of one cell: the algorithm of
// nonlinear_callaway_evolve (src/solver/NonlinearCallawayCollision.cpp:251-285)
// written once for the device kernel (one wavefront per cell) and for a host
// build of the same code.
//
// Every carrier pass is a map-reduce over the cell's slots [begin, end): the
// executor X runs the per-slot body on its lanes (slot begin + lane, + 64, ...)
// and returns the K sums in every lane (X::map_reduce).  The control flow
// between passes (constraint directions, Newton steps, line search) is
// wave-uniform.  Per-slot values that several passes read live in scratch
// arrays indexed by slot: the substep weight a, the metric a n_eq (1 + n_eq),
// up to four constraint directions and the occupation before the collision.
// FP64 throughout (the reference works in long double); parity with the
// reference is to the moment solve's tolerance, not bit for bit.
#pragma once
#include <math.h>
#include <stdint.h>

#if defined(__HIPCC__)
#define PMC_NL_HD __host__ __device__
#else
#define PMC_NL_HD
#endif

namespace pmc_nl {

constexpr double kHbar = 6.582119569e-4;     // eV ps
constexpr double kBoltz = 8.617333262145e-5; // eV/K
constexpr double kFourPi = 4.0 * 3.14159265358979323846;

// Failure codes, in the reference's order of checks.
enum Status : int {
    kOk = 0,
    kNoEnergyConstraint = 1, // "Nonlinear Callaway has no representable energy constraint"
    kSingularJacobian = 2,   // "Singular nonlinear Callaway moment Jacobian"
    kLineSearch = 3,         // "Nonlinear Callaway moment line search failed"
    kNoConvergence = 4,      // "Nonlinear Callaway moment solve did not converge"
    kInvalidOccupation = 5,  // "Nonlinear Callaway occupation is invalid"
    kTemperatureRange = 7    // "Nonlinear Callaway temperature ... is outside the supplied N/U table"
};

// What one cell's collision reads and writes.
struct Data {
    const int* mode;        // slot -> mode word
    int mode_mask;          // active index = word & mode_mask
    double* occ;            // slot -> occupation
    const double* omega;    // active index -> angular frequency (rad/ps) at omega[omega_stride * m]
    int omega_stride;
    const double* neq_bg;   // active index -> n_BE(T_bg) at neq_bg[omega_stride * m]
    const double* q;        // active index -> Cartesian wavevector (3, 1/A)
    int n_active;
    int n_temps;
    const double* temps;    // rate temperature samples (K), increasing
    const double* gn;       // [n_temps * n_active] gamma_N
    const double* gu;       // [n_temps * n_active] gamma_U
    const double* ge;       // [n_temps * n_active] isotope/impurity/defect linewidths
    double* a;              // scratch per slot: substep weight
    double* metric;         // scratch per slot: a n_eq (1 + n_eq)
    double* b;              // scratch: direction j of slot i at b[j * bstride + i]
    double* n0;             // scratch per slot: occupation before the collision
    int64_t bstride;
    double energy_weight;   // eV of a carrier's hw n per (hw n)
};

PMC_NL_HD inline double bose(double x) { return x > 700.0 ? 0.0 : 1.0 / expm1(x); }

// Rate interpolation of PhononMaterial::mode_callaway_rates
// (src/material/PhononMaterialCollisionData.cpp:167-196): lower_bound bracket,
// no clamping outside the samples.
struct RateAt {
    int lo, hi;
    double f;
};
PMC_NL_HD inline int rate_bracket(const Data& D, double T, RateAt& r) {
    const int nt = D.n_temps;
    if (!(T > 0.0) || !isfinite(T) || nt <= 0 || T < D.temps[0] || T > D.temps[nt - 1]) return kTemperatureRange;
    int lo = 0, hi = nt; // first sample >= T
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (D.temps[mid] < T) lo = mid + 1;
        else hi = mid;
    }
    r.hi = lo;
    r.lo = lo > 0 ? lo - 1 : lo;
    r.f = r.lo == r.hi ? 0.0 : (T - D.temps[r.lo]) / (D.temps[r.hi] - D.temps[r.lo]);
    return kOk;
}
PMC_NL_HD inline double rate_of(const Data& D, const double* table, const RateAt& r, int m) {
    const int64_t na = D.n_active;
    return kFourPi * ((1.0 - r.f) * table[r.lo * na + m] + r.f * table[r.hi * na + m]);
}

// One exponential BGK substep (channel, :25-248): n <- (1 - a) n + a n_BE(phi)
// with a = 1 - exp(-rate dt), phi = sum_j lambda_j b_j chosen so the a-weighted
// energy moment (and, for the normal channel, the crystal-momentum moments) of
// the target equal the carriers'.
template <class X>
PMC_NL_HD int channel(X& x, const Data& D, int begin, int end, double T, const RateAt& r, double dt, bool normal) {
    double* const B0 = D.b;
    const int64_t bs = D.bstride;
    // Weights, metric and activity (:34-45).
    double act[1];
    x.map_reduce(begin, end, act, [&](int i, double (&acc)[1]) {
        const int m = D.mode[i] & D.mode_mask;
        const double rate = normal ? rate_of(D, D.gn, r, m) : rate_of(D, D.gu, r, m) + rate_of(D, D.ge, r, m);
        const double a = -expm1(-rate * dt);
        const double e = bose(kHbar * D.omega[D.omega_stride * m] / (kBoltz * T));
        D.a[i] = a;
        D.metric[i] = a * e * (1.0 + e);
        acc[0] += a;
    });
    if (act[0] == 0.0) return kOk;
    // Constraint directions (:46-87): raw vector, normalized, two passes of
    // projections onto the kept directions in the metric, dropped below 1e-12.
    int dim = 0;
    double lam[4] = {0.0, 0.0, 0.0, 0.0};
    const int ndir = normal ? 4 : 1;
    for (int j = 0; j < ndir; ++j) {
        double* const v = B0 + dim * bs;
        double s[1];
        x.map_reduce(begin, end, s, [&](int i, double (&acc)[1]) {
            const int m = D.mode[i] & D.mode_mask;
            const double val = j == 0 ? kHbar * D.omega[D.omega_stride * m] / (kBoltz * T) : D.q[3 * m + (j - 1)];
            v[i] = val;
            acc[0] += D.metric[i] * val * val;
        });
        if (s[0] == 0.0) continue;
        const double start = sqrt(s[0]);
        // The chain v /= start, then for each pass and kept direction k:
        // d = <v, b_k>, v -= d b_k, then |v|.  Each map-reduce applies the
        // pending update to v and forms the next dot product.
        double pend = 0.0;      // coefficient of the pending subtraction
        int pend_dir = -1;      // its direction (-1: the division by start)
        const int nops = 2 * dim;
        for (int op = 0; op <= nops; ++op) {
            const int k = op < nops ? op % dim : -1; // -1: the length
            const double* bk = k >= 0 ? B0 + k * bs : nullptr;
            const double* bp = pend_dir >= 0 ? B0 + pend_dir * bs : nullptr;
            const double p = pend;
            const bool divide = op == 0;
            double d[1];
            x.map_reduce(begin, end, d, [&](int i, double (&acc)[1]) {
                double vi = v[i];
                if (divide) vi /= start;
                else vi -= p * bp[i];
                v[i] = vi;
                acc[0] += D.metric[i] * vi * (bk ? bk[i] : vi);
            });
            if (k >= 0) {
                pend = d[0];
                pend_dir = k;
            } else {
                s[0] = d[0];
            }
        }
        const double len = sqrt(s[0]);
        if (len < 1e-12) continue;
        double none[1];
        x.map_reduce(begin, end, none, [&](int i, double (&)[1]) { v[i] /= len; });
        if (j == 0) lam[0] = start * len;
        ++dim;
    }
    if (dim == 0) return kNoEnergyConstraint;
    // Gradient scales (:91-97).
    double scale[4];
    x.map_reduce(begin, end, scale, [&](int i, double (&acc)[4]) {
        const int m = D.mode[i] & D.mode_mask;
        const double e0 = bose(kHbar * D.omega[D.omega_stride * m] / (kBoltz * T));
        const double w = fmax(D.occ[i], e0);
        for (int jj = 0; jj < 4; ++jj)
            if (jj < dim) acc[jj] += fabs(D.a[i] * B0[jj * bs + i]) * w;
    });
    // Objective, gradient and (packed upper) Hessian of the moment problem
    // (:98-131); the last sum counts slots whose phi is not positive and finite.
    constexpr int kH = 10, kCols = 1 + 4 + kH + 1;
    auto evaluate = [&](const double* l, bool hessian, double (&out)[kCols]) {
        x.map_reduce(begin, end, out, [&](int i, double (&acc)[kCols]) {
            const double a = D.a[i];
            if (a == 0.0) return;
            double bj[4];
            double phi = 0.0;
            for (int jj = 0; jj < 4; ++jj) {
                bj[jj] = jj < dim ? B0[jj * bs + i] : 0.0;
                if (jj < dim) phi += l[jj] * bj[jj];
            }
            if (!(phi > 0.0) || !isfinite(phi)) {
                acc[kCols - 1] += 1.0;
                return;
            }
            const double n = D.occ[i];
            const double e = bose(phi);
            acc[0] += a * (-log(-expm1(-phi)) + n * phi);
            for (int jj = 0; jj < 4; ++jj)
                if (jj < dim) acc[1 + jj] += a * bj[jj] * (n - e);
            if (hessian) {
                int h = 0;
                for (int jj = 0; jj < 4; ++jj)
                    for (int kk = jj; kk < 4; ++kk, ++h)
                        if (kk < dim) acc[5 + h] += a * bj[jj] * bj[kk] * e * (1.0 + e);
            }
        });
    };
    auto grad_error = [&](const double (&v)[kCols]) {
        double err = 0.0;
        for (int jj = 0; jj < dim; ++jj) err = fmax(err, fabs(v[1 + jj]) / fmax(1e-30, scale[jj]));
        return err;
    };
    // Damped Newton (:132-228).
    bool converged = false;
    for (int it = 0; it < 70; ++it) {
        double v0[kCols];
        evaluate(lam, true, v0);
        const double f0 = v0[kCols - 1] > 0.0 ? INFINITY : v0[0];
        const double err = grad_error(v0);
        if (err < 2e-13) {
            converged = true;
            break;
        }
        double M[4][4], dl[4];
        {
            int h = 0;
            for (int jj = 0; jj < 4; ++jj)
                for (int kk = jj; kk < 4; ++kk, ++h) M[jj][kk] = M[kk][jj] = v0[5 + h];
        }
        for (int jj = 0; jj < 4; ++jj) dl[jj] = -v0[1 + jj];
        // Gaussian elimination with partial pivoting.
        for (int col = 0; col < dim; ++col) {
            int piv = col;
            for (int rr = col + 1; rr < dim; ++rr)
                if (fabs(M[rr][col]) > fabs(M[piv][col])) piv = rr;
            if (!(fabs(M[piv][col]) > 1e-30)) return kSingularJacobian;
            if (piv != col) {
                for (int kk = 0; kk < 4; ++kk) {
                    const double t = M[col][kk];
                    M[col][kk] = M[piv][kk];
                    M[piv][kk] = t;
                }
                const double t = dl[col];
                dl[col] = dl[piv];
                dl[piv] = t;
            }
            for (int rr = col + 1; rr < dim; ++rr) {
                const double f = M[rr][col] / M[col][col];
                for (int kk = col; kk < dim; ++kk) M[rr][kk] -= f * M[col][kk];
                dl[rr] -= f * dl[col];
            }
        }
        for (int rr = dim - 1; rr >= 0; --rr) {
            for (int kk = rr + 1; kk < dim; ++kk) dl[rr] -= M[rr][kk] * dl[kk];
            dl[rr] /= M[rr][rr];
        }
        double slope = 0.0;
        for (int jj = 0; jj < dim; ++jj) slope += v0[1 + jj] * dl[jj];
        // Backtracking: sufficient decrease or a 10 % smaller gradient error.
        bool accepted = false;
        for (double t = 1.0; t > 1e-14; t *= 0.5) {
            double trial[4];
            for (int jj = 0; jj < 4; ++jj) trial[jj] = jj < dim ? lam[jj] + t * dl[jj] : 0.0;
            double v1[kCols];
            evaluate(trial, false, v1);
            const bool finite = v1[kCols - 1] == 0.0 && isfinite(v1[0]);
            if (finite && (v1[0] <= f0 + 1e-4 * t * slope || grad_error(v1) < err * 0.9)) {
                for (int jj = 0; jj < 4; ++jj) lam[jj] = trial[jj];
                accepted = true;
                break;
            }
        }
        if (!accepted) return kLineSearch;
    }
    if (!converged) return kNoConvergence;
    // n <- (1 - a) n + a n_BE(phi) (:229-246).
    double bad[1];
    x.map_reduce(begin, end, bad, [&](int i, double (&acc)[1]) {
        const double a = D.a[i];
        if (!(a > 0.0)) return;
        double phi = 0.0;
        for (int jj = 0; jj < 4; ++jj)
            if (jj < dim) phi += lam[jj] * B0[jj * bs + i];
        const double nn = (1.0 - a) * D.occ[i] + a * bose(phi);
        D.occ[i] = nn;
        if (!isfinite(nn) || nn < 0.0) acc[0] += 1.0;
    });
    return bad[0] > 0.0 ? kInvalidOccupation : kOk;
}

// The whole collision of one cell: N(dt/2), R(dt), N(dt/2) at cell temperature
// T with rates at rate_T (:251-285, :316-330).  out[0]: deviational energy
// after, sum hw (n - n_eq(T_bg)) w; out[1]: lifetime residual,
// sum (hw n - hw n_before) w.
template <class X>
PMC_NL_HD int cell(X& x, const Data& D, int begin, int end, double T, double rate_T, double dt, double (&out)[2]) {
    out[0] = out[1] = 0.0;
    if (end <= begin) return kOk;
    RateAt r;
    if (rate_bracket(D, rate_T, r)) return kTemperatureRange;
    double none[1];
    x.map_reduce(begin, end, none, [&](int i, double (&)[1]) { D.n0[i] = D.occ[i]; });
    if (dt > 0.0) {
        int rc = channel(x, D, begin, end, T, r, 0.5 * dt, true);
        if (!rc) rc = channel(x, D, begin, end, T, r, dt, false);
        if (!rc) rc = channel(x, D, begin, end, T, r, 0.5 * dt, true);
        if (rc) return rc;
    }
    x.map_reduce(begin, end, out, [&](int i, double (&acc)[2]) {
        const int m = D.mode[i] & D.mode_mask;
        const double w = D.omega[D.omega_stride * m];
        const double hw = kHbar * w;
        const double n = D.occ[i];
        acc[0] += (kHbar * w * (n - D.neq_bg[D.omega_stride * m])) * D.energy_weight;
        acc[1] += (hw * n - hw * D.n0[i]) * D.energy_weight;
    });
    return kOk;
}

} // namespace pmc_nl
