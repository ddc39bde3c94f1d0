// ORACLE (test infrastructure only): CPU restatement of the reference's individual-species
// lattice diffusion, for the parity tests and the benchmark's CPU baseline.  Never
// linked into the product library.
//
//   oracle_ind_sweep      individual_diffuseX / individual_diffuseY (IndividualSolver.cl:750-1014)
//   oracle_ind_constants  computeDiffusionConstants (IndividualSolver.cl:654-747)
//   oracle_ind_update     IndividualSolver::computeDiffusionConstants (IndividualSolver.cpp:1100-1153:
//                         sortSpecies, computeCellIndices, the constants kernel) followed by the
//                         per-cell time step minimum of InhomogeneousSolver::diffuse
//                         (InhomogeneousSolver.cpp:304-336, 384-414)
//   oracle_ind_run        the host loop of DiffusionModel::oclSolve (DiffusionModel.cpp:1467-1603)
//                         for one individual species and no reactions, with the X/Y sweep order and
//                         time-step updates of InhomogeneousSolver::diffuse (:254-427)
//
// The user hook computeDiffusivityDrift is the identity on the particle's four properties
// (DiffusivityX, DiffusivityY, DriftX, DriftY), the form of the commented default in
// IndividualSolver.cl:690-694.
//
// Time step: the per-cell formula of computeIndividualTimestepX/Y
// (StochasticInhomogeneousDriftDiffusion.cl:130-151) in double with tanh and sinh taken
// from an expm1 by Taylor series / argument reduction (oracle_expm1d), rounded to float; the
// reference evaluates it in float with its OpenCL built-ins (parity unpinned), this is the
// declared contract, evaluated bit-identically by the GPU build.
//
// Pinning: the hop probabilities by the closed-form sweep moments (mean r dt, variance
// 2 D dt) and reflecting/absorbing boundary invariants; the RNG stream (key cell + s*A,
// counter advanced per particle of the cell in sorted order, word x) by cell-locality
// invariants; oracle_expd / oracle_expm1d against libm exp / expm1 and the cell time step
// against the libm-based oracle_inhomogeneous_cell_dt; sort and cell indices by the particle oracle.
// The exact counter order inside a cell is fixed only by reading
// IndividualSolver.cl:766-777 (parity unpinned by execution).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "gpgmp_oracle.h"

namespace {

constexpr uint32_t kKeyBase = 0xdecafbadu;
constexpr uint32_t kCtrHi = 0xf00dcafeu;

// Hop probabilities of individual_diffuseX/Y (IndividualSolver.cl:840-842 / :953-955) with
// the reference's promotions: float operands, double literals.
void hop_probs(float D, float r, float dt, float dx, float &p1, float &p2) {
    const float rrdt = (r * r) * dt;
    const float dx2 = dx * dx;
    const float dxr = dx * r;
    const float ps = (float)(1. - ((double)dt * (2. * (double)D + (double)rrdt)) / (double)dx2);
    const double den = 2. * (double)D + (double)rrdt;
    p1 = (float)((1. - (double)ps) * (1. + (double)dxr / den) / 2.);
    p2 = (float)((double)p1 + (1. - (double)ps) * (1. - (double)dxr / den) / 2.);
}

double expd(double x) {
    const double k = std::rint(x * 1.4426950408889634074);
    const double r = (x - k * 6.93147180369123816490e-01) - k * 1.90821492927058770002e-10;
    // Taylor coefficients 1/n!, n = 14 .. 0, Horner
    static const double c[] = {1.0 / 87178291200.0, 1.0 / 6227020800.0, 1.0 / 479001600.0, 1.0 / 39916800.0,
                               1.0 / 3628800.0,     1.0 / 362880.0,     1.0 / 40320.0,     1.0 / 5040.0,
                               1.0 / 720.0,         1.0 / 120.0,        1.0 / 24.0,        1.0 / 6.0,
                               0.5,                 1.0,                1.0};
    double p = c[0];
    for (int i = 1; i < 15; ++i) p = c[i] + r * p;
    return std::ldexp(p, (int)k);
}

double expm1d(double x) {  // x >= 0
    if (x >= 0.34657359027997264) return expd(x) - 1.0;
    // Taylor coefficients 1/n!, n = 17 .. 2, Horner, then x + x^2 p
    static const double c[] = {1.0 / 355687428096000.0, 1.0 / 20922789888000.0, 1.0 / 1307674368000.0,
                               1.0 / 87178291200.0,     1.0 / 6227020800.0,     1.0 / 479001600.0,
                               1.0 / 39916800.0,        1.0 / 3628800.0,        1.0 / 362880.0,
                               1.0 / 40320.0,           1.0 / 5040.0,           1.0 / 720.0,
                               1.0 / 120.0,             1.0 / 24.0,             1.0 / 6.0,
                               0.5};
    double p = c[0];
    for (int i = 1; i < 16; ++i) p = c[i] + x * p;
    return x + x * (x * p);
}

float ind_cell_dt(float Df, float rf, float hf) {
    const double D = Df, r = rf, h = hf;
    if (r == 0.0) return (float)(h * h / (2.0 * D) * (1.0 - 2.0 / 3.0));
    if (D == 0.0) return (float)(h / std::fabs(r));
    const double eps = h * r / (2.0 * D);
    const double a = std::fabs(eps);
    double th = 1.0, ish2 = 0.0;
    if (!(a > 350.0)) {
        const double e2 = expm1d(2.0 * a), e1 = expm1d(a);
        th = e2 / (e2 + 2.0);
        const double sh = (e1 + e1 / (e1 + 1.0)) * 0.5;
        ish2 = 1.0 / (sh * sh);
    }
    if (eps < 0.0) th = -th;
    double temp = 1.0 + (-1.0 / (th * eps) + ish2);
    if (temp <= 0.0) temp = 1.0 / 3.0;
    return (float)(h / r * th * temp);
}

float ind_min_dt(const float *D, const float *r, int n, float h) {
    float m = INFINITY;
    for (int i = 0; i < n; ++i) m = std::min(m, ind_cell_dt(D[i], r[i], h));
    return m;
}

// OpenCL maxmag(x, y): the argument of larger magnitude, fmax on ties.
float maxmag(float x, float y) {
    const float ax = std::fabs(x), ay = std::fabs(y);
    if (ax > ay) return x;
    if (ay > ax) return y;
    return std::fmax(x, y);
}

}  // namespace

extern "C" void oracle_ind_sweep(const oracle_ind_params *p, uint32_t *keys, const uint32_t *vals,
                                 const uint32_t *left, const uint32_t *right, int n, const float *props, int cap,
                                 int axis, int32_t seed, float dt) {
    const int W = p->W, H = p->H, A = W * H;
    const float dx = oracle_header_float(p->physical_length / (float)W);
    const uint32_t k1 = kKeyBase + (uint32_t)seed;
    const uint32_t step = axis == 0 ? 1u : (uint32_t)W;
    for (int c = 0; c < A; ++c) {
        const int gx = c % W, gy = c / W;
        const uint32_t mlo = axis == 0 ? (gx == 0 ? (uint32_t)p->boundary_mask[0] : 1u)
                                       : (gy == 0 ? (uint32_t)p->boundary_mask[2] : 1u);
        const uint32_t mhi = axis == 0 ? (gx == W - 1 ? (uint32_t)p->boundary_mask[1] : 1u)
                                       : (gy == H - 1 ? (uint32_t)p->boundary_mask[3] : 1u);
        uint32_t ctr = 0;
        for (uint32_t i = left[c]; i <= right[c] && i < (uint32_t)n; ++i) {
            uint32_t key = keys[i];
            if (key != (uint32_t)c) continue;
            const uint32_t idx = vals[i];
            const float D = props[idx + (size_t)axis * cap];
            const float r = props[idx + (size_t)(2 + axis) * cap];
            float p1, p2;
            hop_probs(D, r, dt, dx, p1, p2);
            ++ctr;
            uint32_t w0, w1;
            oracle_threefry2x32_20(ctr, kCtrHi, (uint32_t)c + (uint32_t)p->species * (uint32_t)A, k1, &w0, &w1);
            const float u = oracle_u01(w0);
            const bool up = u <= p1;
            const bool moved = u <= p2;
            if (moved) key = up ? key + step * mhi : key - step * mlo;
            keys[i] = key;
        }
    }
}

extern "C" void oracle_ind_constants(const oracle_ind_params *p, const uint32_t *keys, const uint32_t *vals,
                                     const uint32_t *left, const uint32_t *right, int n, const float *props,
                                     int cap, float *Dx, float *Dy, float *rx, float *ry, int32_t *count) {
    const int A = p->W * p->H;
    for (int c = 0; c < A; ++c) {
        float dx = 0.f, dy = 0.f, qx = 0.f, qy = 0.f;
        int32_t k = 0;
        for (uint32_t i = left[c]; i <= right[c] && i < (uint32_t)n; ++i) {
            if (keys[i] != (uint32_t)c) continue;
            ++k;
            const uint32_t idx = vals[i];
            dx = std::fmax(props[idx], dx);
            dy = std::fmax(props[idx + (size_t)cap], dy);
            qx = maxmag(props[idx + 2 * (size_t)cap], qx);
            qy = maxmag(props[idx + 3 * (size_t)cap], qy);
        }
        Dx[c] = dx;
        Dy[c] = dy;
        rx[c] = qx;
        ry[c] = qy;
        count[c] = k;
    }
}

extern "C" void oracle_ind_update(const oracle_ind_params *p, uint32_t *keys, uint32_t *vals, int n,
                                  const float *props, int cap, uint32_t *left, uint32_t *right, float *Dx,
                                  float *Dy, float *rx, float *ry, int32_t *count, float *dtx, float *dty) {
    const int A = p->W * p->H;
    const float h = oracle_header_float(p->physical_length / (float)p->W);
    oracle_stable_sort_pairs(keys, vals, n);
    oracle_cell_left_right(keys, n, A, left, right);
    oracle_ind_constants(p, keys, vals, left, right, n, props, cap, Dx, Dy, rx, ry, count);
    *dtx = ind_min_dt(Dx, rx, A, h);
    *dty = ind_min_dt(Dy, ry, A, h);
}

extern "C" double oracle_expd(double x) { return expd(x); }
extern "C" double oracle_expm1d(double x) { return expm1d(x); }
extern "C" float oracle_ind_cell_dt(float D, float r, float h) { return ind_cell_dt(D, r, h); }

extern "C" int oracle_ind_run(const oracle_ind_params *p, uint32_t *keys, uint32_t *vals, int n, const float *props,
                              int cap, int steps, uint32_t seed0, int32_t *count, float *sim_time_out) {
    const size_t A = (size_t)p->W * p->H;
    std::vector<uint32_t> left(A), right(A);
    std::vector<float> Dx(A), Dy(A), rx(A), ry(A);
    float dtx, dty, ux, uy;
    oracle_ind_update(p, keys, vals, n, props, cap, left.data(), right.data(), Dx.data(), Dy.data(), rx.data(),
                      ry.data(), count, &dtx, &dty);
    float nx = dtx, ny = dty, sim = 0.f;
    std::vector<int32_t> seeds(steps > 0 ? steps : 1);
    oracle_mt19937_seeds(seed0, steps, seeds.data());
    for (int it = 0; it < steps; ++it) {
        int32_t seed = seeds[it];
        const float mindt = (nx < ny) ? nx : ny;
        if (!(mindt > 0) || std::isinf(mindt)) break;  // nothing can move
        if (nx <= mindt) {
            oracle_ind_sweep(p, keys, vals, left.data(), right.data(), n, props, cap, 0, seed, dtx);
            seed = (int32_t)((uint32_t)seed + 1u);
            oracle_ind_update(p, keys, vals, n, props, cap, left.data(), right.data(), Dx.data(), Dy.data(),
                              rx.data(), ry.data(), count, &ux, &uy);
            dtx = ux;
            nx += dtx;
        }
        if (ny <= mindt) {
            oracle_ind_sweep(p, keys, vals, left.data(), right.data(), n, props, cap, 1, seed, dty);
            seed = (int32_t)((uint32_t)seed + 1u);
            oracle_ind_update(p, keys, vals, n, props, cap, left.data(), right.data(), Dx.data(), Dy.data(),
                              rx.data(), ry.data(), count, &ux, &uy);
            dty = uy;
            ny += dty;
        }
        sim = mindt;
    }
    if (sim_time_out) *sim_time_out = sim;
    return 0;
}
