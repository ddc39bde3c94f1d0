// CPU oracle: RNG, numeric contract, SSA, stochastic diffusion and the host schedule.
// TEST INFRASTRUCTURE ONLY (see gpgmp_oracle.h).  Compiled with -ffp-contract=off.
#include "gpgmp_oracle.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

inline uint32_t rotl32(uint32_t x, int r) { return (x << r) | (x >> (32 - r)); }

}  // namespace

// Threefry-2x32 with 20 rounds, written from Salmon, Moraes, Dror, Shaw,
// "Parallel random numbers: as easy as 1, 2, 3" (SC'11): rotation constants
// {13,15,26,6,17,29,16,24}, key-schedule parity 0x1BD11BDA, a key injection
// before round 0 and after every 4th round (injection s adds ks[s%3],
// ks[(s+1)%3]+s).  The reference calls it as threefry2x32(ctr, key)
// (Random123/threefry.h:733 via Gillespie.cl:69).
// Pinned by: the published known-answer vectors and the vendored Random123
// header compiled under oracle/_ref (tests/test_oracle_rng.py).
extern "C" void oracle_threefry2x32_20(uint32_t c0, uint32_t c1, uint32_t k0, uint32_t k1,
                                       uint32_t *o0, uint32_t *o1) {
    static const int R[8] = {13, 15, 26, 6, 17, 29, 16, 24};
    uint32_t ks[3] = {k0, k1, 0x1BD11BDAu ^ k0 ^ k1};
    uint32_t x0 = c0 + ks[0];
    uint32_t x1 = c1 + ks[1];
    for (int r = 0; r < 20; ++r) {
        x0 += x1;
        x1 = rotl32(x1, R[r % 8]);
        x1 ^= x0;
        if ((r & 3) == 3) {
            uint32_t s = (uint32_t)((r + 1) >> 2);
            x0 += ks[s % 3];
            x1 += ks[(s + 1) % 3] + s;
        }
    }
    *o0 = x0;
    *o1 = x1;
}

// u = word (viewed as int32, the union at Gillespie.cl:45-48) * M_RAN_INVM32 + 0.5,
// evaluated in double (M_RAN_INVM32 is an unsuffixed literal, Solver.cpp:469-471)
// and stored to float (Gillespie.cl:72, StochasticHomogeneousDiffusion.cl:77-79).
// Pinned by: the closed form (N ^ 2^31) / 2^32 rounded to float, and 0/1 edge words.
extern "C" float oracle_u01(uint32_t w) {
    const double inv32 = 2.32830643653869628906e-010;
    return (float)((double)(int32_t)w * inv32 + 0.5);
}

// Double-precision natural log by a fixed sequence of IEEE operations:
// y = m * 2^e with m in [sqrt(1/2), sqrt(2)); s = (m-1)/(m+1);
// log m = 2 atanh s = 2s + 2s*(z/3 + z^2/5 + ... + z^7/15), z = s^2.
// Used for the SSA waiting time log(1.f/r1) (Gillespie.cl:76), whose OpenCL
// built-in is implementation-defined; the result is rounded to float.
// Pinned by: libm log() (error <= 1 ulp of float over sweeps and edge values).
extern "C" double oracle_logd(double y) {
    if (!(y > 0.0)) return (y == 0.0) ? -INFINITY : NAN;
    if (std::isinf(y)) return INFINITY;
    int e = 0;
    double m = std::frexp(y, &e);  // y = m * 2^e, m in [0.5, 1)
    m = m * 2.0;
    e = e - 1;                     // m in [1, 2)
    if (m > 1.4142135623730951) {
        m = m * 0.5;
        e = e + 1;
    }
    const double f = m - 1.0;
    const double s = f / (2.0 + f);
    const double z = s * s;
    double p = 1.0 / 15.0;
    p = 1.0 / 13.0 + z * p;
    p = 1.0 / 11.0 + z * p;
    p = 1.0 / 9.0 + z * p;
    p = 1.0 / 7.0 + z * p;
    p = 1.0 / 5.0 + z * p;
    p = 1.0 / 3.0 + z * p;
    p = z * p;
    const double two_s = 2.0 * s;
    const double logm = two_s + two_s * p;
    const double ln2 = 0.6931471805599453094;
    return (double)e * ln2 + logm;
}

extern "C" float oracle_logf(float y) { return (float)oracle_logd((double)y); }

extern "C" double oracle_expd(double x);

// The kinetic laws' functions (exp, log, log10 in double; pow by binary powering of an
// integral exponent up to 64, else exp(y log x) with the limits at x <= 0), on the oracle's
// own exp/log.  Pinned by: libm exp/log/log10/pow within a few
// ulp, and bit-equality with the HIP library's host evaluation of the same laws (tests).
extern "C" double oracle_law_fn_d(int fn, double x) {
    switch (fn) {
        case 0: return oracle_expd(x);
        case 1: return oracle_logd(x);
        default: return oracle_logd(x) * 0.43429448190325176;
    }
}

extern "C" double oracle_law_pow_d(double x, double y) {
    if (y == 0.0) return 1.0;
    if (x != x || y != y) return x + y;
    if (std::floor(y) == y && std::fabs(y) <= 64.0) {  // binary powering of an integral exponent
        int n = (int)std::fabs(y);
        double r = 1.0, b = x;
        while (n) {
            if (n & 1) r = r * b;
            b = b * b;
            n >>= 1;
        }
        return (y < 0.0) ? 1.0 / r : r;
    }
    if (x == 0.0) return (y > 0.0) ? 0.0 : INFINITY;
    if (x < 0.0) {
        if (std::floor(y) != y) return NAN;
        const double m = oracle_expd(y * oracle_logd(-x));
        const double half = y * 0.5;
        return (std::floor(half) == half) ? m : -m;
    }
    return oracle_expd(y * oracle_logd(x));
}

// Host seed schedule: boost::mt19937 rng(seed0); uniform_int<>(0, INT_MAX) and
// seed_n = draw_n + n (DiffusionModel.cpp:1467-1484).  For a 32-bit engine and
// range 2^31 Boost 1.54's generate_uniform_int uses bucket size 2, so the draw is
// eng() >> 1.  Pinned by: std::mt19937's standard-specified 10000th output
// (4123659995 for the default seed); the bucket mapping is parity unpinned.
extern "C" void oracle_mt19937_seeds(uint32_t seed0, int n, int32_t *out) {
    std::mt19937 eng(seed0);
    for (int i = 0; i < n; ++i) {
        uint32_t draw = (uint32_t)eng() >> 1;
        out[i] = (int32_t)(draw + (uint32_t)i);
    }
}

// Reaction table from rates and stoichiometry (Reaction.cpp:119-138 kinetic law,
// Solver.cpp:843-905 hazards/updates).  Factors follow species index order.
// Pinned by: hand-evaluated laws in tests (int and double literal cases).
extern "C" int oracle_compile_reactions(int S, int R, const float *rates, const int *rs,
                                        const int *ps, oracle_reaction_table *t) {
    if (S < 1 || S > ORACLE_MAX_SPECIES || R < 0 || R > ORACLE_MAX_REACTIONS) return -1;
    std::memset(t, 0, sizeof(*t));
    t->num_species = S;
    t->num_reactions = R;
    for (int i = 0; i < R; ++i) {
        char buf[64];
        std::snprintf(buf, sizeof buf, "%g", (double)rates[i]);
        bool is_int = true;
        for (const char *c = buf; *c; ++c)
            if (!((*c >= '0' && *c <= '9') || *c == '-')) is_int = false;
        t->lit_is_int[i] = is_int ? 1 : 0;
        t->lit_dbl[i] = std::strtod(buf, nullptr);
        t->lit_int[i] = is_int ? (int32_t)std::strtol(buf, nullptr, 10) : 0;
        int nf = 0;
        for (int j = 0; j < S; ++j) {
            for (int m = 0; m < rs[i * S + j]; ++m) {
                if (nf >= ORACLE_MAX_FACTORS) return -2;
                t->factor_species[i][nf] = j;
                t->factor_offset[i][nf] = m;
                ++nf;
            }
            t->net[i][j] = ps[i * S + j] - rs[i * S + j];
        }
        t->n_factors[i] = nf;
    }
    return 0;
}

// Header float constants: printed "%.6e" and re-read as float literals
// (Solver.cpp:494-505, ss << std::scientific).
extern "C" float oracle_header_float(float v) {
    char buf[64];
    std::snprintf(buf, sizeof buf, "%.6e", (double)v);
    return std::strtof(buf, nullptr);
}

// StayProbability = (1.-p0) printed "%.6e" with an f suffix (Solver.cpp:505).
extern "C" float oracle_stay_probability(float p0) {
    char buf[64];
    std::snprintf(buf, sizeof buf, "%.6e", 1.0 - (double)p0);
    return std::strtof(buf, nullptr);
}

// Homogeneous per-species time step p0/((Real)2.*3)*h*h/D in float
// (HomogeneousSolver.cpp:89).  Pinned by: the closed form p0 h^2 / (6 D).
extern "C" float oracle_homogeneous_dt(float p0, float h, float D) {
    float six = (float)2. * 3;
    float a = p0 / six;
    a = a * h;
    a = a * h;
    return a / D;
}

// Per-cell admissible time step for the stochastic inhomogeneous solver
// (StochasticInhomogeneousDriftDiffusion.cl:130-151), evaluated in double on
// the host (the reference's float tanh/sinh are implementation-defined).
// Pinned by: the limits h^2/(6D) (r -> 0) and h/|r| (D = 0).
extern "C" float oracle_inhomogeneous_cell_dt(float Df, float rf, float hf) {
    const double D = Df, r = rf, h = hf;
    double ts;
    if (r == 0.0) {
        ts = h * h / (2.0 * D) * (1.0 - 2.0 / 3.0);
    } else if (D == 0.0) {
        ts = h / std::fabs(r);
    } else {
        const double eps = h * r / (2.0 * D);
        double temp = 1.0 + (-1.0 / (std::tanh(eps) * eps) + 1.0 / (std::sinh(eps) * std::sinh(eps)));
        if (temp <= 0.0) temp = 1.0 / 3.0;
        ts = h / r * std::tanh(eps) * temp;
    }
    return (float)ts;
}

// stochastic_inhomogeneous_fpe per-cell step (StochasticInhomogeneousDriftDiffusion.cl:
// 123-126 for X, 268-271 for Y, both with PhysicalCellWidth): the generated expression
// (Real)StayProbability/6./(diffHere/(deltax*deltax)+fabs(rx/(2.*deltax))) under C's
// promotions -- D/(h*h) in float, |r/(2.*h)| in double, the sum and the quotient in double.
// Pinned by: the pure-diffusion value stay*h^2/(6D) and the pure-drift value stay*h/(3|r|)
// (tests/test_oracle_fpe.py), and the rates' closed forms.
extern "C" float oracle_fpe_cell_dt(float D, float r, float h, float stay) {
    const float dterm = D / (h * h);
    const double den = (double)dterm + std::fabs((double)r / (2. * (double)h));
    return (float)((double)stay / 6. / den);
}

extern "C" float oracle_fpe_dt(const float *D, const float *r, int n, float h, float stay) {
    float m = INFINITY;
    for (int i = 0; i < n; ++i) m = std::min(m, oracle_fpe_cell_dt(D[i], r[i], h, stay));
    return m;
}

// Minimum over cells (computeIndividualTimestepX/Y + reduceBlocks,
// StochasticInhomogeneousDriftDiffusion.cl:88-423).
extern "C" float oracle_inhomogeneous_dt(const float *D, const float *r, int n, float h) {
    float m = INFINITY;
    for (int i = 0; i < n; ++i) m = std::min(m, oracle_inhomogeneous_cell_dt(D[i], r[i], h));
    return m;
}

namespace {

// _h[i+1] = _h[0] + <law>  (Solver.cpp:843-857, law Reaction.cpp:119-138):
// int literal -> int arithmetic then float add; double literal -> double chain.
// A localized reaction's generated hazard is `_h[0] + read_imagef(mask).x * <law>`
// (Solver.cpp:853-854): C evaluates (((w * lit) * X) * (X-1)) ... left to right, so with the
// float mask value w an int literal's chain is float and a double literal's is double.
inline float add_law(const oracle_reaction_table *t, int i, float h0, const int32_t *x, int cell) {
    if ((t->localized >> i) & 1u) {
        const float w = ((t->act[cell] >> i) & 1u) ? 1.0f : 0.0f;
        if (t->lit_is_int[i]) {
            float v = w * (float)t->lit_int[i];
            for (int f = 0; f < t->n_factors[i]; ++f)
                v = v * (float)(x[t->factor_species[i][f]] - t->factor_offset[i][f]);
            return h0 + v;
        }
        double v = (double)w * t->lit_dbl[i];
        for (int f = 0; f < t->n_factors[i]; ++f)
            v = v * (double)(x[t->factor_species[i][f]] - t->factor_offset[i][f]);
        return (float)((double)h0 + v);
    }
    if (t->lit_is_int[i]) {
        uint32_t v = (uint32_t)t->lit_int[i];
        for (int f = 0; f < t->n_factors[i]; ++f)
            v = v * (uint32_t)(x[t->factor_species[i][f]] - t->factor_offset[i][f]);
        return h0 + (float)(int32_t)v;
    }
    double v = t->lit_dbl[i];
    for (int f = 0; f < t->n_factors[i]; ++f)
        v = v * (double)(x[t->factor_species[i][f]] - t->factor_offset[i][f]);
    return (float)((double)h0 + v);
}

int threads_or_default(int n) {
#ifdef _OPENMP
    return n > 0 ? n : omp_get_max_threads();
#else
    (void)n;
    return 1;
#endif
}

}  // namespace

// The cumulative hazards _h[0..R] of one state x in cell `cell` (the hazard block of
// performGillespie, Solver.cpp:843-857).
// Pinned by: laws evaluated by hand in the tests (int/double literals, localized reactions
// with the mask's float chain against the unlocalized int chain, mask 0).
extern "C" void oracle_hazards(const oracle_reaction_table *t, const int32_t *x, int cell, float *h) {
    if (t->hazards_fn) {
        const int W = t->grid_w > 0 ? t->grid_w : 1;
        t->hazards_fn(x, cell % W, cell / W, t->localized ? (uint32_t)t->act[cell] : 0xffffffffu, h);
        return;
    }
    h[0] = 0.f;
    for (int i = 0; i < t->num_reactions; ++i) {
        h[i + 1] = add_law(t, i, h[0], x, cell);
        h[0] = h[i + 1];
    }
}

// performGillespie (Gillespie.cl:4-106) for every cell; the kernel is launched
// over (W,H) with key {cell, 0xdecafbad+seed}, counter {++n, 0xf00dcafe}.
// Pinned by: birth-death Poisson moments, A+B steady state, brute force on tiny grids.
extern "C" void oracle_ssa(int32_t *state, int W, int H, const oracle_reaction_table *t,
                           int32_t seed, float deltat, int nthreads) {
    const int A = W * H, S = t->num_species, R = t->num_reactions;
    if (R == 0) return;
    const double inv32 = 2.32830643653869628906e-010;
#pragma omp parallel for schedule(dynamic, 256) num_threads(threads_or_default(nthreads))
    for (int c = 0; c < A; ++c) {
        int32_t x[ORACLE_MAX_SPECIES];
        for (int s = 0; s < S; ++s) x[s] = state[c + s * A];
        const uint32_t k0 = (uint32_t)c, k1 = 0xdecafbadu + (uint32_t)seed;
        uint32_t ctr = 0;
        float h[ORACLE_MAX_REACTIONS + 1];
        float time = 0.f;
        while (time < deltat) {
            if (t->hazards_fn) {
                t->hazards_fn(x, c % W, c / W, t->localized ? (uint32_t)t->act[c] : 0xffffffffu, h);
            } else {
                h[0] = 0.f;
                for (int i = 0; i < R; ++i) {
                    h[i + 1] = add_law(t, i, h[0], x, c);
                    h[0] = h[i + 1];
                }
            }
            if (h[0] == 0) break;
            if (ctr == (1u << 24)) break;  // same runaway guard as the HIP kernel
            ++ctr;
            uint32_t w0, w1;
            oracle_threefry2x32_20(ctr, 0xf00dcafeu, k0, k1, &w0, &w1);
            const float r1 = (float)((double)(int32_t)w0 * inv32 + 0.5);
            const float r2 = (float)(((double)(int32_t)w1 * inv32 + 0.5) * (double)h[0]);
            const float dt = oracle_logf(1.f / r1) / h[0];
            time += dt;
            if (time < deltat) {
                for (int i = 0; i < R; ++i) {
                    int idx = (i == 0) ? (r2 < h[1]) : ((r2 >= h[i]) * (r2 < h[i + 1]));
                    if (idx)
                        for (int s = 0; s < S; ++s) x[s] += t->net[i][s];
                }
            }
        }
        for (int s = 0; s < S; ++s) state[c + s * A] = x[s];
    }
}

// homogeneous_stochastic_diffuse + homogeneous_stochastic_updateState
// (StochasticHomogeneousDiffusion.cl:8-124) and, when the source mask asks for it,
// homogeneous_stochastic_source_boundary (:128-183) as sequenced by
// HomogeneousSolver::diffuse (HomogeneousSolver.cpp:132-156; boundary 0 takes its
// count from sourceMask[1], as there).  RNG key {cell, 0xdecafbad+seed} carries no
// species offset.
// Pinned by: conservation (reflective), per-axis variance 2 p0/6 h^2, heat kernel.
extern "C" void oracle_hom_diffuse(int32_t *state, int W, int H, int species, int32_t seed,
                                   float stay, const int *bm, const int *sm, int nthreads) {
    const int A = W * H;
    int32_t *st = state + (size_t)species * A;
    std::vector<int32_t> lost((size_t)A * 4);
    const float pi = (float)((1. - (double)stay) / (2. * 3));
    const float p1 = stay + pi, p2 = p1 + pi, p3 = p2 + pi, p4 = p3 + pi;
    const uint32_t k1 = 0xdecafbadu + (uint32_t)seed;
#pragma omp parallel for schedule(dynamic, 256) num_threads(threads_or_default(nthreads))
    for (int c = 0; c < A; ++c) {
        const int gx = c % W, gy = c / W;
        int mask[4];
        mask[0] = (gx == 0) ? bm[0] : 1;
        mask[1] = (gx == W - 1) ? bm[1] : 1;
        mask[2] = (gy == 0) ? bm[2] : 1;
        mask[3] = (gy == H - 1) ? bm[3] : 1;
        int l[4] = {0, 0, 0, 0};
        uint32_t ctr = 0, w0 = 0, w1 = 0;
        const int n = st[c];
        for (int i = 0; i < n; ++i) {
            float u;
            if (!(i % 2)) {
                ++ctr;
                oracle_threefry2x32_20(ctr, 0xf00dcafeu, (uint32_t)c, k1, &w0, &w1);
                u = oracle_u01(w0);
            } else {
                u = oracle_u01(w1);
            }
            int dir = 3;
            dir -= (u <= p3);
            dir -= (u <= p2);
            dir -= (u <= p1);
            l[dir] += mask[dir] * (u > stay && u <= p4);
        }
        for (int d = 0; d < 4; ++d) lost[(size_t)c * 4 + d] = l[d];
        st[c] -= l[0] + l[1] + l[2] + l[3];
    }
#pragma omp parallel for schedule(static) num_threads(threads_or_default(nthreads))
    for (int c = 0; c < A; ++c) {
        const int gx = c % W, gy = c / W;
        int v = st[c];
        if (gx > 0) v += lost[(size_t)(c - 1) * 4 + 1];
        if (gx < W - 1) v += lost[(size_t)(c + 1) * 4 + 0];
        if (gy > 0) v += lost[(size_t)(c - W) * 4 + 3];
        if (gy < H - 1) v += lost[(size_t)(c + W) * 4 + 2];
        st[c] = v;
    }
    // source boundaries, in the order 0,1,2,3 with the counts the host passes
    const int counts[4] = {sm[1], sm[1], sm[2], sm[3]};
    for (int b = 0; b < 4; ++b) {
        if (sm[b] <= 0) continue;
        const int len = (b < 2) ? H : W;
        for (int d = 0; d < len; ++d) {
            int gid;
            switch (b) {
                case 0: gid = d * W; break;
                case 1: gid = (W - 1) + d * W; break;
                case 2: gid = d; break;
                default: gid = d + (H - 1) * W; break;
            }
            uint32_t ctr = 0, w0 = 0, w1 = 0;
            for (int i = 0; i < counts[b]; ++i) {
                float u;
                if (!(i % 2)) {
                    ++ctr;
                    oracle_threefry2x32_20(ctr, 0xf00dcafeu, (uint32_t)d, k1, &w0, &w1);
                    u = oracle_u01(w0);
                } else {
                    u = oracle_u01(w1);
                }
                st[gid] += (u > stay && u <= p1);
            }
        }
    }
}

// One X or Y sweep with the inhomogeneous (fpe = false) or Fokker-Planck hop thresholds.
static void sweep_impl(int32_t *state, int W, int H, int species, int axis, int32_t seed, float dt,
                       const float *Dax, const float *rax, float dxh, const int *bm, int nthreads, bool fpe) {
    const int A = W * H;
    int32_t *st = state + (size_t)species * A;
    const float *Ds = Dax + (size_t)species * A;
    const float *rs = rax + (size_t)species * A;
    std::vector<int32_t> lost((size_t)A * 2);
    const uint32_t k1 = 0xdecafbadu + (uint32_t)seed;
#pragma omp parallel for schedule(dynamic, 256) num_threads(threads_or_default(nthreads))
    for (int c = 0; c < A; ++c) {
        const int gx = c % W, gy = c / W;
        int mask[4];
        mask[0] = (gx == 0) ? bm[0] : 1;
        mask[1] = (gx == W - 1) ? bm[1] : 1;
        mask[2] = (gy == 0) ? bm[2] : 1;
        mask[3] = (gy == H - 1) ? bm[3] : 1;
        const float D = Ds[c], r = rs[c];
        float p1, p2;
        if (fpe) {
            // transition rates (…DriftDiffusion.cl:191-197 / 341-347, stored float4) and the
            // thresholds p1 = rate_plus*dt, p2 = p1 + rate_minus*dt in float (:491-500 / :626-633)
            const float dD = D / (dxh * dxh);
            const float up = (float)((double)dD + (double)r / (2. * (double)dxh));
            const float down = (float)((double)dD - (double)r / (2. * (double)dxh));
            p1 = up * dt;
            p2 = p1 + down * dt;
        } else {
            const float rrdt = (r * r) * dt;
            const float dx2 = dxh * dxh;
            const float dxr = dxh * r;
            const float ps = (float)(1. - ((double)dt * (2. * (double)D + (double)rrdt)) / (double)dx2);
            const double den = 2. * (double)D + (double)rrdt;
            p1 = (float)((1. - (double)ps) * (1. + (double)dxr / den) / 2.);
            p2 = (float)((double)p1 + (1. - (double)ps) * (1. - (double)dxr / den) / 2.);
        }
        const int base = axis ? 2 : 0;
        int l[4] = {0, 0, 0, 0};
        uint32_t ctr = 0, w0 = 0, w1 = 0;
        const int n = st[c];
        const uint32_t k0 = (uint32_t)(c + species * A);
        for (int i = 0; i < n; ++i) {
            float u;
            if (!(i % 2)) {
                ++ctr;
                oracle_threefry2x32_20(ctr, 0xf00dcafeu, k0, k1, &w0, &w1);
                u = oracle_u01(w0);
            } else {
                u = oracle_u01(w1);
            }
            int dir = base + (u <= p1);
            l[dir] += mask[dir] * (u <= p2);
        }
        lost[(size_t)c * 2 + 0] = l[base];
        lost[(size_t)c * 2 + 1] = l[base + 1];
        st[c] -= l[0] + l[1] + l[2] + l[3];
    }
#pragma omp parallel for schedule(static) num_threads(threads_or_default(nthreads))
    for (int c = 0; c < A; ++c) {
        const int gx = c % W, gy = c / W;
        int v = st[c];
        if (axis == 0) {
            if (gx > 0) v += lost[(size_t)(c - 1) * 2 + 1];
            if (gx < W - 1) v += lost[(size_t)(c + 1) * 2 + 0];
        } else {
            if (gy > 0) v += lost[(size_t)(c - W) * 2 + 1];
            if (gy < H - 1) v += lost[(size_t)(c + W) * 2 + 0];
        }
        st[c] = v;
    }
}

// inhomogeneous_diffuseX/Y + inhomogeneous_updateStateX/Y
// (StochasticInhomogeneousDriftDiffusion.cl:426-723); key {cell + s*A, 0xdecafbad+seed};
// probabilities at :518-520 / :648-650 with the reference's float/double promotions.
// Pinned by: conservation, per-sweep mean r dt and variance 2 D dt closed forms.
extern "C" void oracle_inhom_sweep(int32_t *state, int W, int H, int species, int axis, int32_t seed, float dt,
                                   const float *Dax, const float *rax, float dxh, const int *bm, int nthreads) {
    sweep_impl(state, W, H, species, axis, seed, dt, Dax, rax, dxh, bm, nthreads, false);
}

// The same sweep of the stochastic_inhomogeneous_fpe solver (the GPGMP_SOLVER_STOCHASTIC_
// INHOMOGENEOUS_FPE branches of inhomogeneous_diffuseX/Y, :491-500, :626-633).
// Pinned by: conservation, the per-sweep mean hop balance (D/h^2 + r/2h - D/h^2 + r/2h) dt h
// = r dt and second moment 2 D dt (tests/test_oracle_fpe.py).
extern "C" void oracle_fpe_sweep(int32_t *state, int W, int H, int species, int axis, int32_t seed, float dt,
                                 const float *Dax, const float *rax, float dxh, const int *bm, int nthreads) {
    sweep_impl(state, W, H, species, axis, seed, dt, Dax, rax, dxh, bm, nthreads, true);
}

// DiffusionModel::oclSolve host loop (DiffusionModel.cpp:1414-1620) with the
// solver semantics of HomogeneousSolver::diffuse (HomogeneousSolver.cpp:132-164)
// and InhomogeneousSolver::diffuse / getNextDiffusionTime
// (InhomogeneousSolver.cpp:238-427).  No output or alert events.
// Pinned by: the per-kernel checks above plus the statistical end-to-end tests.
// DiffusionModel::oclSolve's loop (DiffusionModel.cpp:1479-1603) with its output events
// (:1519-1572): an output time before the next diffusion makes the iteration a reaction
// step up to it without diffusion, and the state is written once the clock has reached it
// (the next iteration, of length zero); the loop stops once simTime reaches max_time.  The
// initial state is written first (:1442).  interval / max_time = infinity disables them.
// dumps: [max_dumps][S][A]; returns the number of writes (stored or not).
// skip: the first `skip` iterations advance the schedule and the seed stream only (the
// state passed in is the state after them, e.g. a snapshot of an earlier run); the other
// `steps` iterations run.
static int run_loop(const oracle_model *m, const oracle_reaction_table *t, int32_t *state, int skip, int steps,
                    uint32_t seed0, int nthreads, float interval, float max_time, int32_t *dumps, float *dump_times,
                    int max_dumps, int64_t *iterations, float *sim_time_out) {
    steps += skip;
    const int S = m->S, A = m->W * m->H;
    const float h = oracle_header_float(m->physical_length / (float)m->W);
    const float hy = oracle_header_float(m->physical_length / (float)m->H);
    const float stay = oracle_stay_probability(m->p0);
    std::vector<float> dtx(S), dty(S), nx(S), ny(S), dth(S), nh(S);
    if (!m->inhomogeneous) {
        for (int s = 0; s < S; ++s) {
            dth[s] = (m->diffusion[s] > 0) ? oracle_homogeneous_dt(m->p0, m->physical_length / (float)m->W,
                                                                   m->diffusion[s])
                                           : m->max_time;
            nh[s] = dth[s];
        }
    } else {
        for (int s = 0; s < S; ++s) {
            if (m->inhomogeneous == 2) {
                dtx[s] = oracle_fpe_dt(m->Dx + (size_t)s * A, m->rx + (size_t)s * A, A, h, stay);
                dty[s] = oracle_fpe_dt(m->Dy + (size_t)s * A, m->ry + (size_t)s * A, A, h, stay);
            } else {
                dtx[s] = oracle_inhomogeneous_dt(m->Dx + (size_t)s * A, m->rx + (size_t)s * A, A, h);
                dty[s] = oracle_inhomogeneous_dt(m->Dy + (size_t)s * A, m->ry + (size_t)s * A, A, h);
            }
            nx[s] = dtx[s];
            ny[s] = dty[s];
        }
    }
    auto next_times = [&](std::vector<float> &out) {
        out.resize(S);
        for (int s = 0; s < S; ++s) {
            if (!m->inhomogeneous) out[s] = nh[s];
            else out[s] = (nx[s] < ny[s]) ? nx[s] : ny[s];
        }
    };
    std::vector<float> next;
    next_times(next);
    std::vector<int32_t> seeds(steps > 0 ? steps : 1);
    oracle_mt19937_seeds(seed0, steps, seeds.data());
    float simTime = 0.f, outputLast = 0.f;
    int writes = 0;
    auto write = [&](float time) {
        if (writes < max_dumps) {
            std::copy(state, state + (size_t)S * A, dumps + (size_t)writes * S * A);
            dump_times[writes] = time;
        }
        ++writes;
    };
    if (std::isfinite(interval) || std::isfinite(max_time)) write(0.f);
    int n = 0;
    for (; n < steps && simTime < max_time; ++n) {
        int32_t seed = seeds[n];
        float mindt = next[0];
        int minspec = 0;
        for (int i = 1; i < S; ++i)
            if (next[i] < mindt) {
                mindt = next[i];
                minspec = i;
            }
        const float mindtOutput = outputLast + interval;
        const bool outputEvent = mindtOutput < mindt;
        if (outputEvent) mindt = mindtOutput;
        const float deltat = mindt - simTime;
        const bool run = n >= skip;
        if (t->num_reactions > 0) {
            if (run) oracle_ssa(state, m->W, m->H, t, seed, deltat, nthreads);
            seed = (int32_t)((uint32_t)seed + 1u);
        }
        if (outputEvent) {
            if (simTime >= outputLast + interval) {
                write(mindt);
                outputLast += interval;
            }
            simTime = mindt;
            continue;
        }
        for (int i = 0; i < S; ++i) {
            if (!(i == minspec || next[i] <= mindt)) continue;
            if (!(next[i] > 0)) continue;
            if (!m->inhomogeneous) {
                if (run)
                    oracle_hom_diffuse(state, m->W, m->H, i, seed, stay, m->boundary_mask, m->source_mask,
                                       nthreads);
                nh[i] += dth[i];
            } else {
                if (nx[i] <= mindt) {
                    if (run)
                        sweep_impl(state, m->W, m->H, i, 0, seed, dtx[i], m->Dx, m->rx, h, m->boundary_mask,
                                   nthreads, m->inhomogeneous == 2);
                    seed = (int32_t)((uint32_t)seed + 1u);
                    nx[i] += dtx[i];
                }
                if (ny[i] <= mindt) {
                    if (run)
                        sweep_impl(state, m->W, m->H, i, 1, seed, dty[i], m->Dy, m->ry, hy, m->boundary_mask,
                                   nthreads, m->inhomogeneous == 2);
                    seed = (int32_t)((uint32_t)seed + 1u);
                    ny[i] += dty[i];
                }
            }
        }
        next_times(next);
        simTime = mindt;
    }
    if (iterations) *iterations = n;
    if (sim_time_out) *sim_time_out = simTime;
    return writes;
}

extern "C" int oracle_run_output(const oracle_model *m, const oracle_reaction_table *t, int32_t *state,
                                 int steps, uint32_t seed0, int nthreads, float interval, float max_time,
                                 int32_t *dumps, float *dump_times, int max_dumps, int64_t *iterations,
                                 float *sim_time_out) {
    return run_loop(m, t, state, 0, steps, seed0, nthreads, interval, max_time, dumps, dump_times, max_dumps,
                    iterations, sim_time_out);
}

extern "C" int oracle_run(const oracle_model *m, const oracle_reaction_table *t, int32_t *state,
                          int steps, uint32_t seed0, int nthreads, float *sim_time_out) {
    run_loop(m, t, state, 0, steps, seed0, nthreads, INFINITY, INFINITY, nullptr, nullptr, 0, nullptr, sim_time_out);
    return 0;
}

// The loop from iteration `skip` on: state is the state after `skip` iterations.
extern "C" int oracle_run_from(const oracle_model *m, const oracle_reaction_table *t, int32_t *state, int skip,
                               int steps, uint32_t seed0, int nthreads, float *sim_time_out) {
    run_loop(m, t, state, skip, steps, seed0, nthreads, INFINITY, INFINITY, nullptr, nullptr, 0, nullptr,
             sim_time_out);
    return 0;
}
