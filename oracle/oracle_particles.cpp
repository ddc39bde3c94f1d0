// CPU oracle: individual-particle sort, cell offsets and the particle step.
// TEST INFRASTRUCTURE ONLY (see gpgmp_oracle.h).  Compiled with -ffp-contract=off.
#include "gpgmp_oracle.h"

#include <algorithm>
#include <cmath>
#include <numeric>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

// Stable sort of (key, value) pairs by the 32-bit key, the order the reference's
// LSD radix sort produces for KEY_INDEX_PAIR(key, id) sorted on the low 32 bits
// (IndividualSolver.cl:20-30, IndividualSolver.cpp:673-808); padding/dead keys
// 0xffffffff sort last.  Pinned by: std::stable_sort is the definition of stability.
extern "C" void oracle_stable_sort_pairs(uint32_t *keys, uint32_t *vals, int n) {
    std::vector<int> idx(n);
    std::iota(idx.begin(), idx.end(), 0);
    std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return keys[a] < keys[b]; });
    std::vector<uint32_t> k2(n), v2(n);
    for (int i = 0; i < n; ++i) {
        k2[i] = keys[idx[i]];
        v2[i] = vals[idx[i]];
    }
    std::copy(k2.begin(), k2.end(), keys);
    std::copy(v2.begin(), v2.end(), vals);
}

// CSR offsets off[c] = #(keys < c), c = 0..ncells (histogram + exclusive scan).
// Keys >= ncells (dead) are not counted.  Pinned by: std::lower_bound per cell.
extern "C" void oracle_cell_offsets(const uint32_t *k, int n, int ncells, uint32_t *off) {
    std::vector<uint32_t> hist((size_t)ncells + 1, 0);
    for (int i = 0; i < n; ++i)
        if (k[i] < (uint32_t)ncells) hist[k[i]]++;
    uint32_t run = 0;
    for (int c = 0; c <= ncells; ++c) {
        off[c] = run;
        if (c < ncells) run += hist[c];
    }
}

// computeCellIndices (IndividualSolver.cl:564-619): left[c] = first index of cell c,
// right[c] = last index (inclusive); an empty cell gets left = right = the index of
// the next particle (0 before the first key, n after the last).
// Pinned by: brute-force scan of the sorted keys in tests.
extern "C" void oracle_cell_left_right(const uint32_t *k, int n, int ncells, uint32_t *left,
                                       uint32_t *right) {
    std::vector<uint32_t> off((size_t)ncells + 1);
    oracle_cell_offsets(k, n, ncells, off.data());
    for (int c = 0; c < ncells; ++c) {
        left[c] = off[c];
        right[c] = (off[c + 1] > off[c]) ? off[c + 1] - 1 : off[c];
    }
}

namespace {

// sin(pi t) in double: t reduced to [0,2), folded to r in [0,1/2], Taylor to x^21.
// Pinned by: libm sin(M_PI*t) within 1e-15 absolute.
double sinpi_d(double t) {
    double t2 = t - 2.0 * std::floor(t * 0.5);
    double sign = 1.0;
    if (t2 >= 1.0) {
        sign = -1.0;
        t2 = t2 - 1.0;
    }
    double r = (t2 > 0.5) ? (1.0 - t2) : t2;
    const double x = 3.141592653589793116 * r;
    const double z = x * x;
    // 1/(2k+1)! for k = 10..1, Horner in z
    double p = 1.0 / 51090942171709440000.0;
    p = -1.0 / 121645100408832000.0 + z * p;
    p = 1.0 / 355687428096000.0 + z * p;
    p = -1.0 / 1307674368000.0 + z * p;
    p = 1.0 / 6227020800.0 + z * p;
    p = -1.0 / 39916800.0 + z * p;
    p = 1.0 / 362880.0 + z * p;
    p = -1.0 / 5040.0 + z * p;
    p = 1.0 / 120.0 + z * p;
    p = -1.0 / 6.0 + z * p;
    return sign * (x + x * (z * p));
}

// Gaussian pair by the Box-Muller transform on one Threefry call, key {id, seed}, counter
// {step, 0}: u1, u2 = (w + 1/2) 2^-32, g = sqrt(-2 ln u1) (cos 2 pi u2, sin 2 pi u2) with
// cos(2 pi u2) = sin(pi (2 u2 + 1/2)).  Pinned by: mean 0, variance 1 and independence of the
// pair (test_oracle_particles), libm for log and sin.
void gauss2(uint32_t id, uint32_t seed, uint32_t step, double *gx, double *gy) {
    const double inv32 = 2.32830643653869628906e-010;
    uint32_t w0, w1;
    oracle_threefry2x32_20(step, 0u, id, seed, &w0, &w1);
    const double u1 = ((double)w0 + 0.5) * inv32;
    const double u2 = ((double)w1 + 0.5) * inv32;
    const double r = std::sqrt(-2.0 * oracle_logd(u1));
    *gx = r * sinpi_d(2.0 * u2 + 0.5);
    *gy = r * sinpi_d(2.0 * u2);
}

}  // namespace

extern "C" void oracle_gauss2(uint32_t id, uint32_t seed, uint32_t step, double *gx, double *gy) {
    gauss2(id, seed, step, gx, gy);
}

namespace {

inline float wrap_pos(double xd, double L) {
    xd = xd - L * std::floor(xd / L);
    float xf = (float)xd;
    if (xf >= (float)L) xf = xf - (float)L;
    return xf;
}

int nthr(int n) {
#ifdef _OPENMP
    return n > 0 ? n : omp_get_max_threads();
#else
    (void)n;
    return 1;
#endif
}

}  // namespace

// One step of the individual-particle workload (no reference counterpart; nearest
// analogues: individual_diffuseX/Y IndividualSolver.cl:750-1014 for the move and
// maintainPopulation :1017-1882 for removal):
//  1. move: x += v_x dt + sqrt(2 D dt) g_x, v = A(-sin(2 pi y/L), sin(2 pi x/L)),
//     periodic on [0,L), key = floor(y)*side + floor(x);
//  2. stable sort by key (payload: x, y, id);
//  3. offsets off[c] = #(key < c);
//  4. A+A->0: in every cell the disjoint consecutive pairs (off[c]+2j, off[c]+2j+1)
//     are removed when u01(threefry({step, j}, {c, ~seed})) < p;
//  5. order-preserving compaction; off recomputed for the survivors.
// Returns the new particle count.  Pinned by: per-step invariants (sortedness,
// stability, offsets vs lower_bound, conservation without annihilation, displacement
// variance 2 D dt per axis).
extern "C" int oracle_particle_step(const oracle_particle_params *p, int step, int n, float *x,
                                    float *y, uint32_t *id, uint32_t *keys, uint32_t *off, int nt) {
    const int side = p->side;
    const double L = (double)side;
    const int A = side * side;
    const double sigma = std::sqrt(2.0 * (double)p->D * (double)p->dt);
    const double dt = p->dt, amp = p->vortex_amp;
    std::vector<uint32_t> k(n);
#pragma omp parallel for schedule(static) num_threads(nthr(nt))
    for (int i = 0; i < n; ++i) {
        double gx, gy;
        gauss2(id[i], p->seed, (uint32_t)step, &gx, &gy);
        const double xd = x[i], yd = y[i];
        const double vx = -amp * sinpi_d(2.0 * yd / L);
        const double vy = amp * sinpi_d(2.0 * xd / L);
        const float xn = wrap_pos(xd + vx * dt + sigma * gx, L);
        const float yn = wrap_pos(yd + vy * dt + sigma * gy, L);
        x[i] = xn;
        y[i] = yn;
        k[i] = (uint32_t)((int)yn * side + (int)xn);
    }
    std::vector<uint32_t> perm(n);
    std::iota(perm.begin(), perm.end(), 0u);
    std::stable_sort(perm.begin(), perm.end(), [&](uint32_t a, uint32_t b) { return k[a] < k[b]; });
    std::vector<float> xs(n), ys(n);
    std::vector<uint32_t> ids(n), ks(n);
    for (int i = 0; i < n; ++i) {
        xs[i] = x[perm[i]];
        ys[i] = y[perm[i]];
        ids[i] = id[perm[i]];
        ks[i] = k[perm[i]];
    }
    std::vector<uint32_t> of((size_t)A + 1);
    oracle_cell_offsets(ks.data(), n, A, of.data());
    std::vector<uint8_t> keep(n, 1);
#pragma omp parallel for schedule(static) num_threads(nthr(nt))
    for (int c = 0; c < A; ++c) {
        const uint32_t b = of[c], e = of[c + 1];
        const uint32_t pairs = (e - b) / 2;
        for (uint32_t j = 0; j < pairs; ++j) {
            uint32_t w0, w1;
            oracle_threefry2x32_20((uint32_t)step, j, (uint32_t)c, ~p->seed, &w0, &w1);
            if (oracle_u01(w0) < p->p_annihilate) {
                keep[b + 2 * j] = 0;
                keep[b + 2 * j + 1] = 0;
            }
        }
    }
    int m = 0;
    for (int i = 0; i < n; ++i) {
        if (!keep[i]) continue;
        x[m] = xs[i];
        y[m] = ys[i];
        id[m] = ids[i];
        keys[m] = ks[i];
        ++m;
    }
    oracle_cell_offsets(keys, m, A, off);
    return m;
}
