// CPU oracle: individual-particle sort, cell offsets and the particle step.
// TEST INFRASTRUCTURE ONLY (see gpgmp_oracle.h).  Compiled with -ffp-contract=off.
#include "gpgmp_oracle.h"

#include <algorithm>
#include <cmath>
#include <cstring>
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

// log(u) in float for a normal u > 0: u = 2^e m, m in (sqrt(1/2), sqrt(2)], f = m - 1,
// s = f / (2 + f), log(m) = 2 atanh(s) to s^9, Horner in s^2 with std::fma.
// Pinned by: libm log within 3e-7 relative over the Box-Muller domain (test_gauss2_is_box_muller).
float log_pf(float u) {
    uint32_t b;
    std::memcpy(&b, &u, 4);
    int e = (int)((b >> 23) & 0xffu) - 127;
    b = (b & 0x007fffffu) | 0x3f800000u;
    float m;
    std::memcpy(&m, &b, 4);
    if (m > 1.41421356f) {
        m = m * 0.5f;
        e = e + 1;
    }
    const float f = m - 1.0f;
    const float s = f / (2.0f + f);
    const float z = s * s;
    float p = 1.0f / 9.0f;
    p = std::fma(z, p, 1.0f / 7.0f);
    p = std::fma(z, p, 1.0f / 5.0f);
    p = std::fma(z, p, 1.0f / 3.0f);
    const float two_s = 2.0f * s;
    return std::fma((float)e, 0.693147182f, std::fma(two_s, z * p, two_s));
}

// sin(pi t) in float: t reduced to [0,2), folded to r in [0,1/2], x = pi r, Taylor to x^13 in
// Horner form with std::fma (float overloads, correctly rounded as the GPU's).
// Pinned by: libm sin(M_PI*t) within 1e-6 (test_gauss2_is_box_muller, both phases).
float sinpi_pf(float t) {
    float t2 = t - 2.0f * std::floor(t * 0.5f);
    float sign = 1.0f;
    if (t2 >= 1.0f) {
        sign = -1.0f;
        t2 = t2 - 1.0f;
    }
    const float r = (t2 > 0.5f) ? (1.0f - t2) : t2;
    const float x = 3.14159274f * r;
    const float z = x * x;
    float p = 1.0f / 6227020800.0f;
    p = std::fma(z, p, -1.0f / 39916800.0f);
    p = std::fma(z, p, 1.0f / 362880.0f);
    p = std::fma(z, p, -1.0f / 5040.0f);
    p = std::fma(z, p, 1.0f / 120.0f);
    p = std::fma(z, p, -1.0f / 6.0f);
    return sign * std::fma(x, z * p, x);
}

// Gaussian pair by the Box-Muller transform on one Threefry call, key {id, seed}, counter
// {step, 0}, in float: u1, u2 = (w / 2^9 + 1/2) 2^-23, g = sqrt(-2 ln u1) (cos 2 pi u2,
// sin 2 pi u2) with cos(2 pi u2) = sin(pi (2 u2 + 1/2)).  Pinned by: libm on the same words,
// mean 0, variance 1 and independence of the pair (test_oracle_particles).
void gauss2(uint32_t id, uint32_t seed, uint32_t step, float *gx, float *gy) {
    uint32_t w0, w1;
    oracle_threefry2x32_20(step, 0u, id, seed, &w0, &w1);
    const float u1 = ((float)(w0 >> 9) + 0.5f) * 0x1p-23f;
    const float u2 = ((float)(w1 >> 9) + 0.5f) * 0x1p-23f;
    const float r = std::sqrt(-2.0f * log_pf(u1));
    *gx = r * sinpi_pf(2.0f * u2 + 0.5f);
    *gy = r * sinpi_pf(2.0f * u2);
}

}  // namespace

extern "C" void oracle_gauss2(uint32_t id, uint32_t seed, uint32_t step, float *gx, float *gy) {
    gauss2(id, seed, step, gx, gy);
}

namespace {

// v folded into [0, L): v - L floor(v / L) in float, then one fold each way for the rounding.
inline float wrap_pos(float v, float L) {
    const float q = std::floor(v / L);
    v = v - L * q;
    if (v < 0.0f) v = v + L;
    if (v >= L) v = v - L;
    return v;
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
//  1. move, in float: x += v_x dt + sqrt(2 D dt) g_x, v = A(-sin(2 pi y/L), sin(2 pi x/L)),
//     periodic on [0,L), key = floor(y)*side + floor(x);
//  2. stable sort by key (payload: x, y, id);
//  3. offsets off[c] = #(key < c);
//  4. A+A->0: in every cell the disjoint consecutive pairs (off[c]+2j, off[c]+2j+1)
//     are removed when u01(w) < p, w = word (j mod 2) of threefry({step, j / 2}, {c, ~seed})
//     (one Threefry block serves two pairs);
//  5. order-preserving compaction; off recomputed for the survivors.
// Returns the new particle count.  Pinned by: per-step invariants (sortedness,
// stability, offsets vs lower_bound, conservation without annihilation, displacement
// variance 2 D dt per axis).
extern "C" int oracle_particle_step(const oracle_particle_params *p, int step, int n, float *x,
                                    float *y, uint32_t *id, uint32_t *keys, uint32_t *off, int nt) {
    const int side = p->side;
    const float L = (float)side;
    const int A = side * side;
    const float sigma = std::sqrt(2.0f * p->D * p->dt);
    const float dt = p->dt, amp = p->vortex_amp;
    std::vector<uint32_t> k(n);
#pragma omp parallel for schedule(static) num_threads(nthr(nt))
    for (int i = 0; i < n; ++i) {
        float gx, gy;
        gauss2(id[i], p->seed, (uint32_t)step, &gx, &gy);
        const float xo = x[i], yo = y[i];
        const float vx = -amp * sinpi_pf((2.0f * yo) / L);
        const float vy = amp * sinpi_pf((2.0f * xo) / L);
        const float xn = wrap_pos(xo + vx * dt + sigma * gx, L);
        const float yn = wrap_pos(yo + vy * dt + sigma * gy, L);
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
            oracle_threefry2x32_20((uint32_t)step, j >> 1, (uint32_t)c, ~p->seed, &w0, &w1);
            if (oracle_u01((j & 1u) ? w1 : w0) < p->p_annihilate) {
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
