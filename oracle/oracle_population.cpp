// ORACLE (test infrastructure only): CPU restatement of the reference's population
// maintenance of an individual (particle-based) species after a reaction step, for the
// parity tests.  Never linked into the product library.
//
//   oracle_ind_maintain   IndividualSolver::maintainPopulation (IndividualSolver.cpp:458-671)
//                         with its kernels maintainPopulation (IndividualSolver.cl:1017-1131),
//                         segmentedScan (IndividualSolver.cpp:830-950), computeReductionOffsets
//                         (IndividualSolver.cl:1299-1334), scanIndividuals / scanIndividualsBlocks
//                         (:1344-1499), scatterIndividuals (:1501-1581), updateNewParticles
//                         (:1135-1293), writeLostParticles (:1584-1593) and copyTempToKeys
//                         (:1596-1599).
//
// The list is the reference's: keys[cap] (cell of each particle, UINT_MAX for dead ones)
// and vals[cap] (particle id); positions [0, n) hold the live particles sorted by cell,
// [n, cap) the pool of dead ids; props[4][cap] are indexed by particle id.
//
// What the reference's kernels compute, stated as the list they leave behind:
//  1. per cell c with is = its particle count and goal = the species' count after the
//     reactions: if is > goal, is - goal particles are killed, each picked as
//     left[c] + (uint)(u * is) with u = (int)w * 2^-32 + 0.5 in double and w the x then y
//     words of successive Threefry calls keyed {c, 0xdecafbad + seed}, counter {1, 2, ...,
//     0xf00dcafe}, redrawn while the pick is already dead (IndividualSolver.cl:1059-1086);
//  2. if is < goal, goal - is new particles are created in c;
//  3. the new list is, cell by cell, c's new particles followed by its survivors in their
//     old order; new particles take their ids from the front of the pool in cell order
//     (updateNewParticles' indexPool); the killed particles (key UINT_MAX, id kept) follow
//     the live list in their old list order (scatterIndividuals' lostParticles,
//     writeLostParticles), then the unused rest of the old pool;
//  4. a new particle's properties start as its recycled id's stored values and pass
//     through the model's newIndividuals method, here the rule `mode/a/b` per property
//     (0 keep, 1 the constant a, 2 a + b * RANDOM evaluated in double, stored as float)
//     with RANDOM the reference's macro: x then y words of Threefry calls keyed
//     {c, 0xdecafbad + seed + 1}, one stream per cell over all its new particles
//     (IndividualSolver.cl:1147-1158, 1262-1282).
//
// Deviations, documented:
//  * live particles whose key lies outside the grid (absorbing boundaries) are carried
//    after the last cell in their order; the reference leaves their reduction offsets
//    unset (IndividualSolver.cl:1040-1049 writes them only for particles inside a cell);
//  * a pick with u * is rounded up to is (possible only for is > 2^21 particles in one
//    cell) is clamped to is - 1; the reference would read the next cell's particle;
//  * the reference drops the new particles of trailing empty cells when the last listed
//    particle is killed (scatterIndividuals tests the killed key, IndividualSolver.cl:1569);
//    here they are created like any others;
//  * OpenCL may contract a + b * RANDOM into an fma; the contract here is no contraction
//    (parity unpinned for that rounding).
//
// Pinning: a literal restatement of the reference's multi-kernel procedure (flags,
// segmented scan as python/gpgmp/helpers.py:26-94 checks it, reduction offsets, inclusive
// scans, scatter, updateNewParticles' index arithmetic) in the tests on random small lists;
// the reference's own consistency check (IndividualSolver.cpp:1320-1387: ids unique,
// keys sorted, per-cell count == goal); uniformity of the kill picks over a cell's
// particles (chi-square) and of RANDOM.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "gpgmp_oracle.h"

namespace {

constexpr uint32_t kKeyBase = 0xdecafbadu;
constexpr uint32_t kCtrHi = 0xf00dcafeu;
constexpr uint32_t kDead = 0xffffffffu;

// The reference's uniform in double: (int)w * M_RAN_INVM32 + 0.5, M_RAN_INVM32 = 2^-32
// (Solver.cpp:471), exact in double.
double u01_double(uint32_t w) { return (double)(int32_t)w * 0x1p-32 + 0.5; }

// x, y, x, y, ... of successive Threefry calls with counter 1, 2, ... (the RNG pattern of
// maintainPopulation's kill loop and of the RANDOM macro).
struct PairStream {
    uint32_t k0, k1, ri = 0, ctr = 0, w0 = 0, w1 = 0;
    PairStream(uint32_t a, uint32_t b) : k0(a), k1(b) {}
    uint32_t next() {
        uint32_t w;
        if (!(ri % 2)) {
            ++ctr;
            oracle_threefry2x32_20(ctr, kCtrHi, k0, k1, &w0, &w1);
            w = w0;
        } else {
            w = w1;
        }
        ++ri;
        return w;
    }
};

}  // namespace

extern "C" int oracle_ind_maintain(const oracle_ind_params *p, uint32_t *keys, uint32_t *vals, int n, int cap,
                                   float *props, const int32_t *goal, int32_t seed, const oracle_new_rule *rule,
                                   int *n_out) {
    const int A = p->W * p->H;
    for (int c = 0; c < A; ++c)
        if (goal[c] < 0) return -2;
    std::vector<uint32_t> off((size_t)A + 1);
    oracle_cell_offsets(keys, n, A, off.data());
    std::vector<uint32_t> nnew(A, 0);
    uint64_t total_new = 0;
    for (int c = 0; c < A; ++c) {
        const uint32_t is = off[c + 1] - off[c], g = (uint32_t)goal[c];
        if (g > is) total_new += g - is;
    }
    if (total_new > (uint64_t)(cap - n)) return -1;  // "Required number of individuals ... greater than maximum"

    std::vector<uint8_t> killed((size_t)n, 0);
    const uint32_t k1 = kKeyBase + (uint32_t)seed;
    for (int c = 0; c < A; ++c) {
        const uint32_t left = off[c], is = off[c + 1] - off[c], g = (uint32_t)goal[c];
        if (is > g) {
            PairStream rs((uint32_t)c, k1);
            for (uint32_t k = 0; k < is - g; ++k) {
                uint32_t kp;
                do {
                    uint32_t r = (uint32_t)(u01_double(rs.next()) * (double)is);
                    if (r >= is) r = is - 1;
                    kp = left + r;
                } while (killed[kp]);
                killed[kp] = 1;
            }
        } else {
            nnew[c] = g - is;
        }
    }

    std::vector<uint32_t> nk((size_t)cap), nv((size_t)cap);
    size_t w = 0, pool = (size_t)n;
    const uint32_t k1new = kKeyBase + (uint32_t)seed + 1u;
    for (int c = 0; c < A; ++c) {
        if (nnew[c]) {
            PairStream rs((uint32_t)c, k1new);
            for (uint32_t j = 0; j < nnew[c]; ++j) {
                const uint32_t id = vals[pool++];
                if (rule) {
                    for (int q = 0; q < 4; ++q) {
                        float &v = props[id + (size_t)q * cap];
                        if (rule->mode[q] == 1) v = (float)rule->a[q];
                        if (rule->mode[q] == 2) v = (float)(rule->a[q] + rule->b[q] * u01_double(rs.next()));
                    }
                }
                nk[w] = (uint32_t)c;
                nv[w] = id;
                ++w;
            }
        }
        for (uint32_t i = off[c]; i < off[c + 1]; ++i)
            if (!killed[i]) {
                nk[w] = keys[i];
                nv[w] = vals[i];
                ++w;
            }
    }
    for (uint32_t i = off[A]; i < (uint32_t)n; ++i) {  // live keys outside the grid
        nk[w] = keys[i];
        nv[w] = vals[i];
        ++w;
    }
    const int n_new = (int)w;
    for (int i = 0; i < n; ++i)
        if (killed[i]) {
            nk[w] = kDead;
            nv[w] = vals[i];
            ++w;
        }
    for (; pool < (size_t)cap; ++pool) {
        nk[w] = kDead;
        nv[w] = vals[pool];
        ++w;
    }
    std::memcpy(keys, nk.data(), (size_t)cap * 4);
    std::memcpy(vals, nv.data(), (size_t)cap * 4);
    if (n_out) *n_out = n_new;
    return 0;
}
