// Stochastic GMP kernels for gfx950: Gillespie SSA and stochastic diffusion
// splitting (homogeneous, and directionally split inhomogeneous drift-diffusion).
//
// One kernel, k_gmp_step, does the work of a host-loop launch for any number of
// realizations (grids).  Per cell it
//   1. gathers the outflow the neighbours computed in the previous launch
//      (homogeneous_stochastic_updateState, StochasticHomogeneousDiffusion.cl:105-124;
//      inhomogeneous_updateStateX/Y, StochasticInhomogeneousDriftDiffusion.cl:687-723),
//      plus homogeneous source-boundary inflow (:128-183);
//   2. runs the Gillespie direct method over [0, deltat) (performGillespie,
//      Gillespie.cl:4-106) with the species counts held in VGPRs;
//   3. computes this launch's outflow for the due species
//      (homogeneous_stochastic_diffuse :8-103; inhomogeneous_diffuseX/Y :426-685),
//      subtracts it and writes it to the outflow buffer.
// Deferring step 3's gather into the next launch removes the reference's separate
// updateState launches: a host-loop iteration is one launch (homogeneous) or two
// (inhomogeneous: X sweeps, then Y sweeps).  RNG keys never contain a batch or
// tile offset, so results do not depend on the launch shape or batch layout.
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <algorithm>

#include "gpgmp_numeric.h"
#include "stochastic_types.h"

namespace gpgmp {

enum : int { MODE_HOM = 0, MODE_INH_X = 1, MODE_INH_Y = 2 };

template <int S>
__device__ __forceinline__ int32_t sel(const int32_t (&x)[S], int j) {
    int32_t v = x[0];
#pragma unroll
    for (int k = 1; k < S; ++k) v = (j == k) ? x[k] : v;
    return v;
}

// _h[i+1] = _h[0] + <law_i>; int literal -> int product then float add,
// double literal -> double product chain rounded once (Solver.cpp:849-857).
template <int S>
__device__ __forceinline__ float add_law(const RxTable &T, int i, float h0, const int32_t (&x)[S]) {
    const int nf = T.nf[i];
    if ((T.lit_is_int >> i) & 1u) {
        uint32_t v = (uint32_t)T.lit_int[i];
        for (int f = 0; f < nf; ++f) v *= (uint32_t)(sel(x, T.fs[i][f]) - (int32_t)T.fo[i][f]);
        return h0 + (float)(int32_t)v;
    }
    double v = T.lit_dbl[i];
    for (int f = 0; f < nf; ++f) v = v * (double)(sel(x, T.fs[i][f]) - (int32_t)T.fo[i][f]);
    return (float)((double)h0 + v);
}

// Cumulative hazards _h[1..R] and the total _h[0] (Solver.cpp:843-857).
template <int S>
__device__ __forceinline__ void hazards(const RxTable &T, const int32_t (&x)[S], float (&h)[kMaxReactions + 1]) {
    h[0] = 0.f;
#pragma unroll
    for (int i = 0; i < kMaxReactions; ++i) {
        if (i < T.R) {
            h[i + 1] = add_law(T, i, h[0], x);
            h[0] = h[i + 1];
        }
    }
}

// Reaction selection and stoichiometric update (Solver.cpp:867-903): reaction 0 fires
// when r2 < h[1], reaction i when h[i] <= r2 < h[i+1]; nothing fires when rounding
// puts r2 at or above h[R].
template <int S>
__device__ __forceinline__ void fire(const RxTable &T, const float (&h)[kMaxReactions + 1], float r2,
                                     int32_t (&x)[S]) {
#pragma unroll
    for (int i = 0; i < kMaxReactions; ++i) {
        if (i < T.R) {
            const bool f = (i == 0) ? (r2 < h[1]) : ((r2 >= h[i]) && (r2 < h[i + 1]));
#pragma unroll
            for (int j = 0; j < S; ++j) x[j] += f ? (int32_t)T.net[i][j] : 0;
        }
    }
}

// performGillespie for one cell (Gillespie.cl:41-99): key {cell, 0xdecafbad+seed},
// counter {++n, 0xf00dcafe}; r1 -> waiting time, r2 (times h0 in double) -> reaction.
// Runs from (time, ctr) until deltat, or until `budget` events have been drawn: returns
// false when it stops early (time < deltat), so the cell can be resumed bit-exactly.
template <int S>
__device__ __forceinline__ bool ssa_cell(int32_t (&x)[S], uint32_t cell, int32_t seed, float deltat,
                                         const RxTable &T, uint32_t budget, float &time, uint32_t &ctr) {
    const uint32_t k1 = kKeyBase + (uint32_t)seed;
    while (time < deltat) {
        float h[kMaxReactions + 1];
        hazards<S>(T, x, h);
        if (h[0] == 0.f) return true;
        if (ctr == kMaxEventsPerCall) return true;  // runaway guard (never reached by sane models)
        if (ctr >= budget) return false;
        ++ctr;
        const u32x2 w = threefry2x32_20(ctr, kCtrHi, cell, k1);
        const float r1 = u01(w.x);
        const float r2 = (float)(u01_d(w.y) * (double)h[0]);
        const float dt = log_f(1.f / r1) / h[0];
        time += dt;
        if (time < deltat) fire<S>(T, h, r2, x);
    }
    return true;
}

// Per-particle categorical hop of homogeneous_stochastic_diffuse
// (StochasticHomogeneousDiffusion.cl:68-92): counts of u in (stay,p1], (p1,p2],
// (p2,p3], (p3,p4] are the hops to -x, +x, -y, +y.
__device__ __forceinline__ void hom_outflow(int n, uint32_t k0, uint32_t k1, const HomProbs &P,
                                            int32_t (&cnt)[4]) {
    int32_t le0 = 0, le1 = 0, le2 = 0, le3 = 0, le4 = 0;
    uint32_t ctr = 0;
    for (int i = 0; i < n; i += 2) {
        ++ctr;
        const u32x2 w = threefry2x32_20(ctr, kCtrHi, k0, k1);
        float u = u01(w.x);
        le0 += u <= P.stay;
        le1 += u <= P.p1;
        le2 += u <= P.p2;
        le3 += u <= P.p3;
        le4 += u <= P.p4;
        if (i + 1 < n) {
            u = u01(w.y);
            le0 += u <= P.stay;
            le1 += u <= P.p1;
            le2 += u <= P.p2;
            le3 += u <= P.p3;
            le4 += u <= P.p4;
        }
    }
    cnt[0] = le1 - le0;
    cnt[1] = le2 - le1;
    cnt[2] = le3 - le2;
    cnt[3] = le4 - le3;
}

// Per-particle Bernoulli hop of inhomogeneous_diffuseX/Y
// (StochasticInhomogeneousDriftDiffusion.cl:524-547): dir = +[u <= p1], counted
// when u <= p2.  pr = (min(p1,p2), p2): plus = #(u <= pr.x), minus = #(u <= pr.y) - plus.
__device__ __forceinline__ void inh_outflow(int n, uint32_t k0, uint32_t k1, float2 pr,
                                            int32_t &minus, int32_t &plus) {
    int32_t a = 0, b = 0;
    uint32_t ctr = 0;
    for (int i = 0; i < n; i += 2) {
        ++ctr;
        const u32x2 w = threefry2x32_20(ctr, kCtrHi, k0, k1);
        float u = u01(w.x);
        a += u <= pr.x;
        b += u <= pr.y;
        if (i + 1 < n) {
            u = u01(w.y);
            a += u <= pr.x;
            b += u <= pr.y;
        }
    }
    plus = a;
    minus = b - a;
}

// homogeneous_stochastic_source_boundary (StochasticHomogeneousDiffusion.cl:128-183)
// for one boundary position d: `count` draws with key {d, 0xdecafbad+seed}, each adding
// a particle when stay < u <= p1.
__device__ __forceinline__ int32_t source_inflow(int count, uint32_t d, uint32_t k1,
                                                 const HomProbs &P) {
    int32_t add = 0;
    uint32_t ctr = 0;
    for (int i = 0; i < count; i += 2) {
        ++ctr;
        const u32x2 w = threefry2x32_20(ctr, kCtrHi, d, k1);
        float u = u01(w.x);
        add += (u > P.stay) && (u <= P.p1);
        if (i + 1 < count) {
            u = u01(w.y);
            add += (u > P.stay) && (u <= P.p1);
        }
    }
    return add;
}

// Block -> tile with XCD-aware placement: hardware dispatches block b to XCD b % 8,
// so XCD k gets a contiguous run of tiles and neighbour gathers stay in its L2.
__device__ __forceinline__ int xcd_tile(int b, int ntiles) {
    const int per = (ntiles + 7) >> 3;
    const int t = (b & 7) * per + (b >> 3);
    return t;
}

template <int S, int MODE>
__global__ __launch_bounds__(256) void k_gmp_step(
    int32_t *__restrict__ state, const int32_t *__restrict__ lost_in, int32_t *__restrict__ lost_out,
    const float2 *__restrict__ prob, const RealInfo *__restrict__ reals,
    const Tile *__restrict__ tiles, int ntiles, const LaunchPlan *__restrict__ plan,
    const int32_t *__restrict__ seeds_cur, const int32_t *__restrict__ seeds_prev, RxTable T,
    HomProbs P, int4 bm, int cpt, HotList hot) {
    // the grid has 8*ceil(ntiles/8) blocks so that the remap is a bijection
    const int t = xcd_tile((int)blockIdx.x, ntiles);
    if (t >= ntiles) return;
    const Tile tile = tiles[t];
    const RealInfo ri = reals[tile.r];
    const LaunchPlan pl = *plan;
    const int32_t seed = seeds_cur[tile.r];
    const int W = ri.W, H = ri.H;
    const int A = W * H;
    constexpr int ND = (MODE == MODE_HOM) ? 4 : 2;
    int32_t *st = state + ri.state_off;
    const int32_t *lin = lost_in + ri.lost_off;
    int32_t *lout = lost_out + ri.lost_off;
    const bool any_src = (ri.src[0] | ri.src[1] | ri.src[2] | ri.src[3]) != 0;

    for (int k = 0; k < cpt; ++k) {
        const int c = tile.cell0 + k * (int)blockDim.x + (int)threadIdx.x;
        if (c >= A) break;
        const int gx = c % W, gy = c / W;
        int32_t x[S];
        // loads: every species for launches with an SSA, only touched species otherwise
        const uint32_t touched = pl.gather_mask | pl.out_mask;
#pragma unroll
        for (int s = 0; s < S; ++s)
            x[s] = (pl.deltat > 0.f || ((touched >> s) & 1u)) ? st[(size_t)s * A + c] : 0;

        // 1. gather the previous launch's outflow
        if (pl.gather_mask) {
#pragma unroll
            for (int s = 0; s < S; ++s) {
                if (!((pl.gather_mask >> s) & 1u)) continue;
                const int32_t *l = lin + (size_t)s * ND * A;
                int32_t v = x[s];
                if (MODE == MODE_HOM) {
                    if (gx > 0) v += l[1 * (size_t)A + c - 1];
                    if (gx < W - 1) v += l[0 * (size_t)A + c + 1];
                    if (gy > 0) v += l[3 * (size_t)A + c - W];
                    if (gy < H - 1) v += l[2 * (size_t)A + c + W];
                } else if (MODE == MODE_INH_X) {  // gathers the Y outflow of the previous launch
                    if (gy > 0) v += l[1 * (size_t)A + c - W];
                    if (gy < H - 1) v += l[0 * (size_t)A + c + W];
                } else {                          // gathers the X outflow of the previous launch
                    if (gx > 0) v += l[1 * (size_t)A + c - 1];
                    if (gx < W - 1) v += l[0 * (size_t)A + c + 1];
                }
                if (MODE == MODE_HOM && any_src) {
                    const uint32_t k1s = kKeyBase + (uint32_t)seeds_prev[tile.r] + (uint32_t)pl.gather_off;
                    if (ri.src[0] > 0 && gx == 0) v += source_inflow(ri.src[0], (uint32_t)gy, k1s, P);
                    if (ri.src[1] > 0 && gx == W - 1) v += source_inflow(ri.src[1], (uint32_t)gy, k1s, P);
                    if (ri.src[2] > 0 && gy == 0) v += source_inflow(ri.src[2], (uint32_t)gx, k1s, P);
                    if (ri.src[3] > 0 && gy == H - 1) v += source_inflow(ri.src[3], (uint32_t)gx, k1s, P);
                }
                x[s] = v;
            }
        }

        // 2. reactions.  A lane stops after hot.budget events and hands its cell to the
        //    hot-cell kernel, so a few busy cells do not hold a whole wave (and the GPU).
        bool deferred = false;
        if (MODE != MODE_INH_Y && pl.deltat > 0.f && T.R > 0) {
            float time = 0.f;
            uint32_t ctr = 0;
            const bool fin = ssa_cell<S>(x, (uint32_t)c, seed, pl.deltat, T, hot.budget, time, ctr);
            const uint64_t dm = __ballot(!fin);
            if (dm) {
                const int lane = (int)(threadIdx.x & 63);
                const int leader = __ffsll((unsigned long long)dm) - 1;
                uint32_t base = 0;
                if (lane == leader) base = atomicAdd(hot.count, (uint32_t)__popcll(dm));
                base = __shfl(base, leader, 64);
                if (!fin) {
                    const uint32_t idx = base + (uint32_t)__popcll(dm & ((1ull << lane) - 1ull));
                    if (idx < hot.cap) {
                        hot.e[idx] = HotEntry{tile.r, c, time, ctr};
                        deferred = true;
                    } else {  // list full: finish the cell here
                        ssa_cell<S>(x, (uint32_t)c, seed, pl.deltat, T, 0xffffffffu, time, ctr);
                    }
                }
            }
        }

        // 3. outflow of the due species (a deferred cell's comes from the hot-cell kernel)
        if (pl.out_mask && !deferred) {
            int m0 = (gx == 0) ? bm.x : 1;
            int m1 = (gx == W - 1) ? bm.y : 1;
            int m2 = (gy == 0) ? bm.z : 1;
            int m3 = (gy == H - 1) ? bm.w : 1;
#pragma unroll
            for (int s = 0; s < S; ++s) {
                if (!((pl.out_mask >> s) & 1u)) continue;
                const uint32_t k1 = kKeyBase + (uint32_t)seed + (uint32_t)pl.out_off[s];
                int32_t *l = lout + (size_t)s * ND * A;
                if (MODE == MODE_HOM) {
                    int32_t cnt[4];
                    hom_outflow(x[s], (uint32_t)c, k1, P, cnt);
                    const int32_t l0 = m0 * cnt[0], l1 = m1 * cnt[1], l2 = m2 * cnt[2], l3 = m3 * cnt[3];
                    l[0 * (size_t)A + c] = l0;
                    l[1 * (size_t)A + c] = l1;
                    l[2 * (size_t)A + c] = l2;
                    l[3 * (size_t)A + c] = l3;
                    x[s] -= l0 + l1 + l2 + l3;
                } else {
                    const float2 pr = prob[ri.prob_off + (size_t)s * A + c];
                    int32_t mi, pl_;
                    inh_outflow(x[s], (uint32_t)(c + s * A), k1, pr, mi, pl_);
                    const int32_t mm = (MODE == MODE_INH_X) ? m0 : m2;
                    const int32_t mp = (MODE == MODE_INH_X) ? m1 : m3;
                    const int32_t lm = mm * mi, lp = mp * pl_;
                    l[0 * (size_t)A + c] = lm;
                    l[1 * (size_t)A + c] = lp;
                    x[s] -= lm + lp;
                }
            }
        }

#pragma unroll
        for (int s = 0; s < S; ++s)
            if (pl.deltat > 0.f || ((touched >> s) & 1u)) st[(size_t)s * A + c] = x[s];
    }
}

__device__ __forceinline__ float readlane_f(float v, int j) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), j));
}
__device__ __forceinline__ double readlane_d(double v, int j) {
    const uint64_t b = (uint64_t)__double_as_longlong(v);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)b, j);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(b >> 32), j);
    return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}
__device__ __forceinline__ int32_t wave_sum(int32_t v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Hot-cell kernel: one wave per deferred cell.  The SSA continues from (time, ctr);
// its random words and waiting-time logarithms do not depend on the state, so the 64
// lanes compute them for the next 64 events at once (lane l: counter ctr+1+l) and the
// serial chain per event keeps only the hazards, r2, the division by h0 and the
// update.  Then the cell's outflow for the due species is computed with the particle
// loop split over the lanes (call k by lane k % 64) and the counts summed -- the same
// integers as the serial loop.  The last block to finish resets the list.
template <int S, int MODE>
__global__ __launch_bounds__(64) void k_gmp_hot(int32_t *__restrict__ state, int32_t *__restrict__ lost_out,
                                                const float2 *__restrict__ prob, const RealInfo *__restrict__ reals,
                                                const LaunchPlan *__restrict__ plan,
                                                const int32_t *__restrict__ seeds_cur, RxTable T, HomProbs P,
                                                int4 bm, HotList hot) {
    const int lane = (int)threadIdx.x;
    const uint32_t total = hot.count[0];
    if (total == 0) return;
    const uint32_t nh = total < hot.cap ? total : hot.cap;
    const LaunchPlan pl = *plan;
    constexpr int ND = (MODE == MODE_HOM) ? 4 : 2;
    for (uint32_t i = blockIdx.x; i < nh; i += gridDim.x) {
        const HotEntry he = hot.e[i];
        const RealInfo ri = reals[he.r];
        const int32_t seed = seeds_cur[he.r];
        const int W = ri.W, H = ri.H, A = W * H, c = he.c;
        int32_t *st = state + ri.state_off;
        int32_t x[S];
#pragma unroll
        for (int s = 0; s < S; ++s) x[s] = __builtin_amdgcn_readfirstlane(st[(size_t)s * A + c]);
        float time = he.time;
        uint32_t ctr = he.ctr;
        const uint32_t k1 = kKeyBase + (uint32_t)seed;
        bool done = false;
        while (!done) {
            const u32x2 w = threefry2x32_20(ctr + 1u + (uint32_t)lane, kCtrHi, (uint32_t)c, k1);
            const float L = log_f(1.f / u01(w.x));
            const double U2 = u01_d(w.y);
            for (int j = 0; j < 64; ++j) {
                float h[kMaxReactions + 1];
                hazards<S>(T, x, h);
                if (h[0] == 0.f || ctr == kMaxEventsPerCall) {
                    done = true;
                    break;
                }
                ++ctr;
                const float r2 = (float)(readlane_d(U2, j) * (double)h[0]);
                const float dt = readlane_f(L, j) / h[0];
                time += dt;
                if (!(time < pl.deltat)) {
                    done = true;
                    break;
                }
                fire<S>(T, h, r2, x);
            }
        }
        if (pl.out_mask) {
            const int gx = c % W, gy = c / W;
            const int m0 = (gx == 0) ? bm.x : 1, m1 = (gx == W - 1) ? bm.y : 1;
            const int m2 = (gy == 0) ? bm.z : 1, m3 = (gy == H - 1) ? bm.w : 1;
            int32_t *lo = lost_out + ri.lost_off;
#pragma unroll
            for (int s = 0; s < S; ++s) {
                if (!((pl.out_mask >> s) & 1u)) continue;
                const uint32_t k1o = kKeyBase + (uint32_t)seed + (uint32_t)pl.out_off[s];
                const int n = x[s];
                const int calls = (n + 1) >> 1;
                int32_t *l = lo + (size_t)s * ND * A;
                if (MODE == MODE_HOM) {
                    int32_t le0 = 0, le1 = 0, le2 = 0, le3 = 0, le4 = 0;
                    for (int k = lane; k < calls; k += 64) {
                        const u32x2 ww = threefry2x32_20((uint32_t)k + 1u, kCtrHi, (uint32_t)c, k1o);
                        float u = u01(ww.x);
                        le0 += u <= P.stay; le1 += u <= P.p1; le2 += u <= P.p2; le3 += u <= P.p3; le4 += u <= P.p4;
                        if (2 * k + 1 < n) {
                            u = u01(ww.y);
                            le0 += u <= P.stay; le1 += u <= P.p1; le2 += u <= P.p2; le3 += u <= P.p3; le4 += u <= P.p4;
                        }
                    }
                    le0 = wave_sum(le0); le1 = wave_sum(le1); le2 = wave_sum(le2); le3 = wave_sum(le3);
                    le4 = wave_sum(le4);
                    const int32_t l0 = m0 * (le1 - le0), l1 = m1 * (le2 - le1), l2 = m2 * (le3 - le2),
                                  l3 = m3 * (le4 - le3);
                    if (lane == 0) {
                        l[0 * (size_t)A + c] = l0;
                        l[1 * (size_t)A + c] = l1;
                        l[2 * (size_t)A + c] = l2;
                        l[3 * (size_t)A + c] = l3;
                    }
                    x[s] -= l0 + l1 + l2 + l3;
                } else {
                    const float2 pr = prob[ri.prob_off + (size_t)s * A + c];
                    int32_t a = 0, b = 0;
                    for (int k = lane; k < calls; k += 64) {
                        const u32x2 ww = threefry2x32_20((uint32_t)k + 1u, kCtrHi, (uint32_t)(c + s * A), k1o);
                        float u = u01(ww.x);
                        a += u <= pr.x;
                        b += u <= pr.y;
                        if (2 * k + 1 < n) {
                            u = u01(ww.y);
                            a += u <= pr.x;
                            b += u <= pr.y;
                        }
                    }
                    a = wave_sum(a);
                    b = wave_sum(b);
                    const int32_t lm = m0 * (b - a), lp = m1 * a;
                    if (lane == 0) {
                        l[0 * (size_t)A + c] = lm;
                        l[1 * (size_t)A + c] = lp;
                    }
                    x[s] -= lm + lp;
                }
            }
        }
        if (lane == 0) {
#pragma unroll
            for (int s = 0; s < S; ++s) st[(size_t)s * A + c] = x[s];
        }
    }
    // every block read the count before adding itself here, so the last one may reset
    if (lane == 0) {
        __threadfence();
        const uint32_t fin = atomicAdd(&hot.count[1], 1u);
        if (fin == gridDim.x - 1) {
            atomicExch(&hot.count[0], 0u);
            atomicExch(&hot.count[1], 0u);
            __threadfence();
        }
    }
}

// Inhomogeneous hop probabilities (StochasticInhomogeneousDriftDiffusion.cl:518-520 /
// :648-650) with the reference's float/double promotions; stored as (min(p1,p2), p2).
// They depend only on the fields and the sweep's dt, so they are tabulated once.
__global__ void k_inh_probs(float2 *__restrict__ out, const float *__restrict__ D,
                            const float *__restrict__ r, float dt, float dx, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float Di = D[i], ri = r[i];
    const float rrdt = (ri * ri) * dt;
    const float dx2 = dx * dx;
    const float dxr = dx * ri;
    const float ps = (float)(1. - ((double)dt * (2. * (double)Di + (double)rrdt)) / (double)dx2);
    const double den = 2. * (double)Di + (double)rrdt;
    const float p1 = (float)((1. - (double)ps) * (1. + (double)dxr / den) / 2.);
    const float p2 = (float)((double)p1 + (1. - (double)ps) * (1. - (double)dxr / den) / 2.);
    out[i] = make_float2((p1 <= p2) ? p1 : p2, p2);
}

// Per-cell admissible time step of the stochastic inhomogeneous solver
// (computeIndividualTimestepX/Y, StochasticInhomogeneousDriftDiffusion.cl:130-151), in
// double, and its minimum over the cells: block minima, then one block over the
// minima (the reference reduces in a single work-group, which caps grids at 512^2,
// InhomogeneousSolver.cpp:63-85).  Min is order-independent, so the result is exact.
__device__ __forceinline__ float cell_dt(float Df, float rf, float hf) {
    const double D = Df, r = rf, h = hf;
    double ts;
    if (r == 0.0) {
        ts = h * h / (2.0 * D) * (1.0 - 2.0 / 3.0);
    } else if (D == 0.0) {
        ts = h / fabs(r);
    } else {
        const double eps = h * r / (2.0 * D);
        const double th = tanh(eps), sh = sinh(eps);
        double temp = 1.0 + (-1.0 / (th * eps) + 1.0 / (sh * sh));
        if (temp <= 0.0) temp = 1.0 / 3.0;
        ts = h / r * th * temp;
    }
    return (float)ts;
}

__device__ __forceinline__ float block_min(float v, float *smem) {
    for (int off = 32; off > 0; off >>= 1) v = fminf(v, __shfl_down(v, off, 64));
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) smem[wave] = v;
    __syncthreads();
    float m = smem[0];
    for (int w = 1; w < (int)(blockDim.x >> 6); ++w) m = fminf(m, smem[w]);
    return m;
}

__global__ __launch_bounds__(256) void k_min_dt(const float *__restrict__ D, const float *__restrict__ r, int n,
                                                float h, float *__restrict__ block_out) {
    __shared__ float sm[4];
    float v = __builtin_inff();
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        v = fminf(v, cell_dt(D[i], r[i], h));
    const float m = block_min(v, sm);
    if (threadIdx.x == 0) block_out[blockIdx.x] = m;
}

__global__ __launch_bounds__(256) void k_min_reduce(const float *__restrict__ in, int n, float *__restrict__ out) {
    __shared__ float sm[4];
    float v = __builtin_inff();
    for (int i = threadIdx.x; i < n; i += blockDim.x) v = fminf(v, in[i]);
    const float m = block_min(v, sm);
    if (threadIdx.x == 0) *out = m;
}

}  // namespace gpgmp

// ---------------------------------------------------------------------------------------
// Host-side launchers (called from the solver host code in this library).
#include "stochastic_launch.h"

namespace gpgmp {

template <int S, int MODE>
static hipError_t launch_s(const StepLaunch &a) {
    const int grid = 8 * ((a.ntiles + 7) / 8);
    if (a.ev_start) {
        // start/stop timestamps taken by the dispatch packet itself
        hipExtLaunchKernelGGL((k_gmp_step<S, MODE>), dim3(grid), dim3(a.block), 0, a.stream, a.ev_start,
                              a.ev_stop, 0, a.state, a.lost_in, a.lost_out, a.prob, a.reals, a.tiles,
                              a.ntiles, a.plan, a.seeds_cur, a.seeds_prev, *a.T, *a.P, a.bm, a.cpt, a.hot);
    } else {
        hipLaunchKernelGGL((k_gmp_step<S, MODE>), dim3(grid), dim3(a.block), 0, a.stream, a.state,
                           a.lost_in, a.lost_out, a.prob, a.reals, a.tiles, a.ntiles, a.plan, a.seeds_cur,
                           a.seeds_prev, *a.T, *a.P, a.bm, a.cpt, a.hot);
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || MODE == MODE_INH_Y) return e;
    // the cells that used up their event budget
    if (a.ev_hot_start) {
        hipExtLaunchKernelGGL((k_gmp_hot<S, MODE>), dim3(a.hot_grid), dim3(64), 0, a.stream, a.ev_hot_start,
                              a.ev_hot_stop, 0, a.state, a.lost_out, a.prob, a.reals, a.plan, a.seeds_cur, *a.T,
                              *a.P, a.bm, a.hot);
    } else {
        hipLaunchKernelGGL((k_gmp_hot<S, MODE>), dim3(a.hot_grid), dim3(64), 0, a.stream, a.state, a.lost_out,
                           a.prob, a.reals, a.plan, a.seeds_cur, *a.T, *a.P, a.bm, a.hot);
    }
    return hipGetLastError();
}

template <int MODE>
static hipError_t launch_m(const StepLaunch &a) {
    switch (a.T->S) {
        case 1: return launch_s<1, MODE>(a);
        case 2: return launch_s<2, MODE>(a);
        case 3: return launch_s<3, MODE>(a);
        case 4: return launch_s<4, MODE>(a);
        case 5: return launch_s<5, MODE>(a);
        case 6: return launch_s<6, MODE>(a);
        case 7: return launch_s<7, MODE>(a);
        case 8: return launch_s<8, MODE>(a);
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_gmp_step(const StepLaunch &a) {
    if (a.ntiles <= 0) return hipSuccess;
    if (a.block <= 0 || a.block > 256 || (a.block % 64) != 0 || a.cpt <= 0) return hipErrorInvalidValue;
    switch (a.mode) {
        case MODE_HOM: return launch_m<MODE_HOM>(a);
        case MODE_INH_X: return launch_m<MODE_INH_X>(a);
        case MODE_INH_Y: return launch_m<MODE_INH_Y>(a);
        default: return hipErrorInvalidValue;
    }
}

// blocks: scratch of at least 1024 floats; out: one device float
hipError_t launch_min_dt(const float *D, const float *r, int n, float h, float *blocks, float *out,
                         hipStream_t stream) {
    const int nb = n > 0 ? std::min(1024, (n + 255) / 256) : 1;
    hipLaunchKernelGGL(k_min_dt, dim3(nb), dim3(256), 0, stream, D, r, n, h, blocks);
    hipLaunchKernelGGL(k_min_reduce, dim3(1), dim3(256), 0, stream, blocks, nb, out);
    return hipGetLastError();
}

hipError_t launch_inh_probs(float2 *out, const float *D, const float *r, float dt, float dx, int n,
                            hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_inh_probs, dim3((n + 255) / 256), dim3(256), 0, stream, out, D, r, dt, dx, n);
    return hipGetLastError();
}

}  // namespace gpgmp
