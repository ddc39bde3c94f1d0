// Stochastic GMP kernels for gfx950 (device code, included by the per-mode
// translation units): Gillespie SSA and stochastic diffusion splitting (homogeneous,
// and directionally split inhomogeneous drift-diffusion).
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
#pragma once

#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <algorithm>

#include "gpgmp_numeric.h"
#include "law.h"
#include "stochastic_launch.h"
#include "stochastic_types.h"
#include "wave_ops.h"

// Two instantiations of every kernel below: the plain one (int and double rate literals of the
// reaction table) and the extended one for models with localized reactions, float literals or
// custom laws (law programs) -- needs_extended_laws().  The plain kernels carry none of the
// extended code, so they keep their registers; GPGMP_EXTENDED_LAWS selects which one a
// translation unit builds, each in its own inline namespace so their symbols stay apart.
#ifndef GPGMP_EXTENDED_LAWS
#define GPGMP_EXTENDED_LAWS 0
#endif

namespace gpgmp {
#if GPGMP_EXTENDED_LAWS
inline namespace extended_laws {
#else
inline namespace table_laws {
#endif

// MODE_INH_XS: the X sweeps alone (no SSA code, so the sweep-only kernel keeps a small
// register footprint), launched after the SSA launch when the two are split.
enum : int { MODE_HOM = 0, MODE_INH_X = 1, MODE_INH_Y = 2, MODE_INH_XS = 3 };
constexpr bool mode_x(int m) { return m == MODE_INH_X || m == MODE_INH_XS; }
constexpr bool mode_ssa(int m) { return m == MODE_HOM || m == MODE_INH_X; }



template <int S>
__device__ __forceinline__ int32_t sel(const int32_t (&x)[S], int j) {
    int32_t v = x[0];
#pragma unroll
    for (int k = 1; k < S; ++k) v = (j == k) ? x[k] : v;
    return v;
}

// The cell a hazard is evaluated in: its active-reaction bits (RxTable::act) and, for the
// law programs, its grid coordinates (GridX, GridY of the generated kernel).
struct CellCtx {
    uint32_t act;
    int32_t gx, gy;
};

// _h[i+1] = _h[0] + <law_i>; int literal -> int product then float add, double literal ->
// double product chain rounded once (Solver.cpp:849-857).
#if GPGMP_EXTENDED_LAWS
// Extended: a float literal (a user parameter, `const Real k`) -> float product chain; a
// localized reaction's law is `w * <law_i>` with the float mask value w of the cell
// (Solver.cpp:853-854): the chain then starts at w, so an int literal's product is formed in
// float ((w * lit) * X * (X - 1) ...) and a double literal's in double from (double)w * lit; a
// custom law of another shape is its program (law.h), which forms the whole assignment
// `_h[0] + [w *] <law>`.
template <int S>
__device__ __forceinline__ float add_law(const RxTable &T, int i, float h0, const int32_t (&x)[S], const CellCtx &cx) {
    if ((T.custom >> i) & 1u)
        return law_call<S>(T.prog + T.prog_off[i], T.prog_type[i], x, LawCell{cx.gx, cx.gy, cx.act, h0});
    const int nf = T.nf[i];
    const bool loc = (T.localized >> i) & 1u;
    const float w = ((cx.act >> i) & 1u) ? 1.f : 0.f;
    if ((T.lit_is_int >> i) & 1u) {
        if (loc) {
            float v = w * (float)T.lit_int[i];
            for (int f = 0; f < nf; ++f) v = v * (float)(sel(x, T.fs[i][f]) - (int32_t)T.fo[i][f]);
            return h0 + v;
        }
        uint32_t v = (uint32_t)T.lit_int[i];
        for (int f = 0; f < nf; ++f) v *= (uint32_t)(sel(x, T.fs[i][f]) - (int32_t)T.fo[i][f]);
        return h0 + (float)(int32_t)v;
    }
    if ((T.lit_is_float >> i) & 1u) {
        float v = loc ? w * T.lit_flt[i] : T.lit_flt[i];
        for (int f = 0; f < nf; ++f) v = v * (float)(sel(x, T.fs[i][f]) - (int32_t)T.fo[i][f]);
        return h0 + v;
    }
    double v = loc ? (double)w * T.lit_dbl[i] : T.lit_dbl[i];
    for (int f = 0; f < nf; ++f) v = v * (double)(sel(x, T.fs[i][f]) - (int32_t)T.fo[i][f]);
    return (float)((double)h0 + v);
}
#else
template <int S>
__device__ __forceinline__ float add_law(const RxTable &T, int i, float h0, const int32_t (&x)[S], const CellCtx &) {
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
#endif

// Cumulative hazards _h[1..R] and the total _h[0] (Solver.cpp:843-857).
// Compact tables: every law of the first kCompactReactions is formed at once in straight-line
// code.  A factor past the law's nf is 1, which leaves the int and double product chains
// exactly as the general path's; reactions i >= R leave the cumulative sum unchanged.  The
// table entries sit at constant offsets of the kernel argument (scalar registers), and
// the independent products overlap: a lone wave's event costs a fraction of the table
// walk with per-factor branches and loads.
template <int S>
__device__ __forceinline__ void hazards_compact(const RxTable &T, const int32_t (&x)[S],
                                                float (&h)[kMaxReactions + 1]) {
    int32_t d[kCompactReactions][kCompactFactors];
#pragma unroll
    for (int f = 0; f < kCompactFactors; ++f)
#pragma unroll
        for (int i = 0; i < kCompactReactions; ++i)
            d[i][f] = (f < (int)T.nf[i]) ? sel(x, T.fs[i][f]) - (int32_t)T.fo[i][f] : 1;
    double vd[kCompactReactions];
#pragma unroll
    for (int i = 0; i < kCompactReactions; ++i) vd[i] = T.lit_dbl[i];
#pragma unroll
    for (int f = 0; f < kCompactFactors; ++f)
#pragma unroll
        for (int i = 0; i < kCompactReactions; ++i) vd[i] = vd[i] * (double)d[i][f];
    uint32_t vi[kCompactReactions] = {};
    if (T.lit_is_int) {
#pragma unroll
        for (int i = 0; i < kCompactReactions; ++i) vi[i] = (uint32_t)T.lit_int[i];
#pragma unroll
        for (int f = 0; f < kCompactFactors; ++f)
#pragma unroll
            for (int i = 0; i < kCompactReactions; ++i) vi[i] *= (uint32_t)d[i][f];
    }
    float vf[kCompactReactions] = {};
    if (GPGMP_EXTENDED_LAWS && T.lit_is_float) {  // an unused factor is 1.f: v * 1.f is v
#pragma unroll
        for (int i = 0; i < kCompactReactions; ++i) vf[i] = T.lit_flt[i];
#pragma unroll
        for (int f = 0; f < kCompactFactors; ++f)
#pragma unroll
            for (int i = 0; i < kCompactReactions; ++i) vf[i] = vf[i] * (float)d[i][f];
    }
    float t = 0.f;
#pragma unroll
    for (int i = 0; i < kCompactReactions; ++i) {
        const float hi = ((T.lit_is_int >> i) & 1u)                            ? t + (float)(int32_t)vi[i]
                         : (GPGMP_EXTENDED_LAWS && ((T.lit_is_float >> i) & 1u)) ? t + vf[i]
                                                                                 : (float)((double)t + vd[i]);
        t = (i < T.R) ? hi : t;
        h[i + 1] = t;
    }
    h[0] = t;
}

template <int S>
__device__ __forceinline__ void hazards(const RxTable &T, const int32_t (&x)[S], float (&h)[kMaxReactions + 1],
                                        const CellCtx &act) {
    if (T.compact) {
        hazards_compact<S>(T, x, h);
        return;
    }
    h[0] = 0.f;
#if GPGMP_EXTENDED_LAWS
#pragma unroll 1
#else
#pragma unroll
#endif
    for (int i = 0; i < kMaxReactions; ++i) {
        if (i < T.R) {
            h[i + 1] = add_law(T, i, h[0], x, act);
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
    if (T.compact) {
#pragma unroll
        for (int i = 0; i < kCompactReactions; ++i) {
            const bool f = (i < T.R) & ((i == 0) ? (r2 < h[1]) : ((r2 >= h[i]) & (r2 < h[i + 1])));
#pragma unroll
            for (int j = 0; j < S; ++j) x[j] += f ? (int32_t)T.net[i][j] : 0;
        }
        return;
    }
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
// Runs from (time, ctr) until deltat, or until `budget` events have been drawn and fewer
// than `below` lanes of the wave still run (a wave of busy cells keeps going; a lone
// busy cell stops holding its wave): returns false when it stops early (time < deltat),
// so the cell can be resumed bit-exactly; `rate` is then the total hazard h0 of the state
// it stopped in (h0 * (deltat - time) estimates the events it still has to draw).
template <int S, bool COMPACT>
__device__ __forceinline__ bool ssa_cell_t(int32_t (&x)[S], uint32_t cell, int32_t seed, float deltat,
                                           const RxTable &T, uint32_t budget, int below, float &time,
                                           uint32_t &ctr, float &rate, const CellCtx &act) {
    const uint32_t k1 = kKeyBase + (uint32_t)seed;
    if (!(time < deltat)) return true;
    // The draws of an event depend on its counter only, so they are formed one event
    // ahead (one extra draw per call, never used).  With a compact table the next event's
    // Threefry and log chain and this event's hazards are one branch-free block; the
    // empty asm below needs the next draws before the loop's exits, so the compiler keeps
    // them in that block (instead of sinking them past the exits) and interleaves the two
    // independent chains.
    u32x2 w = threefry2x32_20(ctr + 1u, kCtrHi, cell, k1);
    float L = log_f_ge1(1.f / u01(w.x));  // log(1/r1)
    double U2 = u01_d(w.y);
    while (time < deltat) {
        w = threefry2x32_20(ctr + 2u, kCtrHi, cell, k1);
        const float Ln = log_f_ge1(1.f / u01(w.x));
        const double U2n = u01_d(w.y);
        float h[kMaxReactions + 1];
        if constexpr (COMPACT) hazards_compact<S>(T, x, h);
        else hazards<S>(T, x, h, act);
        asm volatile("" ::"v"(Ln), "v"(U2n));
        if (h[0] == 0.f) return true;
        if (ctr == kMaxEventsPerCall) return true;  // runaway guard (never reached by sane models)
        if (ctr >= budget && __popcll(__ballot(1)) < below) {
            rate = h[0];
            return false;
        }
        ++ctr;
        const float r2 = (float)(U2 * (double)h[0]);
        const float dt = L / h[0];
        time += dt;
        if (time < deltat) fire<S>(T, h, r2, x);
        L = Ln;
        U2 = U2n;
    }
    return true;
}

template <int S>
__device__ __forceinline__ bool ssa_cell(int32_t (&x)[S], uint32_t cell, int32_t seed, float deltat,
                                         const RxTable &T, uint32_t budget, int below, float &time,
                                         uint32_t &ctr, float &rate, const CellCtx &act) {
    return T.compact ? ssa_cell_t<S, true>(x, cell, seed, deltat, T, budget, below, time, ctr, rate, act)
                     : ssa_cell_t<S, false>(x, cell, seed, deltat, T, budget, below, time, ctr, rate, act);
}

// The hazards' view of cell c of a W-wide grid, gc its index among the cells of all
// realizations (RxTable::act).
__device__ __forceinline__ CellCtx cell_ctx(const RxTable &T, uint64_t gc, int c, int W) {
#if GPGMP_EXTENDED_LAWS
    return CellCtx{T.localized ? (uint32_t)T.act[gc] : 0xffffu, c % W, c / W};
#else
    (void)T, (void)gc, (void)c, (void)W;
    return CellCtx{0xffffu, 0, 0};
#endif
}

// Lane groups.  With G > 1, G consecutive lanes of a wave (G divides 64) share one
// cell: the random words and logarithms of G successive SSA events are drawn at once
// (lane g: counter ctr+1+g) and broadcast through DPP, the serial event chain runs in
// all lanes of the group, and the particle loops split their Threefry calls over the
// lanes (call k by lane k % G) with the counts summed -- integer sums, so the results
// are the same bits as with one lane per cell.
constexpr int qperm(int a, int b, int c, int d) { return a | (b << 2) | (c << 4) | (d << 6); }

template <int G, int J>
__device__ __forceinline__ int32_t grp_bcast(int32_t v) {
    static_assert(G == 1 || G == 2 || G == 4, "lane groups of 1, 2 or 4");
    if constexpr (G == 1) return v;
    else if constexpr (G == 2) return __builtin_amdgcn_update_dpp(0, v, qperm(J, J, 2 + J, 2 + J), 0xf, 0xf, false);
    else return __builtin_amdgcn_update_dpp(0, v, qperm(J, J, J, J), 0xf, 0xf, false);
}

template <int G, int J = 0>
__device__ __forceinline__ void grp_gather(float v, double d, float (&vs)[G], double (&ds)[G]) {
    if constexpr (J < G) {
        vs[J] = __int_as_float(grp_bcast<G, J>(__float_as_int(v)));
        const uint64_t b = (uint64_t)__double_as_longlong(d);
        const uint32_t lo = (uint32_t)grp_bcast<G, J>((int32_t)(uint32_t)b);
        const uint32_t hi = (uint32_t)grp_bcast<G, J>((int32_t)(uint32_t)(b >> 32));
        ds[J] = __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
        grp_gather<G, J + 1>(v, d, vs, ds);
    }
}

template <int G>
__device__ __forceinline__ int32_t grp_sum(int32_t v) {
    if constexpr (G >= 2) v += __builtin_amdgcn_update_dpp(0, v, qperm(1, 0, 3, 2), 0xf, 0xf, false);
    if constexpr (G >= 4) v += __builtin_amdgcn_update_dpp(0, v, qperm(2, 3, 0, 1), 0xf, 0xf, false);
    return v;
}

// performGillespie for one cell with G lanes (identical event sequence to ssa_cell).
template <int S, int G>
__device__ __forceinline__ bool ssa_cell_g(int32_t (&x)[S], uint32_t cell, int32_t seed, float deltat,
                                           const RxTable &T, uint32_t budget, int below, float &time,
                                           uint32_t &ctr, float &rate, const CellCtx &act) {
    if constexpr (G == 1) {
        return ssa_cell<S>(x, cell, seed, deltat, T, budget, below, time, ctr, rate, act);
    } else {
        const uint32_t k1 = kKeyBase + (uint32_t)seed;
        const uint32_t g = threadIdx.x & (G - 1);
        if (!(time < deltat)) return true;
        while (true) {
            const u32x2 w = threefry2x32_20(ctr + 1u + g, kCtrHi, cell, k1);
            float Ls[G];
            double Us[G];
            grp_gather<G>(log_f(1.f / u01(w.x)), u01_d(w.y), Ls, Us);
#pragma unroll
            for (int j = 0; j < G; ++j) {
                float h[kMaxReactions + 1];
                hazards<S>(T, x, h, act);
                if (h[0] == 0.f) return true;
                if (ctr == kMaxEventsPerCall) return true;  // runaway guard
                if (ctr >= budget && __popcll(__ballot(1)) < below) {
                    rate = h[0];
                    return false;
                }
                ++ctr;
                const float r2 = (float)(Us[j] * (double)h[0]);
                const float dt = Ls[j] / h[0];
                time += dt;
                if (!(time < deltat)) return true;
                fire<S>(T, h, r2, x);
            }
        }
    }
}

// Per-particle categorical hop of homogeneous_stochastic_diffuse
// (StochasticHomogeneousDiffusion.cl:68-92): counts of u in (stay,p1], (p1,p2],
// (p2,p3], (p3,p4] are the hops to -x, +x, -y, +y.  Call k (counter k+1) serves
// particles 2k (word x) and 2k+1 (word y).
template <int G>
__device__ __forceinline__ void hom_outflow(int n, uint32_t k0, uint32_t k1, const HomProbs &P,
                                            int32_t (&cnt)[4]) {
    int32_t le0 = 0, le1 = 0, le2 = 0, le3 = 0, le4 = 0;
    const int calls = (n + 1) >> 1;
    for (int k = (int)(threadIdx.x & (G - 1)); k < calls; k += G) {
        const u32x2 w = threefry2x32_20((uint32_t)k + 1u, kCtrHi, k0, k1);
        float u = u01(w.x);
        le0 += u <= P.stay;
        le1 += u <= P.p1;
        le2 += u <= P.p2;
        le3 += u <= P.p3;
        le4 += u <= P.p4;
        if (2 * k + 1 < n) {
            u = u01(w.y);
            le0 += u <= P.stay;
            le1 += u <= P.p1;
            le2 += u <= P.p2;
            le3 += u <= P.p3;
            le4 += u <= P.p4;
        }
    }
    le0 = grp_sum<G>(le0);
    le1 = grp_sum<G>(le1);
    le2 = grp_sum<G>(le2);
    le3 = grp_sum<G>(le3);
    le4 = grp_sum<G>(le4);
    cnt[0] = le1 - le0;
    cnt[1] = le2 - le1;
    cnt[2] = le3 - le2;
    cnt[3] = le4 - le3;
}

// Per-particle Bernoulli hop of inhomogeneous_diffuseX/Y
// (StochasticInhomogeneousDriftDiffusion.cl:524-547): dir = +[u <= p1], counted
// when u <= p2.  With (min(p1,p2), p2) as exact integer thresholds (k_inh_probs):
// plus = #(w <= x), minus = #(w <= y) - plus, then the per-cell none-mask fix-up.
__device__ __forceinline__ void hop_fixup(const int4 &pr, int32_t &a, int32_t &b) {
    if (pr.z & 1) a = 0;
    if (pr.z & 2) b = 0;
}

template <int G>
__device__ __forceinline__ void inh_outflow(int n, uint32_t k0, uint32_t k1, int4 pr, int32_t &minus,
                                            int32_t &plus) {
    int32_t a = 0, b = 0;
    const int calls = (n + 1) >> 1;
    for (int k = (int)(threadIdx.x & (G - 1)); k < calls; k += G) {
        const u32x2 w = threefry2x32_20((uint32_t)k + 1u, kCtrHi, k0, k1);
        a += (int32_t)w.x <= pr.x;
        b += (int32_t)w.x <= pr.y;
        if (2 * k + 1 < n) {
            a += (int32_t)w.y <= pr.x;
            b += (int32_t)w.y <= pr.y;
        }
    }
    a = grp_sum<G>(a);
    b = grp_sum<G>(b);
    hop_fixup(pr, a, b);
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


// Outflow of species s from cell c for this launch's sweep, computed by a group of G
// lanes (all lanes get the counts; the group's first lane writes them).  Returns the
// particles that left.
template <int MODE, int G>
__device__ __forceinline__ int32_t cell_outflow(int32_t n, int s, int c, int W, int H, int A, int32_t seed,
                                                const LaunchPlan &pl, int4 bm, const RealInfo &ri,
                                                const int4 *__restrict__ prob, int32_t *__restrict__ lout,
                                                const HomProbs &P) {
    constexpr int ND = (MODE == MODE_HOM) ? 4 : 2;
    const bool lead = (threadIdx.x & (G - 1)) == 0;
    const int gx = c % W, gy = c / W;
    const int m0 = (gx == 0) ? bm.x : 1, m1 = (gx == W - 1) ? bm.y : 1;
    const int m2 = (gy == 0) ? bm.z : 1, m3 = (gy == H - 1) ? bm.w : 1;
    const uint32_t k1 = kKeyBase + (uint32_t)seed + (uint32_t)pl.out_off[s];
    int32_t *l = lout + (size_t)s * ND * A;
    if constexpr (MODE == MODE_HOM) {
        int32_t cnt[4];
        hom_outflow<G>(n, (uint32_t)c, k1, P, cnt);
        const int32_t l0 = m0 * cnt[0], l1 = m1 * cnt[1], l2 = m2 * cnt[2], l3 = m3 * cnt[3];
        if (lead) {
            l[0 * (size_t)A + c] = l0;
            l[1 * (size_t)A + c] = l1;
            l[2 * (size_t)A + c] = l2;
            l[3 * (size_t)A + c] = l3;
        }
        return l0 + l1 + l2 + l3;
    } else {
        const int4 pr = prob[ri.prob_off + (size_t)s * A + c];
        int32_t mi, pl_;
        inh_outflow<G>(n, (uint32_t)(c + s * A), k1, pr, mi, pl_);
        const int32_t mm = mode_x(MODE) ? m0 : m2;
        const int32_t mp = mode_x(MODE) ? m1 : m3;
        const int32_t lm = mm * mi, lp = mp * pl_;
        if (lead) {
            l[0 * (size_t)A + c] = lm;
            l[1 * (size_t)A + c] = lp;
        }
        return lm + lp;
    }
}

// Flattened outflow of one species for the cells of a wave (NC = 64/LPC cells, cell q
// owned by lane q*LPC): the cells' Threefry calls are laid end to end and every lane
// takes an equal slice of them (lane L: calls [L*T/64, (L+1)*T/64) of the wave's T), so
// one busy cell no longer sets the wave's time.  A lane walks its slice cell by cell,
// adding its partial counts into the wave's LDS counters when it leaves a cell; the
// owners then turn the totals into the outflow.  Counts are integer sums, so they equal
// the serial loop's.  Only the lanes of one wave touch its LDS slice.
struct FlatLds {
    int32_t start[4][65];       // exclusive scan of the calls per cell, [NC] = total
    int32_t n[4][64];           // particles per cell
    int32_t cell[4][64];        // the cell of each slot
    int2 pr[4][64];             // inhomogeneous thresholds of (min(p1,p2), p2) per cell
    uint32_t cnt[4][64][5];     // counters per cell
};

__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// One cursor of a lane's slice of the flattened calls: the cell it is in, the call
// index inside that cell and the partial counts, added into the wave's LDS counters when
// the cursor leaves the cell.
template <int MODE>
struct FlatWalker {
    int qq;
    uint32_t end, k, k0;
    int32_t n;
    int2 pr;
    uint32_t c[5];

    __device__ __forceinline__ void load(const FlatLds &L, int wave, uint32_t koff) {
        n = L.n[wave][qq];
        pr = (MODE != MODE_HOM) ? L.pr[wave][qq] : make_int2(0, 0);
        k0 = (uint32_t)L.cell[wave][qq] + koff;
#pragma unroll
        for (int j = 0; j < 5; ++j) c[j] = 0u;
    }
    // position at flattened call f: the cell q with st[q] <= f < st[q+1]
    __device__ __forceinline__ void start(uint32_t f, const int32_t *st, const FlatLds &L, int wave, int NC,
                                          uint32_t koff) {
        int a = 0, b = NC;
        while (b - a > 1) {
            const int m = (a + b) >> 1;
            if ((uint32_t)st[m] <= f) a = m;
            else b = m;
        }
        qq = a;
        end = (uint32_t)st[qq + 1];
        k = f - (uint32_t)st[qq];
        load(L, wave, koff);
    }
    __device__ __forceinline__ void flush(FlatLds &L, int wave) {
        atomicAdd(&L.cnt[wave][qq][0], c[0]);
        atomicAdd(&L.cnt[wave][qq][1], c[1]);
        if (MODE == MODE_HOM) {
            atomicAdd(&L.cnt[wave][qq][2], c[2]);
            atomicAdd(&L.cnt[wave][qq][3], c[3]);
            atomicAdd(&L.cnt[wave][qq][4], c[4]);
        }
    }
    // f reached the end of the current cell: move to the next cell with calls
    __device__ __forceinline__ void advance(uint32_t f, const int32_t *st, FlatLds &L, int wave, uint32_t koff) {
        flush(L, wave);
        ++qq;
        while ((uint32_t)st[qq + 1] == f) ++qq;
        end = (uint32_t)st[qq + 1];
        k = 0;
        load(L, wave, koff);
    }
    // particles 2k (word x) and 2k+1 (word y, if the cell has it)
    __device__ __forceinline__ void count(const u32x2 &w, bool valid, const HomProbs &P) {
        const bool two = valid && (int32_t)(2 * k + 1) < n;
        if (MODE == MODE_HOM) {
            const float u = u01(w.x), v = u01(w.y);
            c[0] += (valid && u <= P.stay) + (two && v <= P.stay);
            c[1] += (valid && u <= P.p1) + (two && v <= P.p1);
            c[2] += (valid && u <= P.p2) + (two && v <= P.p2);
            c[3] += (valid && u <= P.p3) + (two && v <= P.p3);
            c[4] += (valid && u <= P.p4) + (two && v <= P.p4);
        } else {
            const int32_t u = (int32_t)w.x, v = (int32_t)w.y;
            c[0] += (valid && u <= pr.x) + (two && v <= pr.x);
            c[1] += (valid && u <= pr.y) + (two && v <= pr.y);
        }
    }
};

template <int MODE, int LPC>
__device__ __forceinline__ void wave_flat_outflow(int32_t &xs, bool deferred, int c, int s, int W, int H, int A,
                                                  int32_t seed, const LaunchPlan &pl, int4 bm, const RealInfo &ri,
                                                  const int4 *__restrict__ prob, int32_t *__restrict__ lout,
                                                  const HomProbs &P, FlatLds &L) {
    constexpr int NC = 64 / LPC;
    constexpr int NCNT = (MODE == MODE_HOM) ? 5 : 2;
    constexpr int ND = (MODE == MODE_HOM) ? 4 : 2;
    const int lane = (int)(threadIdx.x & 63), wave = (int)(threadIdx.x >> 6);
    const int q = lane / LPC;
    const bool own = (lane % LPC) == 0;
    const uint32_t k1 = kKeyBase + (uint32_t)seed + (uint32_t)pl.out_off[s];
    const uint32_t koff = (MODE == MODE_HOM) ? 0u : (uint32_t)(s * A);
    int32_t *st = L.start[wave];
    const int32_t calls = (own && !deferred) ? ((xs + 1) >> 1) : 0;
    const int32_t incl = wave_incl_scan_i32(calls);
    const uint32_t total = (uint32_t)__builtin_amdgcn_readlane(incl, 63);
    int32_t none = 0;  // the cell's none-mask (thresholds below every draw)
    if (own) {
        st[q] = incl - calls;
        L.n[wave][q] = xs;
        L.cell[wave][q] = c;
        if (MODE != MODE_HOM) {
            const int4 pr4 = prob[ri.prob_off + (size_t)s * A + c];
            L.pr[wave][q] = make_int2(pr4.x, pr4.y);
            none = pr4.z;
        }
#pragma unroll
        for (int j = 0; j < NCNT; ++j) L.cnt[wave][q][j] = 0u;
    }
    if (lane == 0) st[NC] = (int32_t)total;
    wave_lds_sync();
    const uint32_t lo = (uint32_t)(((uint64_t)lane * total) >> 6);
    const uint32_t hi = (uint32_t)(((uint64_t)(lane + 1) * total) >> 6);
    // one walker per lane over its slice (with 8 cells per wave, 8 waves per SIMD keep
    // enough Threefry chains in flight; a second interleaved walker per lane measured slower)
    if (lo < hi) {
        FlatWalker<MODE> wa;
        wa.start(lo, st, L, wave, NC, koff);
        for (uint32_t f = lo; f < hi; ++f) {
            if (f == wa.end) wa.advance(f, st, L, wave, koff);
            const u32x2 ra = threefry2x32_20(wa.k + 1u, kCtrHi, wa.k0, k1);
            wa.count(ra, true, P);
            ++wa.k;
        }
        wa.flush(L, wave);
    }
    wave_lds_sync();
    if (own && !deferred) {
        const int gx = c % W, gy = c / W;
        const int m0 = (gx == 0) ? bm.x : 1, m1 = (gx == W - 1) ? bm.y : 1;
        const int m2 = (gy == 0) ? bm.z : 1, m3 = (gy == H - 1) ? bm.w : 1;
        int32_t *l = lout + (size_t)s * ND * A;
        const uint32_t *cn = L.cnt[wave][q];
        if (MODE == MODE_HOM) {
            const int32_t l0 = m0 * (int32_t)(cn[1] - cn[0]), l1 = m1 * (int32_t)(cn[2] - cn[1]);
            const int32_t l2 = m2 * (int32_t)(cn[3] - cn[2]), l3 = m3 * (int32_t)(cn[4] - cn[3]);
            l[0 * (size_t)A + c] = l0;
            l[1 * (size_t)A + c] = l1;
            l[2 * (size_t)A + c] = l2;
            l[3 * (size_t)A + c] = l3;
            xs -= l0 + l1 + l2 + l3;
        } else {
            const int32_t mm = mode_x(MODE) ? m0 : m2;
            const int32_t mp = mode_x(MODE) ? m1 : m3;
            int32_t a = (int32_t)cn[0], b = (int32_t)cn[1];
            hop_fixup(make_int4(0, 0, none, 0), a, b);
            const int32_t lm = mm * (b - a), lp = mp * a;
            l[0 * (size_t)A + c] = lm;
            l[1 * (size_t)A + c] = lp;
            xs -= lm + lp;
        }
    }
    wave_lds_sync();  // the slice is reused by the next species / cell round
}

// Events above which a cell counts as busy for the next step (one speculative batch):
// such cells are run whole by worker waves of the next SSA launch.
constexpr uint32_t kPredictEvents = 64;

// The hot list of one parity: with predicted busy cells two lists alternate by step (the
// SSA launch of step i reads list (i-1)&1 and fills list i&1).
__device__ __forceinline__ HotList hot_view(const HotList &h, int parity) {
    HotList v = h;
    v.e = h.e + (size_t)parity * h.cap;
    v.count = h.count + 2 * parity;
    return v;
}

// Appends one entry per lane group whose lanes want it (the group's first lane writes
// it) to the list (e, count, cap); returns true in the lanes whose entry found room.  Every
// lane of the wave calls it.
__device__ __forceinline__ bool push_hot(bool want, int g, HotEntry *e, uint32_t *count, uint32_t cap,
                                         const HotEntry &he) {
    const uint64_t dm = __ballot(want && g == 0);
    if (!dm) return false;
    const int lane = (int)(threadIdx.x & 63);
    const int leader = __ffsll((unsigned long long)dm) - 1;
    uint32_t base = 0;
    if (lane == leader) base = atomicAdd(count, (uint32_t)__popcll(dm));
    base = __shfl(base, leader, 64);
    if (!want) return false;
    const int first = lane - g;
    const uint32_t idx = base + (uint32_t)__popcll(dm & ((1ull << first) - 1ull));
    if (idx >= cap) return false;
    if (g == 0) e[idx] = he;
    return true;
}

// One tile of a step launch: G lanes per cell for the SSA; the outflow by GS lanes per
// cell (GS = G, or 4 with G = 1), or flattened over the wave (GS = 0).
template <int S, int MODE, int G, int GS>
__device__ __forceinline__ void step_tile(
    int t, int32_t *__restrict__ state, const int32_t *__restrict__ lost_in, int32_t *__restrict__ lost_out,
    const int4 *__restrict__ prob, const RealInfo *__restrict__ reals,
    const Tile *__restrict__ tiles, const LaunchPlan *__restrict__ plan,
    const int32_t *__restrict__ seeds_cur, const int32_t *__restrict__ seeds_prev, const RxTable &T,
    const HomProbs &P, int4 bm, int cpt, const HotList &hot_all, const OneReal &one, FlatLds &flat_lds) {
    Tile tile;
    RealInfo ri;
    if (one.per > 0) {
        tile.r = 0;
        tile.cell0 = t * one.per;
        ri = one.r;
    } else {
        tile = tiles[t];
        ri = reals[tile.r];
    }
    const LaunchPlan pl = *plan;
    if (MODE == MODE_INH_Y && pl.reset_hot && t == 0 && threadIdx.x == 0) {
        const HotList v = hot_view(hot_all, pl.hot_parity);
        v.count[0] = 0u;
        v.count[1] = 0u;
        if (hot_all.lane_count) hot_all.lane_count[0] = 0u;
    }
    const bool predict = mode_ssa(MODE) && pl.predict && hot_all.flag;
    const HotList hot = hot_view(hot_all, predict ? pl.hot_parity : 0);
    const uint8_t *pred_prev = predict ? hot_all.pred + (size_t)(pl.hot_parity ^ 1) * hot_all.pred_stride : nullptr;
    uint8_t *pred_next = predict ? hot_all.pred + (size_t)pl.hot_parity * hot_all.pred_stride : nullptr;
    const int32_t seed = seeds_cur[tile.r];
    const int W = ri.W, H = ri.H;
    const int A = W * H;
    constexpr int ND = (MODE == MODE_HOM) ? 4 : 2;
    int32_t *st = state + ri.state_off;
    const int32_t *lin = lost_in + ri.lost_off;
    int32_t *lout = lost_out + ri.lost_off;
    const bool any_src = (ri.src[0] | ri.src[1] | ri.src[2] | ri.src[3]) != 0;
    const int g = (int)(threadIdx.x & (G - 1));        // lane in the cell's group
    const int cpb = (int)blockDim.x / G;               // cells per block and round

    // flattened sweeps (64/G cells per wave) may spread a wave's cells over the grid
    // (stride A/(64/G), a bijection when 64/G divides A), so that waves covering dense and
    // sparse regions even out; per-cell results do not depend on the mapping (RNG keys are
    // per cell)
    constexpr bool kSpread = (GS == 0 && G >= 4);
    constexpr int NCW = 64 / G;
    const int stride = (kSpread && pl.interleave && A % NCW == 0) ? A / NCW : 1;
    for (int k = 0; k < cpt; ++k) {
        const int lc = tile.cell0 + k * cpb + (int)threadIdx.x / G;
        if (lc >= A) break;
        const int c = (stride > 1) ? (lc % NCW) * stride + lc / NCW : lc;
        // a cell busy last step is run whole (gather + SSA) by a worker wave of this launch
        if (predict && pred_prev[ri.cell_off + c]) {
            if (g == 0) hot.flag[ri.cell_off + c] = 0;
            continue;
        }
        const int gx = c % W, gy = c / W;
        int32_t x[S];
        // loads: every species for launches with an SSA, only touched species otherwise
        const uint32_t touched = pl.gather_mask | pl.out_mask;
#pragma unroll
        for (int s = 0; s < S; ++s)
            x[s] = (pl.deltat > 0.f || ((touched >> s) & 1u)) ? st[(size_t)s * A + c] : 0;

        // 1. gather the previous launch's outflow.  Every neighbour word is loaded from a
        //    clamped (in-grid) address and the edge lanes add 0 instead, so the loads of all
        //    species go out together rather than one divergent branch and wait each; the
        //    integer sum is the same in any order.
        if (pl.gather_mask) {
            const int cxm = (gx > 0) ? c - 1 : c, cxp = (gx < W - 1) ? c + 1 : c;
            const int cym = (gy > 0) ? c - W : c, cyp = (gy < H - 1) ? c + W : c;
#pragma unroll
            for (int s = 0; s < S; ++s) {
                if (!((pl.gather_mask >> s) & 1u)) continue;
                const int32_t *l = lin + (size_t)s * ND * A;
                int32_t v = x[s];
                if (MODE == MODE_HOM) {
                    const int32_t a0 = l[1 * (size_t)A + cxm], a1 = l[0 * (size_t)A + cxp];
                    const int32_t a2 = l[3 * (size_t)A + cym], a3 = l[2 * (size_t)A + cyp];
                    v += ((gx > 0) ? a0 : 0) + ((gx < W - 1) ? a1 : 0) + ((gy > 0) ? a2 : 0) +
                         ((gy < H - 1) ? a3 : 0);
                } else if (mode_x(MODE)) {  // gathers the Y outflow of the previous launch
                    const int32_t a2 = l[1 * (size_t)A + cym], a3 = l[0 * (size_t)A + cyp];
                    v += ((gy > 0) ? a2 : 0) + ((gy < H - 1) ? a3 : 0);
                } else {                          // gathers the X outflow of the previous launch
                    const int32_t a0 = l[1 * (size_t)A + cxm], a1 = l[0 * (size_t)A + cxp];
                    v += ((gx > 0) ? a0 : 0) + ((gx < W - 1) ? a1 : 0);
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

        // 2. reactions.  A cell stops after hot.budget events and is handed to the
        //    hot-cell kernel, so a few busy cells do not hold a whole wave (and the GPU).
        //    An X-sweep launch that runs beside the hot cells skips the flagged ones.
        bool deferred = false;
        if (MODE == MODE_INH_XS && hot.flag) deferred = hot.flag[ri.cell_off + c] != 0;
        if constexpr (mode_ssa(MODE)) {
          if (pl.deltat > 0.f && T.R > 0) {
            float time = 0.f, rate = 0.f;
            uint32_t ctr = 0;
            const CellCtx act = cell_ctx(T, ri.cell_off + (uint64_t)c, c, W);
            const bool fin = ssa_cell_g<S, G>(x, (uint32_t)c, seed, pl.deltat, T, hot.budget, hot.below_active,
                                              time, ctr, rate, act);
            // a deferred cell with a modest expected tail goes to the lane list, the others
            // (and any the lane list has no room for) to the wave list; a cell neither list
            // can take is finished here
            const HotEntry he{tile.r, c, time, ctr};
            const bool to_lane = !fin && hot.lane_e != nullptr && rate * (pl.deltat - time) <= hot.lane_events;
            const bool in_lane = push_hot(to_lane, g, hot.lane_e, hot.lane_count, hot.lane_cap, he);
            const bool in_wave = push_hot(!fin && !in_lane, g, hot.e, hot.count, hot.cap, he);
            deferred = in_lane || in_wave;
            if (!fin && !deferred)
                ssa_cell_g<S, G>(x, (uint32_t)c, seed, pl.deltat, T, 0xffffffffu, 0, time, ctr, rate, act);
          }
        }

        if (mode_ssa(MODE) && hot.flag && g == 0) hot.flag[ri.cell_off + c] = deferred ? 1 : 0;
        if (predict && g == 0) pred_next[ri.cell_off + c] = deferred ? 1 : 0;

        // 3. outflow of the due species (a deferred cell's comes from the hot-cell kernel)
        if (pl.out_mask) {
            if constexpr (GS == 0) {
#pragma unroll
                for (int s = 0; s < S; ++s) {
                    if (!((pl.out_mask >> s) & 1u)) continue;
                    wave_flat_outflow<MODE, G>(x[s], deferred, c, s, W, H, A, seed, pl, bm, ri, prob, lout, P,
                                               flat_lds);
                }
            } else if constexpr (G == 1 && GS == 4) {
                // the wave's 64 cells in 4 rounds of 16, cell 16r+q by lanes 4q..4q+3;
                // the counts go back to the lane that owns the cell
                const int lane = (int)(threadIdx.x & 63);
                const int base = c - lane;
#pragma unroll
                for (int s = 0; s < S; ++s) {
                    if (!((pl.out_mask >> s) & 1u)) continue;
                    int32_t sub = 0;
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const int src = 16 * r + (lane >> 2);
                        const int32_t xs = __shfl(x[s], src, 64);
                        const int dfr = __shfl((int)deferred, src, 64);
                        int32_t res = 0;
                        if (!dfr)
                            res = cell_outflow<MODE, 4>(xs, s, base + src, W, H, A, seed, pl, bm, ri, prob, lout, P);
                        const int32_t back = __shfl(res, 4 * (lane & 15), 64);
                        if ((lane >> 4) == r) sub = back;
                    }
                    x[s] -= sub;
                }
            } else if (!deferred) {
#pragma unroll
                for (int s = 0; s < S; ++s) {
                    if (!((pl.out_mask >> s) & 1u)) continue;
                    x[s] -= cell_outflow<MODE, G>(x[s], s, c, W, H, A, seed, pl, bm, ri, prob, lout, P);
                }
            }
        }

        // a deferred cell's state belongs to the hot-cell work once the SSA launch is done
        if (g == 0 && !(MODE == MODE_INH_XS && deferred)) {
#pragma unroll
            for (int s = 0; s < S; ++s)
                if (pl.deltat > 0.f || ((touched >> s) & 1u)) st[(size_t)s * A + c] = x[s];
        }
    }
}

template <int S, int MODE, int G, int GS>
__global__ __launch_bounds__(256) void k_gmp_step(
    int32_t *__restrict__ state, const int32_t *__restrict__ lost_in, int32_t *__restrict__ lost_out,
    const int4 *__restrict__ prob, const RealInfo *__restrict__ reals,
    const Tile *__restrict__ tiles, int ntiles, const LaunchPlan *__restrict__ plan,
    const int32_t *__restrict__ seeds_cur, const int32_t *__restrict__ seeds_prev, RxTable T,
    HomProbs P, int4 bm, int cpt, HotList hot, OneReal one) {
    // the grid has 8*ceil(ntiles/8) blocks so that the remap is a bijection
    __shared__ FlatLds flat_lds;
    const int t = xcd_tile((int)blockIdx.x, ntiles);
    if (t >= ntiles) return;
    step_tile<S, MODE, G, GS>(t, state, lost_in, lost_out, prob, reals, tiles, plan, seeds_cur, seeds_prev, T, P,
                              bm, cpt, hot, one, flat_lds);
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
__device__ __forceinline__ int32_t wave_sum(int32_t v) { return wave_sum_i32(v); }

// Mask of the reactions whose selection interval holds r2 -- the conditions of fire;
// cumulative hazards never decrease, so at most one bit is set.
template <int S>
__device__ __forceinline__ uint32_t fired(const RxTable &T, const float (&h)[kMaxReactions + 1], float r2) {
    uint32_t m = 0;
    if (T.compact) {
#pragma unroll
        for (int i = 0; i < kCompactReactions; ++i) {
            const bool f = (i < T.R) & ((i == 0) ? (r2 < h[1]) : ((r2 >= h[i]) & (r2 < h[i + 1])));
            m |= f ? (1u << i) : 0u;
        }
        return m;
    }
#pragma unroll
    for (int i = 0; i < kMaxReactions; ++i) {
        if (i < T.R) {
            const bool f = (i == 0) ? (r2 < h[1]) : ((r2 >= h[i]) && (r2 < h[i + 1]));
            m |= f ? (1u << i) : 0u;
        }
    }
    return m;
}

// Net change of species s by the reactions of mask m (the update fire applies).
__device__ __forceinline__ int32_t net_of(const RxTable &T, uint32_t m, int s) {
    int32_t d = 0;
    if (T.compact) {  // fired sets no bit at or above R
#pragma unroll
        for (int i = 0; i < kCompactReactions; ++i) d += ((m >> i) & 1u) ? (int32_t)T.net[i][s] : 0;
        return d;
    }
#pragma unroll
    for (int i = 0; i < kMaxReactions; ++i)
        if (i < T.R) d += ((m >> i) & 1u) ? (int32_t)T.net[i][s] : 0;
    return d;
}

__device__ __forceinline__ int32_t wave_excl_scan(int32_t v) { return wave_incl_scan_i32(v) - v; }

// Continues performGillespie (Gillespie.cl:41-99) of one cell from (time, ctr) with the
// 64 lanes of a wave, speculatively: lane j draws event ctr+1+j and selects its reaction
// with the hazards of the batch's first state; an exclusive scan of the selected net
// changes gives every lane the state before its event; each lane then recomputes the
// hazards of that state and its reaction.  Up to the first lane m whose reaction differs
// from the speculation (or whose total hazard is 0) every state is exact, and lane m's
// own reaction is exact too.  The recomputed reactions become the next speculation and
// the pass repeats (each pass fixes at least one more event; a handful usually fix the
// whole batch), reusing the batch's random words and logarithms.  The float clock before
// every event of the exact prefix is then formed lane-parallel but in the serial order of
// additions (t += log(1/r1)/h0), and the batch stops where ssa_cell would (t >= deltat,
// zero hazard, runaway guard), so the committed events are the serial chain's.
constexpr int kSpecPasses = 8;

template <int S>
__device__ __forceinline__ void ssa_wave_spec(int32_t (&x)[S], uint32_t cell, uint32_t k1, float deltat,
                                              const RxTable &T, float &time, uint32_t &ctr, const CellCtx &act) {
    const int lane = (int)(threadIdx.x & 63);
    if (!(time < deltat)) return;
    while (true) {
        float h[kMaxReactions + 1];
        hazards<S>(T, x, h, act);
        if (h[0] == 0.f) return;
        if (ctr == kMaxEventsPerCall) return;
        const u32x2 w = threefry2x32_20(ctr + 1u + (uint32_t)lane, kCtrHi, cell, k1);
        const float L = log_f(1.f / u01(w.x));
        const double U2 = u01_d(w.y);
        uint32_t ms = fired<S>(T, h, (float)(U2 * (double)h[0]));
        int32_t xj[S];
        float hj[kMaxReactions + 1];
        uint32_t ma = 0;
        uint64_t zb = 0;
        int last = 63;
        for (int pass = 0; pass < kSpecPasses; ++pass) {
#pragma unroll
            for (int s = 0; s < S; ++s) xj[s] = x[s] + wave_excl_scan(net_of(T, ms, s));
            hazards<S>(T, xj, hj, act);
            ma = fired<S>(T, hj, (float)(U2 * (double)hj[0]));
            const bool zero = hj[0] == 0.f;
            zb = __ballot(zero);
            const uint64_t bad = __ballot(zero || ma != ms);
            if (!bad) {
                last = 63;
                break;
            }
            last = __ffsll((unsigned long long)bad) - 1;  // lanes 0..last exact
            if ((zb >> last) & 1ull) break;              // the chain ends at a zero hazard
            ms = ma;
        }
        // the float clock before each event, in the serial order: lane j accumulates
        // d_0, d_1, ..., d_{j-1} one after the other, so t_j is bit-identical to the
        // serial chain's clock (no reassociation)
        const float dt = L / hj[0];
        float t = time;
#pragma unroll
        for (int k = 0; k < 63; ++k) {
            const float dk = readlane_f(dt, k);
            t = (lane > k) ? t + dk : t;
        }
        const float t_after = t + dt;
        // stop before event j: zero total hazard, the runaway guard, or the clock reaching
        // deltat (the event is then not fired)
        const uint64_t upto = (last == 63) ? ~0ull : ((2ull << last) - 1ull);
        uint64_t stopm = (__ballot(!(t_after < deltat)) | zb) & upto;
        const uint32_t room = kMaxEventsPerCall - ctr;  // ctr < kMaxEventsPerCall here
        if (room <= (uint32_t)last) stopm |= 1ull << room;
        if (stopm) {
            const int stop = __ffsll((unsigned long long)stopm) - 1;
#pragma unroll
            for (int s = 0; s < S; ++s) x[s] = __builtin_amdgcn_readlane(xj[s], stop);
            return;
        }
        time = readlane_f(t_after, last);
#pragma unroll
        for (int s = 0; s < S; ++s) x[s] = __builtin_amdgcn_readlane(xj[s] + net_of(T, ma, s), last);
        ctr += (uint32_t)last + 1u;
    }
}

// Deferred busy cells, one wave per cell: worker `wk` of `nwk` takes list entries wk,
// wk + nwk, ...  The SSA continues from (time, ctr) with ssa_wave_spec under the SSA
// launch's plan `pss`.  Then the cell's outflow for the due species of `pout` is computed
// with the particle loop split over the lanes (call k by lane k % 64) and the counts
// summed -- the same integers as the serial loop.  The last worker to finish resets the
// list.
// finish: the last worker to finish empties the list (its own kernel); in the X-sweep
// launch of a split step the following Y launch empties it instead (one thread), so the
// workers do not serialise on a shared counter.
template <int S, int MODE>
__device__ __forceinline__ void hot_cells(int wk, int nwk, int32_t *__restrict__ state, int32_t *__restrict__ lost_out,
                                          const int4 *__restrict__ prob, const RealInfo *__restrict__ reals,
                                          const LaunchPlan &pss, const LaunchPlan &pl,
                                          const int32_t *__restrict__ seeds_cur, const RxTable &T, const HomProbs &P,
                                          int4 bm, const HotList &hot, bool finish) {
    const int lane = (int)(threadIdx.x & 63);
    const uint32_t total = hot.count[0];
    if (total == 0) return;
    const uint32_t nh = total < hot.cap ? total : hot.cap;
    constexpr int ND = (MODE == MODE_HOM) ? 4 : 2;
    for (uint32_t i = (uint32_t)wk; i < nh; i += (uint32_t)nwk) {
        const HotEntry he = hot.e[i];
        if (he.time < 0.f) continue;  // a predicted cell, already run whole by the SSA launch
        const RealInfo ri = reals[he.r];
        const int32_t seed = seeds_cur[he.r];
        const int W = ri.W, H = ri.H, A = W * H, c = he.c;
        int32_t *st = state + ri.state_off;
        int32_t x[S];
#pragma unroll
        for (int s = 0; s < S; ++s) x[s] = __builtin_amdgcn_readfirstlane(st[(size_t)s * A + c]);
        float time = he.time;
        uint32_t ctr = he.ctr;
        ssa_wave_spec<S>(x, (uint32_t)c, kKeyBase + (uint32_t)seed, pss.deltat, T, time, ctr,
                         cell_ctx(T, ri.cell_off + (uint64_t)c, c, W));
        // the whole count is known now: busy cells are predicted for the next step
        if (pss.predict && lane == 0)
            hot.pred[(size_t)pss.hot_parity * hot.pred_stride + ri.cell_off + c] = ctr > kPredictEvents ? 1 : 0;
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
                    const int4 pr = prob[ri.prob_off + (size_t)s * A + c];
                    int32_t a = 0, b = 0;
                    for (int k = lane; k < calls; k += 64) {
                        const u32x2 ww = threefry2x32_20((uint32_t)k + 1u, kCtrHi, (uint32_t)(c + s * A), k1o);
                        a += (int32_t)ww.x <= pr.x;
                        b += (int32_t)ww.x <= pr.y;
                        if (2 * k + 1 < n) {
                            a += (int32_t)ww.y <= pr.x;
                            b += (int32_t)ww.y <= pr.y;
                        }
                    }
                    a = wave_sum(a);
                    b = wave_sum(b);
                    hop_fixup(pr, a, b);
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
    // every worker read the count before adding itself here, so the last one may reset
    if (finish && lane == 0) {
        __threadfence();
        const uint32_t fin = atomicAdd(&hot.count[1], 1u);
        if (fin == (uint32_t)nwk - 1u) {
            atomicExch(&hot.count[0], 0u);
            atomicExch(&hot.count[1], 0u);
            __threadfence();
        }
    }
}

// Deferred cells of the lane list, one per lane: worker wave wk of nwk takes the entries
// [64 (wk + j nwk), 64 (wk + j nwk + 1)), j = 0, 1, ...; each lane resumes its cell's SSA
// from (time, ctr) to deltat with the one-lane event loop (no further deferral: the list
// holds the cells with a modest expected tail), then the wave computes the X outflow of
// its cells for the due species of `pl`, flattened over its lanes.  The lane list serves
// one realization: hot cells run inside the X-sweep launch of an inhomogeneous solver,
// which is never batched.
template <int S>
__device__ __forceinline__ void hot_lane_cells(int wk, int nwk, int32_t *__restrict__ state,
                                               int32_t *__restrict__ lost_out, const int4 *__restrict__ prob,
                                               const RealInfo &ri, const LaunchPlan &pss, const LaunchPlan &pl,
                                               int32_t seed, const RxTable &T, const HomProbs &P, int4 bm,
                                               const HotList &hot, FlatLds &L) {
    const int lane = (int)(threadIdx.x & 63);
    const uint32_t total = __builtin_amdgcn_readfirstlane(hot.lane_count[0]);
    const uint32_t nh = total < hot.lane_cap ? total : hot.lane_cap;
    const int W = ri.W, H = ri.H, A = W * H;
    int32_t *st = state + ri.state_off;
    for (uint32_t b = (uint32_t)wk * 64u; b < nh; b += (uint32_t)nwk * 64u) {
        const uint32_t i = b + (uint32_t)lane;
        const bool valid = i < nh;
        const HotEntry he = valid ? hot.lane_e[i] : HotEntry{0, 0, 0.f, 0u};
        const int c = he.c;
        int32_t x[S];
#pragma unroll
        for (int s = 0; s < S; ++s) x[s] = valid ? st[(size_t)s * A + c] : 0;
        if (valid) {
            float time = he.time, rate = 0.f;
            uint32_t ctr = he.ctr;
            (void)ssa_cell<S>(x, (uint32_t)c, seed, pss.deltat, T, 0xffffffffu, 0, time, ctr, rate,
                              cell_ctx(T, ri.cell_off + (uint64_t)c, c, W));
        }
        if (pl.out_mask) {
#pragma unroll
            for (int s = 0; s < S; ++s) {
                if (!((pl.out_mask >> s) & 1u)) continue;
                wave_flat_outflow<MODE_INH_X, 1>(x[s], !valid, c, s, W, H, A, seed, pl, bm, ri, prob, lost_out + ri.lost_off,
                                                 P, L);
            }
        }
        if (valid) {
#pragma unroll
            for (int s = 0; s < S; ++s) st[(size_t)s * A + c] = x[s];
        }
    }
}

// Predicted busy cells (SSA launch of a split step): the cells of the other parity's list
// whose prediction flag is set (more than kPredictEvents events last step) are run whole
// by one wave each -- the gather of the Y outflow and the SSA from t = 0 with the
// speculative batches -- while the tile lanes skip exactly those cells.  A cell still busy
// goes into this step's list as a predicted entry (time -1) with its flag set, so the
// next SSA launch runs it the same way.  The flag of a cell deferred by a tile lane is set
// by that lane and decided by the X-sweep launch's worker once the cell's count is known.
template <int S>
__device__ __forceinline__ void pred_cells(int wk, int nwk, int32_t *__restrict__ state,
                                           const int32_t *__restrict__ lost_in, const RealInfo *__restrict__ reals,
                                           const LaunchPlan &pl, const int32_t *__restrict__ seeds_cur,
                                           const RxTable &T, const HotList &hot_all) {
    const int lane = (int)(threadIdx.x & 63);
    const HotList prev = hot_view(hot_all, pl.hot_parity ^ 1);
    const HotList cur = hot_view(hot_all, pl.hot_parity);
    uint8_t *pred_next = hot_all.pred + (size_t)pl.hot_parity * hot_all.pred_stride;
    const uint8_t *pred_prev = hot_all.pred + (size_t)(pl.hot_parity ^ 1) * hot_all.pred_stride;
    const uint32_t total = prev.count[0];
    const uint32_t nh = total < prev.cap ? total : prev.cap;
    for (uint32_t i = (uint32_t)wk; i < nh; i += (uint32_t)nwk) {
        const HotEntry he = prev.e[i];
        const RealInfo ri = reals[he.r];
        if (!pred_prev[ri.cell_off + he.c]) continue;  // listed, but not predicted: a tile lane runs it
        const int32_t seed = seeds_cur[he.r];
        const int W = ri.W, H = ri.H, A = W * H, c = he.c;
        const int gy = c / W;
        int32_t *st = state + ri.state_off;
        const int32_t *lin = lost_in + ri.lost_off;
        int32_t x[S];
#pragma unroll
        for (int s = 0; s < S; ++s) {
            int32_t v = st[(size_t)s * A + c];
            if ((pl.gather_mask >> s) & 1u) {  // the Y outflow of the previous launch
                const int32_t *l = lin + (size_t)s * 2 * A;
                if (gy > 0) v += l[1 * (size_t)A + c - W];
                if (gy < H - 1) v += l[0 * (size_t)A + c + W];
            }
            x[s] = __builtin_amdgcn_readfirstlane(v);
        }
        float time = 0.f;
        uint32_t ctr = 0;
        if (pl.deltat > 0.f && T.R > 0)
            ssa_wave_spec<S>(x, (uint32_t)c, kKeyBase + (uint32_t)seed, pl.deltat, T, time, ctr,
                             cell_ctx(T, ri.cell_off + (uint64_t)c, c, W));
        if (lane == 0) {
#pragma unroll
            for (int s = 0; s < S; ++s) st[(size_t)s * A + c] = x[s];
            bool busy = ctr > kPredictEvents;
            if (busy) {
                const uint32_t idx = atomicAdd(cur.count, 1u);
                if (idx < cur.cap) cur.e[idx] = HotEntry{he.r, c, -1.f, 0u};
                else busy = false;
            }
            pred_next[ri.cell_off + c] = busy ? 1 : 0;
        }
    }
}

// Hot-cell kernel after the step launch (one wave per block).
template <int S, int MODE>
__global__ __launch_bounds__(64) void k_gmp_hot(int32_t *__restrict__ state, int32_t *__restrict__ lost_out,
                                                const int4 *__restrict__ prob, const RealInfo *__restrict__ reals,
                                                const LaunchPlan *__restrict__ plan,
                                                const int32_t *__restrict__ seeds_cur, RxTable T, HomProbs P,
                                                int4 bm, HotList hot) {
    const LaunchPlan pl = *plan;
    hot_cells<S, MODE>((int)blockIdx.x, (int)gridDim.x, state, lost_out, prob, reals, pl, pl, seeds_cur, T, P, bm,
                       hot, true);
}

// X sweeps and the deferred busy cells in one launch, so the cells' serial event chains
// run beside the sweeps instead of before them.  Blocks [0, nlane_blocks) are the lane-list
// workers (64 cells per wave), blocks [nlane_blocks, nlane_blocks + nhot_blocks) the
// wave-list workers (one cell per wave); both continue the SSA under the SSA launch's plan
// `plan_ssa`, then compute their cells' X outflow.  The other blocks are the X-sweep tiles
// (64/G cells per wave, flattened), which skip the cells flagged by the SSA launch.  The
// kinds touch disjoint cells.  Both worker counts are multiples of 8 blocks, so the tiles
// keep their XCD placement.  The lane and
// wave workers read the reaction table from two copies of it (T2 and T): with a single
// argument the compiler shares its loads between the two kinds of workers, and the kernel
// then needs 217 VGPRs and spills (two copies: 126, 4 waves per SIMD as before).
template <int S, int G>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k_gmp_xs_hot(
    int32_t *__restrict__ state, const int32_t *__restrict__ lost_in, int32_t *__restrict__ lost_out,
    const int4 *__restrict__ prob, const RealInfo *__restrict__ reals,
    const Tile *__restrict__ tiles, int ntiles, const LaunchPlan *__restrict__ plan,
    const LaunchPlan *__restrict__ plan_ssa, const int32_t *__restrict__ seeds_cur,
    const int32_t *__restrict__ seeds_prev, RxTable T, RxTable T2, HomProbs P, int4 bm, int cpt, HotList hot,
    int nlane_blocks, int nhot_blocks, OneReal one) {
    __shared__ FlatLds flat_lds;
    const int wpb = (int)(blockDim.x >> 6);
    if ((int)blockIdx.x < nlane_blocks) {
        const LaunchPlan pss = *plan_ssa, pl = *plan;
        hot_lane_cells<S>((int)blockIdx.x * wpb + (int)(threadIdx.x >> 6), nlane_blocks * wpb, state, lost_out, prob,
                          reals[0], pss, pl, seeds_cur[0], T2, P, bm, hot, flat_lds);
        return;
    }
    if ((int)blockIdx.x < nlane_blocks + nhot_blocks) {
        const LaunchPlan pss = *plan_ssa, pl = *plan;
        const int par = pss.predict ? pss.hot_parity : 0;
        hot_cells<S, MODE_INH_X>(((int)blockIdx.x - nlane_blocks) * wpb + (int)(threadIdx.x >> 6), nhot_blocks * wpb,
                                 state, lost_out, prob, reals, pss, pl, seeds_cur, T, P, bm, hot_view(hot, par), false);
        return;
    }
    const int t = xcd_tile((int)blockIdx.x - nlane_blocks - nhot_blocks, ntiles);
    if (t >= ntiles) return;
    step_tile<S, MODE_INH_XS, G, 0>(t, state, lost_in, lost_out, prob, reals, tiles, plan, seeds_cur, seeds_prev,
                                    T, P, bm, cpt, hot, one, flat_lds);
}


// SSA launch of a split step with predicted busy cells: blocks [0, npred_blocks) are the
// worker waves of pred_cells, the others the SSA tiles (one lane per cell), which skip the
// predicted cells.  npred_blocks is a multiple of 8, so the tiles keep their XCD placement.
template <int S>
__global__ __launch_bounds__(256) void k_gmp_ssa_pred(
    int32_t *__restrict__ state, const int32_t *__restrict__ lost_in, int32_t *__restrict__ lost_out,
    const int4 *__restrict__ prob, const RealInfo *__restrict__ reals,
    const Tile *__restrict__ tiles, int ntiles, const LaunchPlan *__restrict__ plan,
    const int32_t *__restrict__ seeds_cur, const int32_t *__restrict__ seeds_prev, RxTable T, HomProbs P, int4 bm,
    int cpt, HotList hot, int npred_blocks) {
    __shared__ FlatLds flat_lds;
    if ((int)blockIdx.x < npred_blocks) {
        const int wpb = (int)(blockDim.x >> 6);
        const LaunchPlan pl = *plan;
        pred_cells<S>((int)blockIdx.x * wpb + (int)(threadIdx.x >> 6), npred_blocks * wpb, state, lost_in, reals, pl,
                      seeds_cur, T, hot);
        return;
    }
    const int t = xcd_tile((int)blockIdx.x - npred_blocks, ntiles);
    if (t >= ntiles) return;
    step_tile<S, MODE_INH_X, 1, 0>(t, state, lost_in, lost_out, prob, reals, tiles, plan, seeds_cur, seeds_prev, T,
                                   P, bm, cpt, hot, OneReal{}, flat_lds);
}

template <int S>
static hipError_t launch_ssa_pred(const StepLaunch &a) {
    const int wpb = a.block / 64;
    const int npred = 8 * ((a.hot_grid + 8 * wpb - 1) / (8 * wpb));
    const int grid = npred + 8 * ((a.ntiles + 7) / 8);
    if (a.ev_start) {
        hipExtLaunchKernelGGL((k_gmp_ssa_pred<S>), dim3(grid), dim3(a.block), 0, a.stream, a.ev_start, a.ev_stop, 0,
                              a.state, a.lost_in, a.lost_out, a.prob, a.reals, a.tiles, a.ntiles, a.plan,
                              a.seeds_cur, a.seeds_prev, *a.T, *a.P, a.bm, a.cpt, a.hot, npred);
    } else {
        hipLaunchKernelGGL((k_gmp_ssa_pred<S>), dim3(grid), dim3(a.block), 0, a.stream, a.state, a.lost_in,
                           a.lost_out, a.prob, a.reals, a.tiles, a.ntiles, a.plan, a.seeds_cur, a.seeds_prev, *a.T,
                           *a.P, a.bm, a.cpt, a.hot, npred);
    }
    return hipGetLastError();
}

template <int S, int MODE, int G, int GS>
static hipError_t launch_sg(const StepLaunch &a) {
    const int grid = 8 * ((a.ntiles + 7) / 8);
    if (a.ev_start) {
        // start/stop timestamps taken by the dispatch packet itself
        hipExtLaunchKernelGGL((k_gmp_step<S, MODE, G, GS>), dim3(grid), dim3(a.block), 0, a.stream, a.ev_start,
                              a.ev_stop, 0, a.state, a.lost_in, a.lost_out, a.prob, a.reals, a.tiles,
                              a.ntiles, a.plan, a.seeds_cur, a.seeds_prev, *a.T, *a.P, a.bm, a.cpt, a.hot, a.one);
    } else {
        hipLaunchKernelGGL((k_gmp_step<S, MODE, G, GS>), dim3(grid), dim3(a.block), 0, a.stream, a.state,
                           a.lost_in, a.lost_out, a.prob, a.reals, a.tiles, a.ntiles, a.plan, a.seeds_cur,
                           a.seeds_prev, *a.T, *a.P, a.bm, a.cpt, a.hot, a.one);
    }
    return hipGetLastError();
}

template <int S, int G>
static hipError_t launch_xs_hot_g(const StepLaunch &a) {
    const int wpb = a.block / 64;
    // both worker counts are multiples of 8 blocks, so the tiles keep their XCD placement
    const int nlane = a.hot.lane_e ? 8 * ((a.lane_grid + 8 * wpb - 1) / (8 * wpb)) : 0;
    const int nhot = 8 * ((a.hot_grid + 8 * wpb - 1) / (8 * wpb));
    const int grid = nlane + nhot + 8 * ((a.ntiles + 7) / 8);
    if (a.ev_start) {
        hipExtLaunchKernelGGL((k_gmp_xs_hot<S, G>), dim3(grid), dim3(a.block), 0, a.stream, a.ev_start, a.ev_stop, 0,
                              a.state, a.lost_in, a.lost_out, a.prob, a.reals, a.tiles, a.ntiles, a.plan,
                              a.plan_ssa, a.seeds_cur, a.seeds_prev, *a.T, *a.T, *a.P, a.bm, a.cpt, a.hot, nlane,
                              nhot, a.one);
    } else {
        hipLaunchKernelGGL((k_gmp_xs_hot<S, G>), dim3(grid), dim3(a.block), 0, a.stream, a.state, a.lost_in,
                           a.lost_out, a.prob, a.reals, a.tiles, a.ntiles, a.plan, a.plan_ssa, a.seeds_cur,
                           a.seeds_prev, *a.T, *a.T, *a.P, a.bm, a.cpt, a.hot, nlane, nhot, a.one);
    }
    return hipGetLastError();
}

template <int S>
static hipError_t launch_xs_hot(const StepLaunch &a) {
    return (a.grp == 8) ? launch_xs_hot_g<S, 8>(a) : launch_xs_hot_g<S, 4>(a);
}

template <int S, int MODE>
static hipError_t launch_s(const StepLaunch &a) {
    hipError_t e;
    if constexpr (MODE == MODE_INH_XS) {
        if (a.hot.flag) {
            if ((a.grp != 4 && a.grp != 8) || a.grp_sweep != 0 || !a.plan_ssa) return hipErrorInvalidValue;
            return launch_xs_hot<S>(a);
        }
    }
    if constexpr (MODE == MODE_INH_X) {
        if (a.predict) {
            if (!a.hot.flag || a.grp != 1 || a.grp_sweep != 0 || !a.hot.pred) return hipErrorInvalidValue;
            return launch_ssa_pred<S>(a);
        }
    }
    if (a.grp == 8) {  // flattened sweeps over 8 cells per wave: launches without an SSA
        if constexpr (mode_ssa(MODE)) return hipErrorInvalidValue;
        else return (a.grp_sweep == 0) ? launch_sg<S, MODE, 8, 0>(a) : hipErrorInvalidValue;
    }
    if (a.grp != 1 && a.grp != 4) return hipErrorInvalidValue;
    if (a.grp_sweep == 0) e = (a.grp == 4) ? launch_sg<S, MODE, 4, 0>(a) : launch_sg<S, MODE, 1, 0>(a);
    else if (a.grp == 4) e = launch_sg<S, MODE, 4, 4>(a);
    else if constexpr (!mode_ssa(MODE)) e = launch_sg<S, MODE, 1, 1>(a);  // no SSA: lanes = grp
    else if (a.grp_sweep == 4) e = launch_sg<S, MODE, 1, 4>(a);
    else e = launch_sg<S, MODE, 1, 1>(a);
    if constexpr (!mode_ssa(MODE)) {
        return e;
    } else {
        // the hot cells run inside the X-sweep launch when the SSA launch flags them
        if (e != hipSuccess || a.hot.flag) return e;
        // the cells that used up their event budget
        if (a.ev_hot_start) {
            hipExtLaunchKernelGGL((k_gmp_hot<S, MODE>), dim3(a.hot_grid), dim3(64), 0, a.stream, a.ev_hot_start,
                                  a.ev_hot_stop, 0, a.state, a.lost_out, a.prob, a.reals, a.plan, a.seeds_cur,
                                  *a.T, *a.P, a.bm, a.hot);
        } else {
            hipLaunchKernelGGL((k_gmp_hot<S, MODE>), dim3(a.hot_grid), dim3(64), 0, a.stream, a.state,
                               a.lost_out, a.prob, a.reals, a.plan, a.seeds_cur, *a.T, *a.P, a.bm, a.hot);
        }
        return hipGetLastError();
    }
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

}  // inline namespace
}  // namespace gpgmp
