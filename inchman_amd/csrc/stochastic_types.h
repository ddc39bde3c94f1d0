// Shared host/device types of the stochastic GMP kernels.
#pragma once

#include <stdint.h>

namespace gpgmp {

constexpr int kMaxSpecies = 8;
constexpr int kMaxReactions = 16;
constexpr int kMaxFactors = 6;
// Small models (at most 4 reactions of at most 3 factors each) take a branch-free form of
// the hazard and selection code in the event loops.
constexpr int kCompactReactions = 4;
constexpr int kCompactFactors = 3;
// Upper bound on SSA events per cell per reaction step: a runaway model (hazards so
// large that t += dt stops advancing in float) ends instead of spinning forever.
constexpr uint32_t kMaxEventsPerCall = 1u << 24;

// Reaction table compiled from the model (the reference generates the same
// arithmetic as kernel source, Solver.cpp:843-905 and Reaction.cpp:119-138).
// Passed to kernels by value so every field is wave-uniform (scalar loads).
struct RxTable {
    int32_t S, R;
    uint32_t lit_is_int;               // bit i: rate literal of reaction i is an int literal
    uint32_t lit_is_float;             // bit i: ... a float (a user parameter, `const Real k`)
    int32_t lit_int[kMaxReactions];
    float lit_flt[kMaxReactions];
    double lit_dbl[kMaxReactions];     // every literal's value as a double
    uint8_t nf[kMaxReactions];         // number of falling-factorial factors
    uint8_t fs[kMaxReactions][kMaxFactors];  // factor species
    uint8_t fo[kMaxReactions][kMaxFactors];  // factor offset m of (X - m)
    int8_t net[kMaxReactions][kMaxSpecies];  // product - reactant stoichiometry
    int32_t compact;  // R <= kCompactReactions and every nf <= kCompactFactors
    // Localized reactions (Reaction::addCompartment, Reaction.cpp:209-235): bit i set when
    // reaction i carries a compartment mask; act[cell] (cells of all realizations, cell_off
    // + c) has bit i set where reaction i's mask w(x, y) is 1.  act is read only when
    // localized != 0, and compact is then 0.
    uint32_t localized;
    const uint16_t *act;
    // Custom laws that are not of the mass-action shape (law.h): bit i set when reaction i's
    // hazard `_h[0] + [mask *] <law>` is the program at prog + prog_off[i] (result type
    // prog_type[i]); compact is then 0.
    uint32_t custom;
    const uint32_t *prog;
    uint16_t prog_off[kMaxReactions];
    uint8_t prog_type[kMaxReactions];
    // Deterministic right-hand sides as programs (det_custom != 0: a model with custom laws):
    // per species the RK4 derivative `0. +- n. * [mask *] <law> ...` (Solver.cpp:918-955) and
    // the alpha-QSS production and destruction sums (Solver.cpp:960-1010).
    int32_t det_custom;
    int32_t grid_w;  // deterministic kernels: grid width (GridX / GridY of the programs)
    uint16_t det_rk[kMaxSpecies], det_q[kMaxSpecies], det_d[kMaxSpecies];
    uint8_t det_rk_t[kMaxSpecies], det_q_t[kMaxSpecies], det_d_t[kMaxSpecies];
};

// One realization (grid) of a launch.  The plain solver is a batch of one.
struct RealInfo {
    uint64_t state_off;   // int32 elements: state[state_off + s*A + cell]
    uint64_t lost_off;    // int32 elements into each lost buffer: + (s*ndir + d)*A + cell
    uint64_t prob_off;    // float2 elements into the inhomogeneous probability tables
    uint64_t cell_off;    // cells of the realizations before this one
    int32_t W, H;
    int32_t src[4];       // source-boundary counts per boundary (0 = none)
};

// Work of one launch, shared by all realizations of the launch (they follow one
// schedule).  Seeds are offsets from the realization's host-loop seed of the step
// (seed_n of DiffusionModel.cpp:1482): the SSA uses seed_n itself, the diffusion
// sweeps seed_n + offset.  deltat <= 0 disables the SSA.
struct alignas(16) LaunchPlan {
    float deltat;
    uint32_t gather_mask;   // species whose pending outflow of the previous launch is gathered
    uint32_t out_mask;      // species whose outflow is computed in this launch
    int32_t gather_off;     // source-boundary seed offset, relative to the previous step's seed
    int32_t out_off[kMaxSpecies];
    int32_t interleave;     // flattened sweeps, n = 64/G cells per wave: wave p's slot q is
                            // cell q*(A/n) + p (spread over the grid) instead of n*p + q
    int32_t predict;        // split steps: the SSA launch runs last step's busy cells in
                            // worker waves (hot lists alternate by step parity)
    int32_t hot_parity;     // this step's hot list (0 or 1) when predicting
    int32_t reset_hot;      // Y launch of a split step: empty hot list hot_parity (one thread)
};

// A tile: cells [cell0, cell0 + cells) of realization r.
struct Tile {
    int32_t r;
    int32_t cell0;
};

// A launch over a single realization (per > 0): tile t is cells [t*per, (t+1)*per) of
// realization 0, computed instead of loaded, and the realization's info comes by value, so
// a wave starts its state loads without two dependent loads (tile, then realization).
struct OneReal {
    RealInfo r;
    int32_t per;
};

// A cell whose SSA used up its per-lane event budget in k_gmp_step: the hot-cell
// kernel continues it from (time, ctr) and then computes its outflow.
struct HotEntry {
    int32_t r;      // realization
    int32_t c;      // cell in the realization
    float time;     // SSA clock reached
    uint32_t ctr;   // Threefry counter reached
};

struct HotList {
    HotEntry *e;
    uint32_t *count;  // [0] entries pushed, [1] blocks of the hot kernel finished
    uint32_t cap;
    uint32_t budget;  // SSA events a lane runs before it may defer its cell
    int32_t below_active;  // ... which it does once fewer lanes of its wave still run
    // per cell (cell_off + c): 1 if the cell is in the list.  Non-null only when the
    // hot cells run inside the X-sweep launch: the SSA launch then writes every cell's
    // flag and the sweep waves skip the flagged cells.
    uint8_t *flag;
    // predicted busy cells: two lists (e[0..cap), e[cap..2cap); count[2p], count[2p+1])
    // and pred[p * pred_stride + cell_off + c] = 1 if the cell is in list p
    uint8_t *pred;
    uint32_t pred_stride;
    // Lane list (hot cells fused into the X-sweep launch, one realization, no prediction):
    // a deferred cell whose expected remaining events h0 * (deltat - time) are at most
    // lane_events goes here instead of e and is resumed by one lane of a worker wave
    // (64 such cells per wave) -- the cheapest VALU work per event for the many modest
    // tails of a dense step -- while the busier ones keep one wave each.  Null: no lane list.
    HotEntry *lane_e;
    uint32_t *lane_count;  // [0] entries pushed
    uint32_t lane_cap;
    float lane_events;
};

// Homogeneous diffusion thresholds (StochasticHomogeneousDiffusion.cl:58-62)
struct HomProbs {
    float stay, p1, p2, p3, p4;
};

}  // namespace gpgmp
