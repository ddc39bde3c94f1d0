// Drives gpgmp_hip::HipSolver the way DiffusionModel::oclSolve drives gpgmp::Solver
// (DiffusionModel.cpp:1467-1603: seeds from boost::mt19937 + uniform_int<>(0, INT_MAX),
// minimum next-diffusion time, doReactions, diffuse of the due species) and checks the
// counts against the library's fused host loop (gpgmp_solver_run) from the same state.
// Exit status 0 when they agree bit for bit.
#include <climits>
#include <cstdio>
#include <random>
#include <vector>

#include "gpgmp_hip_solver.hpp"

int main(int argc, char **argv) {
    const int W = 48, H = 32, S = 3, steps = argc > 1 ? std::atoi(argv[1]) : 300;
    gpgmp_reaction_desc rx[4] = {};
    rx[0].rate = 1.0f; rx[0].product_stoich[0] = 1;                                  // 0 -> A
    rx[1].rate = 1.0f; rx[1].product_stoich[1] = 1;                                  // 0 -> B
    rx[2].rate = 0.01f; rx[2].reactant_stoich[0] = 1; rx[2].reactant_stoich[1] = 1;  // A + B -> C
    rx[2].product_stoich[2] = 1;
    rx[3].rate = 0.1f; rx[3].reactant_stoich[2] = 1;                                 // C -> 0
    gpgmp_model_desc d = {};
    d.width = W; d.height = H; d.physical_length = (float)W; d.p0 = 0.15f;
    d.num_species = S; d.diffusion[0] = 1.0f; d.diffusion[1] = 0.5f; d.diffusion[2] = 0.1f;
    d.num_reactions = 4; d.reactions = rx; d.max_time = 1e30f;
    d.solver_type = GPGMP_STOCHASTIC_HOMOGENEOUS;
    std::vector<int> init((size_t)S * W * H, 0);
    unsigned lcg = 12345u;
    for (int i = 0; i < 2 * W * H; ++i) { lcg = lcg * 1664525u + 1013904223u; init[i] = 40 + (int)(lcg >> 28); }
    const unsigned seed0 = 77;
    // (1) the reference's host loop through the adapter
    std::vector<int> state = init;
    gpgmp_hip::HipSolver solver(d, &state);
    std::mt19937 rng(seed0);
    std::vector<float> dt = solver.getNextDiffusionTime();
    float simTime = 0.f;
    for (int n = 0; n < steps; ++n) {
        int seed = (int)((unsigned)(rng() >> 1) + (unsigned)n);
        float mindt = dt[0];
        int minspec = 0;
        for (int i = 1; i < S; ++i)
            if (dt[i] < mindt) { mindt = dt[i]; minspec = i; }
        solver.doReactions(mindt - simTime, seed);
        for (int i = 0; i < S; ++i)
            if (i == minspec || dt[i] <= mindt)
                if (dt[i] > 0) solver.diffuse(i, mindt, seed);
        dt = solver.getNextDiffusionTime();
        simTime = mindt;
    }
    solver.synchronizeStateBufferFromDevice();
    // (2) the fused device loop
    std::vector<int> state2 = init;
    gpgmp_hip::HipSolver fused(d, &state2);
    const float t2 = fused.run(steps, seed0);
    fused.synchronizeStateBufferFromDevice();
    long long diff = 0, total = 0;
    for (size_t i = 0; i < state.size(); ++i) { diff += state[i] != state2[i]; total += state[i]; }
    std::printf("host loop vs fused: %lld differing counts, t %.6f vs %.6f, total %lld\n", diff, simTime, t2, total);
    return (diff == 0 && simTime == t2) ? 0 : 1;
}
