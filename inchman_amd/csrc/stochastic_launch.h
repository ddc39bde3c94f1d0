// Host-side launch interface of the stochastic GMP kernels.
#pragma once

#include <hip/hip_runtime.h>

#include "stochastic_types.h"

namespace gpgmp {

struct StepLaunch {
    int mode;               // 0 homogeneous, 1 inhomogeneous X sweeps (+SSA), 2 inhomogeneous Y sweeps,
                            // 3 inhomogeneous X sweeps without SSA (+ the deferred cells when
                            // hot.flag is set)
    int32_t *state;
    const int32_t *lost_in;
    int32_t *lost_out;
    const int4 *prob;       // inhomogeneous hop thresholds of this launch's axis
    const RealInfo *reals;
    const Tile *tiles;
    int ntiles;
    const LaunchPlan *plan; // device pointer, one plan for all realizations
    const int32_t *seeds_cur;   // [num realizations] seed_n of this step
    const int32_t *seeds_prev;  // [num realizations] seed of the previous step (source boundaries)
    const RxTable *T;
    const HomProbs *P;
    int4 bm;                // boundary masks -x, +x, -y, +y
    int block;              // threads per block (64..256, multiple of 64)
    int cpt;                // cells per thread
    int grp = 1;            // lanes per cell of the SSA (1 or 4)
    int grp_sweep = 1;      // outflow: lanes per cell (4, or grp), or 0 = flattened over the wave
    hipStream_t stream;
    hipEvent_t ev_start = nullptr;  // optional: timestamps of this dispatch
    hipEvent_t ev_stop = nullptr;
    HotList hot{};                  // deferred busy cells (modes with an SSA)
    int hot_grid = 1024;            // waves of the hot-cell kernel
    int lane_grid = 0;              // lane-list worker waves (X sweeps with the hot cells)
    const LaunchPlan *plan_ssa = nullptr;  // X sweeps with the hot cells: the SSA launch's plan
    int predict = 0;        // SSA launch of a split step: run last step's busy cells in worker waves
    hipEvent_t ev_hot_start = nullptr;  // optional: timestamps of the hot-cell dispatch
    hipEvent_t ev_hot_stop = nullptr;
    OneReal one{};          // one.per > 0: a single realization, tiles of one.per cells computed
};

hipError_t launch_gmp_step(const StepLaunch &a);
// per-mode instantiations (stochastic_hom.hip, stochastic_inh_x.hip, stochastic_inh_y.hip)
hipError_t launch_gmp_hom(const StepLaunch &a);
hipError_t launch_gmp_inh_x(const StepLaunch &a);
hipError_t launch_gmp_inh_y(const StepLaunch &a);
hipError_t launch_gmp_inh_xs(const StepLaunch &a);
// the extended instances: localized reactions, float literals, law programs (stochastic_*_cl.hip)
hipError_t launch_gmp_hom_cl(const StepLaunch &a);
hipError_t launch_gmp_inh_x_cl(const StepLaunch &a);
hipError_t launch_gmp_inh_xs_cl(const StepLaunch &a);
hipError_t launch_min_dt(const float *D, const float *r, int n, float h, float *blocks, float *out,
                         hipStream_t stream);
hipError_t launch_inh_probs(int4 *out, const float *D, const float *r, float dt, float dx, int n, int fpe,
                            hipStream_t stream);

}  // namespace gpgmp
