// Host-side launch interface of the individual-species kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace gpgmp {

// one hop sweep (axis 0: X, 1: Y) over the sorted particles (properties in sorted order);
// off = CSR cell offsets [A+1]
hipError_t launch_ind_sweep(int axis, uint32_t *keys, const uint32_t *off, const float *props, int cap, int W, int H,
                            uint32_t key_off, uint32_t k1, float dt, float dx, int4 bm, hipStream_t s);
// per-cell max diffusivities, max-magnitude drifts and counts
hipError_t launch_ind_constants(const uint32_t *off, const float *props, int cap, int A, float *Dx, float *Dy,
                                float *rx, float *ry, int32_t *count, hipStream_t s);
// carry ids and properties into the order of the last sort (perm: source positions, reset to
// 0..n-1 for the next sort)
hipError_t launch_ind_permute(uint32_t *perm, const uint32_t *ids, uint32_t *ids_out, const float *props,
                              float *props_out, int n, int cap, hipStream_t s);
// minimum per-cell time step into *out (device float); blk: nb floats of scratch
hipError_t launch_ind_dt(const float *D, const float *r, int A, float h, float *blk, int nb, float *out,
                         hipStream_t s);

// ---- population maintenance (IndividualSolver::maintainPopulation) ----
// tiles of cells per block; `tiles` holds 3 * mp_tiles(A) words
int mp_tiles(int A);
// kills, per-cell new/kill counts and the scanned tile sums; out[0..4] = total goal, total
// new, total kills, first index outside the grid (off[A]), 1 if a goal was negative (*err
// is zero before and after: it must start zeroed)
hipError_t launch_ind_mp_count(const uint32_t *off, const int32_t *goal, int A, uint8_t *dead, uint32_t *nnew,
                               uint32_t *nkill, uint32_t *tiles, uint32_t *err, uint32_t *out, uint32_t k1,
                               hipStream_t s);
struct MaintainArgs {
    int A, cap;
    uint32_t n, np, total_goal, total_new, total_kill, first_outside;
    const uint32_t *off, *nnew, *nkill, *tiles, *keys, *ids;
    const int32_t *goal;
    const uint8_t *dead;
    const float *props;
    uint32_t *keys2, *ids2, *pos, *off2, *pool_ids;
    float *props2, *pool_props;
    int use_rule;
    int32_t mode[4];
    double a[4], b[4];
    uint32_t k1_new;
};
// new particles from the pool, survivors, killed particles onto the pool, new offsets
hipError_t launch_ind_mp_apply(const MaintainArgs &m, hipStream_t s);

}  // namespace gpgmp
