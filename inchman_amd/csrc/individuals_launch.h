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

}  // namespace gpgmp
