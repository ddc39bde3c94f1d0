// Individual-species lattice diffusion for gfx950: the reference's particle-based
// species (IndividualSolver), whose particles carry their own diffusivity and drift and
// hop between cells with the inhomogeneous drift-diffusion probabilities.
//
// Particles are cell keys sorted stably (the radix sort of particles.hip), with their
// particle index and four properties kept in the same order (k_ind_permute after every
// sort), so every kernel reads them coalesced; off[c] .. off[c+1] is cell c's range.  One thread per cell walks its
// range in sorted order, as individual_diffuseX/Y (IndividualSolver.cl:750-1014) walk
// left[c] .. right[c]: RNG key {cell + species*A, 0xdecafbad+seed}, counter advanced per
// particle of the cell, word x.  The reference rewrites keys in place while an empty
// cell's loop reads the first key of the next cell, so a particle hopping -x into an
// empty cell can be hopped again by that cell's work-item depending on scheduling; here a
// cell owns exactly the particles that were in it when the sweep started.
#include <hip/hip_runtime.h>

#include "gpgmp_numeric.h"
#include "individuals_launch.h"

namespace gpgmp {

// Hop probabilities (IndividualSolver.cl:840-842 / :953-955), float operands with the
// reference's double literals; the same arithmetic as the stochastic sweeps' tables.
__device__ __forceinline__ void ind_hop_probs(float D, float r, float dt, float dx, float &p1, float &p2) {
    const float rrdt = (r * r) * dt;
    const float dx2 = dx * dx;
    const float dxr = dx * r;
    const float ps = (float)(1. - ((double)dt * (2. * (double)D + (double)rrdt)) / (double)dx2);
    const double den = 2. * (double)D + (double)rrdt;
    p1 = (float)((1. - (double)ps) * (1. + (double)dxr / den) / 2.);
    p2 = (float)((double)p1 + (1. - (double)ps) * (1. - (double)dxr / den) / 2.);
}

// individual_diffuseX (AXIS 0) / individual_diffuseY (AXIS 1): a hop moves the key by
// -/+1 (X) or -/+W (Y) times the direction's boundary mask (0 reflective: the particle
// stays; 1 absorbing: the key leaves the cell range or wraps, as in the reference).
template <int AXIS>
__global__ __launch_bounds__(256) void k_ind_sweep(uint32_t *__restrict__ keys, const uint32_t *__restrict__ off,
                                                   const float *__restrict__ props, int cap, int W, int H,
                                                   uint32_t key_off, uint32_t k1, float dt, float dx, int4 bm) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= W * H) return;
    const int gx = c % W, gy = c / W;
    const uint32_t mlo = AXIS == 0 ? (gx == 0 ? (uint32_t)bm.x : 1u) : (gy == 0 ? (uint32_t)bm.z : 1u);
    const uint32_t mhi = AXIS == 0 ? (gx == W - 1 ? (uint32_t)bm.y : 1u) : (gy == H - 1 ? (uint32_t)bm.w : 1u);
    const uint32_t step = AXIS == 0 ? 1u : (uint32_t)W;
    const uint32_t b = off[c], e = off[c + 1];
    uint32_t ctr = 0;
    for (uint32_t i = b; i < e; ++i) {
        float p1, p2;
        ind_hop_probs(props[i + (size_t)AXIS * cap], props[i + (size_t)(2 + AXIS) * cap], dt, dx, p1, p2);
        ++ctr;
        const u32x2 w = threefry2x32_20(ctr, kCtrHi, (uint32_t)c + key_off, k1);
        const float u = u01(w.x);
        uint32_t key = (uint32_t)c;
        if (u <= p2) key = (u <= p1) ? key + step * mhi : key - step * mlo;
        keys[i] = key;
    }
}

// After a sort of (key, position) pairs: carry the particle indices and properties into
// the sorted order (perm[i] = the position the i-th sorted particle came from), and reset
// the positions for the next sort.
__global__ __launch_bounds__(256) void k_ind_permute(uint32_t *__restrict__ perm, const uint32_t *__restrict__ ids,
                                                     uint32_t *__restrict__ ids_out,
                                                     const float *__restrict__ props, float *__restrict__ props_out,
                                                     int n, int cap) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t j = perm[i];
    ids_out[i] = ids[j];
#pragma unroll
    for (int k = 0; k < 4; ++k) props_out[i + (size_t)k * cap] = props[j + (size_t)k * cap];
    perm[i] = (uint32_t)i;
}

// OpenCL maxmag(x, y): the argument of larger magnitude, fmax on ties.
__device__ __forceinline__ float maxmag(float x, float y) {
    const float ax = fabsf(x), ay = fabsf(y);
    return ax > ay ? x : (ay > ax ? y : fmaxf(x, y));
}

// computeDiffusionConstants (IndividualSolver.cl:654-747) with the identity drift /
// diffusivity hook: per cell the largest diffusivities, the largest-magnitude drifts and
// the particle count (written into the species' state).
__global__ __launch_bounds__(256) void k_ind_constants(const uint32_t *__restrict__ off,
                                                       const float *__restrict__ props, int cap, int A,
                                                       float *__restrict__ Dx, float *__restrict__ Dy,
                                                       float *__restrict__ rx, float *__restrict__ ry,
                                                       int32_t *__restrict__ count) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= A) return;
    float dx = 0.f, dy = 0.f, qx = 0.f, qy = 0.f;
    const uint32_t b = off[c], e = off[c + 1];
    for (uint32_t i = b; i < e; ++i) {
        dx = fmaxf(props[i], dx);
        dy = fmaxf(props[i + (size_t)cap], dy);
        qx = maxmag(props[i + 2 * (size_t)cap], qx);
        qy = maxmag(props[i + 3 * (size_t)cap], qy);
    }
    Dx[c] = dx;
    Dy[c] = dy;
    rx[c] = qx;
    ry[c] = qy;
    count[c] = (int32_t)(e - b);
}

// Minimum per-cell time step (ind_cell_dt): block minima, then one block over them.
// min is exact and order-independent.
__device__ __forceinline__ float block_min_f(float v, float *sm) {
    for (int o = 32; o > 0; o >>= 1) v = fminf(v, __shfl_down(v, o, 64));
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) sm[wave] = v;
    __syncthreads();
    float m = sm[0];
    for (int w = 1; w < (int)(blockDim.x >> 6); ++w) m = fminf(m, sm[w]);
    return m;
}

__global__ __launch_bounds__(256) void k_ind_dt_blocks(const float *__restrict__ D, const float *__restrict__ r, int A,
                                                       float h, float *__restrict__ blk) {
    __shared__ float sm[4];
    float v = __builtin_inff();
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < A; i += gridDim.x * blockDim.x)
        v = fminf(v, ind_cell_dt(D[i], r[i], h));
    const float m = block_min_f(v, sm);
    if (threadIdx.x == 0) blk[blockIdx.x] = m;
}

__global__ __launch_bounds__(256) void k_ind_dt_reduce(const float *__restrict__ blk, int nb, float *__restrict__ out) {
    __shared__ float sm[4];
    float v = __builtin_inff();
    for (int i = threadIdx.x; i < nb; i += blockDim.x) v = fminf(v, blk[i]);
    const float m = block_min_f(v, sm);
    if (threadIdx.x == 0) *out = m;
}

hipError_t launch_ind_sweep(int axis, uint32_t *keys, const uint32_t *off, const float *props, int cap, int W, int H,
                            uint32_t key_off, uint32_t k1, float dt, float dx, int4 bm, hipStream_t s) {
    const int A = W * H;
    const dim3 g((A + 255) / 256), b(256);
    if (axis == 0)
        hipLaunchKernelGGL(k_ind_sweep<0>, g, b, 0, s, keys, off, props, cap, W, H, key_off, k1, dt, dx, bm);
    else
        hipLaunchKernelGGL(k_ind_sweep<1>, g, b, 0, s, keys, off, props, cap, W, H, key_off, k1, dt, dx, bm);
    return hipGetLastError();
}

hipError_t launch_ind_constants(const uint32_t *off, const float *props, int cap, int A, float *Dx, float *Dy,
                                float *rx, float *ry, int32_t *count, hipStream_t s) {
    hipLaunchKernelGGL(k_ind_constants, dim3((A + 255) / 256), dim3(256), 0, s, off, props, cap, A, Dx, Dy, rx, ry,
                       count);
    return hipGetLastError();
}

hipError_t launch_ind_permute(uint32_t *perm, const uint32_t *ids, uint32_t *ids_out, const float *props,
                              float *props_out, int n, int cap, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_ind_permute, dim3((n + 255) / 256), dim3(256), 0, s, perm, ids, ids_out, props, props_out, n,
                       cap);
    return hipGetLastError();
}

hipError_t launch_ind_dt(const float *D, const float *r, int A, float h, float *blk, int nb, float *out,
                         hipStream_t s) {
    hipLaunchKernelGGL(k_ind_dt_blocks, dim3(nb), dim3(256), 0, s, D, r, A, h, blk);
    hipLaunchKernelGGL(k_ind_dt_reduce, dim3(1), dim3(256), 0, s, blk, nb, out);
    return hipGetLastError();
}

}  // namespace gpgmp
