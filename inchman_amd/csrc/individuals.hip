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
#include "wave_ops.h"

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

// ---- population maintenance after a reaction step (IndividualSolver::maintainPopulation) ----
//
// The live particles [0, n) are sorted by cell with their ids and properties in the same
// order; the pool of dead particles is a stack (pool_ids / pool_props, top = front of the
// reference's dead list at positions n, n+1, ... of its key list).  One maintenance:
//   k_ind_mp_cells    per cell: kills (the reference's pick loop), new-particle and kill
//                     counts, per-tile sums of (goal, new, kills)
//   k_ind_mp_tiles    exclusive scan of the tile sums, totals, the first out-of-grid index
//   k_ind_mp_new      per cell: its new particles pop ids and properties off the pool top
//                     in cell order and take the newIndividuals rule; written first in
//                     the cell's new range
//   k_ind_mp_scatter  per cell: survivors after the new particles, killed particles pushed
//                     onto the pool in list order (the first killed ends on top), new CSR
//                     offsets
//   k_ind_mp_outside  live particles with keys outside the grid, after the last cell
// The cells are tiles of MP_T consecutive cells, one block each; offsets come from a
// block-wide scan plus the tile's scanned base, so no per-cell scan array is stored.
constexpr int MP_T = 256;

__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) { return wave_incl_scan_u32(v); }

// exclusive scan of one value per thread over the MP_T-thread block; sm: MP_T/64 words
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t *sm, uint32_t &total) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint32_t inc = wave_incl_scan(v);
    if (lane == 63) sm[w] = inc;
    __syncthreads();
    uint32_t base = 0, tot = 0;
#pragma unroll
    for (int i = 0; i < MP_T / 64; ++i) {
        base += (i < w) ? sm[i] : 0u;
        tot += sm[i];
    }
    __syncthreads();
    total = tot;
    return base + inc - v;
}

// x, y, x, y, ... of successive Threefry calls with counter 1, 2, ... (maintainPopulation's
// kill loop and the RANDOM macro, IndividualSolver.cl:1059-1075, 1262-1282)
struct PairStream {
    uint32_t k0, k1, ri, ctr;
    u32x2 w;
    __device__ PairStream(uint32_t a, uint32_t b) : k0(a), k1(b), ri(0), ctr(0), w{0, 0} {}
    __device__ __forceinline__ uint32_t next() {
        uint32_t r;
        if (!(ri & 1u)) {
            ++ctr;
            w = threefry2x32_20(ctr, kCtrHi, k0, k1);
            r = w.x;
        } else {
            r = w.y;
        }
        ++ri;
        return r;
    }
};

__global__ __launch_bounds__(MP_T) void k_ind_mp_cells(const uint32_t *__restrict__ off, const int32_t *__restrict__ goal,
                                                       int A, uint8_t *__restrict__ dead, uint32_t *__restrict__ nnew,
                                                       uint32_t *__restrict__ nkill, uint32_t *__restrict__ tiles,
                                                       int nt, uint32_t *__restrict__ err, uint32_t k1) {
    __shared__ uint32_t sm[MP_T / 64];
    const int c = blockIdx.x * MP_T + threadIdx.x;
    uint32_t g = 0, nw = 0, nk = 0;
    if (c < A) {
        int32_t gi = goal[c];
        if (gi < 0) {
            atomicOr(err, 1u);
            gi = 0;
        }
        g = (uint32_t)gi;
        const uint32_t b = off[c], is = off[c + 1] - b;
        for (uint32_t i = 0; i < is; ++i) dead[b + i] = 0;
        if (is > g) {
            nk = is - g;
            // the reference's pick: left + (uint)(u * is), u in double, redrawn while the
            // pick is already dead; a pick rounded up to is (only for is > 2^21) is clamped
            PairStream rs((uint32_t)c, k1);
            for (uint32_t k = 0; k < nk; ++k) {
                uint32_t r;
                do {
                    r = (uint32_t)(u01_d(rs.next()) * (double)is);
                    r = r < is ? r : is - 1u;
                } while (dead[b + r]);
                dead[b + r] = 1;
            }
        } else {
            nw = g - is;
        }
        nnew[c] = nw;
        nkill[c] = nk;
    }
    uint32_t t0, t1, t2;
    (void)block_excl_scan(g, sm, t0);
    (void)block_excl_scan(nw, sm, t1);
    (void)block_excl_scan(nk, sm, t2);
    if (threadIdx.x == 0) {
        tiles[blockIdx.x] = t0;
        tiles[nt + blockIdx.x] = t1;
        tiles[2 * nt + blockIdx.x] = t2;
    }
}

// one block: exclusive scans of the three tile-sum rows in place; out[0..2] = totals,
// out[3] = off[A] (first particle outside the grid), out[4] = the error flag (then reset)
__global__ __launch_bounds__(MP_T) void k_ind_mp_tiles(uint32_t *__restrict__ tiles, int nt,
                                                       const uint32_t *__restrict__ off, int A,
                                                       uint32_t *__restrict__ err, uint32_t *__restrict__ out) {
    __shared__ uint32_t sm[MP_T / 64];
    for (int row = 0; row < 3; ++row) {
        uint32_t *t = tiles + (size_t)row * nt;
        uint32_t carry = 0;
        for (int b = 0; b < nt; b += MP_T) {
            const int i = b + threadIdx.x;
            const uint32_t v = i < nt ? t[i] : 0u;
            uint32_t tot;
            const uint32_t ex = block_excl_scan(v, sm, tot);
            if (i < nt) t[i] = carry + ex;
            carry += tot;
        }
        if (threadIdx.x == 0) out[row] = carry;
    }
    if (threadIdx.x == 0) {
        out[3] = off[A];
        out[4] = *err;  // negative-goal flag of this maintenance, cleared for the next
        *err = 0u;
    }
}

struct NewRule {
    int32_t mode[4];
    double a[4], b[4];
};

__global__ __launch_bounds__(MP_T) void k_ind_mp_new(const int32_t *__restrict__ goal, const uint32_t *__restrict__ nnew,
                                                     const uint32_t *__restrict__ tiles, int nt, int A,
                                                     const uint32_t *__restrict__ pool_ids,
                                                     const float *__restrict__ pool_props, uint32_t np, int cap,
                                                     uint32_t *__restrict__ keys2, uint32_t *__restrict__ ids2,
                                                     float *__restrict__ props2, uint32_t *__restrict__ pos,
                                                     NewRule rule, int use_rule, uint32_t k1) {
    __shared__ uint32_t sm[MP_T / 64];
    const int c = blockIdx.x * MP_T + threadIdx.x;
    const uint32_t g = c < A ? (uint32_t)max(goal[c], 0) : 0u;
    const uint32_t nw = c < A ? nnew[c] : 0u;
    uint32_t t;
    const uint32_t base = tiles[blockIdx.x] + block_excl_scan(g, sm, t);
    const uint32_t q0 = tiles[nt + blockIdx.x] + block_excl_scan(nw, sm, t);
    if (c >= A || nw == 0) return;
    PairStream rs((uint32_t)c, k1);
    for (uint32_t j = 0; j < nw; ++j) {
        const uint32_t src = np - 1u - (q0 + j), dst = base + j;
        keys2[dst] = (uint32_t)c;
        ids2[dst] = pool_ids[src];
        pos[dst] = dst;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            float v = pool_props[src + (size_t)k * cap];
            if (use_rule) {
                if (rule.mode[k] == 1) v = (float)rule.a[k];
                if (rule.mode[k] == 2) v = (float)(rule.a[k] + rule.b[k] * u01_d(rs.next()));
            }
            props2[dst + (size_t)k * cap] = v;
        }
    }
}

__global__ __launch_bounds__(MP_T) void k_ind_mp_scatter(
    const uint32_t *__restrict__ off, const uint8_t *__restrict__ dead, const int32_t *__restrict__ goal,
    const uint32_t *__restrict__ nnew, const uint32_t *__restrict__ nkill, const uint32_t *__restrict__ tiles, int nt,
    int A, const uint32_t *__restrict__ ids, const float *__restrict__ props, int cap, uint32_t *__restrict__ keys2,
    uint32_t *__restrict__ ids2, float *__restrict__ props2, uint32_t *__restrict__ pos,
    uint32_t *__restrict__ pool_ids, float *__restrict__ pool_props, uint32_t np_popped, uint32_t nkill_total,
    uint32_t *__restrict__ off2) {
    __shared__ uint32_t sm[MP_T / 64];
    const int c = blockIdx.x * MP_T + threadIdx.x;
    const uint32_t g = c < A ? (uint32_t)max(goal[c], 0) : 0u;
    const uint32_t nk = c < A ? nkill[c] : 0u;
    uint32_t t;
    const uint32_t base = tiles[blockIdx.x] + block_excl_scan(g, sm, t);
    const uint32_t kb = tiles[2 * nt + blockIdx.x] + block_excl_scan(nk, sm, t);
    if (c >= A) return;
    off2[c] = base;
    if (c == A - 1) off2[A] = base + g;
    const uint32_t b = off[c], e = off[c + 1];
    uint32_t dst = base + nnew[c], r = kb;
    for (uint32_t i = b; i < e; ++i) {
        if (dead[i]) {
            // killed particles in list order onto the pool, the first killed on top
            const uint32_t pp = np_popped + (nkill_total - 1u - r);
            ++r;
            pool_ids[pp] = ids[i];
#pragma unroll
            for (int k = 0; k < 4; ++k) pool_props[pp + (size_t)k * cap] = props[i + (size_t)k * cap];
        } else {
            keys2[dst] = (uint32_t)c;
            ids2[dst] = ids[i];
            pos[dst] = dst;
#pragma unroll
            for (int k = 0; k < 4; ++k) props2[dst + (size_t)k * cap] = props[i + (size_t)k * cap];
            ++dst;
        }
    }
}

__global__ __launch_bounds__(256) void k_ind_mp_outside(uint32_t first, uint32_t n, uint32_t dst0,
                                                        const uint32_t *__restrict__ keys,
                                                        const uint32_t *__restrict__ ids,
                                                        const float *__restrict__ props, int cap,
                                                        uint32_t *__restrict__ keys2, uint32_t *__restrict__ ids2,
                                                        float *__restrict__ props2, uint32_t *__restrict__ pos) {
    const uint32_t i = first + blockIdx.x * 256u + threadIdx.x;
    if (i >= n) return;
    const uint32_t dst = dst0 + (i - first);
    keys2[dst] = keys[i];
    ids2[dst] = ids[i];
    pos[dst] = dst;
#pragma unroll
    for (int k = 0; k < 4; ++k) props2[dst + (size_t)k * cap] = props[i + (size_t)k * cap];
}

int mp_tiles(int A) { return (A + MP_T - 1) / MP_T; }

hipError_t launch_ind_mp_count(const uint32_t *off, const int32_t *goal, int A, uint8_t *dead, uint32_t *nnew,
                               uint32_t *nkill, uint32_t *tiles, uint32_t *err, uint32_t *out, uint32_t k1,
                               hipStream_t s) {
    const int nt = mp_tiles(A);
    hipLaunchKernelGGL(k_ind_mp_cells, dim3(nt), dim3(MP_T), 0, s, off, goal, A, dead, nnew, nkill, tiles, nt, err, k1);
    hipLaunchKernelGGL(k_ind_mp_tiles, dim3(1), dim3(MP_T), 0, s, tiles, nt, off, A, err, out);
    return hipGetLastError();
}

hipError_t launch_ind_mp_apply(const MaintainArgs &m, hipStream_t s) {
    const int nt = mp_tiles(m.A);
    NewRule rule{};
    for (int k = 0; k < 4; ++k) {
        rule.mode[k] = m.mode[k];
        rule.a[k] = m.a[k];
        rule.b[k] = m.b[k];
    }
    if (m.total_new)
        hipLaunchKernelGGL(k_ind_mp_new, dim3(nt), dim3(MP_T), 0, s, m.goal, m.nnew, m.tiles, nt, m.A, m.pool_ids,
                           m.pool_props, m.np, m.cap, m.keys2, m.ids2, m.props2, m.pos, rule, m.use_rule, m.k1_new);
    hipLaunchKernelGGL(k_ind_mp_scatter, dim3(nt), dim3(MP_T), 0, s, m.off, m.dead, m.goal, m.nnew, m.nkill, m.tiles,
                       nt, m.A, m.ids, m.props, m.cap, m.keys2, m.ids2, m.props2, m.pos, m.pool_ids, m.pool_props,
                       m.np - m.total_new, m.total_kill, m.off2);
    if (m.n > m.first_outside) {
        const uint32_t cnt = m.n - m.first_outside;
        hipLaunchKernelGGL(k_ind_mp_outside, dim3((cnt + 255) / 256), dim3(256), 0, s, m.first_outside, m.n,
                           m.total_goal, m.keys, m.ids, m.props, m.cap, m.keys2, m.ids2, m.props2, m.pos);
    }
    return hipGetLastError();
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
