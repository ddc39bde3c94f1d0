// Individual-particle kernels for gfx950: stable LSD radix sort of (key, value) pairs,
// cell offsets, and the particle step (move, sort, offsets, in-cell A+A->0, compaction).
//
// The reference sorts KEY_INDEX_PAIR(key, id) by the low 32 bits with 4-bit digits,
// warp-synchronous 32-lane split scans and a 3-kernel exclusive scan per digit pass
// (IndividualSolver.cl:36-553, IndividualSolver.cpp:673-808).  Here a pass uses 8-bit
// digits and wave64 ranking: the lanes holding the same digit are found with eight
// __ballot calls, a lane's rank is the popcount of its lower peers, and the per-wave
// digit counts are combined in LDS in (round, wave, lane) order, which is the original
// index order -- so every pass is stable and the result equals std::stable_sort.
#include <hip/hip_runtime.h>

#include "gpgmp_numeric.h"
#include "wave_ops.h"

namespace gpgmp {

constexpr int RS_THREADS = 256;
constexpr int RS_ITEMS = 16;
constexpr int RS_TILE = RS_THREADS * RS_ITEMS;  // 4096 keys per tile
constexpr int RS_WAVES = RS_THREADS / 64;

__device__ __forceinline__ int load_n(const int *d_n, int n_host) { return d_n ? *d_n : n_host; }

// Per-tile digit histogram of one pass: hist[d * ntiles + tile].
__global__ __launch_bounds__(RS_THREADS) void k_radix_hist(const uint32_t *__restrict__ keys, const int *d_n,
                                                            int n_host, int shift, uint32_t *__restrict__ hist,
                                                            int ntiles) {
    __shared__ uint32_t cnt[256];
    const int n = load_n(d_n, n_host);
    cnt[threadIdx.x] = 0;
    __syncthreads();
    const size_t base = (size_t)blockIdx.x * RS_TILE;
#pragma unroll 4
    for (int j = 0; j < RS_ITEMS; ++j) {
        const size_t i = base + (size_t)j * RS_THREADS + threadIdx.x;
        if (i < (size_t)n) atomicAdd(&cnt[(keys[i] >> shift) & 255u], 1u);
    }
    __syncthreads();
    hist[(size_t)threadIdx.x * ntiles + blockIdx.x] = cnt[threadIdx.x];
}

// Block-wide exclusive scan of one value per thread (256 threads); returns the total.
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t *smem /*[RS_WAVES]*/, uint32_t &total) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint32_t x = wave_incl_scan_u32(v);
    if (lane == 63) smem[wave] = x;
    __syncthreads();
    uint32_t wbase = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < RS_WAVES; ++w) {
        const uint32_t s = smem[w];
        if (w < wave) wbase += s;
        tot += s;
    }
    __syncthreads();
    total = tot;
    return wbase + x - v;
}

// Exclusive scan of each digit's row hist[d][0..ntiles) (one block per digit, in
// chunks of 256), writing the digit's total.
__global__ __launch_bounds__(RS_THREADS) void k_radix_scan_rows(uint32_t *__restrict__ hist, int ntiles,
                                                                 uint32_t *__restrict__ digit_total) {
    __shared__ uint32_t sm[RS_WAVES];
    uint32_t *row = hist + (size_t)blockIdx.x * ntiles;
    uint32_t carry = 0;
    for (int c0 = 0; c0 < ntiles; c0 += RS_THREADS) {
        const int i = c0 + threadIdx.x;
        const uint32_t v = (i < ntiles) ? row[i] : 0u;
        uint32_t tot;
        const uint32_t ex = block_exclusive_scan(v, sm, tot);
        if (i < ntiles) row[i] = carry + ex;
        carry += tot;
    }
    if (threadIdx.x == 0) digit_total[blockIdx.x] = carry;
}

// Exclusive scan of the 256 digit totals -> digit bases (one block).
__global__ __launch_bounds__(RS_THREADS) void k_radix_scan_totals(uint32_t *__restrict__ digit_total) {
    __shared__ uint32_t sm[RS_WAVES];
    uint32_t tot;
    const uint32_t v = digit_total[threadIdx.x];
    const uint32_t ex = block_exclusive_scan(v, sm, tot);
    digit_total[threadIdx.x] = ex;
}

// Stable scatter of one pass.  Items are taken in rounds j of 256 consecutive keys
// (index tile*4096 + j*256 + thread), so (round, wave, lane) order is index order.
__global__ __launch_bounds__(RS_THREADS) void k_radix_scatter(
    const uint32_t *__restrict__ kin, const uint32_t *__restrict__ vin, uint32_t *__restrict__ kout,
    uint32_t *__restrict__ vout, const int *d_n, int n_host, int shift, const uint32_t *__restrict__ hist,
    const uint32_t *__restrict__ digit_base, int ntiles) {
    __shared__ uint32_t run[256];
    __shared__ uint32_t base[256];
    __shared__ uint32_t wc[RS_WAVES][256];
    const int n = load_n(d_n, n_host);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    run[threadIdx.x] = 0;
    base[threadIdx.x] = digit_base[threadIdx.x] + hist[(size_t)threadIdx.x * ntiles + blockIdx.x];
    const size_t tbase = (size_t)blockIdx.x * RS_TILE;
    for (int j = 0; j < RS_ITEMS; ++j) {
        const size_t i = tbase + (size_t)j * RS_THREADS + threadIdx.x;
        const bool valid = i < (size_t)n;
        const uint32_t key = valid ? kin[i] : 0u;
        const uint32_t val = valid ? vin[i] : 0u;
        const uint32_t d = (key >> shift) & 255u;
#pragma unroll
        for (int w = 0; w < RS_WAVES; ++w) wc[w][threadIdx.x] = 0;
        __syncthreads();
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const bool bit = (d >> b) & 1u;
            const uint64_t m = __ballot(bit);
            peers &= bit ? m : ~m;
        }
        const uint32_t rank = (uint32_t)__popcll(peers & lt);
        if (valid && rank == 0) wc[wave][d] = (uint32_t)__popcll(peers);
        __syncthreads();
        {
            const uint32_t r = run[threadIdx.x];
            uint32_t acc = r;
#pragma unroll
            for (int w = 0; w < RS_WAVES; ++w) {
                const uint32_t c = wc[w][threadIdx.x];
                wc[w][threadIdx.x] = acc;
                acc += c;
            }
            run[threadIdx.x] = acc;
        }
        __syncthreads();
        if (valid) {
            const uint32_t pos = base[d] + wc[wave][d] + rank;
            kout[pos] = key;
            vout[pos] = val;
        }
        __syncthreads();
    }
}

// off[c] = #(key < c) for c in [0, ncells] by binary search of the sorted keys.
__global__ void k_lower_bound(const uint32_t *__restrict__ keys, const int *d_n, int n_host, int ncells,
                              uint32_t *__restrict__ off) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c > ncells) return;
    const int n = load_n(d_n, n_host);
    int lo = 0, hi = n;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (keys[mid] < (uint32_t)c) lo = mid + 1;
        else hi = mid;
    }
    off[c] = (uint32_t)lo;
}

// computeCellIndices semantics (IndividualSolver.cl:564-619) from the offsets.
__global__ void k_left_right(const uint32_t *__restrict__ off, int ncells, uint32_t *__restrict__ left,
                             uint32_t *__restrict__ right) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncells) return;
    const uint32_t a = off[c], b = off[c + 1];
    left[c] = a;
    right[c] = (b > a) ? b - 1 : a;
}

// ---- particle step ---------------------------------------------------------------------
struct PartConsts {
    int side;
    float L;
    double sigma;  // sqrt(2 D dt)
    double dt;
    double amp;
    float p_ann;
    uint32_t seed;
};

__device__ __forceinline__ float wrap_pos(double xd, double L) {
    xd = xd - L * floor(xd / L);
    float xf = (float)xd;
    if (xf >= (float)L) xf = xf - (float)L;
    return xf;
}

// Move: x += v_x dt + sigma g_x with v = A(-sin(2 pi y/L), sin(2 pi x/L)), periodic;
// writes the cell key and the identity payload for the sort.
__global__ void k_move(float *__restrict__ x, float *__restrict__ y, const uint32_t *__restrict__ id,
                       uint32_t *__restrict__ keys, uint32_t *__restrict__ vals, const int *d_n, int step,
                       PartConsts p) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= *d_n) return;
    double gx, gy;
    gauss2(id[i], p.seed, (uint32_t)step, &gx, &gy);
    const double xd = x[i], yd = y[i];
    const double L = p.L;
    const double vx = -p.amp * sinpi_d(2.0 * yd / L);
    const double vy = p.amp * sinpi_d(2.0 * xd / L);
    const float xn = wrap_pos(xd + vx * p.dt + p.sigma * gx, L);
    const float yn = wrap_pos(yd + vy * p.dt + p.sigma * gy, L);
    x[i] = xn;
    y[i] = yn;
    keys[i] = (uint32_t)((int)yn * p.side + (int)xn);
    vals[i] = (uint32_t)i;
}

// Gather of the positions and identities into sorted order.
__global__ void k_gather(const uint32_t *__restrict__ perm, const float *__restrict__ x, const float *__restrict__ y,
                         const uint32_t *__restrict__ id, float *__restrict__ xs, float *__restrict__ ys,
                         uint32_t *__restrict__ ids, const int *d_n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= *d_n) return;
    const uint32_t j = perm[i];
    xs[i] = x[j];
    ys[i] = y[j];
    ids[i] = id[j];
}

// A+A->0: the disjoint consecutive pairs (off[c]+2j, off[c]+2j+1) of cell c are removed
// when u01(threefry({step, j}, {c, ~seed})) < p; keep[i] = 1 for survivors.
__global__ void k_annihilate(const uint32_t *__restrict__ off, int ncells, uint32_t *__restrict__ keep, int step,
                             PartConsts p) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncells) return;
    const uint32_t b = off[c], e = off[c + 1];
    const uint32_t pairs = (e - b) >> 1;
    for (uint32_t j = 0; j < pairs; ++j) {
        const u32x2 w = threefry2x32_20((uint32_t)step, j, (uint32_t)c, ~p.seed);
        const uint32_t k = (u01(w.x) < p.p_ann) ? 0u : 1u;
        keep[b + 2 * j] = k;
        keep[b + 2 * j + 1] = k;
    }
    if ((e - b) & 1u) keep[e - 1] = 1u;
}

// Per-tile survivor counts (compaction = a stable 1-bit partition).
__global__ __launch_bounds__(RS_THREADS) void k_tile_count(const uint32_t *__restrict__ keep, const int *d_n,
                                                            uint32_t *__restrict__ tile_sum) {
    __shared__ uint32_t sm[RS_WAVES];
    const int n = *d_n;
    const size_t base = (size_t)blockIdx.x * RS_TILE;
    uint32_t s = 0;
    for (int j = 0; j < RS_ITEMS; ++j) {
        const size_t i = base + (size_t)j * RS_THREADS + threadIdx.x;
        if (i < (size_t)n) s += keep[i];
    }
    uint32_t tot;
    (void)block_exclusive_scan(s, sm, tot);
    if (threadIdx.x == 0) tile_sum[blockIdx.x] = tot;
}

// Exclusive scan of tile sums (one block, any count) and the new particle count.
__global__ __launch_bounds__(RS_THREADS) void k_scan_tiles(uint32_t *__restrict__ tile_sum, int ntiles,
                                                            int *__restrict__ d_n_out) {
    __shared__ uint32_t sm[RS_WAVES];
    uint32_t carry = 0;
    for (int c0 = 0; c0 < ntiles; c0 += RS_THREADS) {
        const int i = c0 + threadIdx.x;
        const uint32_t v = (i < ntiles) ? tile_sum[i] : 0u;
        uint32_t tot;
        const uint32_t ex = block_exclusive_scan(v, sm, tot);
        if (i < ntiles) tile_sum[i] = carry + ex;
        carry += tot;
    }
    if (threadIdx.x == 0) *d_n_out = (int)carry;
}

__global__ __launch_bounds__(RS_THREADS) void k_compact(
    const uint32_t *__restrict__ keep, const uint32_t *__restrict__ tile_off, const int *d_n,
    const float *__restrict__ x, const float *__restrict__ y, const uint32_t *__restrict__ id,
    const uint32_t *__restrict__ key, float *__restrict__ xo, float *__restrict__ yo, uint32_t *__restrict__ ido,
    uint32_t *__restrict__ keyo) {
    __shared__ uint32_t sm[RS_WAVES];
    const int n = *d_n;
    const size_t base = (size_t)blockIdx.x * RS_TILE;
    uint32_t run = tile_off[blockIdx.x];
    for (int j = 0; j < RS_ITEMS; ++j) {
        const size_t i = base + (size_t)j * RS_THREADS + threadIdx.x;
        const uint32_t k = (i < (size_t)n) ? keep[i] : 0u;
        uint32_t tot;
        const uint32_t ex = block_exclusive_scan(k, sm, tot);
        if (k) {
            const uint32_t o = run + ex;
            xo[o] = x[i];
            yo[o] = y[i];
            ido[o] = id[i];
            keyo[o] = key[i];
        }
        run += tot;
    }
}

// End of a step: count the particles the step processed, then install the new count.
__global__ void k_finish_step(const int *n_next, int *n, unsigned long long *processed) {
    *processed += (unsigned long long)*n;
    *n = *n_next;
}

}  // namespace gpgmp

// ---------------------------------------------------------------------------------------
#include "particles_launch.h"

namespace gpgmp {

int radix_tiles(int n) { return (n + RS_TILE - 1) / RS_TILE; }

hipError_t radix_sort_pairs(uint32_t *keys, uint32_t *vals, uint32_t *ktmp, uint32_t *vtmp, const int *d_n,
                            int n_ub, int key_bits, uint32_t *hist, uint32_t *digit_total, hipStream_t s,
                            bool *result_in_tmp) {
    const int nt = std::max(radix_tiles(n_ub), 1);
    uint32_t *ki = keys, *vi = vals, *ko = ktmp, *vo = vtmp;
    int passes = 0;
    for (int shift = 0; shift < key_bits; shift += 8) {
        hipLaunchKernelGGL(k_radix_hist, dim3(nt), dim3(RS_THREADS), 0, s, ki, d_n, n_ub, shift, hist, nt);
        hipLaunchKernelGGL(k_radix_scan_rows, dim3(256), dim3(RS_THREADS), 0, s, hist, nt, digit_total);
        hipLaunchKernelGGL(k_radix_scan_totals, dim3(1), dim3(RS_THREADS), 0, s, digit_total);
        hipLaunchKernelGGL(k_radix_scatter, dim3(nt), dim3(RS_THREADS), 0, s, ki, vi, ko, vo, d_n, n_ub, shift, hist,
                           digit_total, nt);
        std::swap(ki, ko);
        std::swap(vi, vo);
        ++passes;
    }
    *result_in_tmp = (passes & 1) != 0;
    return hipGetLastError();
}

hipError_t launch_lower_bound(const uint32_t *keys, const int *d_n, int n_host, int ncells, uint32_t *off,
                              hipStream_t s) {
    hipLaunchKernelGGL(k_lower_bound, dim3((ncells + 1 + 255) / 256), dim3(256), 0, s, keys, d_n, n_host, ncells, off);
    return hipGetLastError();
}

hipError_t launch_left_right(const uint32_t *off, int ncells, uint32_t *left, uint32_t *right, hipStream_t s) {
    hipLaunchKernelGGL(k_left_right, dim3((ncells + 255) / 256), dim3(256), 0, s, off, ncells, left, right);
    return hipGetLastError();
}

static PartConsts consts(const ParticleStepArgs &a) {
    PartConsts p;
    p.side = a.side;
    p.L = (float)a.side;
    p.sigma = __builtin_sqrt(2.0 * (double)a.D * (double)a.dt);
    p.dt = a.dt;
    p.amp = a.vortex_amp;
    p.p_ann = a.p_annihilate;
    p.seed = a.seed;
    return p;
}

hipError_t particle_step(const ParticleStepArgs &a, int step, ParticleBuffers &b, hipStream_t s) {
    const PartConsts p = consts(a);
    const int n_ub = b.n_ub;
    const int ncells = a.side * a.side;
    const dim3 gN((n_ub + 255) / 256 > 0 ? (n_ub + 255) / 256 : 1), blk(256);
    // 1. move
    hipLaunchKernelGGL(k_move, gN, blk, 0, s, b.x, b.y, b.id, b.key, b.val, b.d_n, step, p);
    // 2. stable sort of (cell key, index)
    bool in_tmp = false;
    hipError_t e = radix_sort_pairs(b.key, b.val, b.key2, b.val2, b.d_n, n_ub, a.key_bits, b.hist, b.digit_total, s,
                                    &in_tmp);
    if (e != hipSuccess) return e;
    uint32_t *sk = in_tmp ? b.key2 : b.key;
    uint32_t *sv = in_tmp ? b.val2 : b.val;
    hipLaunchKernelGGL(k_gather, gN, blk, 0, s, sv, b.x, b.y, b.id, b.x2, b.y2, b.id2, b.d_n);
    // 3. offsets
    hipLaunchKernelGGL(k_lower_bound, dim3((ncells + 1 + 255) / 256), dim3(256), 0, s, sk, b.d_n, 0, ncells, b.off);
    // 4. annihilation flags
    hipLaunchKernelGGL(k_annihilate, dim3((ncells + 255) / 256), dim3(256), 0, s, b.off, ncells, b.keep, step, p);
    // 5. compaction back into the primary buffers
    const int nt = std::max(radix_tiles(n_ub), 1);
    hipLaunchKernelGGL(k_tile_count, dim3(nt), dim3(RS_THREADS), 0, s, b.keep, b.d_n, b.tile_sum);
    hipLaunchKernelGGL(k_scan_tiles, dim3(1), dim3(RS_THREADS), 0, s, b.tile_sum, nt, b.d_n2);
    hipLaunchKernelGGL(k_compact, dim3(nt), dim3(RS_THREADS), 0, s, b.keep, b.tile_sum, b.d_n, b.x2, b.y2, b.id2, sk,
                       b.x, b.y, b.id, b.key_out);
    hipLaunchKernelGGL(k_finish_step, dim3(1), dim3(1), 0, s, b.d_n2, b.d_n, b.processed);
    return hipGetLastError();
}

}  // namespace gpgmp
