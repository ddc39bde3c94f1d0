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

// Payload words carried with the keys through a pass (NP of them).
struct RadixPayload {
    const uint32_t *in[3];
    uint32_t *out[3];
};

// Stable scatter of one pass, staged in LDS.  A tile of 4096 keys (16 per thread, in registers
// with their payload) is first counted by digit, then ranked in rounds j of 256 consecutive
// keys (index tile*4096 + j*256 + thread, so (round, wave, lane) order is index order): the
// lanes of a wave holding the same digit are found with eight __ballot calls, a key's rank
// is the popcount of its lower peers plus the counts of the earlier waves and rounds, and it
// is written to LDS at its digit's place in the tile.  The tile, now sorted by digit, is
// written out in LDS order: consecutive threads write consecutive places of one digit's run,
// so the stores coalesce instead of landing one word at a time all over the output.
template <int NP>
__global__ __launch_bounds__(RS_THREADS) void k_radix_scatter(
    const uint32_t *__restrict__ kin, uint32_t *__restrict__ kout, RadixPayload pl, const int *d_n, int n_host,
    int shift, const uint32_t *__restrict__ hist, const uint32_t *__restrict__ digit_base, int ntiles) {
    __shared__ uint32_t skey[RS_TILE];
    __shared__ uint32_t spay[NP][RS_TILE];
    __shared__ uint32_t tstart[256];  // the tile's first place of each digit
    __shared__ uint32_t gbase[256];   // output place of the tile's first key of each digit
    __shared__ uint32_t run[256];
    __shared__ uint32_t wc[RS_WAVES][256];
    __shared__ uint32_t sm[RS_WAVES];
    const int n = load_n(d_n, n_host);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    const size_t tbase = (size_t)blockIdx.x * RS_TILE;
    uint32_t key[RS_ITEMS], pay[NP][RS_ITEMS];
    run[threadIdx.x] = 0;
    gbase[threadIdx.x] = digit_base[threadIdx.x] + hist[(size_t)threadIdx.x * ntiles + blockIdx.x];
    __syncthreads();
#pragma unroll
    for (int j = 0; j < RS_ITEMS; ++j) {
        const size_t i = tbase + (size_t)j * RS_THREADS + threadIdx.x;
        const bool valid = i < (size_t)n;
        key[j] = valid ? kin[i] : 0u;
#pragma unroll
        for (int q = 0; q < NP; ++q) pay[q][j] = valid ? pl.in[q][i] : 0u;
        if (valid) atomicAdd(&run[(key[j] >> shift) & 255u], 1u);
    }
    __syncthreads();
    {
        uint32_t tot;
        const uint32_t c = run[threadIdx.x];
        tstart[threadIdx.x] = block_exclusive_scan(c, sm, tot);
        run[threadIdx.x] = 0;
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < RS_ITEMS; ++j) {
        const size_t i = tbase + (size_t)j * RS_THREADS + threadIdx.x;
        const bool valid = i < (size_t)n;
        const uint32_t d = (key[j] >> shift) & 255u;
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
            const uint32_t lp = tstart[d] + wc[wave][d] + rank;
            skey[lp] = key[j];
#pragma unroll
            for (int q = 0; q < NP; ++q) spay[q][lp] = pay[q][j];
        }
        __syncthreads();  // wc is cleared by the next round
    }
    const int cnt = (n - (int)tbase < RS_TILE) ? n - (int)tbase : RS_TILE;
    for (int lp = threadIdx.x; lp < cnt; lp += RS_THREADS) {
        const uint32_t k = skey[lp];
        const uint32_t d = (k >> shift) & 255u;
        const uint32_t pos = gbase[d] + ((uint32_t)lp - tstart[d]);
        kout[pos] = k;
#pragma unroll
        for (int q = 0; q < NP; ++q) pl.out[q][pos] = spay[q][lp];
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
// writes the cell key for the sort.
__global__ void k_move(float *__restrict__ x, float *__restrict__ y, const uint32_t *__restrict__ id,
                       uint32_t *__restrict__ keys, const int *d_n, int step,
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

// ---- particle step on a list already sorted by cell ---------------------------------------
// After a step the list is sorted by cell (stably), with CSR offsets.  A particle of the
// next step moves from its cell o to a cell c that is, but for a vanishing fraction, one of
// the 3x3 torus neighbours of o, so the stable order by the new cell is fixed by counts:
// with n(o->c) the particles of o moving to c, a particle of o is at rank
//   r = sum over the sources o' of c with key(o') < key(o) of n(o'->c)
//     + #(earlier particles of o moving to c) + #(particles from outside the 3x3 into c
//       with a smaller index)
// among the particles of c, which is the position std::stable_sort gives it (the list is
// sorted by old key, so a smaller source key means a smaller index).  Particles that leave
// the 3x3 neighbourhood go on a far list and are ranked the same way.  No radix pass and no
// gather: every particle is read and written once, in nearly sequential order.
//
// Direction codes: d = (dy + 1) * 3 + (dx + 1) for a move by (dx, dy) in {-1, 0, 1}^2 on the
// torus (side >= 3), kFar otherwise.
constexpr int kFar = 9;

struct FarEntry {
    uint32_t i;  // index in the old order
    uint32_t c;  // new cell
};

__device__ __forceinline__ int torus_step(int a, int b, int L) {
    int d = b - a;
    if (d == L - 1) d = -1;
    else if (d == -(L - 1)) d = 1;
    return d;
}

__device__ __forceinline__ int wrap1(int v, int L) {  // v in [-1, L]
    v += (v < 0) ? L : 0;
    return v - ((v >= L) ? L : 0);
}

// cell (cx, cy) moved by direction d
__device__ __forceinline__ uint32_t moved_cell(int cx, int cy, int d, int L) {
    return (uint32_t)(wrap1(cy + (d / 3 - 1), L) * L + wrap1(cx + (d % 3 - 1), L));
}

// 1. move (as k_move) and classify the move; far moves are listed (order-free atomics: the
//    list is only ever counted, never walked in order)
__global__ void k_coh_move(const float *__restrict__ x, const float *__restrict__ y, const uint32_t *__restrict__ id,
                           float *__restrict__ x2, float *__restrict__ y2, uint8_t *__restrict__ dir, const int *d_n,
                           int step, PartConsts p, uint32_t *__restrict__ far_cnt, FarEntry *__restrict__ far_list,
                           uint32_t *__restrict__ far_in) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= *d_n) return;
    double gx, gy;
    gauss2(id[i], p.seed, (uint32_t)step, &gx, &gy);
    const float xo = x[i], yo = y[i];
    const double xd = xo, yd = yo;
    const double L = p.L;
    const double vx = -p.amp * sinpi_d(2.0 * yd / L);
    const double vy = p.amp * sinpi_d(2.0 * xd / L);
    const float xn = wrap_pos(xd + vx * p.dt + p.sigma * gx, L);
    const float yn = wrap_pos(yd + vy * p.dt + p.sigma * gy, L);
    x2[i] = xn;
    y2[i] = yn;
    const int side = p.side;
    const int cx = (int)xn, cy = (int)yn;
    const int dx = torus_step((int)xo, cx, side);  // the old cell is the old position's (the key)
    const int dy = torus_step((int)yo, cy, side);
    const bool near = dx >= -1 && dx <= 1 && dy >= -1 && dy <= 1;
    dir[i] = (uint8_t)(near ? (dy + 1) * 3 + (dx + 1) : kFar);
    if (!near) {
        const uint32_t c = (uint32_t)(cy * side + cx);
        const uint32_t k = atomicAdd(far_cnt, 1u);
        far_list[k] = FarEntry{(uint32_t)i, c};
        atomicAdd(&far_in[c], 1u);
    }
}

// 2. per old cell: its particles by direction, cnt[d][o] (every entry written)
__global__ void k_coh_count(const uint32_t *__restrict__ off, const uint8_t *__restrict__ dir, int A,
                            uint32_t *__restrict__ cnt) {
    const int o = blockIdx.x * blockDim.x + threadIdx.x;
    if (o >= A) return;
    const uint32_t b = off[o], e = off[o + 1];
    uint32_t c0 = 0, c1 = 0, c2 = 0, c3 = 0, c4 = 0, c5 = 0, c6 = 0, c7 = 0, c8 = 0;
    for (uint32_t i = b; i < e; ++i) {
        const uint32_t d = dir[i];
        c0 += d == 0; c1 += d == 1; c2 += d == 2; c3 += d == 3; c4 += d == 4;
        c5 += d == 5; c6 += d == 6; c7 += d == 7; c8 += d == 8;
    }
    const size_t sA = (size_t)A;
    cnt[0 * sA + o] = c0; cnt[1 * sA + o] = c1; cnt[2 * sA + o] = c2;
    cnt[3 * sA + o] = c3; cnt[4 * sA + o] = c4; cnt[5 * sA + o] = c5;
    cnt[6 * sA + o] = c6; cnt[7 * sA + o] = c7; cnt[8 * sA + o] = c8;
}

__device__ __forceinline__ bool pair_killed(int step, uint32_t j, uint32_t c, const PartConsts &p) {
    const u32x2 w = threefry2x32_20((uint32_t)step, j, c, ~p.seed);
    return u01(w.x) < p.p_ann;
}

// 3. per new cell: the particles arriving from each source (pre[d][c] = the arrivals from
//    the sources with a smaller key than direction d's), the total, the A+A->0 pair draws
//    (first 32 pairs as a mask) and the survivors, summed per 256-cell tile
__global__ __launch_bounds__(256) void k_coh_cells(const uint32_t *__restrict__ cnt, uint32_t *__restrict__ far_in,
                                                   int A, int side, int step, PartConsts p, uint32_t *__restrict__ pre,
                                                   uint32_t *__restrict__ tot, uint32_t *__restrict__ kmask,
                                                   uint32_t *__restrict__ surv, uint32_t *__restrict__ tile_sum,
                                                   uint32_t *__restrict__ far_cnt_next) {
    __shared__ uint32_t sm[RS_WAVES];
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c == 0) *far_cnt_next = 0u;  // the far list of the next step starts empty
    uint32_t S = 0;
    if (c < A) {
        const int cy = c / side, cx = c - cy * side;
        uint32_t key[9], n[9];
        uint32_t T = 0;
#pragma unroll
        for (int d = 0; d < 9; ++d) {
            key[d] = moved_cell(cx, cy, 8 - d, side);  // the source whose move d lands on c
            n[d] = cnt[(size_t)d * A + key[d]];
            T += n[d];
        }
#pragma unroll
        for (int d = 0; d < 9; ++d) {
            uint32_t q = 0;
#pragma unroll
            for (int e = 0; e < 9; ++e) q += (key[e] < key[d]) ? n[e] : 0u;
            pre[(size_t)d * A + c] = q;
        }
        const uint32_t fin = far_in[c];
        if (fin) far_in[c] = 0u;  // counted: clear for the next step
        T += fin;
        uint32_t mask = 0, kills = 0;
        const uint32_t pairs = T >> 1;
        for (uint32_t j = 0; j < pairs; ++j) {
            if (pair_killed(step, j, (uint32_t)c, p)) {
                ++kills;
                if (j < 32) mask |= 1u << j;
            }
        }
        S = T - 2 * kills;
        tot[c] = T;
        kmask[c] = mask;
        surv[c] = S;
    }
    uint32_t total;
    (void)block_exclusive_scan(S, sm, total);
    if (threadIdx.x == 0) tile_sum[blockIdx.x] = total;
}

// 4. exclusive scan of the tile sums (one block of 1024 threads, 4 tiles per thread per
//    round); the new count goes to the other count slot
__global__ __launch_bounds__(1024) void k_coh_scan(uint32_t *__restrict__ tile_sum, int ntiles,
                                                    const int *__restrict__ n_cur, int *__restrict__ n_next,
                                                    unsigned long long *processed) {
    __shared__ uint32_t wsum[16];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t carry = 0;
    for (int c0 = 0; c0 < ntiles; c0 += 4096) {
        const int i0 = c0 + 4 * (int)threadIdx.x;
        uint32_t v[4], t = 0;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            v[k] = (i0 + k < ntiles) ? tile_sum[i0 + k] : 0u;
            t += v[k];
        }
        const uint32_t incl = wave_incl_scan_u32(t);
        if (lane == 63) wsum[wave] = incl;
        __syncthreads();
        uint32_t wbase = 0, all = 0;
#pragma unroll
        for (int w = 0; w < 16; ++w) {
            const uint32_t sw = wsum[w];
            wbase += (w < wave) ? sw : 0u;
            all += sw;
        }
        __syncthreads();
        uint32_t run = carry + wbase + incl - t;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            if (i0 + k < ntiles) tile_sum[i0 + k] = run;
            run += v[k];
        }
        carry += all;
    }
    if (threadIdx.x == 0) {
        *processed += (unsigned long long)*n_cur;
        *n_next = (int)carry;
    }
}

// 5. survivors' CSR offsets off[c] (and off[A] = the total)
__global__ __launch_bounds__(256) void k_coh_offsets(const uint32_t *__restrict__ surv,
                                                     const uint32_t *__restrict__ tile_off, int A,
                                                     uint32_t *__restrict__ off) {
    __shared__ uint32_t sm[RS_WAVES];
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t S = (c < A) ? surv[c] : 0u;
    uint32_t total;
    const uint32_t ex = block_exclusive_scan(S, sm, total);
    if (c < A) off[c] = tile_off[blockIdx.x] + ex;
    if (c == A - 1) off[A] = tile_off[blockIdx.x] + ex + S;
}

// 6. every particle to its place in the new order; the killed pairs drop out.  A particle's
//    rank among the earlier particles of its cell with its move class is counted over the
//    move classes of the block's particles and the 256 before them, staged in LDS (a cell
//    reaching further back counts the rest from global memory).  A thread takes CS_PER
//    particles (a block CS_BLOCK * CS_PER consecutive ones) and issues their loads together:
//    the kernel is bound by memory latency, not bandwidth, so each wave keeps several
//    particles' loads in flight.
constexpr int CS_BLOCK = 256, CS_PER = 4, CS_SPAN = CS_BLOCK * CS_PER, CS_BACK = 256;

__global__ __launch_bounds__(CS_BLOCK) void k_coh_scatter(
    const float *__restrict__ x2, const float *__restrict__ y2, const uint32_t *__restrict__ id,
    const uint32_t *__restrict__ key, const uint8_t *__restrict__ dir, const uint32_t *__restrict__ off_old,
    const uint32_t *__restrict__ cnt, const uint32_t *__restrict__ pre, const uint32_t *__restrict__ tot,
    const uint32_t *__restrict__ kmask, const uint32_t *__restrict__ off_new, const uint32_t *__restrict__ far_cnt,
    const FarEntry *__restrict__ far_list, const int *d_n, int A, int side, int step, PartConsts p,
    float *__restrict__ xo, float *__restrict__ yo, uint32_t *__restrict__ ido, uint32_t *__restrict__ keyo,
    uint32_t cap) {
    __shared__ uint8_t sdir[CS_BACK + CS_SPAN];
    const int n = *d_n;
    const int b0 = blockIdx.x * CS_SPAN;
    const int w0 = b0 - CS_BACK;  // first index staged
    for (int k = threadIdx.x; k < CS_BACK + CS_SPAN; k += CS_BLOCK) {
        const int j = w0 + k;
        sdir[k] = (j >= 0 && j < n) ? dir[j] : (uint8_t)kFar;
    }
    __syncthreads();
    const uint32_t nf = *far_cnt;
    float xv[CS_PER], yv[CS_PER];
    uint32_t idv[CS_PER], o[CS_PER], d[CS_PER], c[CS_PER];
    // every load that depends only on the index, then every one that depends only on the
    // cells, so a wave waits on two rounds of memory latency for its CS_PER particles
#pragma unroll
    for (int q = 0; q < CS_PER; ++q) {
        const int i = b0 + q * CS_BLOCK + (int)threadIdx.x;
        const bool v = i < n;
        xv[q] = v ? x2[i] : 0.f;
        yv[q] = v ? y2[i] : 0.f;
        idv[q] = v ? id[i] : 0u;
        o[q] = v ? key[i] : 0u;
        d[q] = v ? sdir[CS_BACK + q * CS_BLOCK + threadIdx.x] : (uint32_t)kFar;
    }
    int s0[CS_PER];
    uint32_t pr[CS_PER], T[CS_PER], m[CS_PER], base[CS_PER];
#pragma unroll
    for (int q = 0; q < CS_PER; ++q) {
        const int oy = (int)(o[q] / (uint32_t)side), ox = (int)o[q] - oy * side;
        c[q] = (d[q] < 9) ? moved_cell(ox, oy, (int)d[q], side) : (uint32_t)((int)yv[q] * side + (int)xv[q]);
        s0[q] = (int)off_old[o[q]];
        pr[q] = (d[q] < 9) ? pre[(size_t)d[q] * A + c[q]] : 0u;
        T[q] = tot[c[q]];
        m[q] = kmask[c[q]];
        base[q] = off_new[c[q]];
    }
#pragma unroll
    for (int q = 0; q < CS_PER; ++q) {
        const int i = b0 + q * CS_BLOCK + (int)threadIdx.x;
        if (i >= n) continue;
        uint32_t r;
        if (d[q] < 9) {
            uint32_t k = 0;
            for (int j = (s0[q] > w0 ? s0[q] : w0); j < i; ++j) k += sdir[j - w0] == d[q];
            for (int j = s0[q]; j < w0; ++j) k += dir[j] == d[q];
            r = pr[q] + k;
        } else {
            const int cy = (int)c[q] / side, cx = (int)c[q] - cy * side;
            r = 0;
            for (int e = 0; e < 9; ++e) {
                const uint32_t s = moved_cell(cx, cy, 8 - e, side);
                if (s < o[q]) r += cnt[(size_t)e * A + s];
            }
        }
        for (uint32_t f = 0; f < nf; ++f) {
            const FarEntry fe = far_list[f];
            r += (fe.c == c[q] && fe.i < (uint32_t)i) ? 1u : 0u;
        }
        const uint32_t j = r >> 1;
        const bool partner = (r | 1u) < T[q];
        bool killed = false;
        if (partner) killed = (j < 32) ? ((m[q] >> j) & 1u) != 0 : pair_killed(step, j, c[q], p);
        if (killed) continue;
        uint32_t before = (uint32_t)__popc(m[q] & ((j < 32) ? ((1u << j) - 1u) : 0xffffffffu));
        for (uint32_t jj = 32; jj < j; ++jj) before += pair_killed(step, jj, c[q], p) ? 1u : 0u;
        const uint32_t pos = base[q] + r - 2u * before;
        if (pos >= cap) continue;  // never for a consistent list; keeps every store inside the buffers
        xo[pos] = xv[q];
        yo[pos] = yv[q];
        ido[pos] = idv[q];
        keyo[pos] = c[q];
    }
}

}  // namespace gpgmp

// ---------------------------------------------------------------------------------------
#include "particles_launch.h"

namespace gpgmp {

int radix_tiles(int n) { return (n + RS_TILE - 1) / RS_TILE; }

// LSD passes of 8-bit digits over key_bits bits, ping-ponging the keys and np payload arrays
// between the given and the tmp buffers; *result_in_tmp tells where the result is.
hipError_t radix_sort_payload(uint32_t *keys, uint32_t *ktmp, uint32_t *const *pay, uint32_t *const *ptmp, int np,
                              const int *d_n, int n_ub, int key_bits, uint32_t *hist, uint32_t *digit_total,
                              hipStream_t s, bool *result_in_tmp) {
    const int nt = std::max(radix_tiles(n_ub), 1);
    uint32_t *ki = keys, *ko = ktmp;
    uint32_t *pi[3] = {nullptr, nullptr, nullptr}, *po[3] = {nullptr, nullptr, nullptr};
    for (int q = 0; q < np; ++q) {
        pi[q] = pay[q];
        po[q] = ptmp[q];
    }
    int passes = 0;
    for (int shift = 0; shift < key_bits; shift += 8) {
        hipLaunchKernelGGL(k_radix_hist, dim3(nt), dim3(RS_THREADS), 0, s, ki, d_n, n_ub, shift, hist, nt);
        hipLaunchKernelGGL(k_radix_scan_rows, dim3(256), dim3(RS_THREADS), 0, s, hist, nt, digit_total);
        hipLaunchKernelGGL(k_radix_scan_totals, dim3(1), dim3(RS_THREADS), 0, s, digit_total);
        RadixPayload pl;
        for (int q = 0; q < 3; ++q) {
            pl.in[q] = pi[q];
            pl.out[q] = po[q];
        }
        if (np == 1)
            hipLaunchKernelGGL(k_radix_scatter<1>, dim3(nt), dim3(RS_THREADS), 0, s, ki, ko, pl, d_n, n_ub, shift, hist,
                               digit_total, nt);
        else if (np == 3)
            hipLaunchKernelGGL(k_radix_scatter<3>, dim3(nt), dim3(RS_THREADS), 0, s, ki, ko, pl, d_n, n_ub, shift, hist,
                               digit_total, nt);
        else
            return hipErrorInvalidValue;
        std::swap(ki, ko);
        for (int q = 0; q < np; ++q) std::swap(pi[q], po[q]);
        ++passes;
    }
    *result_in_tmp = (passes & 1) != 0;
    return hipGetLastError();
}

hipError_t radix_sort_pairs(uint32_t *keys, uint32_t *vals, uint32_t *ktmp, uint32_t *vtmp, const int *d_n,
                            int n_ub, int key_bits, uint32_t *hist, uint32_t *digit_total, hipStream_t s,
                            bool *result_in_tmp) {
    uint32_t *pay[1] = {vals}, *ptmp[1] = {vtmp};
    return radix_sort_payload(keys, ktmp, pay, ptmp, 1, d_n, n_ub, key_bits, hist, digit_total, s, result_in_tmp);
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

hipError_t particle_step_sorted(const ParticleStepArgs &a, int step, const SortedStepBuffers &b, hipStream_t s) {
    const PartConsts p = consts(a);
    const int A = a.side * a.side;
    const dim3 blk(256);
    const dim3 gN((b.n_ub + 255) / 256 > 0 ? (b.n_ub + 255) / 256 : 1), gA((A + 255) / 256);
    const int ntiles = (A + 255) / 256;
    hipLaunchKernelGGL(k_coh_move, gN, blk, 0, s, b.x, b.y, b.id, b.x2, b.y2, b.dir, b.n_cur, step, p, b.far_cnt,
                       (FarEntry *)b.far_list, b.far_in);
    hipLaunchKernelGGL(k_coh_count, gA, blk, 0, s, b.off, b.dir, A, b.cnt);
    hipLaunchKernelGGL(k_coh_cells, gA, blk, 0, s, b.cnt, b.far_in, A, a.side, step, p, b.pre, b.tot, b.kmask, b.surv,
                       b.tile_sum, b.far_cnt_next);
    hipLaunchKernelGGL(k_coh_scan, dim3(1), dim3(1024), 0, s, b.tile_sum, ntiles, b.n_cur, b.n_next, b.processed);
    hipLaunchKernelGGL(k_coh_offsets, gA, blk, 0, s, b.surv, b.tile_sum, A, b.off_new);
    hipLaunchKernelGGL(k_coh_scatter, dim3((b.n_ub + CS_SPAN - 1) / CS_SPAN > 0 ? (b.n_ub + CS_SPAN - 1) / CS_SPAN : 1),
                       dim3(CS_BLOCK), 0, s, b.x2, b.y2, b.id, b.key, b.dir, b.off, b.cnt, b.pre, b.tot, b.kmask,
                       b.off_new, b.far_cnt, (const FarEntry *)b.far_list, b.n_cur, A, a.side, step, p, b.xo, b.yo,
                       b.ido, b.keyo, (uint32_t)b.capacity);
    return hipGetLastError();
}

size_t far_entry_bytes() { return sizeof(FarEntry); }

hipError_t particle_step(const ParticleStepArgs &a, int step, ParticleBuffers &b, hipStream_t s) {
    const PartConsts p = consts(a);
    const int n_ub = b.n_ub;
    const int ncells = a.side * a.side;
    const dim3 gN((n_ub + 255) / 256 > 0 ? (n_ub + 255) / 256 : 1), blk(256);
    // 1. move (positions in place) and cell keys
    hipLaunchKernelGGL(k_move, gN, blk, 0, s, b.x, b.y, b.id, b.key, b.d_n, step, p);
    // 2. stable sort by cell key carrying the positions and ids; an odd number of passes
    //    (an all-zero digit pass above the key is a stable copy) leaves the result in the
    //    second buffers, from which the compaction writes back into the first ones
    int bits = a.key_bits;
    if ((((bits + 7) / 8) & 1) == 0) bits += 8;
    bool in_tmp = false;
    uint32_t *pay[3] = {(uint32_t *)b.x, (uint32_t *)b.y, b.id};
    uint32_t *ptmp[3] = {(uint32_t *)b.x2, (uint32_t *)b.y2, b.id2};
    hipError_t e = radix_sort_payload(b.key, b.key2, pay, ptmp, 3, b.d_n, n_ub, bits, b.hist, b.digit_total, s,
                                      &in_tmp);
    if (e != hipSuccess) return e;
    if (!in_tmp) return hipErrorInvalidValue;
    uint32_t *sk = b.key2;
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
