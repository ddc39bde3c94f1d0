// Individual-particle kernels for gfx950: stable LSD radix sort of (key, value) pairs,
// cell offsets, and the particle step (move, sort, offsets, in-cell A+A->0, compaction).
//
// The reference sorts KEY_INDEX_PAIR(key, id) by the low 32 bits with 4-bit digits,
// warp-synchronous 32-lane split scans and a 3-kernel exclusive scan per digit pass
// (IndividualSolver.cl:36-553, IndividualSolver.cpp:673-808).  Here a pass uses 8-bit
// digits and wave64 ranking: the lanes holding the same digit are found with eight
// __ballot calls, a lane's rank is the popcount of its lower peers, and the per-wave
// digit counts are combined in LDS in (wave, round, lane) order over a wave's contiguous
// part of the tile, which is the original index order -- so every pass is stable and the
// result equals std::stable_sort.
#include <hip/hip_runtime.h>

#include "gpgmp_numeric.h"
#include "wave_ops.h"

namespace gpgmp {

constexpr int RS_THREADS = 256;
constexpr int RS_ITEMS = 16;
constexpr int RS_TILE = RS_THREADS * RS_ITEMS;  // 4096 keys per tile
constexpr int RS_WAVES = RS_THREADS / 64;

__device__ __forceinline__ int load_n(const int *d_n, int n_host) { return d_n ? *d_n : n_host; }

// The lanes among `among` whose low NB bits of `key` equal this lane's: one __ballot per bit,
// the lane's own bit keeping the ballot or its complement, in two 32-bit halves (a bit of 1
// keeps m, a bit of 0 keeps ~m = m ^ 0xffffffff).
template <int NB>
__device__ __forceinline__ uint64_t match_low_bits(uint32_t key, uint64_t among) {
    uint32_t lo = (uint32_t)among, hi = (uint32_t)(among >> 32);
#pragma unroll
    for (int b = 0; b < NB; ++b) {
        const uint32_t bit = (key >> b) & 1u;
        const uint64_t m = __ballot(bit);
        const uint32_t f = bit - 1u;
        lo &= (uint32_t)m ^ f;
        hi &= (uint32_t)(m >> 32) ^ f;
    }
    return ((uint64_t)hi << 32) | lo;
}

// Per-tile digit histogram of one pass: hist[d * ntiles + tile].  A full tile of 16-byte
// aligned keys is read as 16-byte words, every load of the thread issued before the counts.
__global__ __launch_bounds__(RS_THREADS) void k_radix_hist(const uint32_t *__restrict__ keys, const int *d_n,
                                                            int n_host, int shift, uint32_t *__restrict__ hist,
                                                            int ntiles) {
    __shared__ uint32_t cnt[256];
    const int n = load_n(d_n, n_host);
    cnt[threadIdx.x] = 0;
    __syncthreads();
    const size_t base = (size_t)blockIdx.x * RS_TILE;
    const bool vec = base + RS_TILE <= (size_t)n && (reinterpret_cast<uintptr_t>(keys) & 15u) == 0;
    if (vec) {
        const uint4 *k4 = reinterpret_cast<const uint4 *>(keys + base);
        uint4 v[RS_ITEMS / 4];
#pragma unroll
        for (int j = 0; j < RS_ITEMS / 4; ++j) v[j] = k4[j * RS_THREADS + threadIdx.x];
#pragma unroll
        for (int j = 0; j < RS_ITEMS / 4; ++j) {
            atomicAdd(&cnt[(v[j].x >> shift) & 255u], 1u);
            atomicAdd(&cnt[(v[j].y >> shift) & 255u], 1u);
            atomicAdd(&cnt[(v[j].z >> shift) & 255u], 1u);
            atomicAdd(&cnt[(v[j].w >> shift) & 255u], 1u);
        }
    } else {
#pragma unroll 4
        for (int j = 0; j < RS_ITEMS; ++j) {
            const size_t i = base + (size_t)j * RS_THREADS + threadIdx.x;
            if (i < (size_t)n) atomicAdd(&cnt[(keys[i] >> shift) & 255u], 1u);
        }
    }
    __syncthreads();
    hist[(size_t)threadIdx.x * ntiles + blockIdx.x] = cnt[threadIdx.x];
}

// Block-wide exclusive scan of one value per thread (NW waves); returns the total.
template <int NW = RS_WAVES>
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t *smem /*[NW]*/, uint32_t &total) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint32_t x = wave_incl_scan_u32(v);
    if (lane == 63) smem[wave] = x;
    __syncthreads();
    uint32_t wbase = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < NW; ++w) {
        const uint32_t s = smem[w];
        if (w < wave) wbase += s;
        tot += s;
    }
    __syncthreads();
    total = tot;
    return wbase + x - v;
}

// Exclusive scan of each digit's row hist[d][0..ntiles) (one block per digit; a thread takes
// RS_ITEMS consecutive entries of a round of RS_TILE), writing the digit's total.
__global__ __launch_bounds__(RS_THREADS) void k_radix_scan_rows(uint32_t *__restrict__ hist, int ntiles,
                                                                 uint32_t *__restrict__ digit_total) {
    __shared__ uint32_t sm[RS_WAVES];
    uint32_t *row = hist + (size_t)blockIdx.x * ntiles;
    uint32_t carry = 0;
    for (int c0 = 0; c0 < ntiles; c0 += RS_TILE) {
        const int i0 = c0 + RS_ITEMS * (int)threadIdx.x;
        uint32_t v[RS_ITEMS], t = 0;
#pragma unroll
        for (int k = 0; k < RS_ITEMS; ++k) {
            v[k] = (i0 + k < ntiles) ? row[i0 + k] : 0u;
            t += v[k];
        }
        uint32_t tot;
        uint32_t run = carry + block_exclusive_scan(t, sm, tot);
#pragma unroll
        for (int k = 0; k < RS_ITEMS; ++k) {
            if (i0 + k < ntiles) row[i0 + k] = run;
            run += v[k];
        }
        carry += tot;
    }
    if (threadIdx.x == 0) digit_total[blockIdx.x] = carry;
}

// Payload words carried with the keys through a pass: NP separate arrays in[q] / out[q], or
// (AOS) one array of NP-word records in[0] / out[0], moved with one load and one store each.
struct RadixPayload {
    const uint32_t *in[3];
    uint32_t *out[3];
};

template <int NP>
struct PayRec {
    uint32_t w[NP];
};

// Stable scatter of one pass, staged in LDS, by blocks of SX_THREADS threads (8 waves, SX_ITEMS
// keys each, so two blocks share a CU's LDS).  Wave w of a block takes the contiguous part
// [w * 64 * SX_ITEMS, (w + 1) * 64 * SX_ITEMS) of the tile, 64 consecutive keys per round, with
// the keys and their payload in registers, and ranks each key without block barriers: the
// lanes of the round holding the same digit are found with eight __ballot calls, and the
// key's rank within the wave is the count of its digit in the wave's earlier rounds (kept in
// the wave's own LDS counters, bumped by the digit group's last lane) plus its lower peers.
// Then one pass over the digits turns the per-wave counts into each wave's offset within its
// digit's run of the tile -- (wave, round, lane) order is index order, so the pass is stable
// -- and the keys go to LDS at their digit's place.  The tile, now sorted by digit, is written
// out in LDS order: consecutive threads write consecutive places of one digit's run, so the
// stores coalesce instead of landing one word at a time all over the output.
constexpr int SX_THREADS = 512;
constexpr int SX_WAVES = SX_THREADS / 64;
constexpr int SX_ITEMS = RS_TILE / SX_THREADS;  // 8 keys per thread
static_assert(SX_ITEMS * SX_THREADS == RS_TILE, "a scatter block covers one tile");

template <int NP, bool AOS>
__global__ __launch_bounds__(SX_THREADS) void k_radix_scatter(
    const uint32_t *__restrict__ kin, uint32_t *__restrict__ kout, RadixPayload pl, const int *d_n, int n_host,
    int shift, const uint32_t *__restrict__ hist, const uint32_t *__restrict__ digit_total, int ntiles) {
    __shared__ uint32_t skey[RS_TILE];
    __shared__ uint32_t spay[NP][RS_TILE];
    __shared__ uint32_t wcnt[SX_WAVES][256];  // per wave: its keys of each digit, then its offset in the run
    __shared__ uint32_t tstart[256];          // the tile's first place of each digit
    __shared__ uint32_t gbase[256];           // output place of the tile's first key of each digit
    __shared__ uint32_t sm[SX_WAVES];
    const int n = load_n(d_n, n_host);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    const size_t tbase = (size_t)blockIdx.x * RS_TILE;
    const size_t wbase = tbase + (size_t)wave * (SX_ITEMS * 64);
#pragma unroll
    for (int k = 0; k < SX_WAVES * 256 / SX_THREADS; ++k) (&wcnt[0][0])[k * SX_THREADS + threadIdx.x] = 0;
    {
        // the digits' first output places: the exclusive scan of the pass's digit totals
        uint32_t all;
        const uint32_t dbase = block_exclusive_scan<SX_WAVES>((threadIdx.x < 256) ? digit_total[threadIdx.x] : 0u,
                                                              sm, all);
        if (threadIdx.x < 256) gbase[threadIdx.x] = dbase + hist[(size_t)threadIdx.x * ntiles + blockIdx.x];
    }
    uint32_t key[SX_ITEMS], pay[NP][SX_ITEMS], rank[SX_ITEMS];
#pragma unroll
    for (int j = 0; j < SX_ITEMS; ++j) {
        const size_t i = wbase + (size_t)j * 64 + lane;
        const bool valid = i < (size_t)n;
        key[j] = valid ? kin[i] : 0u;
        if (AOS) {
            PayRec<NP> r{};
            if (valid) r = reinterpret_cast<const PayRec<NP> *>(pl.in[0])[i];
#pragma unroll
            for (int q = 0; q < NP; ++q) pay[q][j] = r.w[q];
        } else {
#pragma unroll
            for (int q = 0; q < NP; ++q) pay[q][j] = valid ? pl.in[q][i] : 0u;
        }
    }
    __syncthreads();
    uint32_t *cw = wcnt[wave];
#pragma unroll
    for (int j = 0; j < SX_ITEMS; ++j) {
        const bool valid = wbase + (size_t)j * 64 + lane < (size_t)n;
        const uint32_t d = (key[j] >> shift) & 255u;
        const uint64_t peers = match_low_bits<8>(d, __ballot(valid));
        const uint32_t before = cw[d];
        rank[j] = before + (uint32_t)__popcll(peers & lt);
        // every lane has read the count before the group's last lane bumps it (one wave:
        // LDS operations in order)
        if (valid && (peers >> lane) == 1ull) cw[d] = before + (uint32_t)__popcll(peers);
    }
    __syncthreads();
    {
        uint32_t acc = 0;  // thread t < 256: digit t
        if (threadIdx.x < 256) {
#pragma unroll
            for (int w = 0; w < SX_WAVES; ++w) {
                const uint32_t c = wcnt[w][threadIdx.x];
                wcnt[w][threadIdx.x] = acc;
                acc += c;
            }
        }
        uint32_t tot;
        const uint32_t ex = block_exclusive_scan<SX_WAVES>(acc, sm, tot);
        if (threadIdx.x < 256) tstart[threadIdx.x] = ex;
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < SX_ITEMS; ++j) {
        if (wbase + (size_t)j * 64 + lane < (size_t)n) {
            const uint32_t d = (key[j] >> shift) & 255u;
            const uint32_t lp = tstart[d] + cw[d] + rank[j];
            skey[lp] = key[j];
#pragma unroll
            for (int q = 0; q < NP; ++q) spay[q][lp] = pay[q][j];
        }
    }
    __syncthreads();
    const int cnt = (n - (int)tbase < RS_TILE) ? n - (int)tbase : RS_TILE;
    for (int lp = threadIdx.x; lp < cnt; lp += SX_THREADS) {
        const uint32_t k = skey[lp];
        const uint32_t d = (k >> shift) & 255u;
        const uint32_t pos = gbase[d] + ((uint32_t)lp - tstart[d]);
        kout[pos] = k;
        if (AOS) {
            PayRec<NP> r;
#pragma unroll
            for (int q = 0; q < NP; ++q) r.w[q] = spay[q][lp];
            reinterpret_cast<PayRec<NP> *>(pl.out[0])[pos] = r;
        } else {
#pragma unroll
            for (int q = 0; q < NP; ++q) pl.out[q][pos] = spay[q][lp];
        }
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
    float inv_L;  // 1/L, exact when the side is a power of two
    float sigma;  // sqrt(2 D dt), in float
    float dt;
    float amp;
    float p_ann;
    uint32_t seed;
};

// v folded into [0, L): v - L floor(v / L), then one fold each way for the rounding.  POW2
// (the side a power of two): v / L as the product v (1/L), which is then exact, so the bits
// are those of the division.
template <bool POW2>
__device__ __forceinline__ float wrap_pos(float v, const PartConsts &p) {
    const float L = p.L;
    const float q = __builtin_floorf(POW2 ? v * p.inv_L : v / L);
    v = v - L * q;
    if (v < 0.0f) v = v + L;
    if (v >= L) v = v - L;
    return v;
}

// Move, in float: x += v_x dt + sigma g_x with v = A(-sin(2 pi y/L), sin(2 pi x/L)),
// periodic, g the Box-Muller pair of Threefry key {id, seed}, counter {step, 0}.
template <bool POW2>
__device__ __forceinline__ void move_particle(float xo, float yo, uint32_t idv, int step, const PartConsts &p,
                                              float &xn, float &yn) {
    float gx, gy;
    gauss2_f(idv, p.seed, (uint32_t)step, &gx, &gy);
    const float ty = POW2 ? (2.0f * yo) * p.inv_L : (2.0f * yo) / p.L;
    const float tx = POW2 ? (2.0f * xo) * p.inv_L : (2.0f * xo) / p.L;
    // positions in [0, L): ty, tx in [0, 2]
    const float vx = -p.amp * sinpi_pf4(ty);
    const float vy = p.amp * sinpi_pf4(tx);
    xn = wrap_pos<POW2>(xo + vx * p.dt + p.sigma * gx, p);
    yn = wrap_pos<POW2>(yo + vy * p.dt + p.sigma * gy, p);
}

// Element i of a per-particle array through a 32-bit byte offset (i < 2^28, the capacity
// limit): the address is the array's scalar base plus a 32-bit vector offset, with no
// 64-bit address arithmetic per access.
template <typename T>
__device__ __forceinline__ T &at(T *p, uint32_t i) {
    return *reinterpret_cast<T *>(reinterpret_cast<char *>(p) + (i << 2));
}
template <typename T>
__device__ __forceinline__ const T &at(const T *p, uint32_t i) {
    return *reinterpret_cast<const T *>(reinterpret_cast<const char *>(p) + (i << 2));
}

// Move, writing the cell key and the record {x, y, id} that the sort carries, and the sort's
// first-pass histogram (low 8 bits of the key) of the block's tile: one block per tile.
constexpr int MV_THREADS = 1024;

template <bool POW2>
__global__ __launch_bounds__(MV_THREADS) void k_move(const float *__restrict__ x, const float *__restrict__ y,
                                                     const uint32_t *__restrict__ id, uint32_t *__restrict__ keys,
                                                     PayRec<3> *__restrict__ recs, const int *d_n, int step,
                                                     PartConsts p, uint32_t *__restrict__ hist, int ntiles) {
    __shared__ uint32_t cnt[256];
    const int n = *d_n;
    if (threadIdx.x < 256) cnt[threadIdx.x] = 0u;
    __syncthreads();
    const int base = blockIdx.x * RS_TILE;
#pragma unroll
    for (int j = 0; j < RS_TILE / MV_THREADS; ++j) {
        const int i = base + j * MV_THREADS + (int)threadIdx.x;
        if (i < n) {
            float xn, yn;
            const uint32_t idv = at(id, (uint32_t)i);
            move_particle<POW2>(at(x, (uint32_t)i), at(y, (uint32_t)i), idv, step, p, xn, yn);
            const uint32_t k = (uint32_t)((int)yn * p.side + (int)xn);
            at(keys, (uint32_t)i) = k;
            recs[i] = PayRec<3>{{__float_as_uint(xn), __float_as_uint(yn), idv}};
            atomicAdd(&cnt[k & 255u], 1u);
        }
    }
    __syncthreads();
    if (threadIdx.x < 256) hist[(size_t)threadIdx.x * ntiles + blockIdx.x] = cnt[threadIdx.x];
}

// A+A->0 draws: pairs 2m and 2m+1 of cell c take the two words of the Threefry block
// {step, m}, key {c, ~seed}; pair j is removed when u01(its word) < p.
__device__ __forceinline__ bool pair_killed(int step, uint32_t j, uint32_t c, const PartConsts &p) {
    const u32x2 w = threefry2x32_20((uint32_t)step, j >> 1, c, ~p.seed);
    return u01((j & 1u) ? w.y : w.x) < p.p_ann;
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
// gather: every particle is read and written once per kernel, in nearly sequential order.
//
// The move takes tiles of consecutive old cells (64 per wave): a tile's particles are the
// consecutive range off[c0] .. off[c0 + 64), so the per-cell counts and every particle's rank
// among the earlier particles of its cell with its move class stay inside the wave (LDS).
// Each particle then carries one code word to the scatter, which is flat over the particles.
//
// Direction codes: d = (dy + 1) * 3 + (dx + 1) for a move by (dx, dy) in {-1, 0, 1}^2 on the
// torus (side >= 3), kFar otherwise; a near particle's old cell is its new cell moved by 8 - d.
constexpr int kFar = 9;
constexpr int CT = 256;  // old cells per block of the move (64 per wave)

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

// Block -> work item with XCD-aware placement: the hardware dispatches block b to XCD b % 8,
// so XCD k gets the contiguous items [k per, (k + 1) per) and the gathers from neighbouring
// rows, like the partial-line stores into one region, stay in one L2.  The grid is 8 * per
// blocks; a block past the last item returns at once.
__device__ __forceinline__ uint32_t xcd_item(uint32_t b, uint32_t per) { return (b & 7u) * per + (b >> 3); }

// First index of cell c in the list a step leaves behind: the scanned offset of its 256-cell
// tile plus the record's offset within the tile.
__device__ __forceinline__ uint32_t cell_offset(const uint4 *__restrict__ rec, const uint32_t *__restrict__ tile_off,
                                                int c) {
    return tile_off[c >> 8] + rec[c].x;
}

// 1. move the particles of the wave's 64 old cells, 64 consecutive particles per round,
//    classify each move and rank it: among the particles of one old cell with one move class
//    the rank is the count carried from the earlier rounds (LDS, per wave) plus the lower
//    lanes of the round in the same cell run with the same class (a few __ballot calls).  Code
//    word per particle: d | rank << 4 for a near move, kFar | old cell << 4 for a far one.
//    The carried counts end as cnt[d][o] (every entry of the wave's cells written).  Far
//    moves are listed with order-free atomics: the list is only counted, never walked in order.
template <bool POW2>
__global__ __launch_bounds__(CT) void k_cs_move(const float *__restrict__ x, const float *__restrict__ y,
                                                const uint32_t *__restrict__ id, const uint4 *__restrict__ rec,
                                                const uint32_t *__restrict__ tile_off, const int *d_n, int A,
                                                int step, PartConsts p,
                                                float *__restrict__ x2, float *__restrict__ y2,
                                                uint32_t *__restrict__ code, uint32_t *__restrict__ cnt,
                                                uint32_t *__restrict__ far_cnt, FarEntry *__restrict__ far_list,
                                                uint32_t *__restrict__ far_in) {
    __shared__ uint32_t carry[CT / 64][9][64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int cw = blockIdx.x * CT + wave * 64;  // the wave's first cell
    const int ncell = (A - cw < 64) ? A - cw : 64;
    if (ncell <= 0) return;  // wave-uniform; no block barrier follows
    uint32_t(*cr)[64] = carry[wave];
#pragma unroll
    for (int d = 0; d < 9; ++d) cr[d][lane] = 0u;
    const uint32_t n = (uint32_t)*d_n;
    const uint32_t q0 = min(cell_offset(rec, tile_off, cw), n);
    const uint32_t q1 = (cw + ncell < A) ? min(cell_offset(rec, tile_off, cw + ncell), n) : n;
    const uint64_t lower = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    const int side = p.side;
    for (uint32_t b0 = q0; b0 < q1; b0 += 64) {
        const uint32_t i = b0 + (uint32_t)lane;
        const bool valid = i < q1;
        float xo = 0.f, yo = 0.f;
        uint32_t idv = 0u;
        if (valid) {
            xo = at(x, i);
            yo = at(y, i);
            idv = at(id, i);
        }
        float xn, yn;
        move_particle<POW2>(xo, yo, idv, step, p, xn, yn);
        const int ox = (int)xo, oy = (int)yo, cx = (int)xn, cy = (int)yn;
        const int dx = torus_step(ox, cx, side);
        const int dy = torus_step(oy, cy, side);
        const bool near = dx >= -1 && dx <= 1 && dy >= -1 && dy <= 1;
        const uint32_t d = near ? (uint32_t)((dy + 1) * 3 + (dx + 1)) : (uint32_t)kFar;
        const uint32_t o = (uint32_t)(oy * side + ox);
        const uint32_t lo = o - (uint32_t)cw;  // the wave's cell i sits in (below 64 for a consistent list)
        const bool ranked = valid && near && lo < (uint32_t)ncell;
        // the round's lanes are sorted by cell: a lane's cell run starts at the last head
        // (a lane whose cell differs from its lower neighbour's) at or below it
        const uint32_t lo_prev = (uint32_t)__shfl_up((int)lo, 1);
        const uint64_t heads = __ballot(lane == 0 || lo != lo_prev);
        const int s0 = 63 - __clzll(heads & (lower | (1ull << lane)));
        const uint64_t run_lower = lower & ~((1ull << s0) - 1ull);
        const uint64_t same_d = match_low_bits<4>(d, __ballot(ranked));
        uint32_t rank = 0;
        if (ranked) rank = cr[d][lo] + (uint32_t)__popcll(same_d & run_lower);
        // every read of the round precedes its increments (one wave: LDS operations in order)
        if (ranked) atomicAdd(&cr[d][lo], 1u);
        if (valid) {
            at(x2, i) = xn;
            at(y2, i) = yn;
            at(code, i) = ranked ? (d | (rank << 4)) : ((uint32_t)kFar | (o << 4));
            if (!near) {
                const uint32_t c = (uint32_t)(cy * side + cx);
                const uint32_t k = atomicAdd(far_cnt, 1u);
                far_list[k] = FarEntry{i, c};
                atomicAdd(&far_in[c], 1u);
            }
        }
    }
    if (lane < ncell) {
#pragma unroll
        for (int d = 0; d < 9; ++d) cnt[(size_t)d * A + cw + lane] = cr[d][lane];
    }
}

// Exclusive scan in place of ntiles tile sums by one block of NT threads (4096 / NT tiles per
// thread per round); returns the total.
template <int NT>
__device__ __forceinline__ uint32_t scan_tiles(uint32_t *__restrict__ tile_sum, int ntiles, uint32_t *wsum) {
    constexpr int PER = 4096 / NT, NW = NT / 64;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t carry = 0;
    for (int c0 = 0; c0 < ntiles; c0 += 4096) {
        const int i0 = c0 + PER * (int)threadIdx.x;
        uint32_t v[PER], t = 0;
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            v[k] = (i0 + k < ntiles) ? tile_sum[i0 + k] : 0u;
            t += v[k];
        }
        const uint32_t incl = wave_incl_scan_u32(t);
        if (lane == 63) wsum[wave] = incl;
        __syncthreads();
        uint32_t wbase = 0, all = 0;
#pragma unroll
        for (int w = 0; w < NW; ++w) {
            const uint32_t sw = wsum[w];
            wbase += (w < wave) ? sw : 0u;
            all += sw;
        }
        __syncthreads();
        uint32_t run = carry + wbase + incl - t;
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            if (i0 + k < ntiles) tile_sum[i0 + k] = run;
            run += v[k];
        }
        carry += all;
    }
    return carry;
}

// 3. exclusive scan of the survivors' tile sums; the new count goes to the other count slot
__global__ __launch_bounds__(1024) void k_coh_scan(uint32_t *__restrict__ tile_sum, int ntiles,
                                                    const int *__restrict__ n_cur, int *__restrict__ n_next,
                                                    unsigned long long *processed) {
    __shared__ uint32_t wsum[16];
    const uint32_t total = scan_tiles<1024>(tile_sum, ntiles, wsum);
    if (threadIdx.x == 0) {
        *processed += (unsigned long long)*n_cur;
        *n_next = (int)total;
    }
}

// The A+A->0 draws of the pairs of a wave's 64 cells (lane k: cell cell0 + k with `pairs`
// pairs), one Threefry block per two pairs, flattened over the wave: lane k takes the wave's
// blocks k, k + 64, ..., whichever cells they belong to, so a wave runs ceil(blocks / 64)
// draws instead of its busiest cell's count.  Returns the lane's cell's kills and the kill
// mask of its first 32 pairs.  sbeg, spairs, skill, smask: 64 words of LDS per wave.  Every
// thread of the block calls it (it holds block barriers).
__device__ __forceinline__ void wave_pair_draws(uint32_t pairs, uint32_t cell0, int step, const PartConsts &p,
                                                uint32_t *sbeg, uint32_t *spairs, uint32_t *skill,
                                                uint32_t *smask, uint32_t &kills, uint32_t &mask) {
    const int lane = threadIdx.x & 63;
    const uint32_t nb = (pairs + 1) >> 1;
    const uint32_t incl = wave_incl_scan_u32(nb);
    const uint32_t W = (uint32_t)__builtin_amdgcn_readlane((int)incl, 63);
    sbeg[lane] = incl - nb;
    spairs[lane] = pairs;
    skill[lane] = 0u;
    smask[lane] = 0u;
    __syncthreads();
    for (uint32_t it = (uint32_t)lane; it < W; it += 64) {
        int lo = 0, hi = 63;  // the last lane whose cell's blocks start at or before it
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (sbeg[mid] <= it) lo = mid;
            else hi = mid - 1;
        }
        const uint32_t m = it - sbeg[lo];
        const uint32_t j0 = 2 * m;
        const u32x2 w = threefry2x32_20((uint32_t)step, m, cell0 + (uint32_t)lo, ~p.seed);
        const bool k0 = u01(w.x) < p.p_ann;
        const bool k1 = (j0 + 1 < spairs[lo]) && u01(w.y) < p.p_ann;
        if (k0 || k1) {
            atomicAdd(&skill[lo], (uint32_t)k0 + (uint32_t)k1);
            if (j0 < 32) atomicOr(&smask[lo], (k0 ? 1u << j0 : 0u) | (k1 ? 1u << (j0 + 1) : 0u));
        }
    }
    __syncthreads();
    kills = skill[lane];
    mask = smask[lane];
}

// Place of a particle at rank r among the T particles of new cell c (kill mask m of the
// first 32 pairs, survivors' offset base): false when its pair is killed, else the index
// base + r - 2 (killed pairs before it).
__device__ __forceinline__ bool survivor_place(uint32_t r, uint32_t T, uint32_t m, uint32_t base, uint32_t c,
                                               int step, const PartConsts &p, uint32_t &pos) {
    const uint32_t j = r >> 1;
    const bool partner = (r | 1u) < T;
    bool killed = false;
    if (partner) killed = (j < 32) ? ((m >> j) & 1u) != 0 : pair_killed(step, j, c, p);
    if (killed) return false;
    uint32_t before = (uint32_t)__popc(m & ((j < 32) ? ((1u << j) - 1u) : 0xffffffffu));
    for (uint32_t jj = 32; jj < j; ++jj) before += pair_killed(step, jj, c, p) ? 1u : 0u;
    pos = base + r - 2u * before;
    return true;
}

// 2. per new cell: the particles arriving from each source (pre[d][c] = the arrivals from
//    the sources with a smaller key than direction d's; pre[4][c], the particles that stay,
//    goes to the cell record), the total T with the far arrivals,
//    the A+A->0 draws of its T/2 pairs, the kill mask of the first 32 pairs and the
//    survivors, summed per 256-cell tile (the draws flattened over the wave, wave_pair_draws).
__global__ __launch_bounds__(256) void k_cs_cells(const uint32_t *__restrict__ cnt, uint32_t *__restrict__ far_in,
                                                  int A, int side, int step, PartConsts p, uint32_t *__restrict__ pre,
                                                  uint4 *__restrict__ rec, uint32_t *__restrict__ tile_sum,
                                                  uint32_t *__restrict__ far_cnt_next, uint32_t ntiles, uint32_t per) {
    __shared__ uint32_t sm[RS_WAVES];
    __shared__ uint32_t sbeg[RS_WAVES][64];   // the wave's first block of each lane's cell
    __shared__ uint32_t spairs[RS_WAVES][64];
    __shared__ uint32_t skill[RS_WAVES][64];
    __shared__ uint32_t smask[RS_WAVES][64];
    const uint32_t tile = xcd_item(blockIdx.x, per);
    if (tile >= ntiles) return;  // block-uniform, before any barrier
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int c = (int)tile * 256 + threadIdx.x;
    if (c == 0) *far_cnt_next = 0u;  // the far list of the next step starts empty
    uint32_t T = 0, pre4 = 0;
    if (c < A) {
        const int cy = c / side, cx = c - cy * side;
        uint32_t key[9], nn[9];
#pragma unroll
        for (int d = 0; d < 9; ++d) {
            key[d] = moved_cell(cx, cy, 8 - d, side);  // the source whose move d lands on c
            nn[d] = cnt[(size_t)d * A + key[d]];
            T += nn[d];
        }
        uint32_t q[9];
        if (cx >= 1 && cx < side - 1 && cy >= 1 && cy < side - 1) {
            // no source wraps: key[d] = c + (1 - d / 3) side + (1 - d % 3) rises as d falls, so
            // the sources below d's are those of d + 1 .. 8
            uint32_t run = 0;
#pragma unroll
            for (int d = 8; d >= 0; --d) {
                q[d] = run;
                run += nn[d];
            }
        } else {
#pragma unroll
            for (int d = 0; d < 9; ++d) {
                q[d] = 0;
#pragma unroll
                for (int e = 0; e < 9; ++e) q[d] += (key[e] < key[d]) ? nn[e] : 0u;
            }
        }
        pre4 = q[4];  // the particles that stay: kept in the cell record
#pragma unroll
        for (int d = 0; d < 9; ++d)
            if (d != 4) pre[(size_t)d * A + c] = q[d];
        const uint32_t fin = far_in[c];
        if (fin) far_in[c] = 0u;  // counted: clear for the next step
        T += fin;
    }
    uint32_t kills, mask;
    wave_pair_draws(T >> 1, tile * 256u + (uint32_t)(wave * 64), step, p, sbeg[wave], spairs[wave], skill[wave],
                    smask[wave], kills, mask);
    const uint32_t S = T - 2 * kills;
    uint32_t total;
    const uint32_t ex = block_exclusive_scan(S, sm, total);
    if (c < A) rec[c] = make_uint4(ex, T, mask, pre4);  // .x: the survivors before c in its tile
    if (threadIdx.x == 0) tile_sum[tile] = total;
}

// 4. every particle to its place in the new order; the killed pairs drop out.  Flat over the
//    particles: a thread takes SC_PER of them and issues their loads together (the kernel is
//    bound by memory latency, so each wave keeps several particles' loads in flight), and one
//    16-byte load of its new cell's record.  A near particle's rank in its new cell is
//    pre[d][c] (the record's for d = 4) plus its rank within (old cell, d) from the code word; a far one's counts the arrivals from the sources below its old cell.  Both
//    add the far arrivals with a smaller index.
constexpr int SC_BLOCK = 256, SC_PER = 2, SC_SPAN = SC_BLOCK * SC_PER;

__global__ __launch_bounds__(SC_BLOCK) void k_cs_scatter(
    const float *__restrict__ x2, const float *__restrict__ y2, const uint32_t *__restrict__ id,
    const uint32_t *__restrict__ code, const uint32_t *__restrict__ cnt, const uint32_t *__restrict__ pre,
    const uint4 *__restrict__ rec, const uint32_t *__restrict__ tile_off, const uint32_t *__restrict__ far_cnt,
    const FarEntry *__restrict__ far_list, const int *d_n, int A, int side, int step, PartConsts p,
    float *__restrict__ xo, float *__restrict__ yo, uint32_t *__restrict__ ido, uint32_t cap, uint32_t per) {
    const uint32_t n = (uint32_t)*d_n;
    const uint32_t b0 = xcd_item(blockIdx.x, per) * SC_SPAN;
    if (b0 >= n) return;
    const uint32_t nf = *far_cnt;
    float xv[SC_PER], yv[SC_PER];
    uint32_t idv[SC_PER], cd[SC_PER];
#pragma unroll
    for (int q = 0; q < SC_PER; ++q) {
        const uint32_t i = b0 + q * SC_BLOCK + threadIdx.x;
        const bool v = i < n;
        xv[q] = v ? at(x2, i) : 0.f;
        yv[q] = v ? at(y2, i) : 0.f;
        idv[q] = v ? at(id, i) : 0u;
        cd[q] = v ? at(code, i) : (uint32_t)kFar;
    }
    uint32_t c[SC_PER], pr[SC_PER], T[SC_PER], m[SC_PER], base[SC_PER];
#pragma unroll
    for (int q = 0; q < SC_PER; ++q) {
        const uint32_t i = b0 + q * SC_BLOCK + threadIdx.x;
        const bool v = i < n;
        const uint32_t d = cd[q] & 15u;
        c[q] = (uint32_t)((int)yv[q] * side + (int)xv[q]);
        const uint4 rc = v ? rec[c[q]] : make_uint4(0u, 0u, 0u, 0u);
        pr[q] = (v && d < 9) ? ((d == 4) ? rc.w : pre[(size_t)d * A + c[q]]) + (cd[q] >> 4) : 0u;
        T[q] = rc.y;
        m[q] = rc.z;
        base[q] = v ? tile_off[c[q] >> 8] + rc.x : 0u;
    }
#pragma unroll
    for (int q = 0; q < SC_PER; ++q) {
        const uint32_t i = b0 + q * SC_BLOCK + threadIdx.x;
        if (i >= n) continue;
        uint32_t r = pr[q];
        if ((cd[q] & 15u) >= 9) {
            const int cy = (int)yv[q], cx = (int)xv[q];
            const uint32_t o = cd[q] >> 4;
            for (int e = 0; e < 9; ++e) {
                const uint32_t s = moved_cell(cx, cy, 8 - e, side);
                if (s < o) r += cnt[(size_t)e * A + s];
            }
        }
        for (uint32_t f = 0; f < nf; ++f) {
            const FarEntry fe = far_list[f];
            r += (fe.c == c[q] && fe.i < i) ? 1u : 0u;
        }
        uint32_t pos;
        if (!survivor_place(r, T[q], m[q], base[q], c[q], step, p, pos)) continue;
        if (pos >= cap) continue;  // never for a consistent list; keeps every store inside the buffers
        at(xo, pos) = xv[q];
        at(yo, pos) = yv[q];
        at(ido, pos) = idv[q];
    }
}

// ---- after the radix sort (a list in any order, moved and sorted by the new cell) ----------
// off[c] = #(key < c) for c in [0, A] from the sorted keys: particle i writes off[c] = i for the
// cells after the previous particle's key up to its own (the empty cells before it and its
// own cell when it starts one); the thread past the last particle writes the cells after it.
// Every entry is written once; an empty stretch of cells is written by one thread.  A thread
// takes 4 consecutive particles (one 16-byte load where the keys are aligned and all live),
// the key before them from the lower lane.
__global__ __launch_bounds__(256) void k_offsets_sorted(const uint32_t *__restrict__ key, const int *d_n, int A,
                                                        uint32_t *__restrict__ off) {
    const uint32_t n = (uint32_t)*d_n;
    const uint32_t i0 = 4u * (blockIdx.x * blockDim.x + threadIdx.x);
    const uint32_t lane = threadIdx.x & 63u;
    uint32_t k[4];
    if (i0 + 4u <= n && (reinterpret_cast<uintptr_t>(key) & 15u) == 0) {
        const uint4 v = *reinterpret_cast<const uint4 *>(key + i0);
        k[0] = v.x;
        k[1] = v.y;
        k[2] = v.z;
        k[3] = v.w;
    } else {
#pragma unroll
        for (int j = 0; j < 4; ++j) k[j] = (i0 + j < n) ? key[i0 + j] : 0u;
    }
    // the key of particle i0 - 1: the lower lane's last, or loaded by a wave's first lane
    uint32_t prev = (uint32_t)__shfl_up((int)k[3], 1);
    if (lane == 0 && i0 > 0 && i0 - 1 < n) prev = key[i0 - 1];
    if (i0 > n) return;  // after the shuffle: every lane of the wave took part
    const uint32_t a = (uint32_t)A;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const uint32_t i = i0 + j;
        if (i > n) break;
        const uint32_t lo = (i == 0) ? 0u : min(prev, a - 1u) + 1u;
        const uint32_t hi = (i == n) ? a : min(k[j], a);
        for (uint32_t c = lo; c <= hi; ++c) off[c] = i;
        prev = k[j];
    }
}

// per cell: its particles T (from the sorted list's offsets), the A+A->0 draws, the kill mask
// and the survivors, summed per 256-cell tile
__global__ __launch_bounds__(256) void k_rs_cells(const uint32_t *__restrict__ off_sorted, int A, int step,
                                                  PartConsts p, uint4 *__restrict__ rec,
                                                  uint32_t *__restrict__ tile_s, uint32_t ntiles, uint32_t per) {
    __shared__ uint32_t sm[RS_WAVES];
    __shared__ uint32_t sbeg[RS_WAVES][64], spairs[RS_WAVES][64], skill[RS_WAVES][64], smask[RS_WAVES][64];
    const uint32_t tile = xcd_item(blockIdx.x, per);
    if (tile >= ntiles) return;  // block-uniform, before any barrier
    const int wave = threadIdx.x >> 6;
    const int c = (int)tile * 256 + threadIdx.x;
    const uint32_t T = (c < A) ? off_sorted[c + 1] - off_sorted[c] : 0u;
    uint32_t kills, mask;
    wave_pair_draws(T >> 1, tile * 256u + (uint32_t)(wave * 64), step, p, sbeg[wave], spairs[wave], skill[wave],
                    smask[wave], kills, mask);
    const uint32_t S = T - 2 * kills;
    uint32_t total;
    const uint32_t ex = block_exclusive_scan(S, sm, total);
    if (c < A) rec[c] = make_uint4(ex, T, mask, 0u);  // .x: the survivors before c in its tile
    if (threadIdx.x == 0) tile_s[tile] = total;
}

// the survivors of the sorted list to their places (stable compaction), flat over the list:
// a particle's rank in its cell is its index less the cell's first index; its cell comes from
// its position as the move computed its key
__global__ __launch_bounds__(SC_BLOCK) void k_rs_compact(
    const PayRec<3> *__restrict__ recs, const uint32_t *__restrict__ off_sorted,
    const uint4 *__restrict__ rec, const uint32_t *__restrict__ tile_off, const int *d_n, int step, PartConsts p,
    float *__restrict__ xo, float *__restrict__ yo, uint32_t *__restrict__ ido, uint32_t cap, uint32_t per) {
    const uint32_t n = (uint32_t)*d_n;
    const uint32_t b0 = xcd_item(blockIdx.x, per) * SC_SPAN;
    if (b0 >= n) return;
    const uint32_t A = (uint32_t)p.side * (uint32_t)p.side;
    float xv[SC_PER], yv[SC_PER];
    uint32_t idv[SC_PER], c[SC_PER];
#pragma unroll
    for (int q = 0; q < SC_PER; ++q) {
        const uint32_t i = b0 + q * SC_BLOCK + threadIdx.x;
        const bool v = i < n;
        PayRec<3> r{};
        if (v) r = *reinterpret_cast<const PayRec<3> *>(reinterpret_cast<const char *>(recs) + i * 12u);
        xv[q] = __uint_as_float(r.w[0]);
        yv[q] = __uint_as_float(r.w[1]);
        idv[q] = r.w[2];
        c[q] = v ? (uint32_t)((int)yv[q] * p.side + (int)xv[q]) : 0u;
    }
    uint32_t s0[SC_PER], T[SC_PER], m[SC_PER], base[SC_PER];
#pragma unroll
    for (int q = 0; q < SC_PER; ++q) {
        const uint32_t i = b0 + q * SC_BLOCK + threadIdx.x;
        const bool v = i < n && c[q] < A;
        s0[q] = v ? off_sorted[c[q]] : i;
        const uint4 rc = v ? rec[c[q]] : make_uint4(0u, 0u, 0u, 0u);
        T[q] = rc.y;
        m[q] = rc.z;
        base[q] = v ? tile_off[c[q] >> 8] + rc.x : 0u;
    }
#pragma unroll
    for (int q = 0; q < SC_PER; ++q) {
        const uint32_t i = b0 + q * SC_BLOCK + threadIdx.x;
        if (i >= n || c[q] >= A) continue;
        uint32_t pos;
        if (!survivor_place(i - s0[q], T[q], m[q], base[q], c[q], step, p, pos)) continue;
        if (pos >= cap) continue;  // never for a consistent list; keeps every store inside the buffers
        at(xo, pos) = xv[q];
        at(yo, pos) = yv[q];
        at(ido, pos) = idv[q];
    }
}

// Cell keys of a list from its positions (the sorted-list step does not store them).
__global__ void k_keys_from_pos(const float *__restrict__ x, const float *__restrict__ y, const int *d_n, int side,
                                uint32_t *__restrict__ keys) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= *d_n) return;
    keys[i] = (uint32_t)((int)y[i] * side + (int)x[i]);
}

}  // namespace gpgmp

// ---------------------------------------------------------------------------------------
#include "particles_launch.h"

namespace gpgmp {

int radix_tiles(int n) { return (n + RS_TILE - 1) / RS_TILE; }

// LSD passes of 8-bit digits over key_bits bits, ping-ponging the keys and their payload
// between the given and the tmp buffers: one value array (np = 1), or one array of 3-word
// records (np = 3, aos); *result_in_tmp tells where the result is.  hist0_ready: the first
// pass's tile histograms are already in hist.
hipError_t radix_sort_payload(uint32_t *keys, uint32_t *ktmp, uint32_t *const *pay, uint32_t *const *ptmp, int np,
                              bool aos, const int *d_n, int n_ub, int key_bits, uint32_t *hist,
                              uint32_t *digit_total, hipStream_t s, bool *result_in_tmp, bool hist0_ready) {
    const int narr = aos ? 1 : np;
    const int nt = std::max(radix_tiles(n_ub), 1);
    uint32_t *ki = keys, *ko = ktmp;
    uint32_t *pi[3] = {nullptr, nullptr, nullptr}, *po[3] = {nullptr, nullptr, nullptr};
    for (int q = 0; q < narr; ++q) {
        pi[q] = pay[q];
        po[q] = ptmp[q];
    }
    int passes = 0;
    for (int shift = 0; shift < key_bits; shift += 8) {
        if (shift > 0 || !hist0_ready)
            hipLaunchKernelGGL(k_radix_hist, dim3(nt), dim3(RS_THREADS), 0, s, ki, d_n, n_ub, shift, hist, nt);
        hipLaunchKernelGGL(k_radix_scan_rows, dim3(256), dim3(RS_THREADS), 0, s, hist, nt, digit_total);
        RadixPayload pl;
        for (int q = 0; q < 3; ++q) {
            pl.in[q] = pi[q];
            pl.out[q] = po[q];
        }
        if (np == 1 && !aos)
            hipLaunchKernelGGL((k_radix_scatter<1, false>), dim3(nt), dim3(SX_THREADS), 0, s, ki, ko, pl, d_n, n_ub,
                               shift, hist, digit_total, nt);
        else if (np == 3 && aos)
            hipLaunchKernelGGL((k_radix_scatter<3, true>), dim3(nt), dim3(SX_THREADS), 0, s, ki, ko, pl, d_n, n_ub,
                               shift, hist, digit_total, nt);
        else
            return hipErrorInvalidValue;
        std::swap(ki, ko);
        for (int q = 0; q < narr; ++q) std::swap(pi[q], po[q]);
        ++passes;
    }
    *result_in_tmp = (passes & 1) != 0;
    return hipGetLastError();
}

hipError_t radix_sort_pairs(uint32_t *keys, uint32_t *vals, uint32_t *ktmp, uint32_t *vtmp, const int *d_n,
                            int n_ub, int key_bits, uint32_t *hist, uint32_t *digit_total, hipStream_t s,
                            bool *result_in_tmp) {
    uint32_t *pay[1] = {vals}, *ptmp[1] = {vtmp};
    return radix_sort_payload(keys, ktmp, pay, ptmp, 1, false, d_n, n_ub, key_bits, hist, digit_total, s,
                              result_in_tmp, false);
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
    p.inv_L = 1.0f / (float)a.side;
    p.sigma = __builtin_sqrtf(2.0f * a.D * a.dt);
    p.dt = a.dt;
    p.amp = a.vortex_amp;
    p.p_ann = a.p_annihilate;
    p.seed = a.seed;
    return p;
}

static bool pow2_side(int side) { return side > 0 && (side & (side - 1)) == 0; }

hipError_t particle_step_sorted(const ParticleStepArgs &a, int step, const SortedStepBuffers &b, hipStream_t s) {
    const PartConsts p = consts(a);
    const int A = a.side * a.side;
    const int tiles = (A + CT - 1) / CT;
    const dim3 blk(CT);
    // the current list's offsets: b.tile_sum (scanned) plus the records' in-tile offsets, from
    // the step that made the list; both are overwritten only after the move has read them
    hipLaunchKernelGGL((pow2_side(a.side) ? k_cs_move<true> : k_cs_move<false>), dim3(tiles), blk, 0, s, b.x, b.y,
                       b.id, (const uint4 *)b.rec, (const uint32_t *)b.tile_sum, b.n_cur, A, step, p, b.x2, b.y2,
                       b.code, b.cnt, b.far_cnt, (FarEntry *)b.far_list, b.far_in);
    const uint32_t tper = (uint32_t)(tiles + 7) / 8;
    hipLaunchKernelGGL(k_cs_cells, dim3(8 * tper), blk, 0, s, b.cnt, b.far_in, A, a.side, step, p, b.pre, b.rec,
                       b.tile_sum, b.far_cnt_next, (uint32_t)tiles, tper);
    hipLaunchKernelGGL(k_coh_scan, dim3(1), dim3(1024), 0, s, b.tile_sum, tiles, b.n_cur, b.n_next, b.processed);
    const int sblocks = std::max((b.n_ub + SC_SPAN - 1) / SC_SPAN, 1);
    const uint32_t sper = (uint32_t)(sblocks + 7) / 8;
    hipLaunchKernelGGL(k_cs_scatter, dim3(8 * sper), dim3(SC_BLOCK), 0, s, b.x2, b.y2, b.id, b.code, b.cnt, b.pre,
                       (const uint4 *)b.rec, (const uint32_t *)b.tile_sum, b.far_cnt, (const FarEntry *)b.far_list,
                       b.n_cur, A, a.side, step, p, b.xo, b.yo, b.ido, (uint32_t)b.capacity, sper);
    return hipGetLastError();
}

hipError_t launch_keys_from_pos(const float *x, const float *y, const int *d_n, int n_ub, int side, uint32_t *keys,
                                hipStream_t s) {
    if (n_ub <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_keys_from_pos, dim3((n_ub + 255) / 256), dim3(256), 0, s, x, y, d_n, side, keys);
    return hipGetLastError();
}

size_t far_entry_bytes() { return sizeof(FarEntry); }

hipError_t particle_step(const ParticleStepArgs &a, int step, ParticleBuffers &b, hipStream_t s) {
    const PartConsts p = consts(a);
    const int n_ub = b.n_ub;
    const int A = a.side * a.side;
    const dim3 blk(256);
    // 1. move: cell keys, the records {x, y, id} and the first digit's tile histograms
    const int nt = std::max(radix_tiles(n_ub), 1);
    hipLaunchKernelGGL((pow2_side(a.side) ? k_move<true> : k_move<false>), dim3(nt), dim3(MV_THREADS), 0, s, b.x,
                       b.y, b.id, b.key, (PayRec<3> *)b.recA, b.d_n, step, p, b.hist, nt);
    // 2. stable sort by cell key carrying the records; an odd number of passes (an all-zero
    //    digit pass above the key is a stable copy) leaves the result in the second buffers,
    //    from which the compaction writes the list back
    int bits = a.key_bits;
    if ((((bits + 7) / 8) & 1) == 0) bits += 8;
    bool in_tmp = false;
    uint32_t *pay[1] = {b.recA};
    uint32_t *ptmp[1] = {b.recB};
    hipError_t e = radix_sort_payload(b.key, b.key2, pay, ptmp, 3, true, b.d_n, n_ub, bits, b.hist, b.digit_total,
                                      s, &in_tmp, true);
    if (e != hipSuccess) return e;
    if (!in_tmp) return hipErrorInvalidValue;
    // 3. the sorted list's offsets; per cell: A+A->0 draws and survivors with their in-tile
    //    offsets (the records), the tile sums scanned: a survivor cell's first index (this
    //    step's output, the next step's input) is tile offset + record offset; the new count
    //    to d_n2
    const int tiles = (A + 255) / 256;
    const uint32_t tper = (uint32_t)(tiles + 7) / 8;
    hipLaunchKernelGGL(k_offsets_sorted, dim3(n_ub / 1024 + 1), blk, 0, s, b.key2, b.d_n, A, b.off_sorted);
    hipLaunchKernelGGL(k_rs_cells, dim3(8 * tper), blk, 0, s, b.off_sorted, A, step, p, b.rec, b.tile_s,
                       (uint32_t)tiles, tper);
    hipLaunchKernelGGL(k_coh_scan, dim3(1), dim3(1024), 0, s, b.tile_s, tiles, b.d_n, b.d_n2, b.processed);
    // 4. stable compaction back into the primary buffers
    const int cblocks = std::max((n_ub + SC_SPAN - 1) / SC_SPAN, 1);
    const uint32_t cper = (uint32_t)(cblocks + 7) / 8;
    hipLaunchKernelGGL(k_rs_compact, dim3(8 * cper), dim3(SC_BLOCK), 0, s, (const PayRec<3> *)b.recB,
                       b.off_sorted, b.rec, b.tile_s, b.d_n, step, p, b.x, b.y, b.id, (uint32_t)b.capacity, cper);
    return hipGetLastError();
}

}  // namespace gpgmp
