// Host-side launch interface of the particle kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

namespace gpgmp {

struct ParticleStepArgs {
    int side;
    float D, dt, vortex_amp, p_annihilate;
    uint32_t seed;
    int key_bits;  // bits the cell keys occupy (passes = ceil(key_bits / 8))
};

// Device buffers of a particle system (capacity >= every n it holds).
struct ParticleBuffers {
    float *x, *y;                  // the list (read by the move, written by the compaction)
    uint32_t *id, *key, *key2, *hist, *digit_total;
    uint32_t *recA, *recB;         // {x, y, id} records carried by the sort, 3 words per particle
    uint4 *rec;                    // per cell: survivors before it in its tile, particles, kill mask, 0
    uint32_t *tile_s;              // per 256-cell tile: survivors, then their scanned offsets
    uint32_t *off_sorted;          // offsets of the sorted list before the annihilation
    int *d_n, *d_n2;  // live count (device) and its next value
    unsigned long long *processed;  // particle-steps done so far (device)
    int n_ub;         // host upper bound of the live count (grid sizes)
    int capacity;     // elements of every per-particle buffer
};

// Buffers of a step on a list already sorted by cell (the current list x, y, id; its cells'
// first indices are tile_sum[c / 256] + rec[c].x, left by the step that made it; the new
// list goes to xo, yo, ido with its offsets in the same form; neither stores keys).
struct SortedStepBuffers {
    const float *x, *y;
    const uint32_t *id;
    float *x2, *y2;           // moved positions, old order
    uint32_t *code;           // move class and rank (or old cell) per particle
    uint32_t *cnt, *pre;      // [9][A]
    uint4 *rec;               // per new cell: survivors before it in its tile, particles, kill mask, stayers
    uint32_t *tile_sum;       // per 256-cell tile: survivors, then their scanned offsets
    uint32_t *far_cnt, *far_cnt_next, *far_in;
    void *far_list;           // far_entry_bytes() per entry, capacity entries
    const int *n_cur;
    int *n_next;
    unsigned long long *processed;
    float *xo, *yo;
    uint32_t *ido;
    int n_ub;
    int capacity;             // elements of every per-particle buffer
};

int radix_tiles(int n);
size_t far_entry_bytes();
hipError_t particle_step_sorted(const ParticleStepArgs &a, int step, const SortedStepBuffers &b, hipStream_t s);
hipError_t radix_sort_payload(uint32_t *keys, uint32_t *ktmp, uint32_t *const *pay, uint32_t *const *ptmp, int np,
                              bool aos, const int *d_n, int n_ub, int key_bits, uint32_t *hist,
                              uint32_t *digit_total, hipStream_t s, bool *result_in_tmp, bool hist0_ready);
hipError_t radix_sort_pairs(uint32_t *keys, uint32_t *vals, uint32_t *ktmp, uint32_t *vtmp, const int *d_n,
                            int n_ub, int key_bits, uint32_t *hist, uint32_t *digit_total, hipStream_t s,
                            bool *result_in_tmp);
hipError_t launch_lower_bound(const uint32_t *keys, const int *d_n, int n_host, int ncells, uint32_t *off,
                              hipStream_t s);
hipError_t launch_left_right(const uint32_t *off, int ncells, uint32_t *left, uint32_t *right, hipStream_t s);
// A step on a list in any order: move, radix sort, A+A->0, compaction; the survivors' count
// goes to d_n2 and their cells' first indices to tile_s (scanned) + rec[c].x.
hipError_t particle_step(const ParticleStepArgs &a, int step, ParticleBuffers &b, hipStream_t s);
// keys[i] = floor(y) * side + floor(x) for the live particles (n_ub bounds the grid)
hipError_t launch_keys_from_pos(const float *x, const float *y, const int *d_n, int n_ub, int side, uint32_t *keys,
                                hipStream_t s);

}  // namespace gpgmp
