// C ABI of the individual-particle path: radix sort, cell offsets and the particle system.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <memory>

#include "host.h"
#include "particles_launch.h"

using namespace gpgmp;

template <typename F>
static int guarded_p(F &&f) {
    try {
        f();
        return GPGMP_OK;
    } catch (const Error &e) {
        gpgmp_set_last_error(e.what());
        return e.code;
    } catch (const std::bad_alloc &) {
        gpgmp_set_last_error("host allocation failed");
        return GPGMP_E_CAPACITY;
    } catch (const std::exception &e) {
        gpgmp_set_last_error(e.what());
        return GPGMP_E_INVALID;
    }
}

// The sorted-list step packs a far particle's old cell into 28 bits of its code word.
constexpr int kSortedMaxSide = 16384;

static int bits_for(uint64_t ncells) {
    int b = 0;
    while (((uint64_t)1 << b) < ncells) ++b;
    return std::max(b, 1);
}

struct gpgmp_particles {
    gpgmp_particle_params p{};
    int capacity = 0;
    hipStream_t stream = nullptr;
    // current list (x, y, id; key_out and the offsets `off` are rebuilt for get) and the buffers
    // of the general step
    DevBuf<float> x, y, x2, y2;
    DevBuf<uint32_t> id, key, key2, key_out, off, hist, digit_total;
    DevBuf<uint32_t> prec;  // the general step's {x, y, id} records: two arrays of 3 words per particle
    // per-cell work of both steps
    DevBuf<uint4> rec;  // per cell: survivors before it in its tile, particles, kill mask, stayers
    DevBuf<uint32_t> tile_cells, off_sorted;
    // step on the sorted list: the next list (xa, ya, ida) and its per-cell work
    DevBuf<float> xa, ya;
    DevBuf<uint32_t> ida, cnt, pre, far_cnt, far_in;
    DevBuf<uint32_t> code;
    DevBuf<uint8_t> far_list;
    DevBuf<int> dn;  // live count in slot cur, the next count in the other
    int cur = 0;
    int far_slot = 0;
    DevBuf<unsigned long long> processed;  // particle-steps done
    int *n_host = nullptr;  // pinned: [0] synchronous reads, [1] the asynchronous one
    hipEvent_t n_ev = nullptr;
    bool n_pending = false;
    int n_ub = 0;
    bool stepped = false;  // keys/offsets exist once a step has sorted the list
    bool sorted_ok = false;  // moves stay within the 3x3 neighbourhood all but never
    bool keys_stale = false;  // the steps store no keys: key_out is rebuilt from the positions on demand
    ~gpgmp_particles() {
        (void)hipStreamSynchronize(stream);
        if (n_ev) (void)hipEventDestroy(n_ev);
        if (n_host) (void)hipHostFree(n_host);
    }
    ParticleBuffers buffers() {
        ParticleBuffers b;
        b.x = x.get(); b.y = y.get();
        b.id = id.get(); b.key = key.get(); b.key2 = key2.get();
        b.recA = prec.get(); b.recB = prec.get() + (size_t)3 * capacity;
        b.hist = hist.get(); b.digit_total = digit_total.get();
        b.rec = rec.get();
        b.tile_s = tile_cells.get(); b.off_sorted = off_sorted.get();
        b.d_n = dn.get() + cur; b.d_n2 = dn.get() + (cur ^ 1);
        b.processed = processed.get();
        b.n_ub = n_ub;
        b.capacity = capacity;
        return b;
    }
    SortedStepBuffers sorted_buffers() {
        SortedStepBuffers b;
        b.x = x.get(); b.y = y.get(); b.id = id.get();
        b.x2 = x2.get(); b.y2 = y2.get(); b.code = code.get();
        b.cnt = cnt.get(); b.pre = pre.get(); b.rec = rec.get();
        b.tile_sum = tile_cells.get();
        b.far_cnt = far_cnt.get() + far_slot; b.far_cnt_next = far_cnt.get() + (far_slot ^ 1);
        b.far_in = far_in.get(); b.far_list = far_list.get();
        b.n_cur = dn.get() + cur; b.n_next = dn.get() + (cur ^ 1);
        b.processed = processed.get();
        b.xo = xa.get(); b.yo = ya.get(); b.ido = ida.get();
        b.n_ub = n_ub;
        b.capacity = capacity;
        return b;
    }
    int read_n() {
        GPGMP_HIP_CHECK(hipMemcpyAsync(n_host, dn.get() + cur, sizeof(int), hipMemcpyDeviceToHost, stream));
        GPGMP_HIP_CHECK(hipStreamSynchronize(stream));
        n_ub = *n_host;
        n_pending = false;
        return *n_host;
    }
    // The live count only shrinks: grids are sized by an upper bound that a copy in the
    // background lowers now and then, so no step waits for the device.
    void poll_n() {
        if (n_pending && hipEventQuery(n_ev) == hipSuccess) {
            n_ub = std::min(n_ub, n_host[1]);
            n_pending = false;
        }
        if (!n_pending) {
            GPGMP_HIP_CHECK(hipMemcpyAsync(n_host + 1, dn.get() + cur, sizeof(int), hipMemcpyDeviceToHost, stream));
            GPGMP_HIP_CHECK(hipEventRecord(n_ev, stream));
            n_pending = true;
        }
    }
};

extern "C" {

int gpgmp_sort_pairs_device(uint32_t *keys, uint32_t *vals, uint32_t *keys_tmp, uint32_t *vals_tmp, int32_t n,
                            int32_t key_bits, void *stream) {
    return guarded_p([&] {
        GPGMP_REQUIRE(n >= 0, GPGMP_E_INVALID, "n must be non-negative");
        if (n == 0) return;
        GPGMP_REQUIRE(keys && vals && keys_tmp && vals_tmp, GPGMP_E_INVALID, "null buffer");
        GPGMP_REQUIRE(key_bits >= 1 && key_bits <= 32, GPGMP_E_INVALID, "key_bits must be in [1, 32]");
        hipStream_t s = (hipStream_t)stream;
        const int nt = std::max(radix_tiles(n), 1);
        DevBuf<uint32_t> hist((size_t)256 * nt), tot(256);
        bool in_tmp = false;
        GPGMP_HIP_CHECK(radix_sort_pairs(keys, vals, keys_tmp, vals_tmp, nullptr, n, key_bits, hist.get(), tot.get(),
                                         s, &in_tmp));
        if (in_tmp) {
            GPGMP_HIP_CHECK(hipMemcpyAsync(keys, keys_tmp, (size_t)n * 4, hipMemcpyDeviceToDevice, s));
            GPGMP_HIP_CHECK(hipMemcpyAsync(vals, vals_tmp, (size_t)n * 4, hipMemcpyDeviceToDevice, s));
        }
        GPGMP_HIP_CHECK(hipStreamSynchronize(s));  // scratch is freed on return
    });
}

int gpgmp_cell_offsets_device(const uint32_t *sorted_keys, int32_t n, int32_t ncells, uint32_t *off, void *stream) {
    return guarded_p([&] {
        GPGMP_REQUIRE(off && ncells >= 1 && n >= 0 && (n == 0 || sorted_keys), GPGMP_E_INVALID, "bad argument");
        GPGMP_HIP_CHECK(launch_lower_bound(sorted_keys, nullptr, n, ncells, off, (hipStream_t)stream));
    });
}

int gpgmp_cell_left_right_device(const uint32_t *sorted_keys, int32_t n, int32_t ncells, uint32_t *left,
                                 uint32_t *right, void *stream) {
    return guarded_p([&] {
        GPGMP_REQUIRE(left && right && ncells >= 1 && n >= 0 && (n == 0 || sorted_keys), GPGMP_E_INVALID,
                      "bad argument");
        hipStream_t s = (hipStream_t)stream;
        DevBuf<uint32_t> off((size_t)ncells + 1);
        GPGMP_HIP_CHECK(launch_lower_bound(sorted_keys, nullptr, n, ncells, off.get(), s));
        GPGMP_HIP_CHECK(launch_left_right(off.get(), ncells, left, right, s));
        GPGMP_HIP_CHECK(hipStreamSynchronize(s));
    });
}

int gpgmp_particles_create(const gpgmp_particle_params *p, int32_t capacity, void *stream, gpgmp_particles **out) {
    return guarded_p([&] {
        GPGMP_REQUIRE(p && out, GPGMP_E_INVALID, "null argument");
        *out = nullptr;
        GPGMP_REQUIRE(p->side >= 1 && p->side <= 46340, GPGMP_E_INVALID, "side out of range");
        GPGMP_REQUIRE(capacity >= 1 && capacity <= (1 << 28), GPGMP_E_CAPACITY, "capacity out of range");
        GPGMP_REQUIRE(p->D >= 0.f && p->dt >= 0.f && p->p_annihilate >= 0.f && p->p_annihilate <= 1.f,
                      GPGMP_E_INVALID, "bad parameters");
        std::unique_ptr<gpgmp_particles> ps(new gpgmp_particles());
        ps->p = *p;
        ps->capacity = capacity;
        ps->stream = (hipStream_t)stream;  // NULL: the default stream
        const size_t n = capacity;
        for (DevBuf<float> *b : {&ps->x, &ps->y, &ps->x2, &ps->y2}) b->alloc(n);
        for (DevBuf<uint32_t> *b : {&ps->id, &ps->key, &ps->key2, &ps->key_out}) b->alloc(n);
        ps->prec.alloc(6 * n);
        const size_t ncells = (size_t)p->side * p->side;
        ps->off.alloc(ncells + 1);
        ps->off_sorted.alloc(ncells + 1);
        ps->rec.alloc(ncells);
        ps->tile_cells.alloc((ncells + 255) / 256);
        const int nt = std::max(radix_tiles(capacity), 1);
        ps->hist.alloc((size_t)256 * nt);
        ps->digit_total.alloc(256);
        // the step on the sorted list: moves of at most one cell per axis unless a Gaussian
        // displacement of 6 sigma plus the largest drift moves a whole cell (then the far
        // list, exact but slow when long, would fill; those parameters take the radix step)
        const double sigma = std::sqrt(2.0 * (double)p->D * (double)p->dt);
        ps->sorted_ok = p->side >= 3 && p->side <= kSortedMaxSide &&
                        6.0 * sigma + std::fabs((double)p->vortex_amp) * (double)p->dt < 1.0;
        if (ps->sorted_ok) {
            ps->xa.alloc(n);
            ps->ya.alloc(n);
            ps->ida.alloc(n);
            ps->code.alloc(n);
            ps->far_list.alloc(n * far_entry_bytes());
            ps->cnt.alloc(9 * ncells);
            ps->pre.alloc(9 * ncells);
            ps->far_cnt.alloc(2);
            ps->far_in.alloc(ncells);
            GPGMP_HIP_CHECK(hipMemsetAsync(ps->far_in.get(), 0, ncells * sizeof(uint32_t), ps->stream));
            GPGMP_HIP_CHECK(hipMemsetAsync(ps->far_cnt.get(), 0, 2 * sizeof(uint32_t), ps->stream));
        }
        ps->dn.alloc(2);
        ps->processed.alloc(1);
        GPGMP_HIP_CHECK(hipMemsetAsync(ps->processed.get(), 0, sizeof(unsigned long long), ps->stream));
        GPGMP_HIP_CHECK(hipHostMalloc((void **)&ps->n_host, 2 * sizeof(int)));
        ps->n_host[0] = ps->n_host[1] = 0;
        GPGMP_HIP_CHECK(hipEventCreateWithFlags(&ps->n_ev, hipEventDisableTiming));
        GPGMP_HIP_CHECK(hipMemsetAsync(ps->dn.get(), 0, 2 * sizeof(int), ps->stream));
        GPGMP_HIP_CHECK(hipStreamSynchronize(ps->stream));
        *out = ps.release();
    });
}

int gpgmp_particles_destroy(gpgmp_particles *ps) {
    delete ps;
    return GPGMP_OK;
}

int gpgmp_particles_set(gpgmp_particles *ps, const float *x, const float *y, const uint32_t *id, int32_t n) {
    return guarded_p([&] {
        GPGMP_REQUIRE(ps && n >= 0 && (n == 0 || (x && y && id)), GPGMP_E_INVALID, "bad argument");
        GPGMP_REQUIRE(n <= ps->capacity, GPGMP_E_CAPACITY, "more particles than the capacity");
        const float L = (float)ps->p.side;
        for (int i = 0; i < n; ++i)
            GPGMP_REQUIRE(x[i] >= 0.f && x[i] < L && y[i] >= 0.f && y[i] < L, GPGMP_E_INVALID,
                          "positions must lie in [0, side)");
        hipStream_t s = ps->stream;
        if (n) {
            GPGMP_HIP_CHECK(hipMemcpyAsync(ps->x.get(), x, (size_t)n * 4, hipMemcpyHostToDevice, s));
            GPGMP_HIP_CHECK(hipMemcpyAsync(ps->y.get(), y, (size_t)n * 4, hipMemcpyHostToDevice, s));
            GPGMP_HIP_CHECK(hipMemcpyAsync(ps->id.get(), id, (size_t)n * 4, hipMemcpyHostToDevice, s));
        }
        GPGMP_HIP_CHECK(hipStreamSynchronize(s));  // no copy of the count still in flight
        ps->n_pending = false;
        ps->cur = 0;
        *ps->n_host = n;
        GPGMP_HIP_CHECK(hipMemcpyAsync(ps->dn.get(), ps->n_host, sizeof(int), hipMemcpyHostToDevice, s));
        GPGMP_HIP_CHECK(hipStreamSynchronize(s));
        ps->n_ub = n;
        ps->stepped = false;
    });
}

int gpgmp_particles_count(gpgmp_particles *ps, int32_t *n_out) {
    return guarded_p([&] {
        GPGMP_REQUIRE(ps && n_out, GPGMP_E_INVALID, "null argument");
        *n_out = ps->read_n();
    });
}

int gpgmp_particles_get(gpgmp_particles *ps, float *x, float *y, uint32_t *id, uint32_t *keys, uint32_t *off,
                        int32_t *n_out) {
    return guarded_p([&] {
        GPGMP_REQUIRE(ps && n_out, GPGMP_E_INVALID, "null argument");
        GPGMP_REQUIRE(ps->stepped || (!keys && !off), GPGMP_E_STATE, "keys and offsets exist after a step");
        const int n = ps->read_n();
        *n_out = n;
        hipStream_t s = ps->stream;
        const size_t ncells = (size_t)ps->p.side * ps->p.side;
        if (n && ps->keys_stale && (keys || off)) {
            GPGMP_HIP_CHECK(launch_keys_from_pos(ps->x.get(), ps->y.get(), ps->dn.get() + ps->cur, n, ps->p.side,
                                                 ps->key_out.get(), s));
            ps->keys_stale = false;
        }
        if (n) {
            if (x) GPGMP_HIP_CHECK(hipMemcpyAsync(x, ps->x.get(), (size_t)n * 4, hipMemcpyDeviceToHost, s));
            if (y) GPGMP_HIP_CHECK(hipMemcpyAsync(y, ps->y.get(), (size_t)n * 4, hipMemcpyDeviceToHost, s));
            if (id) GPGMP_HIP_CHECK(hipMemcpyAsync(id, ps->id.get(), (size_t)n * 4, hipMemcpyDeviceToHost, s));
            if (keys) GPGMP_HIP_CHECK(hipMemcpyAsync(keys, ps->key_out.get(), (size_t)n * 4, hipMemcpyDeviceToHost, s));
        }
        if (off) {
            // offsets of the stored (compacted, sorted) keys
            GPGMP_HIP_CHECK(launch_lower_bound(ps->key_out.get(), ps->dn.get() + ps->cur, n, (int)ncells, ps->off.get(), s));
            GPGMP_HIP_CHECK(hipMemcpyAsync(off, ps->off.get(), (ncells + 1) * 4, hipMemcpyDeviceToHost, s));
        }
        GPGMP_HIP_CHECK(hipStreamSynchronize(s));
    });
}

int gpgmp_particles_set_sorted_step(gpgmp_particles *ps, int32_t mode) {
    return guarded_p([&] {
        GPGMP_REQUIRE(ps && mode >= 0 && mode <= 2, GPGMP_E_INVALID, "mode must be 0, 1 or 2");
        GPGMP_REQUIRE(mode != 2 || (ps->p.side >= 3 && ps->p.side <= kSortedMaxSide), GPGMP_E_INVALID,
                      "the sorted-list step needs 3 <= side <= 16384");
        GPGMP_HIP_CHECK(hipStreamSynchronize(ps->stream));
        if (mode == 2 && !ps->xa.get()) {
            const size_t n = ps->capacity, ncells = (size_t)ps->p.side * ps->p.side;
            ps->xa.alloc(n);
            ps->ya.alloc(n);
            ps->ida.alloc(n);
            ps->code.alloc(n);
            ps->far_list.alloc(n * far_entry_bytes());
            ps->cnt.alloc(9 * ncells);
            ps->pre.alloc(9 * ncells);
            ps->far_cnt.alloc(2);
            ps->far_in.alloc(ncells);
            GPGMP_HIP_CHECK(hipMemsetAsync(ps->far_in.get(), 0, ncells * sizeof(uint32_t), ps->stream));
            GPGMP_HIP_CHECK(hipMemsetAsync(ps->far_cnt.get(), 0, 2 * sizeof(uint32_t), ps->stream));
            GPGMP_HIP_CHECK(hipStreamSynchronize(ps->stream));
        }
        if (mode == 0) ps->sorted_ok = false;
        if (mode == 2) ps->sorted_ok = true;
        if (mode == 1) {
            const double sigma = std::sqrt(2.0 * (double)ps->p.D * (double)ps->p.dt);
            ps->sorted_ok = ps->xa.get() && ps->p.side >= 3 && ps->p.side <= kSortedMaxSide &&
                            6.0 * sigma + std::fabs((double)ps->p.vortex_amp) * (double)ps->p.dt < 1.0;
        }
    });
}

int gpgmp_particles_processed(gpgmp_particles *ps, uint64_t *particle_steps) {
    return guarded_p([&] {
        GPGMP_REQUIRE(ps && particle_steps, GPGMP_E_INVALID, "null argument");
        unsigned long long v = 0;
        GPGMP_HIP_CHECK(hipMemcpyAsync(&v, ps->processed.get(), sizeof v, hipMemcpyDeviceToHost, ps->stream));
        GPGMP_HIP_CHECK(hipStreamSynchronize(ps->stream));
        *particle_steps = v;
    });
}

int gpgmp_particles_step(gpgmp_particles *ps, int32_t step0, int32_t steps) {
    return guarded_p([&] {
        GPGMP_REQUIRE(ps && steps >= 0, GPGMP_E_INVALID, "bad argument");
        ParticleStepArgs a;
        a.side = ps->p.side;
        a.D = ps->p.D;
        a.dt = ps->p.dt;
        a.vortex_amp = ps->p.vortex_amp;
        a.p_annihilate = ps->p.p_annihilate;
        a.seed = ps->p.seed;
        a.key_bits = bits_for((uint64_t)ps->p.side * ps->p.side);
        for (int k = 0; k < steps; ++k) {
            if (ps->n_ub == 0) break;
            if (ps->stepped && ps->sorted_ok) {
                // the list is sorted by cell with its offsets: counts and one scatter
                GPGMP_HIP_CHECK(particle_step_sorted(a, step0 + k, ps->sorted_buffers(), ps->stream));
                std::swap(ps->x, ps->xa);
                std::swap(ps->y, ps->ya);
                std::swap(ps->id, ps->ida);
                ps->cur ^= 1;
                ps->far_slot ^= 1;
                ps->keys_stale = true;
            } else {
                // a list in any order: radix sort, offsets, annihilation, compaction
                ParticleBuffers b = ps->buffers();
                GPGMP_HIP_CHECK(particle_step(a, step0 + k, b, ps->stream));
                ps->cur ^= 1;
                ps->keys_stale = true;
            }
            ps->stepped = true;
            ps->poll_n();
        }
    });
}

}  // extern "C"
