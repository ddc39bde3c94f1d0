// C ABI of the individual-species lattice diffusion (the reference's IndividualSolver as
// driven by InhomogeneousSolver::diffuse for a particle-based species).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <memory>
#include <vector>

#include "gpgmp_numeric.h"
#include "host.h"
#include "individuals_launch.h"
#include "particles_launch.h"

using namespace gpgmp;

namespace {

template <typename F>
int guarded_i(F &&f) {
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

int key_bits_for(uint64_t cells) {
    int b = 1;
    while (((uint64_t)1 << b) < cells) ++b;
    return b;
}

}  // namespace

struct gpgmp_individuals {
    gpgmp_individuals_desc d{};
    int A = 0, cap = 0, n = 0, key_bits = 32, nb = 1;
    float h = 0.f;
    hipStream_t s = nullptr;
    // keys: cell per particle; pos: positions 0..n-1 sorted along with the keys (the
    // permutation); ids / props: particle index and properties in the keys' order
    DevBuf<uint32_t> keys, pos, keys2, pos2, ids, ids2, off, hist, tot;
    DevBuf<float> props, props2, Dx, Dy, rx, ry, blk, dt_dev;
    DevBuf<int32_t> count;
    // dead-particle pool (a stack: top = the front of the reference's dead list) and the
    // scratch of a population maintenance
    DevBuf<uint32_t> pool_ids, off2, nnew, nkill, mp_tiles, mp_err, mp_out;
    DevBuf<float> pool_props;
    DevBuf<uint8_t> dead;
    DevBuf<int32_t> goal_dev;
    int np = 0;  // pool size; n + np == cap
    int32_t *count_bound = nullptr;  // a device slice that mirrors the per-cell counts (solver state)
    void publish_counts() {
        if (count_bound)
            GPGMP_HIP_CHECK(hipMemcpyAsync(count_bound, count.get(), (size_t)A * 4, hipMemcpyDeviceToDevice, s));
    }
    uint32_t *mp_host = nullptr;  // pinned [8]
    float *dt_host = nullptr;  // pinned [2]
    float dtx = 0.f, dty = 0.f;
    int64_t sweeps = 0;  // hop sweeps done since the particles were set
    bool ready = false;  // particles set and sorted
    ~gpgmp_individuals() {
        (void)hipStreamSynchronize(s);
        if (dt_host) (void)hipHostFree(dt_host);
        if (mp_host) (void)hipHostFree(mp_host);
    }

    // IndividualSolver::maintainPopulation (IndividualSolver.cpp:458-671): make cell c hold
    // exactly goal[c] particles.  Kills follow the reference's pick loop (RNG key
    // {cell, 0xdecafbad + seed}); new particles recycle dead ids from the pool front in cell
    // order and take the newIndividuals rule (RANDOM keyed {cell, 0xdecafbad + seed + 1});
    // the list stays sorted, so only the offsets and constants are rebuilt (no sort).
    // One host synchronisation reads the totals, as the reference reads d_total/d_totalAdded.
    void maintain(const int32_t *goal, bool on_device, int32_t seed, const gpgmp_new_individual_rule *rule) {
        const int32_t *gd = goal;
        if (!on_device) {
            GPGMP_HIP_CHECK(hipMemcpyAsync(goal_dev.get(), goal, (size_t)A * 4, hipMemcpyHostToDevice, s));
            gd = goal_dev.get();
        }
        GPGMP_HIP_CHECK(launch_ind_mp_count(off.get(), gd, A, dead.get(), nnew.get(), nkill.get(), mp_tiles.get(),
                                            mp_err.get(), mp_out.get(), kKeyBase + (uint32_t)seed, s));
        GPGMP_HIP_CHECK(hipMemcpyAsync(mp_host, mp_out.get(), 5 * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
        GPGMP_HIP_CHECK(hipStreamSynchronize(s));
        GPGMP_REQUIRE(mp_host[4] == 0, GPGMP_E_INVALID, "negative goal count");
        const uint32_t total_goal = mp_host[0], total_new = mp_host[1], total_kill = mp_host[2],
                       first_out = mp_host[3];
        GPGMP_REQUIRE(total_new <= (uint32_t)np, GPGMP_E_CAPACITY,
                      "required number of individuals is greater than the capacity");
        const uint32_t n_new = (uint32_t)n - total_kill + total_new;
        GPGMP_REQUIRE(n_new == total_goal + ((uint32_t)n - first_out), GPGMP_E_STATE,
                      "population counts disagree");
        MaintainArgs m{};
        m.A = A;
        m.cap = cap;
        m.n = (uint32_t)n;
        m.np = (uint32_t)np;
        m.total_goal = total_goal;
        m.total_new = total_new;
        m.total_kill = total_kill;
        m.first_outside = first_out;
        m.off = off.get();
        m.nnew = nnew.get();
        m.nkill = nkill.get();
        m.tiles = mp_tiles.get();
        m.keys = keys.get();
        m.ids = ids.get();
        m.goal = gd;
        m.dead = dead.get();
        m.props = props.get();
        m.keys2 = keys2.get();
        m.ids2 = ids2.get();
        m.pos = pos.get();
        m.off2 = off2.get();
        m.pool_ids = pool_ids.get();
        m.props2 = props2.get();
        m.pool_props = pool_props.get();
        m.use_rule = rule != nullptr;
        for (int k = 0; k < 4; ++k) {
            m.mode[k] = rule ? rule->mode[k] : 0;
            m.a[k] = rule ? rule->a[k] : 0.0;
            m.b[k] = rule ? rule->b[k] : 0.0;
        }
        m.k1_new = kKeyBase + (uint32_t)seed + 1u;
        GPGMP_HIP_CHECK(launch_ind_mp_apply(m, s));
        std::swap(keys, keys2);
        std::swap(ids, ids2);
        std::swap(props, props2);
        std::swap(off, off2);
        n = (int)n_new;
        np = np - (int)total_new + (int)total_kill;
        GPGMP_HIP_CHECK(launch_ind_constants(off.get(), props.get(), cap, A, Dx.get(), Dy.get(), rx.get(), ry.get(),
                                             count.get(), s));
        publish_counts();
    }

    // IndividualSolver::computeDiffusionConstants (IndividualSolver.cpp:1100-1153):
    // sortSpecies, computeCellIndices, the constants kernel; then the per-cell time step
    // minimum of the axes in `axes` (bit 0: X, bit 1: Y), read back to the host as the
    // reference does (InhomogeneousSolver.cpp:330-336).
    void update(int axes) {
        if (n > 0) {
            bool in_tmp = false;
            GPGMP_HIP_CHECK(radix_sort_pairs(keys.get(), pos.get(), keys2.get(), pos2.get(), nullptr, n, key_bits,
                                             hist.get(), tot.get(), s, &in_tmp));
            if (in_tmp) {
                std::swap(keys, keys2);
                std::swap(pos, pos2);
            }
            GPGMP_HIP_CHECK(launch_ind_permute(pos.get(), ids.get(), ids2.get(), props.get(), props2.get(), n, cap, s));
            std::swap(ids, ids2);
            std::swap(props, props2);
        }
        GPGMP_HIP_CHECK(launch_lower_bound(keys.get(), nullptr, n, A, off.get(), s));
        GPGMP_HIP_CHECK(launch_ind_constants(off.get(), props.get(), cap, A, Dx.get(), Dy.get(), rx.get(), ry.get(),
                                             count.get(), s));
        publish_counts();
        if (axes & 1) GPGMP_HIP_CHECK(launch_ind_dt(Dx.get(), rx.get(), A, h, blk.get(), nb, dt_dev.get(), s));
        if (axes & 2) GPGMP_HIP_CHECK(launch_ind_dt(Dy.get(), ry.get(), A, h, blk.get(), nb, dt_dev.get() + 1, s));
        if (axes) {
            GPGMP_HIP_CHECK(hipMemcpyAsync(dt_host, dt_dev.get(), 2 * sizeof(float), hipMemcpyDeviceToHost, s));
            GPGMP_HIP_CHECK(hipStreamSynchronize(s));
            if (axes & 1) dtx = dt_host[0];
            if (axes & 2) dty = dt_host[1];
        }
        ready = true;
    }

    // individual_diffuseX/Y (IndividualSolver.cpp:416-456) with `seed`
    void sweep(int axis, float dt, int32_t seed) {
        ++sweeps;
        const int4 bm = make_int4(d.boundary_mask[0], d.boundary_mask[1], d.boundary_mask[2], d.boundary_mask[3]);
        GPGMP_HIP_CHECK(launch_ind_sweep(axis, keys.get(), off.get(), props.get(), cap, d.width, d.height,
                                         (uint32_t)d.species * (uint32_t)A, kKeyBase + (uint32_t)seed, dt, h, bm, s));
    }
};

extern "C" {

int gpgmp_individuals_create(const gpgmp_individuals_desc *desc, int32_t capacity, void *stream,
                             gpgmp_individuals **out) {
    return guarded_i([&] {
        GPGMP_REQUIRE(desc && out, GPGMP_E_INVALID, "null argument");
        *out = nullptr;
        GPGMP_REQUIRE(desc->width >= 1 && desc->height >= 1 && desc->physical_length > 0.f, GPGMP_E_INVALID,
                      "bad grid");
        GPGMP_REQUIRE((int64_t)desc->width * desc->height < ((int64_t)1 << 31), GPGMP_E_CAPACITY, "grid too large");
        GPGMP_REQUIRE(capacity >= 1, GPGMP_E_INVALID, "capacity must be positive");
        GPGMP_REQUIRE(desc->species >= 0, GPGMP_E_INVALID, "species index must be non-negative");
        for (int i = 0; i < 4; ++i)
            GPGMP_REQUIRE(desc->boundary_mask[i] == 0 || desc->boundary_mask[i] == 1, GPGMP_E_INVALID,
                          "boundary masks are 0 (reflective) or 1 (absorbing)");
        std::unique_ptr<gpgmp_individuals> g(new gpgmp_individuals());
        g->d = *desc;
        g->A = desc->width * desc->height;
        g->cap = capacity;
        g->s = (hipStream_t)stream;
        g->h = header_float(desc->physical_length / (float)desc->width);
        // absorbing boundaries move keys out of [0, A) (up to 2^32 - W): sort all 32 bits
        const bool absorbing = desc->boundary_mask[0] | desc->boundary_mask[1] | desc->boundary_mask[2] |
                               desc->boundary_mask[3];
        g->key_bits = absorbing ? 32 : key_bits_for((uint64_t)g->A);
        g->nb = std::min(1024, (g->A + 255) / 256);
        const size_t cp = (size_t)capacity;
        g->keys.alloc(cp);
        g->pos.alloc(cp);
        g->keys2.alloc(cp);
        g->pos2.alloc(cp);
        g->ids.alloc(cp);
        g->ids2.alloc(cp);
        g->hist.alloc((size_t)256 * std::max(radix_tiles(capacity), 1));
        g->tot.alloc(256);
        g->off.alloc((size_t)g->A + 1);
        g->props.alloc(4 * cp);
        g->props2.alloc(4 * cp);
        g->Dx.alloc(g->A);
        g->Dy.alloc(g->A);
        g->rx.alloc(g->A);
        g->ry.alloc(g->A);
        g->count.alloc(g->A);
        g->blk.alloc(g->nb);
        g->dt_dev.alloc(2);
        g->pool_ids.alloc(cp);
        g->pool_props.alloc(4 * cp);
        g->dead.alloc(cp);
        g->off2.alloc((size_t)g->A + 1);
        g->nnew.alloc(g->A);
        g->nkill.alloc(g->A);
        g->goal_dev.alloc(g->A);
        g->mp_tiles.alloc((size_t)3 * mp_tiles(g->A));
        g->mp_err.alloc(1);
        g->mp_out.alloc(8);
        GPGMP_HIP_CHECK(hipMemsetAsync(g->mp_err.get(), 0, sizeof(uint32_t), g->s));
        GPGMP_HIP_CHECK(hipHostMalloc((void **)&g->mp_host, 8 * sizeof(uint32_t)));
        GPGMP_HIP_CHECK(hipHostMalloc((void **)&g->dt_host, 2 * sizeof(float)));
        *out = g.release();
    });
}

int gpgmp_individuals_destroy(gpgmp_individuals *g) {
    delete g;
    return GPGMP_OK;
}

int gpgmp_individuals_set(gpgmp_individuals *g, const uint32_t *cells, const float *props, int32_t n) {
    return guarded_i([&] {
        GPGMP_REQUIRE(g && n >= 0 && (n == 0 || (cells && props)), GPGMP_E_INVALID, "bad argument");
        GPGMP_REQUIRE(n <= g->cap, GPGMP_E_CAPACITY, "more particles than the capacity");
        for (int i = 0; i < n; ++i)
            GPGMP_REQUIRE(cells[i] < (uint32_t)g->A, GPGMP_E_INVALID, "particle cell outside the grid");
        std::vector<uint32_t> ids(n);
        for (int i = 0; i < n; ++i) ids[i] = (uint32_t)i;
        g->n = n;
        g->np = g->cap - n;
        g->sweeps = 0;
        // the pool of unused ids n .. cap-1 with the reference's defaults (diffusivities 1,
        // drifts 0, IndividualSolver.cpp:84-98), id n at the top
        if (g->np > 0) {
            std::vector<uint32_t> pid((size_t)g->np);
            for (int j = 0; j < g->np; ++j) pid[j] = (uint32_t)(g->cap - 1 - j);
            std::vector<float> pp((size_t)g->np);
            GPGMP_HIP_CHECK(hipMemcpyAsync(g->pool_ids.get(), pid.data(), (size_t)g->np * 4, hipMemcpyHostToDevice, g->s));
            for (int k = 0; k < 4; ++k) {
                std::fill(pp.begin(), pp.end(), k < 2 ? 1.0f : 0.0f);
                GPGMP_HIP_CHECK(hipMemcpyAsync(g->pool_props.get() + (size_t)k * g->cap, pp.data(), (size_t)g->np * 4,
                                               hipMemcpyHostToDevice, g->s));
                GPGMP_HIP_CHECK(hipStreamSynchronize(g->s));  // pp is reused
            }
        }
        if (n > 0) {
            GPGMP_HIP_CHECK(hipMemcpyAsync(g->keys.get(), cells, (size_t)n * 4, hipMemcpyHostToDevice, g->s));
            GPGMP_HIP_CHECK(hipMemcpyAsync(g->ids.get(), ids.data(), (size_t)n * 4, hipMemcpyHostToDevice, g->s));
            GPGMP_HIP_CHECK(hipMemcpyAsync(g->pos.get(), ids.data(), (size_t)n * 4, hipMemcpyHostToDevice, g->s));
            for (int k = 0; k < 4; ++k)
                GPGMP_HIP_CHECK(hipMemcpyAsync(g->props.get() + (size_t)k * g->cap, props + (size_t)k * n,
                                               (size_t)n * 4, hipMemcpyHostToDevice, g->s));
        }
        // the solver's first time steps (InhomogeneousSolver.cpp:205-225)
        g->update(3);
    });
}

int gpgmp_individuals_time_steps(gpgmp_individuals *g, float *dt_x, float *dt_y) {
    return guarded_i([&] {
        GPGMP_REQUIRE(g && g->ready, GPGMP_E_STATE, "no particles set");
        if (dt_x) *dt_x = g->dtx;
        if (dt_y) *dt_y = g->dty;
    });
}

int gpgmp_individuals_diffuse(gpgmp_individuals *g, int32_t axis, float dt, int32_t *seed) {
    return guarded_i([&] {
        GPGMP_REQUIRE(g && seed && (axis == 0 || axis == 1), GPGMP_E_INVALID, "bad argument");
        GPGMP_REQUIRE(g->ready, GPGMP_E_STATE, "no particles set");
        g->sweep(axis, dt, *seed);
        *seed = (int32_t)((uint32_t)*seed + 1u);
        g->ready = false;  // keys are unsorted until the next update
    });
}

int gpgmp_individuals_update(gpgmp_individuals *g, int32_t axes, float *dt_x, float *dt_y) {
    return guarded_i([&] {
        GPGMP_REQUIRE(g && axes >= 0 && axes <= 3, GPGMP_E_INVALID, "bad argument");
        g->update(axes);
        if (dt_x) *dt_x = g->dtx;
        if (dt_y) *dt_y = g->dty;
    });
}

int gpgmp_individuals_run(gpgmp_individuals *g, int32_t steps, uint32_t seed0, float *sim_time_out) {
    return guarded_i([&] {
        GPGMP_REQUIRE(g && steps >= 0, GPGMP_E_INVALID, "bad argument");
        GPGMP_REQUIRE(g->ready, GPGMP_E_STATE, "no particles set");
        SeedStream seeds(seed0);
        float nx = g->dtx, ny = g->dty, sim = 0.f;
        for (int it = 0; it < steps; ++it) {
            int32_t seed = seeds.next();
            const float mindt = (nx < ny) ? nx : ny;
            if (!(mindt > 0.f) || std::isinf(mindt)) break;  // nothing can move
            // InhomogeneousSolver::diffuse (InhomogeneousSolver.cpp:254-427): each due
            // axis sweeps with its current step, then the constants and that axis' step
            // are recomputed
            if (nx <= mindt) {
                g->sweep(0, g->dtx, seed);
                seed = (int32_t)((uint32_t)seed + 1u);
                g->update(1);
                nx += g->dtx;
            }
            if (ny <= mindt) {
                g->sweep(1, g->dty, seed);
                seed = (int32_t)((uint32_t)seed + 1u);
                g->update(2);
                ny += g->dty;
            }
            sim = mindt;
        }
        if (sim_time_out) *sim_time_out = sim;
    });
}

int gpgmp_individuals_get(gpgmp_individuals *g, uint32_t *keys, uint32_t *vals, int32_t *count, int32_t *n_out) {
    return guarded_i([&] {
        GPGMP_REQUIRE(g, GPGMP_E_INVALID, "null argument");
        GPGMP_REQUIRE(g->ready, GPGMP_E_STATE, "particles are unsorted: call update first");
        if (keys && g->n) GPGMP_HIP_CHECK(hipMemcpyAsync(keys, g->keys.get(), (size_t)g->n * 4, hipMemcpyDeviceToHost, g->s));
        if (vals && g->n) GPGMP_HIP_CHECK(hipMemcpyAsync(vals, g->ids.get(), (size_t)g->n * 4, hipMemcpyDeviceToHost, g->s));
        if (count) GPGMP_HIP_CHECK(hipMemcpyAsync(count, g->count.get(), (size_t)g->A * 4, hipMemcpyDeviceToHost, g->s));
        GPGMP_HIP_CHECK(hipStreamSynchronize(g->s));
        if (n_out) *n_out = g->n;
    });
}

int gpgmp_individuals_constants(gpgmp_individuals *g, float *Dx, float *Dy, float *rx, float *ry) {
    return guarded_i([&] {
        GPGMP_REQUIRE(g, GPGMP_E_INVALID, "null argument");
        GPGMP_REQUIRE(g->ready, GPGMP_E_STATE, "particles are unsorted: call update first");
        const size_t b = (size_t)g->A * 4;
        if (Dx) GPGMP_HIP_CHECK(hipMemcpyAsync(Dx, g->Dx.get(), b, hipMemcpyDeviceToHost, g->s));
        if (Dy) GPGMP_HIP_CHECK(hipMemcpyAsync(Dy, g->Dy.get(), b, hipMemcpyDeviceToHost, g->s));
        if (rx) GPGMP_HIP_CHECK(hipMemcpyAsync(rx, g->rx.get(), b, hipMemcpyDeviceToHost, g->s));
        if (ry) GPGMP_HIP_CHECK(hipMemcpyAsync(ry, g->ry.get(), b, hipMemcpyDeviceToHost, g->s));
        GPGMP_HIP_CHECK(hipStreamSynchronize(g->s));
    });
}

int gpgmp_individuals_sweeps(gpgmp_individuals *g, int64_t *sweeps) {
    return guarded_i([&] {
        GPGMP_REQUIRE(g && sweeps, GPGMP_E_INVALID, "null argument");
        *sweeps = g->sweeps;
    });
}

int gpgmp_individuals_maintain(gpgmp_individuals *g, const int32_t *goal, int32_t goal_on_device, int32_t seed,
                               const gpgmp_new_individual_rule *rule) {
    return guarded_i([&] {
        GPGMP_REQUIRE(g && goal, GPGMP_E_INVALID, "null argument");
        GPGMP_REQUIRE(g->ready, GPGMP_E_STATE, "particles are unsorted: call update first");
        if (rule)
            for (int k = 0; k < 4; ++k)
                GPGMP_REQUIRE(rule->mode[k] >= 0 && rule->mode[k] <= 2, GPGMP_E_INVALID, "rule modes are 0, 1 or 2");
        g->maintain(goal, goal_on_device != 0, seed, rule);
    });
}

int gpgmp_individuals_bind_counts(gpgmp_individuals *g, int32_t *dev_counts) {
    return guarded_i([&] {
        GPGMP_REQUIRE(g, GPGMP_E_INVALID, "null argument");
        g->count_bound = dev_counts;
        if (g->ready) g->publish_counts();
        GPGMP_HIP_CHECK(hipStreamSynchronize(g->s));
    });
}

int gpgmp_individuals_synchronize(gpgmp_individuals *g) {
    return guarded_i([&] {
        GPGMP_REQUIRE(g, GPGMP_E_INVALID, "null argument");
        GPGMP_HIP_CHECK(hipStreamSynchronize(g->s));
    });
}

int gpgmp_individuals_get_props(gpgmp_individuals *g, float *props) {
    return guarded_i([&] {
        GPGMP_REQUIRE(g && props, GPGMP_E_INVALID, "null argument");
        GPGMP_REQUIRE(g->ready, GPGMP_E_STATE, "particles are unsorted: call update first");
        for (int k = 0; k < 4 && g->n; ++k)
            GPGMP_HIP_CHECK(hipMemcpyAsync(props + (size_t)k * g->n, g->props.get() + (size_t)k * g->cap,
                                           (size_t)g->n * 4, hipMemcpyDeviceToHost, g->s));
        GPGMP_HIP_CHECK(hipStreamSynchronize(g->s));
    });
}

int gpgmp_individuals_get_pool(gpgmp_individuals *g, uint32_t *ids, float *props, int32_t *np_out) {
    return guarded_i([&] {
        GPGMP_REQUIRE(g, GPGMP_E_INVALID, "null argument");
        const size_t np = (size_t)g->np;
        if (ids && np) {
            GPGMP_HIP_CHECK(hipMemcpyAsync(ids, g->pool_ids.get(), np * 4, hipMemcpyDeviceToHost, g->s));
            GPGMP_HIP_CHECK(hipStreamSynchronize(g->s));
            std::reverse(ids, ids + np);  // front (stack top) first
        }
        if (props && np) {
            for (int k = 0; k < 4; ++k)
                GPGMP_HIP_CHECK(hipMemcpyAsync(props + (size_t)k * np, g->pool_props.get() + (size_t)k * g->cap,
                                               np * 4, hipMemcpyDeviceToHost, g->s));
            GPGMP_HIP_CHECK(hipStreamSynchronize(g->s));
            for (int k = 0; k < 4; ++k) std::reverse(props + (size_t)k * np, props + (size_t)(k + 1) * np);
        }
        if (np_out) *np_out = g->np;
    });
}

int gpgmp_individuals_keys_device(gpgmp_individuals *g, uint32_t **keys, uint32_t **vals, uint32_t **off) {
    return guarded_i([&] {
        GPGMP_REQUIRE(g, GPGMP_E_INVALID, "null argument");
        if (keys) *keys = g->keys.get();
        if (vals) *vals = g->ids.get();
        if (off) *off = g->off.get();
    });
}

}  // extern "C"
