// Deterministic entry points: the reference's stencils/RK4 on device buffers, and the
// differentiable Schnakenberg reaction-diffusion solver (forward with checkpoints,
// adjoint with segment recomputation).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <memory>

#include "deterministic_launch.h"
#include "host.h"

using namespace gpgmp;

template <typename F>
static int guarded_det(F &&f) {
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

struct gpgmp_rdd {
    gpgmp_rdd_params p{};
    int steps = 0, ckpt = 0, nck = 0;
    hipStream_t stream = nullptr;
    size_t A = 0;
    DevBuf<float> ck;        // nck checkpoints of (u, v)
    DevBuf<float> seg, seg2; // ckpt recomputed states of a segment, two segments (seg2 on first backward)
    DevBuf<float> tmp;       // ping-pong partner of the caller's buffers
    DevBuf<double> gk_tiles; // per-tile dL/dk accumulators
    int ntiles = 0;
    bool saved = false;
    uint64_t generation = 0;  // saving forward runs so far; the last one is what backward differentiates
    int ns = 4;  // forward steps per launch (temporal blocking); the adjoint takes up to 4
    RddConsts c{};
    DevBuf<float> kdev;  // rates set on the device (gpgmp_rdd_set_rates_device)
    // the backward recomputes the next segment on `side` while the caller's stream walks the
    // current one (default priority, non-blocking; events order the two)
    hipStream_t side = nullptr;
    hipEvent_t ev_start = nullptr, ev_rec[2] = {nullptr, nullptr}, ev_adj[2] = {nullptr, nullptr};
    ~gpgmp_rdd() {
        (void)hipStreamSynchronize(stream);
        if (side) {
            (void)hipStreamSynchronize(side);
            (void)hipStreamDestroy(side);
        }
        for (hipEvent_t e : {ev_start, ev_rec[0], ev_rec[1], ev_adj[0], ev_adj[1]})
            if (e) (void)hipEventDestroy(e);
    }
};

static RddConsts rdd_consts(const gpgmp_rdd_params &p) {
    RddConsts c;
    c.dt = p.dt;
    c.cu = p.D0u / (p.h * p.h);
    c.cv = p.D0v / (p.h * p.h);
    for (int i = 0; i < 4; ++i) c.k[i] = p.k[i];
    return c;
}

extern "C" {

int gpgmp_det_hom_diffuse_device(const float *in, float *out, int32_t W, int32_t H, int32_t S, const float *D_host,
                                 float dt, float h, void *stream) {
    return guarded_det([&] {
        GPGMP_REQUIRE(in && out && D_host && in != out, GPGMP_E_INVALID, "bad buffers (Jacobi needs in != out)");
        GPGMP_REQUIRE(W >= 2 && H >= 2 && S >= 1, GPGMP_E_INVALID, "bad shape");
        const size_t A = (size_t)W * H;
        for (int s = 0; s < S; ++s)
            GPGMP_HIP_CHECK(launch_det_hom_diffuse(in + s * A, out + s * A, W, H, dt * D_host[s] / (h * h),
                                                   (hipStream_t)stream));
    });
}

int gpgmp_det_inhom_sweep_device(const float *in, float *out, int32_t W, int32_t H, int32_t species, int32_t axis,
                                 float dt, const float *D_axis, const float *r_axis, float h, void *stream) {
    return guarded_det([&] {
        GPGMP_REQUIRE(in && out && D_axis && r_axis && in != out, GPGMP_E_INVALID, "bad buffers");
        GPGMP_REQUIRE(W >= 2 && H >= 2 && species >= 0 && (axis == 0 || axis == 1), GPGMP_E_INVALID, "bad shape");
        const size_t o = (size_t)species * W * H;
        GPGMP_HIP_CHECK(launch_det_inhom_sweep(in + o, out + o, D_axis + o, r_axis + o, W, H, axis, dt / (h * h),
                                               dt / (2.f * h), (hipStream_t)stream));
    });
}

int gpgmp_det_rk4_device(const gpgmp_model *m, float *state, float dt, void *stream) {
    return guarded_det([&] {
        GPGMP_REQUIRE(m && state, GPGMP_E_INVALID, "null argument");
        const int A = m->m.desc.width * m->m.desc.height;
        if (m->m.table.R == 0) return;
        GPGMP_REQUIRE(!m->m.table.localized && !has_custom_laws(m->m), GPGMP_E_UNSUPPORTED,
                      "localized reactions and custom laws need a deterministic solver (it owns the mask and "
                      "the law programs)");
        GPGMP_HIP_CHECK(launch_det_rk4(state, A, m->m.table, dt, (hipStream_t)stream));
    });
}

int gpgmp_det_aqss_device(const gpgmp_model *m, float *state, float dt, int32_t flags, void *stream) {
    return guarded_det([&] {
        GPGMP_REQUIRE(m && state, GPGMP_E_INVALID, "null argument");
        GPGMP_REQUIRE((flags & ~GPGMP_AQSS_PADE_FIX) == 0, GPGMP_E_INVALID, "unknown flags");
        const int A = m->m.desc.width * m->m.desc.height;
        if (m->m.table.R == 0) return;
        GPGMP_REQUIRE(!m->m.table.localized && !has_custom_laws(m->m), GPGMP_E_UNSUPPORTED,
                      "localized reactions and custom laws need a deterministic solver (it owns the mask and "
                      "the law programs)");
        GPGMP_HIP_CHECK(launch_det_aqss(state, A, m->m.table, dt, flags & GPGMP_AQSS_PADE_FIX, (hipStream_t)stream));
    });
}

int gpgmp_rdd_create(const gpgmp_rdd_params *p, int32_t steps, int32_t ckpt_every, void *stream, gpgmp_rdd **out) {
    return guarded_det([&] {
        GPGMP_REQUIRE(p && out, GPGMP_E_INVALID, "null argument");
        *out = nullptr;
        GPGMP_REQUIRE(p->width >= 2 && p->height >= 2, GPGMP_E_INVALID, "grid sides must be >= 2");
        GPGMP_REQUIRE(steps >= 0 && ckpt_every >= 1, GPGMP_E_INVALID, "bad steps / checkpoint interval");
        GPGMP_REQUIRE(p->h > 0.f && p->dt >= 0.f, GPGMP_E_INVALID, "bad h / dt");
        std::unique_ptr<gpgmp_rdd> r(new gpgmp_rdd());
        r->p = *p;
        r->steps = steps;
        r->ckpt = std::min(ckpt_every, std::max(steps, 1));
        r->nck = steps > 0 ? (steps + r->ckpt - 1) / r->ckpt : 0;
        r->A = (size_t)p->width * p->height;
        r->stream = (hipStream_t)stream;  // NULL: the default stream
        r->ck.alloc((size_t)std::max(r->nck, 1) * 2 * r->A);
        r->seg.alloc((size_t)r->ckpt * 2 * r->A);
        r->tmp.alloc(2 * r->A);
        r->ntiles = rdd_num_tiles(p->width, p->height);
        r->gk_tiles.alloc((size_t)r->ntiles * 4);
        r->c = rdd_consts(*p);
        r->c.kdev = nullptr;
        const int side = std::min(p->width, p->height);
        r->ns = side >= 5 ? 4 : (side >= 3 ? 2 : 1);
        *out = r.release();
    });
}

int gpgmp_rdd_destroy(gpgmp_rdd *r) {
    delete r;
    return GPGMP_OK;
}

int gpgmp_rdd_set_rates(gpgmp_rdd *r, const float *k) {
    return guarded_det([&] {
        GPGMP_REQUIRE(r && k, GPGMP_E_INVALID, "null argument");
        for (int i = 0; i < 4; ++i) r->p.k[i] = k[i];
        r->c = rdd_consts(r->p);
        r->saved = false;
    });
}

int gpgmp_rdd_set_rates_device(gpgmp_rdd *r, const float *k_dev) {
    return guarded_det([&] {
        GPGMP_REQUIRE(r && k_dev, GPGMP_E_INVALID, "null argument");
        if (!r->kdev.get()) r->kdev.alloc(4);
        GPGMP_HIP_CHECK(hipMemcpyAsync(r->kdev.get(), k_dev, 4 * sizeof(float), hipMemcpyDeviceToDevice, r->stream));
        r->c.kdev = r->kdev.get();
        r->saved = false;
    });
}

int gpgmp_rdd_generation(gpgmp_rdd *r, uint64_t *generation) {
    return guarded_det([&] {
        GPGMP_REQUIRE(r && generation, GPGMP_E_INVALID, "null argument");
        *generation = r->saved ? r->generation : 0;
    });
}

int gpgmp_rdd_set_steps_per_launch(gpgmp_rdd *r, int32_t ns) {
    return guarded_det([&] {
        GPGMP_REQUIRE(r && ns >= 1 && ns <= 4, GPGMP_E_INVALID, "steps per launch must be 1 .. 4");
        GPGMP_REQUIRE(std::min(r->p.width, r->p.height) >= ns + 1, GPGMP_E_INVALID,
                      "ns steps per launch need sides >= ns + 1");
        r->ns = ns;
    });
}

int gpgmp_rdd_forward(gpgmp_rdd *r, const float *D, float *u, float *v, int32_t save) {
    return guarded_det([&] {
        GPGMP_REQUIRE(r && D && u && v, GPGMP_E_INVALID, "null argument");
        const int W = r->p.width, H = r->p.height;
        const size_t A = r->A;
        float *bu[2] = {u, r->tmp.get()};
        float *bv[2] = {v, r->tmp.get() + A};
        int cur = 0;
        int n = 0;
        while (n < r->steps) {
            if (save && n % r->ckpt == 0) {
                float *c = r->ck.get() + (size_t)(n / r->ckpt) * 2 * A;
                GPGMP_HIP_CHECK(hipMemcpyAsync(c, bu[cur], A * sizeof(float), hipMemcpyDeviceToDevice, r->stream));
                GPGMP_HIP_CHECK(hipMemcpyAsync(c + A, bv[cur], A * sizeof(float), hipMemcpyDeviceToDevice, r->stream));
            }
            // ns steps per launch unless a checkpoint or the end falls in between
            int k = std::min(r->ns, r->steps - n);
            if (save) k = std::min(k, r->ckpt - n % r->ckpt);
            GPGMP_HIP_CHECK(launch_rdd_step(bu[cur], bv[cur], D, nullptr, nullptr, bu[cur ^ 1], bv[cur ^ 1], W, H, r->c, k,
                                            r->stream));
            cur ^= 1;
            n += k;
        }
        if (cur) {
            GPGMP_HIP_CHECK(hipMemcpyAsync(u, bu[1], A * sizeof(float), hipMemcpyDeviceToDevice, r->stream));
            GPGMP_HIP_CHECK(hipMemcpyAsync(v, bv[1], A * sizeof(float), hipMemcpyDeviceToDevice, r->stream));
        }
        r->saved = save != 0;
        if (save) ++r->generation;
    });
}

int gpgmp_rdd_backward(gpgmp_rdd *r, uint64_t generation, const float *D, float *lu, float *lv, float *gD,
                       float *gk) {
    return guarded_det([&] {
        GPGMP_REQUIRE(r && D && lu && lv && gD && gk, GPGMP_E_INVALID, "null argument");
        GPGMP_REQUIRE(r->saved, GPGMP_E_STATE, "backward needs a forward run with save != 0");
        GPGMP_REQUIRE(generation == 0 || generation == r->generation, GPGMP_E_STATE,
                      "the saved forward is not the one asked for (another forward ran since)");
        const int W = r->p.width, H = r->p.height;
        const size_t A = r->A;
        GPGMP_HIP_CHECK(launch_fill_f32(gD, A, 0.f, r->stream));
        GPGMP_HIP_CHECK(launch_fill_f64(r->gk_tiles.get(), (size_t)r->ntiles * 4, 0.0, r->stream));
        float *blu[2] = {lu, r->tmp.get()};
        float *blv[2] = {lv, r->tmp.get() + A};
        int cur = 0;
        if (!r->side) {
            r->seg2.alloc((size_t)r->ckpt * 2 * A);
            GPGMP_HIP_CHECK(hipStreamCreateWithFlags(&r->side, hipStreamNonBlocking));
            for (hipEvent_t *e : {&r->ev_start, &r->ev_rec[0], &r->ev_rec[1], &r->ev_adj[0], &r->ev_adj[1]})
                GPGMP_HIP_CHECK(hipEventCreateWithFlags(e, hipEventDisableTiming));
        }
        float *segs[2] = {r->seg.get(), r->seg2.get()};
        // state n of segment c (its checkpoint at n0, the recomputed states after it)
        auto state = [&](int c, int n) -> float * {
            const int n0 = c * r->ckpt;
            return (n == n0) ? r->ck.get() + (size_t)c * 2 * A : segs[c & 1] + (size_t)(n - n0) * 2 * A;
        };
        // recompute the states n0+1 .. n1-1 of segment c on stream st
        auto recompute = [&](int c, hipStream_t st) {
            const int n0 = c * r->ckpt, n1 = std::min(r->steps, n0 + r->ckpt);
            auto su = [&](int n) { return state(c, n); };
            int n = n0;
            while (n + 1 < n1) {
                if (r->ns >= 4 && n + 4 < n1) {  // states n+1 .. n+3 from the launch's LDS, n+4 its output
                    GPGMP_HIP_CHECK(launch_rdd_step(su(n), su(n) + A, D, su(n + 1), su(n + 1) + A, su(n + 4),
                                                    su(n + 4) + A, W, H, r->c, 4, st, nullptr, nullptr, 2 * A));
                    n += 4;
                } else if (r->ns >= 2 && n + 2 < n1) {
                    GPGMP_HIP_CHECK(launch_rdd_step(su(n), su(n) + A, D, su(n + 1), su(n + 1) + A, su(n + 2),
                                                    su(n + 2) + A, W, H, r->c, 2, st));
                    n += 2;
                } else {
                    GPGMP_HIP_CHECK(launch_rdd_step(su(n), su(n) + A, D, nullptr, nullptr, su(n + 1), su(n + 1) + A, W, H,
                                                    r->c, 1, st));
                    n += 1;
                }
            }
        };
        // walk the steps of segment c backwards on the caller's stream
        auto walk = [&](int c) {
            const int n0 = c * r->ckpt, n1 = std::min(r->steps, n0 + r->ckpt);
            auto su = [&](int n) { return state(c, n); };
            int n = n1 - 1;
            while (n >= n0) {
                if (r->ns >= 4 && n - 3 >= n0) {
                    const float *us[4] = {su(n), su(n - 1), su(n - 2), su(n - 3)};
                    const float *vs[4] = {su(n) + A, su(n - 1) + A, su(n - 2) + A, su(n - 3) + A};
                    GPGMP_HIP_CHECK(launch_rdd_adjoint_n(blu[cur], blv[cur], us, vs, D, blu[cur ^ 1], blv[cur ^ 1], gD,
                                                         r->gk_tiles.get(), W, H, r->c, 4, r->stream));
                    n -= 4;
                } else if (r->ns >= 2 && n - 1 >= n0) {
                    GPGMP_HIP_CHECK(launch_rdd_adjoint(blu[cur], blv[cur], su(n), su(n) + A, su(n - 1), su(n - 1) + A, D,
                                                       blu[cur ^ 1], blv[cur ^ 1], gD, r->gk_tiles.get(), W, H, r->c, 2,
                                                       r->stream));
                    n -= 2;
                } else {
                    GPGMP_HIP_CHECK(launch_rdd_adjoint(blu[cur], blv[cur], su(n), su(n) + A, nullptr, nullptr, D,
                                                       blu[cur ^ 1], blv[cur ^ 1], gD, r->gk_tiles.get(), W, H, r->c, 1,
                                                       r->stream));
                    n -= 1;
                }
                cur ^= 1;
            }
        };
        // segment c-1 is recomputed on the side stream while segment c is walked (two segment
        // buffers; a buffer is recomputed again only after the walk that read it)
        GPGMP_HIP_CHECK(hipEventRecord(r->ev_start, r->stream));
        GPGMP_HIP_CHECK(hipStreamWaitEvent(r->side, r->ev_start, 0));
        const int last = r->nck - 1;
        if (last >= 0) {
            recompute(last, r->side);
            GPGMP_HIP_CHECK(hipEventRecord(r->ev_rec[last & 1], r->side));
        }
        for (int c = last; c >= 0; --c) {
            if (c >= 1) {
                if (c + 1 <= last) GPGMP_HIP_CHECK(hipStreamWaitEvent(r->side, r->ev_adj[(c + 1) & 1], 0));
                recompute(c - 1, r->side);
                GPGMP_HIP_CHECK(hipEventRecord(r->ev_rec[(c - 1) & 1], r->side));
            }
            GPGMP_HIP_CHECK(hipStreamWaitEvent(r->stream, r->ev_rec[c & 1], 0));
            walk(c);
            GPGMP_HIP_CHECK(hipEventRecord(r->ev_adj[c & 1], r->stream));
        }
        if (cur) {
            GPGMP_HIP_CHECK(hipMemcpyAsync(lu, blu[1], A * sizeof(float), hipMemcpyDeviceToDevice, r->stream));
            GPGMP_HIP_CHECK(hipMemcpyAsync(lv, blv[1], A * sizeof(float), hipMemcpyDeviceToDevice, r->stream));
        }
        GPGMP_HIP_CHECK(launch_rdd_gk_reduce(r->gk_tiles.get(), r->ntiles, gk, r->stream));
    });
}

}  // extern "C"
