// C ABI of the GPGMP HIP library: status codes at the boundary, C++ inside.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "det_solver.h"
#include "fields_launch.h"
#include "law_compile.h"
#include "host.h"

using namespace gpgmp;

namespace {
thread_local std::string g_last_error;

template <typename F>
int guarded(F &&f) {
    try {
        f();
        return GPGMP_OK;
    } catch (const Error &e) {
        g_last_error = e.what();
        return e.code;
    } catch (const std::bad_alloc &) {
        g_last_error = "host allocation failed";
        return GPGMP_E_CAPACITY;
    } catch (const std::exception &e) {
        g_last_error = e.what();
        return GPGMP_E_INVALID;
    }
}

// The caller's stream; NULL is the default (null) stream, so work is ordered with
// every other blocking stream of the process (torch's default stream included).
struct StreamHolder {
    hipStream_t s = nullptr;
    void init(void *user) { s = (hipStream_t)user; }
};
}  // namespace

// Plain solver: one realization driven through the gpgmp::Solver-shaped calls, or
// through whole host loops (run).
struct gpgmp_solver {
    Model m;
    StreamHolder stream;
    std::unique_ptr<GmpEngine> eng;  // stochastic types
    std::unique_ptr<DetEngine> det;  // deterministic types
    std::vector<float> dtx, dty, nx, ny;
    DevBuf<LaunchPlan> plan_d;
    DevBuf<int32_t> seed_d;
    std::unique_ptr<GmpRunner> runner;
    uint32_t runner_seed0 = 0;
    // output events of the host loop: interval, max_time (infinity = off), and up to
    // max_dumps state snapshots kept on the device with their times
    float out_interval = __builtin_inff(), max_time = __builtin_inff();
    int32_t max_dumps = 0;
    DevBuf<int32_t> dumps;
    std::vector<float> dump_times;
    // individual (particle-based) species, by species index: their counts in the state
    // are the particles' (gpgmp_individuals_bind_counts), reactions are followed by the
    // population maintenance and their diffusion is the particles' hops
    struct Individuals {
        gpgmp_individuals *g;
        bool has_rule;
        gpgmp_new_individual_rule rule;
    };
    std::map<int, Individuals> inds;
    // non-linear diffusivity (InhomogeneousSolver.cpp:304-336,384-417): the drift/diffusivity
    // method's programs, its field parameters [nfields][A], and the fields [4][S][A] (cur:
    // DiffusivityX, DiffusivityY, DriftX, DriftY as the sweeps use them; out: the method's)
    struct Nonlinear {
        DevBuf<uint32_t> code, tab;
        DevBuf<float> params, cur, out;
    };
    std::unique_ptr<Nonlinear> nl;
    void record_dump(float t) {
        const size_t n = eng->state_elems();
        if ((int32_t)dump_times.size() < max_dumps)
            GPGMP_HIP_CHECK(hipMemcpyAsync(dumps.get() + dump_times.size() * n, eng->state(), n * sizeof(int32_t),
                                           hipMemcpyDeviceToDevice, eng->stream()));
        dump_times.push_back(t);
    }
};

struct gpgmp_batch {
    Model m;
    StreamHolder stream;
    std::unique_ptr<GmpEngine> eng;
    std::vector<float> dtx, dty;
    std::vector<uint32_t> seed0s;
    std::unique_ptr<GmpRunner> runner;
};

void gpgmp_set_last_error(const char *msg) { g_last_error = msg ? msg : ""; }

extern "C" {

const char *gpgmp_last_error(void) { return g_last_error.c_str(); }

int gpgmp_version(int32_t *major, int32_t *minor) {
    if (major) *major = 1;
    if (minor) *minor = 0;
    return GPGMP_OK;
}

int gpgmp_model_create(const gpgmp_model_desc *desc, gpgmp_model **out) {
    return guarded([&] {
        GPGMP_REQUIRE(desc && out, GPGMP_E_INVALID, "null argument");
        *out = nullptr;
        std::unique_ptr<gpgmp_model> h(new gpgmp_model{compile_model(*desc)});
        *out = h.release();
    });
}

int gpgmp_model_destroy(gpgmp_model *m) {
    delete m;
    return GPGMP_OK;
}

int gpgmp_model_rate_literals(const gpgmp_model *m, int32_t *is_int, double *value) {
    return guarded([&] {
        GPGMP_REQUIRE(m && is_int && value, GPGMP_E_INVALID, "null argument");
        for (int i = 0; i < m->m.table.R; ++i) {
            is_int[i] = (m->m.table.lit_is_int >> i) & 1u;
            value[i] = m->m.table.lit_dbl[i];
        }
    });
}

int gpgmp_model_add_reaction_compartment(gpgmp_model *m, int32_t reaction, int32_t x0, int32_t y0, int32_t x1,
                                          int32_t y1) {
    return guarded([&] {
        GPGMP_REQUIRE(m, GPGMP_E_INVALID, "null argument");
        GPGMP_REQUIRE(reaction >= 0 && reaction < m->m.table.R, GPGMP_E_INVALID, "no such reaction");
        GPGMP_REQUIRE(m->m.compartments.size() < 4096, GPGMP_E_CAPACITY, "too many compartments");
        m->m.compartments.push_back({reaction, x0, y0, x1, y1});
        m->m.table.localized |= 1u << reaction;
        m->m.table.compact = 0;  // the localized law is formed by the general table walk
    });
}

int gpgmp_model_set_species_ids(gpgmp_model *m, const char *const *ids) {
    return guarded([&] {
        GPGMP_REQUIRE(m && ids, GPGMP_E_INVALID, "null argument");
        std::vector<std::string> v;
        for (int s = 0; s < m->m.desc.num_species; ++s) {
            GPGMP_REQUIRE(ids[s] && ids[s][0], GPGMP_E_INVALID, "empty species id");
            v.push_back(ids[s]);
        }
        m->m.species_ids = v;
    });
}

int gpgmp_model_set_parameter(gpgmp_model *m, const char *name, float value) {
    return guarded([&] {
        GPGMP_REQUIRE(m && name && name[0], GPGMP_E_INVALID, "null argument");
        m->m.params[name] = value;
    });
}

int gpgmp_model_set_kinetic_law(gpgmp_model *m, int32_t reaction, const char *formula) {
    return guarded([&] {
        GPGMP_REQUIRE(m && formula, GPGMP_E_INVALID, "null argument");
        GPGMP_REQUIRE(reaction >= 0 && reaction < m->m.table.R, GPGMP_E_INVALID, "no such reaction");
        Model trial = m->m;
        trial.custom_law.resize((size_t)trial.table.R);
        trial.custom_law[(size_t)reaction] = formula;
        RxTable T = trial.table;
        (void)build_law_programs(trial, T);  // throws with the reason if the formula does not compile
        m->m.custom_law = trial.custom_law;
    });
}

int gpgmp_model_law_hazard(const gpgmp_model *m, int32_t reaction, const int32_t *x, int32_t gx, int32_t gy,
                           float h0, uint32_t act, int32_t how, float *out) {
    return guarded([&] {
        GPGMP_REQUIRE(m && x && out, GPGMP_E_INVALID, "null argument");
        GPGMP_REQUIRE(reaction >= 0 && reaction < m->m.table.R, GPGMP_E_INVALID, "no such reaction");
        GPGMP_REQUIRE(how >= 0 && how <= 2, GPGMP_E_INVALID, "how must be 0, 1 or 2");
        *out = law_hazard_host(m->m, reaction, x, gx, gy, h0, act, how);
    });
}

int gpgmp_model_law_rhs(const gpgmp_model *m, int32_t species, int32_t which, const float *y, int32_t gx,
                        int32_t gy, uint32_t act, float *out) {
    return guarded([&] {
        GPGMP_REQUIRE(m && y && out, GPGMP_E_INVALID, "null argument");
        GPGMP_REQUIRE(species >= 0 && species < m->m.table.S && which >= 0 && which <= 2, GPGMP_E_INVALID,
                      "bad species or expression");
        (void)law_rhs_host(m->m, species, which, y, gx, gy, act, out);
    });
}

int gpgmp_model_eval_drift_diffusivity(const gpgmp_model *m, const char *method, const char *const *field_names,
                                       int32_t nfields, const float *fields, const int32_t *state,
                                       const float *current, float sim_time, float *out, void *stream) {
    return guarded([&] {
        GPGMP_REQUIRE(m && method && state && out, GPGMP_E_INVALID, "null argument");
        GPGMP_REQUIRE(nfields >= 0 && (nfields == 0 || (field_names && fields)), GPGMP_E_INVALID,
                      "fields need their names and values");
        const Model &md = m->m;
        const int W = md.desc.width, H = md.desc.height, S = md.desc.num_species;
        const size_t A = (size_t)W * H;
        LawEnv env = law_env(md, false);
        for (int k = 0; k < nfields; ++k) env.fields.push_back(field_names[k]);
        const std::vector<LawProgram> progs = compile_method(method, env);
        std::vector<uint32_t> code, table;
        for (const LawProgram &p : progs) {
            GPGMP_REQUIRE(code.size() + p.code.size() <= (size_t)1 << 20, GPGMP_E_CAPACITY, "method too long");
            table.push_back((uint32_t)code.size());
            table.push_back((uint32_t)p.type);
            code.insert(code.end(), p.code.begin(), p.code.end());
        }
        hipStream_t st = (hipStream_t)stream;
        DevBuf<uint32_t> dcode(code.size()), dtab(table.size());
        DevBuf<int32_t> dstate((size_t)S * A);
        DevBuf<float> dfields(std::max<size_t>((size_t)nfields * A, 1)), dcur((size_t)4 * S * A),
            dout((size_t)4 * S * A);
        GPGMP_HIP_CHECK(hipMemcpyAsync(dcode.get(), code.data(), code.size() * 4, hipMemcpyHostToDevice, st));
        GPGMP_HIP_CHECK(hipMemcpyAsync(dtab.get(), table.data(), table.size() * 4, hipMemcpyHostToDevice, st));
        GPGMP_HIP_CHECK(hipMemcpyAsync(dstate.get(), state, (size_t)S * A * 4, hipMemcpyHostToDevice, st));
        if (nfields > 0)
            GPGMP_HIP_CHECK(hipMemcpyAsync(dfields.get(), fields, (size_t)nfields * A * 4, hipMemcpyHostToDevice, st));
        if (current)
            GPGMP_HIP_CHECK(hipMemcpyAsync(dcur.get(), current, (size_t)4 * S * A * 4, hipMemcpyHostToDevice, st));
        else
            GPGMP_HIP_CHECK(hipMemsetAsync(dcur.get(), 0, (size_t)4 * S * A * 4, st));
        LawExtra ex{dstate.get(), dfields.get(), dcur.get(), W, H, S, 0.f, 0.f, sim_time};
        const float pm = header_float(md.desc.physical_length);
        GPGMP_HIP_CHECK(launch_drift_fields(dcode.get(), dtab.get(), ex, pm, pm, dout.get(), st));
        GPGMP_HIP_CHECK(hipMemcpyAsync(out, dout.get(), (size_t)4 * S * A * 4, hipMemcpyDeviceToHost, st));
        GPGMP_HIP_CHECK(hipStreamSynchronize(st));
    });
}

int gpgmp_model_method_dependencies(const gpgmp_model *m, const char *method, const char *const *field_names,
                                    int32_t nfields, uint32_t *deps) {
    return guarded([&] {
        GPGMP_REQUIRE(m && method && deps, GPGMP_E_INVALID, "null argument");
        GPGMP_REQUIRE(nfields >= 0 && (nfields == 0 || field_names), GPGMP_E_INVALID, "fields need their names");
        LawEnv env = law_env(m->m, false);
        for (int k = 0; k < nfields; ++k) env.fields.push_back(field_names[k]);
        *deps = method_dependencies(compile_method(method, env), (int)env.species.size());
    });
}

int gpgmp_model_method_host(const gpgmp_model *m, const char *method, const char *const *field_names,
                            int32_t nfields, const float *fields, const int32_t *state, const float *current,
                            float sim_time, float *out) {
    return guarded([&] {
        GPGMP_REQUIRE(m && method && state && current && out, GPGMP_E_INVALID, "null argument");
        const Model &md = m->m;
        const int W = md.desc.width, H = md.desc.height, S = md.desc.num_species;
        const size_t A = (size_t)W * H;
        LawEnv env = law_env(md, false);
        for (int k = 0; k < nfields; ++k) env.fields.push_back(field_names[k]);
        const std::vector<LawProgram> progs = compile_method(method, env);
        const float pm = header_float(md.desc.physical_length);
        const int32_t none[1] = {0};
        for (size_t c = 0; c < A; ++c) {
            const int gx = (int)(c % W), gy = (int)(c / W);
            LawExtra ex{state, fields, current, W, H, S, 0.f, 0.f, sim_time};
            ex.px = (float)((double)(pm / (float)W) * ((double)(float)gx - (double)(float)W / 2.));
            ex.py = (float)((double)(pm / (float)H) * ((double)(float)gy - (double)(float)H / 2.));
            for (size_t k = 0; k < progs.size(); ++k)
                out[k * A + c] = law_to_float(law_run<int32_t, 1>(progs[k].code.data(), none, LawCell{gx, gy, 0u, 0.f}, &ex),
                                              progs[k].type);
        }
    });
}

int gpgmp_model_localized(const gpgmp_model *m, uint32_t *mask) {
    return guarded([&] {
        GPGMP_REQUIRE(m && mask, GPGMP_E_INVALID, "null argument");
        *mask = m->m.table.localized;
    });
}

// ---- plain solver -------------------------------------------------------------------
int gpgmp_solver_create(const gpgmp_model *model, void *stream, uint32_t flags, gpgmp_solver **out) {
    return guarded([&] {
        GPGMP_REQUIRE(model && out, GPGMP_E_INVALID, "null argument");
        *out = nullptr;
        GPGMP_REQUIRE((flags & ~(GPGMP_SOLVER_AQSS_PADE_FIX | GPGMP_SOLVER_REFERENCE_QUIRKS)) == 0, GPGMP_E_INVALID,
                      "unknown solver flags");
        std::unique_ptr<gpgmp_solver> s(new gpgmp_solver());
        s->m = model->m;
        s->m.desc.reactions = s->m.reactions.empty() ? nullptr : s->m.reactions.data();
        s->stream.init(stream);
        if (s->m.deterministic()) {
            s->det.reset(new DetEngine(s->m, flags, s->stream.s));
            s->dtx = s->det->dt_x();
            s->dty = s->det->dt_y();
            *out = s.release();
            return;
        }
        const gpgmp_model_desc &d = s->m.desc;
        const int S = d.num_species;
        s->eng.reset(new GmpEngine(s->m, {d.width}, {d.height}, s->stream.s));
        s->dtx.assign(S, 0.f);
        s->dty.assign(S, 0.f);
        if (!s->m.inhomogeneous()) {
            // HomogeneousSolver.cpp:87-95 (h = physicalLength / gridWidth in float)
            const float h = d.physical_length / (float)d.width;
            for (int i = 0; i < S; ++i)
                s->dtx[i] = (d.diffusion[i] > 0) ? homogeneous_dt(d.p0, h, d.diffusion[i]) : d.max_time;
            s->dty = s->dtx;
        } else {
            // InhomogeneousSolver.cpp:189-212: per species and axis, the minimum admissible
            // step over the cells (both axes use PhysicalCellWidth, as the reference's
            // computeIndividualTimestepY does), the Fokker-Planck form's for the fpe type
            const int A = d.width * d.height;
            const bool fpe = s->m.fpe();
            auto cell_dt = [&](float D, float r) {
                return fpe ? fpe_cell_dt(D, r, s->m.cell_width, s->m.probs.stay)
                           : inhomogeneous_cell_dt(D, r, s->m.cell_width);
            };
            for (int i = 0; i < S; ++i) {
                float mx = INFINITY, my = INFINITY;
                for (int c = 0; c < A; ++c) {
                    mx = std::min(mx, cell_dt(s->m.Dx[(size_t)i * A + c], s->m.rx[(size_t)i * A + c]));
                    my = std::min(my, cell_dt(s->m.Dy[(size_t)i * A + c], s->m.ry[(size_t)i * A + c]));
                }
                s->dtx[i] = mx;
                s->dty[i] = my;
                s->eng->set_probs(0, i, 0, s->m.Dx.data() + (size_t)i * A, s->m.rx.data() + (size_t)i * A, mx,
                                  s->m.cell_width);
                s->eng->set_probs(0, i, 1, s->m.Dy.data() + (size_t)i * A, s->m.ry.data() + (size_t)i * A, my,
                                  s->m.cell_height);
            }
        }
        s->nx = s->dtx;
        s->ny = s->dty;
        s->plan_d.alloc(4);
        s->seed_d.alloc(4);
        *out = s.release();
    });
}

int gpgmp_solver_destroy(gpgmp_solver *s) {
    if (s) {
        s->runner.reset();
        if (s->stream.s) (void)hipStreamSynchronize(s->stream.s);
    }
    delete s;
    return GPGMP_OK;
}

// the stochastic engine of a handle (the calls that only stochastic types have)
static GmpEngine &stoch(gpgmp_solver *s) {
    GPGMP_REQUIRE(s, GPGMP_E_INVALID, "null argument");
    GPGMP_REQUIRE(s->eng != nullptr, GPGMP_E_UNSUPPORTED, "a call of the stochastic solver types");
    return *s->eng;
}

int gpgmp_solver_set_state(gpgmp_solver *s, const int32_t *host) {
    return guarded([&] {
        GPGMP_REQUIRE(s && host, GPGMP_E_INVALID, "null argument");
        GPGMP_REQUIRE(!s->det, GPGMP_E_STATE, "deterministic solvers hold fp32 concentrations");
        s->eng->set_state(0, host);
    });
}

int gpgmp_solver_get_state(gpgmp_solver *s, int32_t *host) {
    return guarded([&] {
        GPGMP_REQUIRE(s && host, GPGMP_E_INVALID, "null argument");
        GPGMP_REQUIRE(!s->det, GPGMP_E_STATE, "deterministic solvers hold fp32 concentrations");
        s->eng->get_state(0, host);
    });
}

int gpgmp_solver_set_state_f32(gpgmp_solver *s, const float *host) {
    return guarded([&] {
        GPGMP_REQUIRE(s && host, GPGMP_E_INVALID, "null argument");
        GPGMP_REQUIRE(s->det != nullptr, GPGMP_E_STATE, "stochastic solvers hold int32 counts");
        s->det->set_state(host);
    });
}

int gpgmp_solver_get_state_f32(gpgmp_solver *s, float *host) {
    return guarded([&] {
        GPGMP_REQUIRE(s && host, GPGMP_E_INVALID, "null argument");
        GPGMP_REQUIRE(s->det != nullptr, GPGMP_E_STATE, "stochastic solvers hold int32 counts");
        s->det->get_state(host);
    });
}

int gpgmp_solver_state_device(gpgmp_solver *s, void **dev_ptr) {
    return guarded([&] {
        GPGMP_REQUIRE(s && dev_ptr, GPGMP_E_INVALID, "null argument");
        *dev_ptr = s->det ? (void *)s->det->state() : (void *)s->eng->state();
    });
}

namespace {
// one API-level launch with a single plan and the given seeds (current, previous)
void api_launch(gpgmp_solver *s, int mode, const LaunchPlan &p, int32_t seed_cur, int32_t seed_prev,
                const int32_t *lin, int32_t *lout) {
    const int32_t seeds[2] = {seed_cur, seed_prev};
    GPGMP_HIP_CHECK(hipMemcpyAsync(s->plan_d.get(), &p, sizeof p, hipMemcpyHostToDevice, s->stream.s));
    GPGMP_HIP_CHECK(hipMemcpyAsync(s->seed_d.get(), seeds, sizeof seeds, hipMemcpyHostToDevice, s->stream.s));
    s->eng->launch(mode, s->plan_d.get(), s->seed_d.get(), s->seed_d.get() + 1, lin, lout);
    // the host staging above is stack memory: keep the call synchronous like the
    // reference's finish() after every launch (Kernel.h:252-253)
    GPGMP_HIP_CHECK(hipStreamSynchronize(s->stream.s));
}
}  // namespace

// status of a C-ABI call made inside a guarded body
static void ok(int rc) {
    if (rc != GPGMP_OK) throw Error(rc, gpgmp_last_error());
}

int gpgmp_solver_do_reactions(gpgmp_solver *s, float dt, int32_t *seed) {
    return guarded([&] {
        GPGMP_REQUIRE(s && seed, GPGMP_E_INVALID, "null argument");
        if (s->det) {  // Solver.cpp:250-266: the integrator; the seed is not used
            s->det->do_reactions(dt);
            GPGMP_HIP_CHECK(hipStreamSynchronize(s->stream.s));
            return;
        }
        if (s->m.table.R > 0) {
            LaunchPlan p{};
            p.deltat = dt;
            api_launch(s, s->m.inhomogeneous() ? 1 : 0, p, *seed, *seed, s->eng->lost(1), s->eng->lost(0));
        }
        *seed = (int32_t)((uint32_t)*seed + 1u);  // Solver.cpp:239
        // Solver.cpp:241-247: maintainPopulation(ii, seed++, d_state) per individual species
        if (!s->inds.empty()) GPGMP_HIP_CHECK(hipStreamSynchronize(s->eng->stream()));
        const int A = s->m.desc.width * s->m.desc.height;
        for (auto &kv : s->inds) {
            ok(gpgmp_individuals_maintain(kv.second.g, s->eng->state() + (size_t)kv.first * A, 1, *seed,
                                          kv.second.has_rule ? &kv.second.rule : nullptr));
            ok(gpgmp_individuals_synchronize(kv.second.g));
            *seed = (int32_t)((uint32_t)*seed + 1u);
        }
    });
}

int gpgmp_solver_attach_individuals(gpgmp_solver *s, int32_t species, gpgmp_individuals *g,
                                    const gpgmp_new_individual_rule *rule) {
    return guarded([&] {
        GPGMP_REQUIRE(s && g, GPGMP_E_INVALID, "null argument");
        GPGMP_REQUIRE(s->m.desc.solver_type == GPGMP_STOCHASTIC_INHOMOGENEOUS || s->m.fpe(), GPGMP_E_UNSUPPORTED,
                      "individual species belong to the stochastic inhomogeneous solvers");
        GPGMP_REQUIRE(species >= 0 && species < s->m.desc.num_species, GPGMP_E_INVALID, "species out of range");
        const int A = s->m.desc.width * s->m.desc.height;
        GPGMP_HIP_CHECK(hipStreamSynchronize(s->eng->stream()));
        ok(gpgmp_individuals_bind_counts(g, s->eng->state() + (size_t)species * A));
        float dx = 0.f, dy = 0.f;  // InhomogeneousSolver.cpp:205-225: the first steps from the particles
        ok(gpgmp_individuals_time_steps(g, &dx, &dy));
        s->dtx[species] = s->nx[species] = dx;
        s->dty[species] = s->ny[species] = dy;
        gpgmp_solver::Individuals a{g, rule != nullptr, {}};
        if (rule) a.rule = *rule;
        s->inds[species] = a;
    });
}

int gpgmp_solver_individuals(gpgmp_solver *s, int32_t species, gpgmp_individuals **out) {
    return guarded([&] {
        GPGMP_REQUIRE(s && out, GPGMP_E_INVALID, "null argument");
        GPGMP_REQUIRE(species >= 0 && species < s->m.desc.num_species, GPGMP_E_INVALID, "species out of range");
        auto it = s->inds.find(species);
        *out = (it == s->inds.end()) ? nullptr : it->second.g;
    });
}

// After a sweep of a species on one axis of a non-linear model, as InhomogeneousSolver::diffuse
// does (InhomogeneousSolver.cpp:304-336 after X, :384-417 after Y): the method re-evaluated on
// the swept state at the current time gives the species' four fields (the other species keep
// theirs); the swept axis' time step becomes the minimum admissible step over the cells
// (computeIndividualTimestepX/Y + reduceBlocks, as at set-up).  The cellular-automaton sweeps
// of both axes then take their probabilities from the current fields and time steps, which
// the reference's diffuseX/Y kernels compute at sweep time; the Fokker-Planck sweeps use the
// rates that only the swept axis' time-step kernel stores (.x/.y for X, .z/.w for Y), times
// that axis' step, so only the swept axis' probabilities change.
static void nonlinear_update(gpgmp_solver *s, int sp, int axis, float t) {
    const gpgmp_model_desc &d = s->m.desc;
    const int W = d.width, H = d.height, S = d.num_species;
    const size_t A = (size_t)W * H;
    hipStream_t st = s->eng->stream();
    gpgmp_solver::Nonlinear &nl = *s->nl;
    LawExtra ex{s->eng->state(), nl.params.get(), nl.cur.get(), W, H, S, 0.f, 0.f, t};
    const float pm = header_float(d.physical_length);
    GPGMP_HIP_CHECK(launch_drift_fields(nl.code.get(), nl.tab.get(), ex, pm, pm, nl.out.get(), st));
    std::vector<float> *host[4] = {&s->m.Dx, &s->m.Dy, &s->m.rx, &s->m.ry};
    for (int f = 0; f < 4; ++f) {
        const size_t off = ((size_t)f * S + sp) * A;
        GPGMP_HIP_CHECK(hipMemcpyAsync(nl.cur.get() + off, nl.out.get() + off, A * sizeof(float),
                                       hipMemcpyDeviceToDevice, st));
        GPGMP_HIP_CHECK(hipMemcpyAsync(host[f]->data() + (size_t)sp * A, nl.out.get() + off, A * sizeof(float),
                                       hipMemcpyDeviceToHost, st));
    }
    GPGMP_HIP_CHECK(hipStreamSynchronize(st));
    const float *Dx = s->m.Dx.data() + (size_t)sp * A, *Dy = s->m.Dy.data() + (size_t)sp * A;
    const float *rx = s->m.rx.data() + (size_t)sp * A, *ry = s->m.ry.data() + (size_t)sp * A;
    const float *D = axis ? Dy : Dx, *r = axis ? ry : rx;
    const bool fpe = s->m.fpe();
    float m = INFINITY;
    for (size_t c = 0; c < A; ++c)
        m = std::min(m, fpe ? fpe_cell_dt(D[c], r[c], s->m.cell_width, s->m.probs.stay)
                            : inhomogeneous_cell_dt(D[c], r[c], s->m.cell_width));
    (axis ? s->dty : s->dtx)[sp] = m;
    if (!fpe || axis == 0) s->eng->set_probs(0, sp, 0, Dx, rx, s->dtx[sp], s->m.cell_width);
    if (!fpe || axis == 1) s->eng->set_probs(0, sp, 1, Dy, ry, s->dty[sp], s->m.cell_height);
}

int gpgmp_loop_seeds(uint32_t seed0, int32_t n, int32_t *out) {
    return guarded([&] {
        GPGMP_REQUIRE(n >= 0 && (n == 0 || out), GPGMP_E_INVALID, "bad argument");
        SeedStream ss(seed0);
        for (int32_t i = 0; i < n; ++i) out[i] = ss.next();
    });
}

int gpgmp_solver_set_nonlinear_method(gpgmp_solver *s, const char *method, const char *const *field_names,
                                      int32_t nfields, const float *fields) {
    return guarded([&] {
        GPGMP_REQUIRE(s && method, GPGMP_E_INVALID, "null argument");
        GPGMP_REQUIRE(nfields >= 0 && (nfields == 0 || (field_names && fields)), GPGMP_E_INVALID,
                      "fields need their names and values");
        GPGMP_REQUIRE(s->eng && (s->m.desc.solver_type == GPGMP_STOCHASTIC_INHOMOGENEOUS || s->m.fpe()),
                      GPGMP_E_UNSUPPORTED, "non-linear diffusivity needs a stochastic inhomogeneous solver");
        const gpgmp_model_desc &d = s->m.desc;
        const int S = d.num_species;
        const size_t A = (size_t)d.width * d.height;
        LawEnv env = law_env(s->m, false);
        for (int k = 0; k < nfields; ++k) env.fields.push_back(field_names[k]);
        const std::vector<LawProgram> progs = compile_method(method, env);
        std::vector<uint32_t> code, table;
        for (const LawProgram &p : progs) {
            GPGMP_REQUIRE(code.size() + p.code.size() <= (size_t)1 << 20, GPGMP_E_CAPACITY, "method too long");
            table.push_back((uint32_t)code.size());
            table.push_back((uint32_t)p.type);
            code.insert(code.end(), p.code.begin(), p.code.end());
        }
        std::unique_ptr<gpgmp_solver::Nonlinear> nl(new gpgmp_solver::Nonlinear());
        hipStream_t st = s->eng->stream();
        nl->code.alloc(std::max<size_t>(code.size(), 1));
        nl->tab.alloc(std::max<size_t>(table.size(), 1));
        nl->params.alloc(std::max<size_t>((size_t)nfields * A, 1));
        nl->cur.alloc((size_t)4 * S * A);
        nl->out.alloc((size_t)4 * S * A);
        GPGMP_HIP_CHECK(hipMemcpyAsync(nl->code.get(), code.data(), code.size() * 4, hipMemcpyHostToDevice, st));
        GPGMP_HIP_CHECK(hipMemcpyAsync(nl->tab.get(), table.data(), table.size() * 4, hipMemcpyHostToDevice, st));
        if (nfields > 0)
            GPGMP_HIP_CHECK(hipMemcpyAsync(nl->params.get(), fields, (size_t)nfields * A * 4, hipMemcpyHostToDevice,
                                           st));
        const std::vector<float> *host[4] = {&s->m.Dx, &s->m.Dy, &s->m.rx, &s->m.ry};
        for (int f = 0; f < 4; ++f)
            GPGMP_HIP_CHECK(hipMemcpyAsync(nl->cur.get() + (size_t)f * S * A, host[f]->data(), (size_t)S * A * 4,
                                           hipMemcpyHostToDevice, st));
        GPGMP_HIP_CHECK(hipStreamSynchronize(st));
        s->nl = std::move(nl);
        s->runner.reset();  // the device loop does not re-evaluate the method
    });
}

int gpgmp_solver_diffuse(gpgmp_solver *s, int32_t species, float t, int32_t *seed) {
    return guarded([&] {
        GPGMP_REQUIRE(s && seed, GPGMP_E_INVALID, "null argument");
        GPGMP_REQUIRE(species >= 0 && species < s->m.desc.num_species, GPGMP_E_INVALID, "species out of range");
        if (s->det) {  // the deterministic stencils; the seed is not used
            s->det->diffuse(species, t);
            GPGMP_HIP_CHECK(hipStreamSynchronize(s->stream.s));
            return;
        }
        const uint32_t bit = 1u << species;
        if (!s->m.inhomogeneous()) {
            // homogeneous_stochastic_diffuse + updateState (+ sources) with the same seed
            LaunchPlan p{};
            p.out_mask = bit;
            api_launch(s, 0, p, *seed, *seed, s->eng->lost(1), s->eng->lost(0));
            LaunchPlan g{};
            g.gather_mask = bit;
            api_launch(s, 0, g, *seed, *seed, s->eng->lost(0), s->eng->lost(1));
            s->nx[species] += s->dtx[species];
            s->ny[species] = s->nx[species];
            return;
        }
        auto it = s->inds.find(species);
        if (it != s->inds.end()) {
            // InhomogeneousSolver::diffuse for an individual species (InhomogeneousSolver.cpp:264-269,
            // 304-343, 355-420): the hop sweep, seed++, then the constants and the new time step,
            // which is the one added to the next diffusion time
            gpgmp_individuals *g = it->second.g;
            GPGMP_HIP_CHECK(hipStreamSynchronize(s->eng->stream()));
            for (int axis = 0; axis < 2; ++axis) {
                float &next = axis ? s->ny[species] : s->nx[species];
                float &dt = axis ? s->dty[species] : s->dtx[species];
                if (!(next <= t)) continue;
                ok(gpgmp_individuals_diffuse(g, axis, dt, seed));
                float dnew = 0.f;
                ok(gpgmp_individuals_update(g, 1 << axis, axis ? nullptr : &dnew, axis ? &dnew : nullptr));
                dt = dnew;
                next += dt;
            }
            ok(gpgmp_individuals_synchronize(g));
            return;
        }
        if (s->nx[species] <= t) {
            LaunchPlan p{};
            p.out_mask = bit;
            api_launch(s, 1, p, *seed, *seed, s->eng->lost(1), s->eng->lost(0));
            LaunchPlan g{};
            g.gather_mask = bit;
            api_launch(s, 2, g, *seed, *seed, s->eng->lost(0), s->eng->lost(1));
            *seed = (int32_t)((uint32_t)*seed + 1u);
            if (s->nl) nonlinear_update(s, species, 0, t);
            s->nx[species] += s->dtx[species];
        }
        if (s->ny[species] <= t) {
            LaunchPlan p{};
            p.out_mask = bit;
            api_launch(s, 2, p, *seed, *seed, s->eng->lost(0), s->eng->lost(1));
            LaunchPlan g{};
            g.gather_mask = bit;
            api_launch(s, 1, g, *seed, *seed, s->eng->lost(1), s->eng->lost(0));
            *seed = (int32_t)((uint32_t)*seed + 1u);
            if (s->nl) nonlinear_update(s, species, 1, t);
            s->ny[species] += s->dty[species];
        }
    });
}

int gpgmp_solver_next_diffusion_time(gpgmp_solver *s, float *out) {
    return guarded([&] {
        GPGMP_REQUIRE(s && out, GPGMP_E_INVALID, "null argument");
        if (s->det) {
            const std::vector<float> nt = s->det->next_diffusion_time();
            std::copy(nt.begin(), nt.end(), out);
            return;
        }
        for (int i = 0; i < s->m.desc.num_species; ++i)
            out[i] = (!s->m.inhomogeneous()) ? s->nx[i] : ((s->nx[i] < s->ny[i]) ? s->nx[i] : s->ny[i]);
    });
}

int gpgmp_solver_time_steps(gpgmp_solver *s, float *dt_x, float *dt_y) {
    return guarded([&] {
        GPGMP_REQUIRE(s && dt_x && dt_y, GPGMP_E_INVALID, "null argument");
        for (int i = 0; i < s->m.desc.num_species; ++i) {
            dt_x[i] = s->dtx[i];
            dt_y[i] = s->dty[i];
        }
    });
}

int gpgmp_solver_run(gpgmp_solver *s, int32_t steps, uint32_t seed0, float *sim_time_out) {
    return guarded([&] {
        GPGMP_REQUIRE(s, GPGMP_E_INVALID, "null argument");
        if (s->det) {  // no seeds in the deterministic loop
            s->det->reset();
            s->det->run(steps);
            if (sim_time_out) *sim_time_out = s->det->sim_time();
            return;
        }
        GPGMP_REQUIRE(s->inds.empty(), GPGMP_E_UNSUPPORTED,
                      "models with individual species run through the per-call API (do_reactions / diffuse)");
        GPGMP_REQUIRE(!s->nl, GPGMP_E_UNSUPPORTED,
                      "non-linear diffusivity runs through the per-call API (do_reactions / diffuse)");
        s->runner.reset();
        s->runner.reset(new GmpRunner(*s->eng, s->dtx, s->dty, {seed0}));
        s->runner_seed0 = seed0;
        if (std::isfinite(s->out_interval) || std::isfinite(s->max_time)) {
            // DiffusionModel::oclSolve writes the initial state before its loop
            // (DiffusionModel.cpp:1442), then at output events
            s->dump_times.clear();
            s->record_dump(0.f);
            s->runner->set_output(s->out_interval, s->max_time, [s](float t) { s->record_dump(t); });
        }
        s->runner->run(steps, 0);
        if (sim_time_out) *sim_time_out = s->runner->sim_time();
    });
}

int gpgmp_solver_run_continue(gpgmp_solver *s, int32_t steps, int32_t graph_chunk, float *sim_time_out) {
    return guarded([&] {
        GPGMP_REQUIRE(s, GPGMP_E_INVALID, "null argument");
        if (s->det) {
            s->det->run(steps);
            if (sim_time_out) *sim_time_out = s->det->sim_time();
            return;
        }
        GPGMP_REQUIRE(s->runner != nullptr, GPGMP_E_STATE, "no run to continue");
        s->runner->run(steps, graph_chunk);
        if (sim_time_out) *sim_time_out = s->runner->sim_time();
    });
}

int gpgmp_solver_skip(gpgmp_solver *s, int32_t steps, float *sim_time_out) {
    return guarded([&] {
        (void)stoch(s);
        GPGMP_REQUIRE(s, GPGMP_E_INVALID, "null argument");
        GPGMP_REQUIRE(s->runner != nullptr, GPGMP_E_STATE, "no run to continue");
        s->runner->skip(steps);
        if (sim_time_out) *sim_time_out = s->runner->sim_time();
    });
}

int gpgmp_solver_prepare_run(gpgmp_solver *s, int32_t steps, int32_t graph_chunk) {
    return guarded([&] {
        (void)stoch(s);
        GPGMP_REQUIRE(s, GPGMP_E_INVALID, "null argument");
        GPGMP_REQUIRE(s->runner != nullptr, GPGMP_E_STATE, "no run to continue");
        s->runner->prepare(steps, graph_chunk);
    });
}

int gpgmp_solver_set_output(gpgmp_solver *s, float interval, float max_time, int32_t max_dumps) {
    return guarded([&] {
        (void)stoch(s);
        GPGMP_REQUIRE(s && max_dumps >= 0, GPGMP_E_INVALID, "bad argument");
        GPGMP_REQUIRE(!std::isnan(interval) && !std::isnan(max_time), GPGMP_E_INVALID, "NaN interval or max_time");
        s->out_interval = interval > 0.f ? interval : __builtin_inff();
        s->max_time = max_time > 0.f ? max_time : __builtin_inff();
        s->max_dumps = max_dumps;
        GPGMP_HIP_CHECK(hipStreamSynchronize(s->eng->stream()));
        s->dumps.alloc((size_t)max_dumps * s->eng->state_elems());
        s->dump_times.clear();
    });
}

int gpgmp_solver_get_output(gpgmp_solver *s, int32_t *counts, float *times, int32_t *n_stored, int32_t *n_total,
                            int64_t *iterations) {
    return guarded([&] {
        (void)stoch(s);
        GPGMP_REQUIRE(s, GPGMP_E_INVALID, "null argument");
        const int32_t total = (int32_t)s->dump_times.size();
        const int32_t stored = total < s->max_dumps ? total : s->max_dumps;
        if (counts && stored)
            GPGMP_HIP_CHECK(hipMemcpyAsync(counts, s->dumps.get(), (size_t)stored * s->eng->state_elems() * 4,
                                           hipMemcpyDeviceToHost, s->eng->stream()));
        GPGMP_HIP_CHECK(hipStreamSynchronize(s->eng->stream()));
        if (times) std::copy(s->dump_times.begin(), s->dump_times.begin() + stored, times);
        if (n_stored) *n_stored = stored;
        if (n_total) *n_total = total;
        if (iterations) *iterations = s->runner ? s->runner->iterations() : 0;
    });
}

int gpgmp_solver_set_predict_busy(gpgmp_solver *s, int32_t on) {
    return guarded([&] {
        (void)stoch(s);
        GPGMP_REQUIRE(s, GPGMP_E_INVALID, "null argument");
        GPGMP_HIP_CHECK(hipStreamSynchronize(s->eng->stream()));
        s->eng->set_predict(on != 0);
    });
}

int gpgmp_solver_set_sweep_spread(gpgmp_solver *s, int32_t on) {
    return guarded([&] {
        (void)stoch(s);
        GPGMP_REQUIRE(s, GPGMP_E_INVALID, "null argument");
        GPGMP_HIP_CHECK(hipStreamSynchronize(s->eng->stream()));
        s->eng->set_spread(on != 0);
    });
}

int gpgmp_solver_set_sweep_wave_cells(gpgmp_solver *s, int32_t cells) {
    return guarded([&] {
        (void)stoch(s);
        GPGMP_REQUIRE(s, GPGMP_E_INVALID, "null argument");
        GPGMP_HIP_CHECK(hipStreamSynchronize(s->eng->stream()));
        s->eng->set_sweep_wave_cells(cells);
    });
}

int gpgmp_solver_set_launch(gpgmp_solver *s, int32_t block, int32_t cpt) {
    return guarded([&] {
        (void)stoch(s);
        GPGMP_REQUIRE(s, GPGMP_E_INVALID, "null argument");
        s->eng->set_launch(block, cpt);
    });
}

int gpgmp_solver_set_lanes_per_cell(gpgmp_solver *s, int32_t ssa_lanes, int32_t sweep_lanes) {
    return guarded([&] {
        (void)stoch(s);
        GPGMP_REQUIRE(s, GPGMP_E_INVALID, "null argument");
        s->eng->set_launch(s->eng->block(), s->eng->cpt(), ssa_lanes, sweep_lanes);
    });
}

int gpgmp_solver_get_launch(gpgmp_solver *s, int32_t *block, int32_t *cpt, int32_t *ssa_lanes, int32_t *sweep_lanes) {
    return guarded([&] {
        (void)stoch(s);
        GPGMP_REQUIRE(s && block && cpt && ssa_lanes && sweep_lanes, GPGMP_E_INVALID, "null argument");
        *block = s->eng->block();
        *cpt = s->eng->cpt();
        *ssa_lanes = s->eng->grp();
        *sweep_lanes = s->eng->grp_sweep();
    });
}

int gpgmp_solver_set_event_budget(gpgmp_solver *s, uint32_t budget, int32_t below_active) {
    return guarded([&] {
        (void)stoch(s);
        GPGMP_REQUIRE(s && budget >= 1 && below_active >= 0 && below_active <= 65, GPGMP_E_INVALID,
                      "bad argument");
        GPGMP_HIP_CHECK(hipStreamSynchronize(s->stream.s));
        s->eng->set_hot_budget(budget, below_active);
    });
}

int gpgmp_solver_set_hot_workers(gpgmp_solver *s, float lane_events, int32_t lane_waves, int32_t wave_waves) {
    return guarded([&] {
        (void)stoch(s);
        GPGMP_REQUIRE(s, GPGMP_E_INVALID, "null argument");
        GPGMP_HIP_CHECK(hipStreamSynchronize(s->stream.s));
        s->eng->set_hot_workers(lane_events, lane_waves, wave_waves);
    });
}

int gpgmp_solver_kernel_times(gpgmp_solver *s, int32_t on, double *ms, int64_t *n, double *bytes) {
    return guarded([&] {
        (void)stoch(s);
        GPGMP_REQUIRE(s, GPGMP_E_INVALID, "null argument");
        GPGMP_REQUIRE(s->runner != nullptr, GPGMP_E_STATE, "no run");
        if (ms && n) s->runner->kernel_times(ms, n, bytes);
        s->runner->set_timing(on);
    });
}

int gpgmp_solver_synchronize(gpgmp_solver *s) {
    return guarded([&] {
        GPGMP_REQUIRE(s, GPGMP_E_INVALID, "null argument");
        GPGMP_HIP_CHECK(hipStreamSynchronize(s->stream.s));
    });
}

// ---- batch ------------------------------------------------------------------------------
int gpgmp_batch_create(const gpgmp_model *tmpl, int32_t n, const int32_t *widths, const int32_t *heights,
                       const uint32_t *seed0s, void *stream, gpgmp_batch **out) {
    return guarded([&] {
        GPGMP_REQUIRE(tmpl && widths && heights && seed0s && out, GPGMP_E_INVALID, "null argument");
        GPGMP_REQUIRE(n >= 1, GPGMP_E_INVALID, "n must be positive");
        *out = nullptr;
        GPGMP_REQUIRE(tmpl->m.desc.solver_type == GPGMP_STOCHASTIC_HOMOGENEOUS, GPGMP_E_UNSUPPORTED,
                      "batches run the stochastic homogeneous solver");
        std::unique_ptr<gpgmp_batch> b(new gpgmp_batch());
        b->m = tmpl->m;
        b->m.desc.reactions = b->m.reactions.empty() ? nullptr : b->m.reactions.data();
        b->stream.init(stream);
        std::vector<int> W(widths, widths + n), H(heights, heights + n);
        int64_t total = 0;
        for (int r = 0; r < n; ++r) total += (int64_t)W[r] * H[r];
        GPGMP_REQUIRE(total * b->m.desc.num_species <= ((int64_t)1 << 31), GPGMP_E_CAPACITY,
                      "batch exceeds 2^31 counts");
        b->eng.reset(new GmpEngine(b->m, W, H, b->stream.s));
        const gpgmp_model_desc &d = b->m.desc;
        const float h = d.physical_length / (float)d.width;
        const int S = d.num_species;
        b->dtx.assign(S, 0.f);
        for (int i = 0; i < S; ++i)
            b->dtx[i] = (d.diffusion[i] > 0) ? homogeneous_dt(d.p0, h, d.diffusion[i]) : d.max_time;
        b->dty = b->dtx;
        b->seed0s.assign(seed0s, seed0s + n);
        *out = b.release();
    });
}

int gpgmp_batch_destroy(gpgmp_batch *b) {
    if (b) b->runner.reset();
    delete b;
    return GPGMP_OK;
}

int gpgmp_batch_set_state(gpgmp_batch *b, int32_t r, const int32_t *host) {
    return guarded([&] {
        GPGMP_REQUIRE(b && host, GPGMP_E_INVALID, "null argument");
        b->eng->set_state(r, host);
    });
}

int gpgmp_batch_get_state(gpgmp_batch *b, int32_t r, int32_t *host) {
    return guarded([&] {
        GPGMP_REQUIRE(b && host, GPGMP_E_INVALID, "null argument");
        b->eng->get_state(r, host);
    });
}

int gpgmp_batch_set_launch(gpgmp_batch *b, int32_t block, int32_t cpt) {
    return guarded([&] {
        GPGMP_REQUIRE(b, GPGMP_E_INVALID, "null argument");
        b->eng->set_launch(block, cpt);
    });
}

int gpgmp_batch_set_lanes_per_cell(gpgmp_batch *b, int32_t ssa_lanes, int32_t sweep_lanes) {
    return guarded([&] {
        GPGMP_REQUIRE(b, GPGMP_E_INVALID, "null argument");
        b->eng->set_launch(b->eng->block(), b->eng->cpt(), ssa_lanes, sweep_lanes);
    });
}

int gpgmp_batch_get_launch(gpgmp_batch *b, int32_t *block, int32_t *cpt, int32_t *ssa_lanes, int32_t *sweep_lanes) {
    return guarded([&] {
        GPGMP_REQUIRE(b && block && cpt && ssa_lanes && sweep_lanes, GPGMP_E_INVALID, "null argument");
        *block = b->eng->block();
        *cpt = b->eng->cpt();
        *ssa_lanes = b->eng->grp();
        *sweep_lanes = b->eng->grp_sweep();
    });
}

int gpgmp_batch_set_event_budget(gpgmp_batch *b, uint32_t budget, int32_t below_active) {
    return guarded([&] {
        GPGMP_REQUIRE(b && budget >= 1 && below_active >= 0 && below_active <= 65, GPGMP_E_INVALID,
                      "bad argument");
        GPGMP_HIP_CHECK(hipStreamSynchronize(b->stream.s));
        b->eng->set_hot_budget(budget, below_active);
    });
}

int gpgmp_batch_run(gpgmp_batch *b, int32_t steps, int32_t graph_chunk) {
    return guarded([&] {
        GPGMP_REQUIRE(b, GPGMP_E_INVALID, "null argument");
        if (!b->runner) b->runner.reset(new GmpRunner(*b->eng, b->dtx, b->dty, b->seed0s));
        b->runner->run(steps, graph_chunk);
    });
}

int gpgmp_batch_prepare_run(gpgmp_batch *b, int32_t steps, int32_t graph_chunk) {
    return guarded([&] {
        GPGMP_REQUIRE(b, GPGMP_E_INVALID, "null argument");
        if (!b->runner) b->runner.reset(new GmpRunner(*b->eng, b->dtx, b->dty, b->seed0s));
        b->runner->prepare(steps, graph_chunk);
    });
}

int gpgmp_batch_reset(gpgmp_batch *b) {
    return guarded([&] {
        GPGMP_REQUIRE(b, GPGMP_E_INVALID, "null argument");
        b->runner.reset();
    });
}

// Project-written autotuner: every candidate (block, cells per thread, lanes per cell)
// runs the same
// trial steps from the same saved state with a fresh schedule; the fastest is kept and
// the state and schedule are restored.  Results never depend on the shape (RNG keys
// carry no tile or batch offset), only the speed does.
int gpgmp_batch_autotune(gpgmp_batch *b, int32_t trial_steps, int32_t *block_out, int32_t *cpt_out) {
    return guarded([&] {
        GPGMP_REQUIRE(b && trial_steps > 0, GPGMP_E_INVALID, "bad argument");
        GmpEngine &e = *b->eng;
        hipStream_t st = b->stream.s;
        const size_t n = e.state_elems();
        DevBuf<int32_t> saved(n);
        GPGMP_HIP_CHECK(hipMemcpyAsync(saved.get(), e.state(), n * sizeof(int32_t), hipMemcpyDeviceToDevice, st));
        GPGMP_HIP_CHECK(hipStreamSynchronize(st));
        static const int blocks[] = {64, 128, 256};
        static const int cpts[] = {1, 2, 4};
        static const int grps[][2] = {{1, 0}, {1, 1}, {1, 4}, {4, 4}, {4, 0}};
        double best = 1e300;
        int bb = e.block(), bc = e.cpt(), bg = e.grp(), bgs = e.grp_sweep();
        hipEvent_t t0, t1;
        GPGMP_HIP_CHECK(hipEventCreate(&t0));
        GPGMP_HIP_CHECK(hipEventCreate(&t1));
        for (const auto &g : grps)
        for (int blk : blocks)
            for (int c : cpts) {
                e.set_launch(blk, c, g[0], g[1]);
                GPGMP_HIP_CHECK(hipMemcpyAsync(e.state(), saved.get(), n * sizeof(int32_t),
                                               hipMemcpyDeviceToDevice, st));
                GmpRunner trial(e, b->dtx, b->dty, b->seed0s);
                GPGMP_HIP_CHECK(hipEventRecord(t0, st));
                trial.run(trial_steps, 0);
                GPGMP_HIP_CHECK(hipEventRecord(t1, st));
                GPGMP_HIP_CHECK(hipEventSynchronize(t1));
                float ms = 0.f;
                GPGMP_HIP_CHECK(hipEventElapsedTime(&ms, t0, t1));
                if (ms < best) {
                    best = ms;
                    bb = blk;
                    bc = c;
                    bg = g[0];
                    bgs = g[1];
                }
            }
        (void)hipEventDestroy(t0);
        (void)hipEventDestroy(t1);
        e.set_launch(bb, bc, bg, bgs);
        GPGMP_HIP_CHECK(hipMemcpyAsync(e.state(), saved.get(), n * sizeof(int32_t), hipMemcpyDeviceToDevice, st));
        GPGMP_HIP_CHECK(hipStreamSynchronize(st));
        if (block_out) *block_out = bb;
        if (cpt_out) *cpt_out = bc;
    });
}

int gpgmp_batch_kernel_times(gpgmp_batch *b, int32_t on, double *ms, int64_t *n, double *bytes) {
    return guarded([&] {
        GPGMP_REQUIRE(b, GPGMP_E_INVALID, "null argument");
        if (!b->runner) b->runner.reset(new GmpRunner(*b->eng, b->dtx, b->dty, b->seed0s));
        if (ms && n) b->runner->kernel_times(ms, n, bytes);
        b->runner->set_timing(on);
    });
}

int gpgmp_batch_total_cells(gpgmp_batch *b, int64_t *cells) {
    return guarded([&] {
        GPGMP_REQUIRE(b && cells, GPGMP_E_INVALID, "null argument");
        *cells = b->eng->total_cells();
    });
}

int gpgmp_min_timestep_device(const float *D, const float *r, int32_t n, float h, float *out_host, void *stream) {
    return guarded([&] {
        GPGMP_REQUIRE(D && r && out_host && n >= 1 && h > 0.f, GPGMP_E_INVALID, "bad argument");
        hipStream_t s = (hipStream_t)stream;
        DevBuf<float> scratch(1025);
        GPGMP_HIP_CHECK(launch_min_dt(D, r, n, h, scratch.get(), scratch.get() + 1024, s));
        GPGMP_HIP_CHECK(hipMemcpyAsync(out_host, scratch.get() + 1024, sizeof(float), hipMemcpyDeviceToHost, s));
        GPGMP_HIP_CHECK(hipStreamSynchronize(s));
    });
}

int gpgmp_batch_synchronize(gpgmp_batch *b) {
    return guarded([&] {
        GPGMP_REQUIRE(b, GPGMP_E_INVALID, "null argument");
        GPGMP_HIP_CHECK(hipStreamSynchronize(b->stream.s));
    });
}

}  // extern "C"
