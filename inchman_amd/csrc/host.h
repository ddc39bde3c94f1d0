// Host-side C++ of the GPGMP HIP library: model compile, schedule, solver classes.
#pragma once

#include <hip/hip_runtime.h>

#include <array>
#include <cstdint>
#include <random>
#include <stdexcept>
#include <string>
#include <functional>
#include <map>
#include <vector>

#include "../../include/gpgmp_hip.h"
#include "stochastic_launch.h"
#include "stochastic_types.h"

namespace gpgmp {

// Error carrying a C-ABI status code (the reference's gpgmp::Error, DiffusionModel.h:68-105,
// becomes a return code at the boundary).
struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

#define GPGMP_HIP_CHECK(expr)                                                                  \
    do {                                                                                       \
        hipError_t e_ = (expr);                                                                \
        if (e_ != hipSuccess)                                                                  \
            throw ::gpgmp::Error(GPGMP_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

#define GPGMP_REQUIRE(cond, code, msg)                  \
    do {                                                \
        if (!(cond)) throw ::gpgmp::Error((code), (msg)); \
    } while (0)

// Owning device allocation.
template <typename T>
class DevBuf {
  public:
    DevBuf() = default;
    explicit DevBuf(size_t n) { alloc(n); }
    ~DevBuf() { release(); }
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    DevBuf(DevBuf &&o) noexcept : p_(o.p_), n_(o.n_) { o.p_ = nullptr; o.n_ = 0; }
    DevBuf &operator=(DevBuf &&o) noexcept {
        if (this != &o) { release(); p_ = o.p_; n_ = o.n_; o.p_ = nullptr; o.n_ = 0; }
        return *this;
    }
    void alloc(size_t n) {
        release();
        if (n == 0) return;
        GPGMP_HIP_CHECK(hipMalloc((void **)&p_, n * sizeof(T)));
        n_ = n;
    }
    void release() {
        if (p_) (void)hipFree(p_);
        p_ = nullptr;
        n_ = 0;
    }
    T *get() const { return p_; }
    size_t size() const { return n_; }

  private:
    T *p_ = nullptr;
    size_t n_ = 0;
};

// Compiled model: the reaction table and the constants the reference bakes into the
// kernel header (Solver.cpp:449-567).
struct Model {
    gpgmp_model_desc desc{};
    std::vector<gpgmp_reaction_desc> reactions;
    std::vector<float> Dx, Dy, rx, ry;  // inhomogeneous fields [S][H][W]
    RxTable table{};
    HomProbs probs{};
    float cell_width = 0.f;   // PhysicalCellWidth as printed into the header
    float cell_height = 0.f;  // PhysicalCellHeight
    // compartments of the localized reactions: {reaction, x0, y0, x1, y1} with inclusive
    // cell bounds (Reaction::addCompartment, Compartment::isInCompartment)
    std::vector<std::array<int32_t, 5>> compartments;
    // names the custom laws use: Species::id per species index, the user parameters
    // (DiffusionModel::setParameter, a Real each), and per reaction its custom law
    // (Reaction::setCustomLaw; empty: the law generated from the stoichiometry)
    std::vector<std::string> species_ids;
    std::map<std::string, float> params;
    std::vector<std::string> custom_law;
    bool inhomogeneous() const {
        return desc.solver_type == GPGMP_STOCHASTIC_INHOMOGENEOUS ||
               desc.solver_type == GPGMP_STOCHASTIC_INHOMOGENEOUS_FPE ||
               desc.solver_type == GPGMP_DETERMINISTIC_INHOMOGENEOUS_RK4 ||
               desc.solver_type == GPGMP_DETERMINISTIC_INHOMOGENEOUS_AQSS;
    }
    // stochastic_inhomogeneous_fpe: the inhomogeneous solver with the Fokker-Planck
    // discretisation's time step and transition rates
    bool fpe() const { return desc.solver_type == GPGMP_STOCHASTIC_INHOMOGENEOUS_FPE; }
    bool deterministic() const {
        return desc.solver_type >= GPGMP_DETERMINISTIC_HOMOGENEOUS_RK4 &&
               desc.solver_type <= GPGMP_DETERMINISTIC_INHOMOGENEOUS_AQSS;
    }
};

Model compile_model(const gpgmp_model_desc &d);
// Applies the model's custom laws to a copy of its reaction table (law.cpp): mass-action
// shaped laws become table entries, the others programs; a model with any custom law also
// gets the deterministic right-hand sides as programs.  Returns the program words (to be
// uploaded; T.prog is set by the caller) -- empty when no reaction needs a program.
std::vector<uint32_t> build_law_programs(const Model &m, RxTable &T);
inline bool has_custom_laws(const Model &m) {
    for (const auto &c : m.custom_law)
        if (!c.empty()) return true;
    return false;
}
// The species id of index s (Species::id), "S<s>" when none was given.
std::string species_id(const Model &m, int s);

// Active-reaction bits per cell of a W x H grid (bit i: reaction i not localized, or the
// cell inside one of its compartments; Reaction::reactionMask, Reaction.cpp:216-235).
void reaction_act(const Model &m, int W, int H, uint16_t *out);
float header_float(float v);
float homogeneous_dt(float p0, float h, float D);
float inhomogeneous_cell_dt(float D, float r, float h);
float fpe_cell_dt(float D, float r, float h, float stay);

// boost::mt19937(seed0) + uniform_int<>(0, INT_MAX), seed_n = draw_n + n
// (DiffusionModel.cpp:1467-1484); the draw is eng() >> 1 for this range.
class SeedStream {
  public:
    explicit SeedStream(uint32_t seed0) : eng_(seed0) {}
    int32_t next() {
        const uint32_t draw = (uint32_t)eng_() >> 1;
        return (int32_t)(draw + (uint32_t)n_++);
    }

  private:
    std::mt19937 eng_;
    int32_t n_ = 0;
};

// One host-loop iteration as decided by DiffusionModel::oclSolve
// (DiffusionModel.cpp:1479-1603) for a stochastic solver.  Seeds are offsets from the
// iteration's seed_n: the SSA uses seed_n and increments it (Solver.cpp:239);
// homogeneous diffusion reuses the incremented seed for every species
// (HomogeneousSolver.cpp:136); each inhomogeneous sweep takes the current seed and
// increments it (InhomogeneousSolver.cpp:278,370).
struct StepDecision {
    float deltat = 0.f;
    uint32_t due_x = 0, due_y = 0;  // homogeneous: due_x holds the diffused species
    int32_t off_x[kMaxSpecies] = {};
    int32_t off_y[kMaxSpecies] = {};
    int32_t off_end = 0;            // seed offset after the iteration
    bool dump = false;              // an output event that writes the state (at dump_time)
    float dump_time = 0.f;
};

// The multi-rate schedule: which species diffuse when.  It depends only on the time
// steps, so it is computed on the host ahead of the launches.
class Schedule {
  public:
    Schedule(bool inhomogeneous, int S, bool has_reactions, const std::vector<float> &dtx,
             const std::vector<float> &dty);
    StepDecision next();
    float sim_time() const { return sim_time_; }
    std::vector<float> next_diffusion_time() const;
    // Output events every `interval` of simulation time (DiffusionModel.cpp:1519-1572) and
    // the loop's stop at max_time (:1479-1480); infinity disables either.
    void set_output(float interval, float max_time) {
        out_interval_ = interval;
        max_time_ = max_time;
    }
    bool finished() const { return !(sim_time_ < max_time_); }

  private:
    bool inh_;
    int S_;
    bool rx_;
    std::vector<float> dtx_, dty_, nx_, ny_;
    float sim_time_ = 0.f;
    float out_interval_ = __builtin_inff(), out_last_ = 0.f, max_time_ = __builtin_inff();
};

// Device engine of the stochastic solvers: a set of realizations of one network.
class GmpEngine {
  public:
    GmpEngine(const Model &m, const std::vector<int> &W, const std::vector<int> &H, hipStream_t stream);

    int num_real() const { return (int)reals_.size(); }
    int S() const { return table_.S; }
    const RealInfo &real(int r) const { return reals_[r]; }
    int64_t total_cells() const { return total_cells_; }
    int32_t *state() const { return state_.get(); }
    size_t state_elems() const { return state_.size(); }
    hipStream_t stream() const { return stream_; }
    void set_stream(hipStream_t s) { stream_ = s; }

    // threads per block, cells per thread, lanes per cell (1 or 4) of the launches
    // with an SSA (modes 0, 1) and of the sweep-only launches (mode 2)
    void set_launch(int block, int cpt, int grp_ssa, int grp_sweep);
    void set_launch(int block, int cpt) { set_launch(block, cpt, grp_, grp_sweep_); }
    // SSA events a lane runs before its cell moves to the hot-cell kernel
    void set_hot_budget(uint32_t b, int below_active) {
        hot_.budget = b;
        hot_.below_active = below_active;
        ++shape_version_;
    }
    // changes with every launch-shape or budget change (captured graphs bake them in)
    uint64_t shape_version() const { return shape_version_; }
    uint32_t hot_budget() const { return hot_.budget; }
    int block() const { return block_; }
    int cpt() const { return cpt_; }
    int grp() const { return grp_; }
    int grp_sweep() const { return grp_sweep_; }
    // inhomogeneous with flattened sweeps: the SSA and the X sweeps in separate launches
    // (the sweeps then run with 8 or 16 cells per wave, sweep_wave_cells())
    bool split_x() const { return inh_ && grp_sweep_ == 0; }
    // split steps continue the deferred cells inside the X-sweep launch
    bool fused_hot() const { return split_x(); }
    // split steps with one SSA lane per cell: the SSA launch runs the previous step's busy
    // cells in worker waves
    bool predict() const { return predict_ && fused_hot() && grp_ == 1; }
    void set_predict(bool on) {
        predict_ = on;
        ++shape_version_;
    }
    void reset_hot();
    // Workers of the deferred cells in the X-sweep launch of a split step: a cell whose
    // expected remaining events are at most lane_events is resumed by one lane (lane_waves
    // worker waves, 64 cells each per round), the others by one wave each (wave_waves
    // worker waves); lane_events <= 0 sends every cell to the wave workers.
    void set_hot_workers(float lane_events, int lane_waves, int wave_waves) {
        GPGMP_REQUIRE(lane_waves >= 8 && wave_waves >= 8 && lane_waves <= 65536 && wave_waves <= 65536,
                      GPGMP_E_INVALID, "worker waves must be in [8, 65536]");
        GPGMP_REQUIRE(!(lane_events != lane_events), GPGMP_E_INVALID, "lane_events is NaN");
        lane_events_ = lane_events;
        lane_waves_ = lane_waves;
        hot_waves_ = wave_waves;
        ++shape_version_;
    }
    float lane_events() const { return lane_events_; }
    int lane_waves() const { return lane_waves_; }
    int hot_waves() const { return hot_waves_; }
    // flattened sweeps spread each wave's n cells over the grid (stride A/n)
    bool spread() const { return spread_; }
    void set_spread(bool on) {
        spread_ = on;
        ++shape_version_;
    }
    // cells per wave of the flattened launches without an SSA (8, the default: twice the
    // waves of 16, which evens out the X-sweep launch beside the busy cells)
    int sweep_wave_cells() const { return sweep_wave_cells_; }
    void set_sweep_wave_cells(int n) {
        GPGMP_REQUIRE(n == 8 || n == 16, GPGMP_E_INVALID, "cells per sweep wave must be 8 or 16");
        sweep_wave_cells_ = n;
        ++shape_version_;
    }

    void set_state(int r, const int32_t *host);
    void get_state(int r, int32_t *host);

    // inhomogeneous probability table of realization r, species s, axis; D, r host [A]
    void set_probs(int r, int s, int axis, const float *D, const float *rr, float dt, float dx);

    // one launch (all device pointers).  mode 0 homogeneous, 1 inhomogeneous X kernel
    // (gather of the Y outflow, SSA, X sweeps), 2 Y kernel, 3 X kernel with the sweep
    // layout and no SSA (the X sweeps of a split step).  fuse (modes 1 and 3 of a split
    // step, plans adjacent): the SSA launch flags its deferred cells instead of launching
    // the hot-cell kernel, and the X-sweep launch continues them.
    void launch(int mode, const LaunchPlan *plan, const int32_t *seeds_cur, const int32_t *seeds_prev,
                const int32_t *lost_in, int32_t *lost_out, hipEvent_t ev_start = nullptr,
                hipEvent_t ev_stop = nullptr, hipEvent_t ev_hot_start = nullptr,
                hipEvent_t ev_hot_stop = nullptr, bool fuse = false);

    int32_t *lost(int i) const { return lost_[i].get(); }

    RxTable table_{};
    HomProbs probs_{};
    int4 bm_{};
    bool inh_ = false;
    bool fpe_ = false;

  private:
    void build_tiles();
    std::vector<RealInfo> reals_;
    DevBuf<uint16_t> act_;  // active-reaction bits per cell (localized reactions only)
    DevBuf<uint32_t> prog_;  // kinetic-law programs (custom laws only)
    DevBuf<int32_t> state_;
    DevBuf<int32_t> lost_[2];
    DevBuf<int4> prob_[2];
    DevBuf<RealInfo> reals_dev_;
    DevBuf<Tile> tiles_dev_[3];  // tiles for 1, 4 and 8 lanes per cell
    DevBuf<HotEntry> hot_e_;
    DevBuf<HotEntry> hot_lane_;
    DevBuf<uint32_t> hot_count_;
    DevBuf<uint8_t> hot_flag_;
    DevBuf<uint8_t> hot_pred_;
    HotList hot_{};
    int ntiles_[3] = {0, 0, 0};
    int block_ = 256, cpt_ = 1, grp_ = 1, grp_sweep_ = 1;
    bool spread_ = true;
    int sweep_wave_cells_ = 8;
    bool predict_ = false;
    float lane_events_ = 16.f;
    int lane_waves_ = 1024;
    int hot_waves_ = 1024;
    uint64_t shape_version_ = 0;
    int64_t total_cells_ = 0;
    hipStream_t stream_;
};

// Runs host-loop iterations of every realization of an engine (one shared schedule,
// one seed stream per realization): uploads chunks of launch plans and seeds, issues
// the launches (optionally as a replayed HIP graph per chunk) and the final gather.
class GmpRunner {
  public:
    GmpRunner(GmpEngine &eng, const std::vector<float> &dtx, const std::vector<float> &dty,
              const std::vector<uint32_t> &seed0s);
    ~GmpRunner();
    void run(int steps, int graph_chunk);
    // Captures the HIP graphs run(steps, graph_chunk) would replay from the current state of
    // the loop, without running anything (the loop's state is unchanged).
    void prepare(int steps, int graph_chunk);
    // Advances the loop by `steps` iterations without running them (nothing pending after).
    void skip(int steps);
    float sim_time() const { return sched_.sim_time(); }
    // Output events and the max_time stop of the reference loop; `dump` is called, after
    // the launches of each output iteration that writes (the state is then fully gathered),
    // with the simulation time of the dump, to enqueue a copy of the state on the stream.
    void set_output(float interval, float max_time, std::function<void(float)> dump) {
        sched_.set_output(interval, max_time);
        dump_ = std::move(dump);
    }
    int64_t iterations() const { return iterations_; }
    // Per-launch HIP-event timing of the eager (non-graph) launches, by kernel mode
    // (0 homogeneous, 1 SSA (+ X sweeps unless split), 2 Y sweeps, 3 hot cells in their own
    // launch, 4 X sweeps + hot cells of a split step); accumulated over run() calls.
    // every = 1: time every launch; every = k > 1: time the launches of kTimingWindow - 1
    // consecutive steps out of every k steps, after one step whose events are not counted
    // (dispatch-tied events on launches that directly follow untimed ones read skewed
    // times, by up to 60% of a launch here; from the second timed step on they agree with
    // the profiler's kernel trace) (0: none)
    void set_timing(int every) {
        timing_ = every > 0;
        every_ = every > 0 ? every : 1;
    }
    void kernel_times(double *ms, int64_t *n, double *bytes) const {
        for (int i = 0; i < 5; ++i) {
            ms[i] = kt_ms_[i];
            n[i] = kt_n_[i];
            if (bytes) bytes[i] = kt_bytes_[i];
        }
    }
    std::vector<float> next_diffusion_time() const { return sched_.next_diffusion_time(); }

  private:
    struct Chunk;
    void issue(Chunk &c, int nlaunch_steps, bool use_graph, bool dry);
    void run_impl(int steps, int graph_chunk, bool dry);
    static constexpr int kMinGraphSteps = 2;
    GmpEngine &eng_;
    Schedule sched_;
    std::vector<SeedStream> seeds_;
    StepDecision prev_{};
    bool has_prev_ = false;
    std::vector<int32_t> prev_seed_;
    int parity_ = 0;
    std::vector<Chunk *> chunks_;
    std::function<void(float)> dump_;
    int64_t iterations_ = 0;
    int64_t plan_step_ = 0;  // steps planned so far (hot-list parity)
    void timed_launch(int mode, const LaunchPlan *plan, const LaunchPlan &hplan, const int32_t *scur,
                      const int32_t *sprev, const int32_t *lin, int32_t *lout, int how);
    void collect_times();
    bool timing_ = false;
    int every_ = 1;
    static constexpr int kTimingWindow = 5;
    int64_t step_no_ = 0;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> ev_pool_;
    std::vector<int> ev_mode_;
    size_t ev_used_ = 0;
    // index 3: the hot-cell kernel that follows every launch with an SSA
    double kt_ms_[5] = {0, 0, 0, 0, 0};
    int64_t kt_n_[5] = {0, 0, 0, 0, 0};
    double kt_bytes_[5] = {0, 0, 0, 0, 0};  // algorithmic HBM bytes of the timed launches
    int chunk_cap_ = 0;
    int graph_steps_ = 0;
};

}  // namespace gpgmp

// Opaque handles of the C ABI that more than one translation unit needs.
struct gpgmp_model {
    gpgmp::Model m;
};

// Sets the calling thread's gpgmp_last_error() message.
void gpgmp_set_last_error(const char *msg);
