// Model compile, host-loop schedule, device engine and runner of the stochastic solvers.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "host.h"

namespace gpgmp {

// Header constants are printed "%.6e" (ss << std::scientific) and re-read as float
// literals by the OpenCL compiler (Solver.cpp:494-505).
float header_float(float v) {
    char buf[64];
    std::snprintf(buf, sizeof buf, "%.6e", (double)v);
    return std::strtof(buf, nullptr);
}

// p0/((Real)2.*GMP_DIMENSIONALITY)*h*h/D evaluated in float (HomogeneousSolver.cpp:89).
float homogeneous_dt(float p0, float h, float D) {
    const float six = (float)2. * 3;
    float a = p0 / six;
    a = a * h;
    a = a * h;
    return a / D;
}

// Admissible per-cell step of the stochastic inhomogeneous solver
// (StochasticInhomogeneousDriftDiffusion.cl:130-151), in double on the host.
float inhomogeneous_cell_dt(float Df, float rf, float hf) {
    const double D = Df, r = rf, h = hf;
    double ts;
    if (r == 0.0) {
        ts = h * h / (2.0 * D) * (1.0 - 2.0 / 3.0);
    } else if (D == 0.0) {
        ts = h / std::fabs(r);
    } else {
        const double eps = h * r / (2.0 * D);
        double temp = 1.0 + (-1.0 / (std::tanh(eps) * eps) + 1.0 / (std::sinh(eps) * std::sinh(eps)));
        if (temp <= 0.0) temp = 1.0 / 3.0;
        ts = h / r * std::tanh(eps) * temp;
    }
    return (float)ts;
}

// Per-cell step of the stochastic_inhomogeneous_fpe solver
// (StochasticInhomogeneousDriftDiffusion.cl:123-126, 268-271):
// (Real)StayProbability/6./(diffHere/(deltax*deltax) + fabs(rx/(2.*deltax))) -- a float
// quotient D/h^2 plus a double |r|/(2h), the header float StayProbability over 6. in double.
float fpe_cell_dt(float Df, float rf, float hf, float stay) {
    const float dD = Df / (hf * hf);
    const double den = (double)dD + std::fabs((double)rf / (2. * (double)hf));
    return (float)((double)stay / 6. / den);
}

void reaction_act(const Model &m, int W, int H, uint16_t *out) {
    uint16_t base = 0;
    for (int i = 0; i < m.table.R; ++i)
        if (!((m.table.localized >> i) & 1u)) base |= (uint16_t)(1u << i);
    for (int64_t c = 0; c < (int64_t)W * H; ++c) out[c] = base;
    for (const auto &cp : m.compartments) {
        const int x0 = std::max(cp[1], 0), y0 = std::max(cp[2], 0);
        const int x1 = std::min(cp[3], W - 1), y1 = std::min(cp[4], H - 1);
        for (int y = y0; y <= y1; ++y)
            for (int x = x0; x <= x1; ++x) out[(int64_t)y * W + x] |= (uint16_t)(1u << cp[0]);
    }
}

Model compile_model(const gpgmp_model_desc &d) {
    Model m;
    m.desc = d;
    const int S = d.num_species, R = d.num_reactions;
    GPGMP_REQUIRE(S >= 1 && S <= kMaxSpecies, GPGMP_E_INVALID, "num_species must be in [1, 8]");
    GPGMP_REQUIRE(R >= 0 && R <= kMaxReactions, GPGMP_E_INVALID, "num_reactions must be in [0, 16]");
    GPGMP_REQUIRE(R == 0 || d.reactions != nullptr, GPGMP_E_INVALID, "reactions is NULL");
    GPGMP_REQUIRE(d.width >= 16 && d.height >= 16 && d.width % 16 == 0 && d.height % 16 == 0,
                  GPGMP_E_INVALID, "grid sides must be positive multiples of 16");
    GPGMP_REQUIRE((int64_t)d.width * d.height <= (int64_t)1 << 28, GPGMP_E_CAPACITY, "grid too large");
    GPGMP_REQUIRE(d.physical_length > 0.f && std::isfinite(d.physical_length), GPGMP_E_INVALID,
                  "physical_length must be positive");
    GPGMP_REQUIRE(d.solver_type >= 0 && d.solver_type <= GPGMP_STOCHASTIC_INHOMOGENEOUS_FPE, GPGMP_E_INVALID,
                  "unknown solver type");
    for (int b = 0; b < 4; ++b)
        GPGMP_REQUIRE(d.boundary_mask[b] == 0 || d.boundary_mask[b] == 1, GPGMP_E_INVALID,
                      "boundary masks must be 0 or 1");
    const int A = d.width * d.height;
    if (R > 0) m.reactions.assign(d.reactions, d.reactions + R);
    m.desc.reactions = m.reactions.empty() ? nullptr : m.reactions.data();

    RxTable &T = m.table;
    std::memset(&T, 0, sizeof T);
    T.S = S;
    T.R = R;
    for (int i = 0; i < R; ++i) {
        const gpgmp_reaction_desc &rd = m.reactions[i];
        GPGMP_REQUIRE(std::isfinite(rd.rate) && rd.rate >= 0.f, GPGMP_E_INVALID,
                      "reaction rates must be finite and non-negative");
        // the law's rate is printed with `ss << rate` (Reaction.cpp:122): %g, 6 digits
        char buf[64];
        std::snprintf(buf, sizeof buf, "%g", (double)rd.rate);
        bool is_int = true;
        for (const char *c = buf; *c; ++c)
            if (!((*c >= '0' && *c <= '9') || *c == '-')) is_int = false;
        if (is_int) {
            T.lit_is_int |= 1u << i;
            T.lit_int[i] = (int32_t)std::strtol(buf, nullptr, 10);
        }
        T.lit_dbl[i] = std::strtod(buf, nullptr);
        int nf = 0;
        for (int j = 0; j < S; ++j) {
            GPGMP_REQUIRE(rd.reactant_stoich[j] >= 0 && rd.product_stoich[j] >= 0, GPGMP_E_INVALID,
                          "stoichiometries must be non-negative");
            for (int q = 0; q < rd.reactant_stoich[j]; ++q) {
                GPGMP_REQUIRE(nf < kMaxFactors, GPGMP_E_INVALID, "reaction order above 6");
                T.fs[i][nf] = (uint8_t)j;
                T.fo[i][nf] = (uint8_t)q;
                ++nf;
            }
            const int net = rd.product_stoich[j] - rd.reactant_stoich[j];
            GPGMP_REQUIRE(net >= -127 && net <= 127, GPGMP_E_INVALID, "stoichiometry out of range");
            T.net[i][j] = (int8_t)net;
        }
        T.nf[i] = (uint8_t)nf;
    }
    T.compact = (R <= kCompactReactions) ? 1 : 0;
    for (int i = 0; i < R; ++i)
        if (T.nf[i] > kCompactFactors) T.compact = 0;
    // StayProbability = (1.-p0) printed "%.6e" (Solver.cpp:505); the thresholds follow
    // StochasticHomogeneousDiffusion.cl:58-62.
    {
        char buf[64];
        std::snprintf(buf, sizeof buf, "%.6e", 1.0 - (double)d.p0);
        const float stay = std::strtof(buf, nullptr);
        const float pi = (float)((1. - (double)stay) / (2. * 3));
        m.probs.stay = stay;
        m.probs.p1 = stay + pi;
        m.probs.p2 = m.probs.p1 + pi;
        m.probs.p3 = m.probs.p2 + pi;
        m.probs.p4 = m.probs.p3 + pi;
    }
    m.cell_width = header_float(d.physical_length / (float)d.width);
    m.cell_height = header_float(d.physical_length / (float)d.height);
    if (m.inhomogeneous()) {
        GPGMP_REQUIRE(d.diffusivity_x && d.diffusivity_y && d.drift_x && d.drift_y, GPGMP_E_INVALID,
                      "inhomogeneous solvers need diffusivity and drift fields");
        const size_t n = (size_t)S * A;
        m.Dx.assign(d.diffusivity_x, d.diffusivity_x + n);
        m.Dy.assign(d.diffusivity_y, d.diffusivity_y + n);
        m.rx.assign(d.drift_x, d.drift_x + n);
        m.ry.assign(d.drift_y, d.drift_y + n);
        for (size_t i = 0; i < n; ++i)
            GPGMP_REQUIRE(std::isfinite(m.Dx[i]) && std::isfinite(m.Dy[i]) && m.Dx[i] >= 0.f &&
                              m.Dy[i] >= 0.f && std::isfinite(m.rx[i]) && std::isfinite(m.ry[i]),
                          GPGMP_E_INVALID, "fields must be finite with non-negative diffusivity");
    } else {
        for (int s = 0; s < S; ++s)
            GPGMP_REQUIRE(std::isfinite(d.diffusion[s]) && d.diffusion[s] >= 0.f, GPGMP_E_INVALID,
                          "diffusion constants must be finite and non-negative");
    }
    m.desc.diffusivity_x = m.Dx.empty() ? nullptr : m.Dx.data();
    m.desc.diffusivity_y = m.Dy.empty() ? nullptr : m.Dy.data();
    m.desc.drift_x = m.rx.empty() ? nullptr : m.rx.data();
    m.desc.drift_y = m.ry.empty() ? nullptr : m.ry.data();
    return m;
}

// ------------------------------------------------------------------------------------
Schedule::Schedule(bool inh, int S, bool has_reactions, const std::vector<float> &dtx,
                   const std::vector<float> &dty)
    : inh_(inh), S_(S), rx_(has_reactions), dtx_(dtx), dty_(dty), nx_(dtx), ny_(dty) {}

std::vector<float> Schedule::next_diffusion_time() const {
    std::vector<float> out(S_);
    for (int s = 0; s < S_; ++s) out[s] = (!inh_) ? nx_[s] : ((nx_[s] < ny_[s]) ? nx_[s] : ny_[s]);
    return out;
}

StepDecision Schedule::next() {
    StepDecision d;
    const std::vector<float> next = next_diffusion_time();
    float mindt = next[0];
    int minspec = 0;
    for (int i = 1; i < S_; ++i)
        if (next[i] < mindt) {
            mindt = next[i];
            minspec = i;
        }
    // an output time before the next diffusion makes this an output event: reactions up
    // to it, no diffusion; the state is written once the clock has reached it, which is
    // the iteration after (DiffusionModel.cpp:1519-1527, 1553-1572)
    const float mindt_out = out_last_ + out_interval_;
    const bool output_event = mindt_out < mindt;
    if (output_event) mindt = mindt_out;
    d.deltat = mindt - sim_time_;
    int32_t off = rx_ ? 1 : 0;  // Solver::doReactions increments the seed
    if (output_event) {
        if (sim_time_ >= out_last_ + out_interval_) {
            d.dump = true;
            d.dump_time = mindt;
            out_last_ += out_interval_;
        }
        d.off_end = off;
        sim_time_ = mindt;
        return d;
    }
    for (int i = 0; i < S_; ++i) {
        if (!(i == minspec || next[i] <= mindt)) continue;
        if (!(next[i] > 0)) continue;
        if (!inh_) {
            d.due_x |= 1u << i;
            d.off_x[i] = off;
            nx_[i] += dtx_[i];
        } else {
            if (nx_[i] <= mindt) {
                d.due_x |= 1u << i;
                d.off_x[i] = off++;
                nx_[i] += dtx_[i];
            }
            if (ny_[i] <= mindt) {
                d.due_y |= 1u << i;
                d.off_y[i] = off++;
                ny_[i] += dty_[i];
            }
        }
    }
    d.off_end = off;
    sim_time_ = mindt;
    return d;
}

// ------------------------------------------------------------------------------------
GmpEngine::GmpEngine(const Model &m, const std::vector<int> &W, const std::vector<int> &H,
                     hipStream_t stream)
    : table_(m.table), probs_(m.probs), inh_(m.inhomogeneous()), fpe_(m.fpe()), stream_(stream) {
    bm_ = make_int4(m.desc.boundary_mask[0], m.desc.boundary_mask[1], m.desc.boundary_mask[2],
                    m.desc.boundary_mask[3]);
    const int S = table_.S;
    const int nd = inh_ ? 2 : 4;
    uint64_t so = 0, lo = 0, po = 0;
    for (size_t r = 0; r < W.size(); ++r) {
        GPGMP_REQUIRE(W[r] >= 16 && H[r] >= 16 && W[r] % 16 == 0 && H[r] % 16 == 0, GPGMP_E_INVALID,
                      "grid sides must be positive multiples of 16");
        RealInfo ri{};
        ri.W = W[r];
        ri.H = H[r];
        ri.state_off = so;
        ri.lost_off = lo;
        ri.prob_off = po;
        ri.cell_off = (uint64_t)total_cells_;
        // source counts per boundary as HomogeneousSolver::diffuse passes them
        // (HomogeneousSolver.cpp:141-156: boundary 0 takes sourceMask[1])
        const int32_t *sm = m.desc.source_mask;
        if (!inh_) {
            ri.src[0] = sm[0] > 0 ? sm[1] : 0;
            ri.src[1] = sm[1] > 0 ? sm[1] : 0;
            ri.src[2] = sm[2] > 0 ? sm[2] : 0;
            ri.src[3] = sm[3] > 0 ? sm[3] : 0;
        }
        const uint64_t A = (uint64_t)W[r] * H[r];
        so += A * S;
        lo += A * S * nd;
        po += A * S;
        total_cells_ += (int64_t)A;
        reals_.push_back(ri);
    }
    {
        const std::vector<uint32_t> code = build_law_programs(m, table_);
        table_.prog = nullptr;
        if (!code.empty()) {
            prog_.alloc(code.size());
            GPGMP_HIP_CHECK(hipMemcpy(prog_.get(), code.data(), code.size() * 4, hipMemcpyHostToDevice));
            table_.prog = prog_.get();
        }
    }
    if (table_.localized) {
        std::vector<uint16_t> act((size_t)total_cells_);
        for (size_t r = 0; r < reals_.size(); ++r)
            reaction_act(m, reals_[r].W, reals_[r].H, act.data() + reals_[r].cell_off);
        act_.alloc(act.size());
        GPGMP_HIP_CHECK(hipMemcpy(act_.get(), act.data(), act.size() * sizeof(uint16_t), hipMemcpyHostToDevice));
        table_.act = act_.get();
    } else {
        table_.act = nullptr;
    }
    state_.alloc(so);
    lost_[0].alloc(lo);
    lost_[1].alloc(lo);
    if (inh_) {
        prob_[0].alloc(po);
        prob_[1].alloc(po);
    }
    // deferred busy cells: at most one entry per cell of the launch, capped at 2^20
    hot_.cap = (uint32_t)std::min<int64_t>(total_cells_, (int64_t)1 << 20);
    // deferral of busy cells: after 8 events, once fewer than 24 lanes of the wave run
    // (measured best on the 256x256 Schnakenberg run among budgets 2..64, thresholds 4..65)
    hot_.budget = 8;
    hot_.below_active = 24;
    hot_e_.alloc(2 * (size_t)hot_.cap);  // two lists when busy cells are predicted
    hot_lane_.alloc((size_t)hot_.cap);
    hot_count_.alloc(8);
    hot_flag_.alloc((size_t)total_cells_);
    hot_pred_.alloc(2 * (size_t)total_cells_);
    hot_.e = hot_e_.get();
    hot_.count = hot_count_.get();
    hot_.flag = nullptr;
    hot_.pred = hot_pred_.get();
    hot_.pred_stride = (uint32_t)total_cells_;
    hot_.lane_count = hot_count_.get() + 4;  // the lane list's count (entries only when fused)
    GPGMP_HIP_CHECK(hipMemsetAsync(hot_flag_.get(), 0, (size_t)total_cells_, stream_));
    reset_hot();
    reals_dev_.alloc(reals_.size());
    GPGMP_HIP_CHECK(hipMemcpyAsync(reals_dev_.get(), reals_.data(), reals_.size() * sizeof(RealInfo),
                                   hipMemcpyHostToDevice, stream_));
    // zero the outflow buffers once so that a stale read can never see garbage
    GPGMP_HIP_CHECK(hipMemsetAsync(lost_[0].get(), 0, lo * sizeof(int32_t), stream_));
    GPGMP_HIP_CHECK(hipMemsetAsync(lost_[1].get(), 0, lo * sizeof(int32_t), stream_));
    GPGMP_HIP_CHECK(hipMemsetAsync(state_.get(), 0, so * sizeof(int32_t), stream_));
    // the SSA launches keep one lane per cell (the least VALU work per event); the
    // particle loops of each wave's cells are flattened over all its lanes (measured best
    // on the 256x256 Schnakenberg run; the batch autotuner may choose otherwise)
    grp_ = 1;
    grp_sweep_ = 0;
    build_tiles();
}

void GmpEngine::build_tiles() {
    for (int k = 0; k < 3; ++k) {
        std::vector<Tile> tiles;
        const int per = block_ / (k == 0 ? 1 : (k == 1 ? 4 : 8)) * cpt_;
        for (size_t r = 0; r < reals_.size(); ++r) {
            const int A = reals_[r].W * reals_[r].H;
            for (int c = 0; c < A; c += per) tiles.push_back(Tile{(int32_t)r, (int32_t)c});
        }
        ntiles_[k] = (int)tiles.size();
        tiles_dev_[k].alloc(tiles.size());
        GPGMP_HIP_CHECK(hipMemcpyAsync(tiles_dev_[k].get(), tiles.data(), tiles.size() * sizeof(Tile),
                                       hipMemcpyHostToDevice, stream_));
    }
    GPGMP_HIP_CHECK(hipStreamSynchronize(stream_));
}

void GmpEngine::set_launch(int block, int cpt, int grp, int grp_sweep) {
    GPGMP_REQUIRE(block >= 64 && block <= 256 && block % 64 == 0, GPGMP_E_INVALID,
                  "block must be 64, 128, 192 or 256");
    GPGMP_REQUIRE(cpt >= 1 && cpt <= 16, GPGMP_E_INVALID, "cells per thread must be in [1, 16]");
    GPGMP_REQUIRE((grp == 1 || grp == 4) && (grp_sweep == 0 || grp_sweep == 1 || grp_sweep == 4), GPGMP_E_INVALID,
                  "lanes per cell must be 1 or 4 (SSA) and 0, 1 or 4 (sweeps)");
    ++shape_version_;
    grp_ = grp;
    grp_sweep_ = grp_sweep;
    if (block == block_ && cpt == cpt_) return;
    GPGMP_HIP_CHECK(hipStreamSynchronize(stream_));
    block_ = block;
    cpt_ = cpt;
    build_tiles();
}

void GmpEngine::set_state(int r, const int32_t *host) {
    GPGMP_REQUIRE(r >= 0 && r < num_real(), GPGMP_E_INVALID, "realization index out of range");
    const RealInfo &ri = reals_[r];
    const size_t n = (size_t)ri.W * ri.H * S();
    for (size_t i = 0; i < n; ++i) GPGMP_REQUIRE(host[i] >= 0, GPGMP_E_INVALID, "negative count");
    GPGMP_HIP_CHECK(hipMemcpyAsync(state_.get() + ri.state_off, host, n * sizeof(int32_t),
                                   hipMemcpyHostToDevice, stream_));
    GPGMP_HIP_CHECK(hipStreamSynchronize(stream_));
}

void GmpEngine::get_state(int r, int32_t *host) {
    GPGMP_REQUIRE(r >= 0 && r < num_real(), GPGMP_E_INVALID, "realization index out of range");
    const RealInfo &ri = reals_[r];
    const size_t n = (size_t)ri.W * ri.H * S();
    GPGMP_HIP_CHECK(hipMemcpyAsync(host, state_.get() + ri.state_off, n * sizeof(int32_t),
                                   hipMemcpyDeviceToHost, stream_));
    GPGMP_HIP_CHECK(hipStreamSynchronize(stream_));
}

void GmpEngine::set_probs(int r, int s, int axis, const float *D, const float *rr, float dt, float dx) {
    const RealInfo &ri = reals_[r];
    const int A = ri.W * ri.H;
    DevBuf<float> dD(A), dr(A);
    GPGMP_HIP_CHECK(hipMemcpyAsync(dD.get(), D, A * sizeof(float), hipMemcpyHostToDevice, stream_));
    GPGMP_HIP_CHECK(hipMemcpyAsync(dr.get(), rr, A * sizeof(float), hipMemcpyHostToDevice, stream_));
    GPGMP_HIP_CHECK(launch_inh_probs(prob_[axis].get() + ri.prob_off + (size_t)s * A, dD.get(), dr.get(),
                                     dt, dx, A, fpe_ ? 1 : 0, stream_));
    GPGMP_HIP_CHECK(hipStreamSynchronize(stream_));
}

void GmpEngine::launch(int mode, const LaunchPlan *plan, const int32_t *seeds_cur,
                       const int32_t *seeds_prev, const int32_t *lost_in, int32_t *lost_out,
                       hipEvent_t ev_start, hipEvent_t ev_stop, hipEvent_t ev_hot_start,
                       hipEvent_t ev_hot_stop, bool fuse) {
    StepLaunch a{};
    a.mode = mode;
    a.state = state_.get();
    a.lost_in = lost_in;
    a.lost_out = lost_out;
    a.prob = inh_ ? prob_[mode == 2 ? 1 : 0].get() : nullptr;
    a.reals = reals_dev_.get();
    if (mode == 2 || mode == 3) {  // no SSA: 16 (or 8) cells per wave unless one lane per cell is asked for
        a.grp = (grp_sweep_ == 1) ? 1 : ((grp_sweep_ == 0 && sweep_wave_cells_ == 8) ? 8 : 4);
        a.grp_sweep = (grp_sweep_ == 0) ? 0 : a.grp;
    } else {
        a.grp = grp_;
        a.grp_sweep = (grp_sweep_ == 0) ? 0 : (grp_ == 4 ? 4 : grp_sweep_);
    }
    const int tk = (a.grp == 8) ? 2 : (a.grp == 4 ? 1 : 0);
    a.tiles = tiles_dev_[tk].get();
    a.ntiles = ntiles_[tk];
    if (reals_.size() == 1) {  // the tiles of build_tiles(), computed in the kernel
        a.one.r = reals_[0];
        a.one.per = block_ / (tk == 0 ? 1 : (tk == 1 ? 4 : 8)) * cpt_;
    }
    a.plan = plan;
    a.seeds_cur = seeds_cur;
    a.seeds_prev = seeds_prev;
    a.T = &table_;
    a.P = &probs_;
    a.bm = bm_;
    a.block = block_;
    a.cpt = cpt_;
    a.stream = stream_;
    a.ev_start = ev_start;
    a.ev_stop = ev_stop;
    a.hot = hot_;
    a.hot_grid = (int)std::min<uint32_t>(hot_.cap, (uint32_t)hot_waves_);
    if (fuse) {
        GPGMP_REQUIRE(fused_hot() && (mode == 1 || mode == 3), GPGMP_E_INVALID, "hot cells fuse into split steps only");
        GPGMP_REQUIRE(num_real() == 1, GPGMP_E_INVALID, "hot cells fuse into single-realization launches only");
        // the SSA launch flags its deferred cells; the X-sweep launch (plan = SSA plan + 1)
        // continues them beside the sweeps of the others
        a.hot.flag = hot_flag_.get();
        if (!predict() && lane_events_ > 0.f) {  // modest tails to the lane list
            a.hot.lane_e = hot_lane_.get();
            a.hot.lane_cap = hot_.cap;
            a.hot.lane_events = lane_events_;
            a.lane_grid = lane_waves_;
        }
        if (mode == 3) a.plan_ssa = plan - 1;
        if (mode == 1) a.predict = predict() ? 1 : 0;
    }
    a.ev_hot_start = ev_hot_start;
    a.ev_hot_stop = ev_hot_stop;
    GPGMP_HIP_CHECK(launch_gmp_step(a));
}

// empty hot lists and no predicted cells (a run starts from here)
void GmpEngine::reset_hot() {
    GPGMP_HIP_CHECK(hipMemsetAsync(hot_count_.get(), 0, 8 * sizeof(uint32_t), stream_));
    GPGMP_HIP_CHECK(hipMemsetAsync(hot_pred_.get(), 0, 2 * (size_t)total_cells_, stream_));
}

// ------------------------------------------------------------------------------------
struct GmpRunner::Chunk {
    int cap = 0, nreal = 0;
    LaunchPlan *hplan = nullptr;   // pinned [3*cap + 1]
    int32_t *hseed = nullptr;      // pinned [(cap + 1) * nreal]
    DevBuf<LaunchPlan> dplan;
    DevBuf<int32_t> dseed;
    hipEvent_t copied = nullptr;
    // Captured graphs of this chunk's buffers.  A graph bakes in the launches of steps
    // [from, steps) of the chunk and, for homogeneous models, the outflow buffers those
    // steps read and write, which alternate with the step's parity: an output event can
    // end a chunk after an odd number of steps, so a graph is only reused at the parity
    // it was captured at.
    struct Graph {
        hipGraphExec_t exec = nullptr;
        int steps = 0, from = 0, parity = 0;
        uint64_t version = 0;  // engine shape version the graph was captured with
        uint64_t used = 0;     // last use (eviction order)
    };
    static constexpr size_t kMaxGraphs = 8;
    std::vector<Graph> graphs;
    uint64_t uses = 0;
    Chunk(int cap_, int nreal_) : cap(cap_), nreal(nreal_) {
        GPGMP_HIP_CHECK(hipHostMalloc((void **)&hplan, sizeof(LaunchPlan) * (3 * cap + 1)));
        GPGMP_HIP_CHECK(hipHostMalloc((void **)&hseed, sizeof(int32_t) * (size_t)(cap + 1) * nreal));
        dplan.alloc(3 * cap + 1);
        dseed.alloc((size_t)(cap + 1) * nreal);
        GPGMP_HIP_CHECK(hipEventCreateWithFlags(&copied, hipEventDisableTiming));
    }
    Graph *find(int steps, int from, int parity, uint64_t version) {
        for (Graph &g : graphs)
            if (g.steps == steps && g.from == from && g.parity == parity && g.version == version) {
                g.used = ++uses;
                return &g;
            }
        return nullptr;
    }
    void insert(const Graph &g) {
        if (graphs.size() == kMaxGraphs) {
            auto lru = std::min_element(graphs.begin(), graphs.end(),
                                        [](const Graph &a, const Graph &b) { return a.used < b.used; });
            (void)hipGraphExecDestroy(lru->exec);
            graphs.erase(lru);
        }
        graphs.push_back(g);
        graphs.back().used = ++uses;
    }
    ~Chunk() {
        for (Graph &g : graphs) (void)hipGraphExecDestroy(g.exec);
        if (copied) {
            (void)hipEventSynchronize(copied);
            (void)hipEventDestroy(copied);
        }
        if (hplan) (void)hipHostFree(hplan);
        if (hseed) (void)hipHostFree(hseed);
    }
};

GmpRunner::GmpRunner(GmpEngine &eng, const std::vector<float> &dtx, const std::vector<float> &dty,
                     const std::vector<uint32_t> &seed0s)
    : eng_(eng), sched_(eng.inh_, eng.S(), eng.table_.R > 0, dtx, dty) {
    GPGMP_REQUIRE((int)seed0s.size() == eng.num_real(), GPGMP_E_INVALID, "one seed0 per realization");
    for (uint32_t s0 : seed0s) seeds_.emplace_back(s0);
    prev_seed_.assign(eng.num_real(), 0);
    eng_.reset_hot();
}

GmpRunner::~GmpRunner() {
    (void)hipStreamSynchronize(eng_.stream());
    for (Chunk *c : chunks_) delete c;
    for (auto &e : ev_pool_) {
        (void)hipEventDestroy(e.first);
        (void)hipEventDestroy(e.second);
    }
}

// Algorithmic HBM bytes of one launch: counts read and written (all species when the
// SSA runs, else the touched ones), ND outflow words read per gathered species, ND
// written per species with outflow, plus the (p1,p2) table of inhomogeneous sweeps and
// the hot-cell flag byte (written by the SSA launch, read by the X sweeps) of split steps.
static double launch_bytes(const GmpEngine &e, int mode, const LaunchPlan &p) {
    const int S = e.S();
    const int nd = (mode == 0) ? 4 : 2;
    const int ng = __builtin_popcount(p.gather_mask), no = __builtin_popcount(p.out_mask);
    const bool ssa = (mode != 2) && p.deltat > 0.f && e.table_.R > 0;
    const int nstate = ssa ? S : __builtin_popcount(p.gather_mask | p.out_mask);
    const double flag = (e.fused_hot() && (mode == 1 || mode == 3)) ? 1.0 : 0.0;
    const double per_cell = 8.0 * nstate + 4.0 * nd * ng + 4.0 * nd * no + (mode != 0 ? 8.0 * no : 0.0) + flag;
    return per_cell * (double)e.total_cells();
}

// how: 0 plain launch; 1 timed (dispatch-tied start/stop events, accumulated); 2 with
// dispatch-tied events that are not accumulated (the first step of a timing window).
void GmpRunner::timed_launch(int mode, const LaunchPlan *plan, const LaunchPlan &hplan, const int32_t *scur,
                             const int32_t *sprev, const int32_t *lin, int32_t *lout, int how) {
    const bool fuse = eng_.fused_hot() && (mode == 1 || mode == 3);
    if (how == 0) {
        eng_.launch(mode, plan, scur, sprev, lin, lout, nullptr, nullptr, nullptr, nullptr, fuse);
        return;
    }
    // returns a copy: growing the pool would invalidate references
    auto next_pair = [&](int m) -> std::pair<hipEvent_t, hipEvent_t> {
        if (ev_used_ == ev_pool_.size()) {
            hipEvent_t a, b;
            GPGMP_HIP_CHECK(hipEventCreate(&a));
            GPGMP_HIP_CHECK(hipEventCreate(&b));
            ev_pool_.emplace_back(a, b);
            ev_mode_.push_back(0);
        }
        ev_mode_[ev_used_] = m;
        return ev_pool_[ev_used_++];
    };
    const bool keep = how == 1;
    const int slot = (mode == 3) ? 4 : mode;
    const std::pair<hipEvent_t, hipEvent_t> e = next_pair(keep ? slot : -1);
    if (keep) kt_bytes_[slot] += launch_bytes(eng_, mode, hplan);
    if (mode == 2 || mode == 3 || fuse) {
        eng_.launch(mode, plan, scur, sprev, lin, lout, e.first, e.second, nullptr, nullptr, fuse);
    } else {
        const std::pair<hipEvent_t, hipEvent_t> h = next_pair(keep ? 3 : -1);
        eng_.launch(mode, plan, scur, sprev, lin, lout, e.first, e.second, h.first, h.second);
    }
}

void GmpRunner::collect_times() {
    if (ev_used_ == 0) return;
    GPGMP_HIP_CHECK(hipStreamSynchronize(eng_.stream()));
    for (size_t i = 0; i < ev_used_; ++i) {
        float ms = 0.f;
        if (ev_mode_[i] < 0) continue;
        GPGMP_HIP_CHECK(hipEventElapsedTime(&ms, ev_pool_[i].first, ev_pool_[i].second));
        kt_ms_[ev_mode_[i]] += ms;
        kt_n_[ev_mode_[i]] += 1;
    }
    ev_used_ = 0;
}

// Issues the launches of one chunk whose plans/seeds are already in its host buffers.
// dry: issues nothing; only captures (and caches) the graph a real issue of this chunk
// would replay.
void GmpRunner::issue(Chunk &c, int nsteps, bool use_graph, bool dry) {
    const int nreal = eng_.num_real();
    const int lps = eng_.inh_ ? (eng_.split_x() ? 3 : 2) : 1;  // launches per step
    if (!dry) {
        GPGMP_HIP_CHECK(hipMemcpyAsync(c.dplan.get(), c.hplan, sizeof(LaunchPlan) * lps * nsteps,
                                       hipMemcpyHostToDevice, eng_.stream()));
        GPGMP_HIP_CHECK(hipMemcpyAsync(c.dseed.get(), c.hseed, sizeof(int32_t) * (size_t)(nsteps + 1) * nreal,
                                       hipMemcpyHostToDevice, eng_.stream()));
        GPGMP_HIP_CHECK(hipEventRecord(c.copied, eng_.stream()));
    }
    auto body = [&](int j0, int j1) {
        for (int j = j0; j < j1; ++j) {
            const int32_t *scur = c.dseed.get() + (size_t)(j + 1) * nreal;
            const int32_t *sprev = c.dseed.get() + (size_t)j * nreal;
            // timing windows: of every every_ steps the first kTimingWindow carry events,
            // and the first step of a window is not counted
            const int64_t sn = step_no_++;
            int how = 0;
            if (timing_) {
                const int64_t r = sn % every_;
                how = every_ == 1 ? 1 : r == 0 ? 2 : r < kTimingWindow ? 1 : 0;
            }
            const int how_last = how;
            if (!eng_.inh_) {
                const int p = (parity_ + j) & 1;
                timed_launch(0, c.dplan.get() + j, c.hplan[j], scur, sprev, eng_.lost(p ^ 1), eng_.lost(p),
                             how_last);
            } else if (eng_.split_x()) {
                timed_launch(1, c.dplan.get() + 3 * j, c.hplan[3 * j], scur, sprev, eng_.lost(1), eng_.lost(0),
                             how);
                timed_launch(3, c.dplan.get() + 3 * j + 1, c.hplan[3 * j + 1], scur, sprev, eng_.lost(1),
                             eng_.lost(0), how);
                timed_launch(2, c.dplan.get() + 3 * j + 2, c.hplan[3 * j + 2], scur, sprev, eng_.lost(0),
                             eng_.lost(1), how_last);
            } else {
                timed_launch(1, c.dplan.get() + 2 * j, c.hplan[2 * j], scur, sprev, eng_.lost(1), eng_.lost(0),
                             how);
                timed_launch(2, c.dplan.get() + 2 * j + 1, c.hplan[2 * j + 1], scur, sprev, eng_.lost(0),
                             eng_.lost(1), how_last);
            }
        }
    };
    // With timing on, the chunk's event-carrying steps run eagerly when they form a prefix
    // of the chunk (a chunk length that is a multiple of the timing period keeps them
    // there) and the rest of the chunk replays a graph; otherwise the chunk runs eagerly.
    // A graph of fewer than kMinGraphSteps steps is not worth its capture.
    int k = 0;
    if (use_graph && timing_) {
        auto timed = [&](int64_t sn) { return every_ == 1 || (sn % every_) < kTimingWindow; };
        while (k < nsteps && timed(step_no_ + k)) ++k;
        for (int j = k; j < nsteps; ++j)
            if (timed(step_no_ + j)) use_graph = false;
    }
    if (!use_graph || nsteps - k < kMinGraphSteps) {
        if (dry) step_no_ += nsteps;
        else body(0, nsteps);
        return;
    }
    if (dry) step_no_ += k;
    else body(0, k);
    // a replay issues no host code, so the step counter of the timing windows is advanced
    // here for the graph's steps
    const int64_t sn_graph = step_no_;
    struct StepRestore {
        int64_t &n;
        int64_t v;
        ~StepRestore() { n = v; }
    } step_restore{step_no_, sn_graph + (nsteps - k)};
    const bool saved_timing = timing_;
    timing_ = false;  // events are not captured into graphs
    struct TimingRestore {
        bool &t;
        bool v;
        ~TimingRestore() { t = v; }
    } timing_restore{timing_, saved_timing};
    // homogeneous steps read and write the outflow buffers of their parity
    const int par = eng_.inh_ ? 0 : parity_;
    Chunk::Graph *gr = c.find(nsteps, k, par, eng_.shape_version());
    if (!gr) {
        // capture on a private stream, then replay on the engine's stream
        hipStream_t cap;
        GPGMP_HIP_CHECK(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking));
        hipStream_t saved = eng_.stream();
        struct Restore {
            GmpEngine &e;
            hipStream_t s;
            ~Restore() { e.set_stream(s); }
        } restore{eng_, saved};
        eng_.set_stream(cap);
        hipGraph_t g = nullptr;
        GPGMP_HIP_CHECK(hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal));
        try {
            body(k, nsteps);
        } catch (...) {
            (void)hipStreamEndCapture(cap, &g);
            if (g) (void)hipGraphDestroy(g);
            (void)hipStreamDestroy(cap);
            throw;
        }
        GPGMP_HIP_CHECK(hipStreamEndCapture(cap, &g));
        eng_.set_stream(saved);
        Chunk::Graph ng;
        const hipError_t ie = hipGraphInstantiate(&ng.exec, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        (void)hipStreamDestroy(cap);
        GPGMP_HIP_CHECK(ie);
        ng.steps = nsteps;
        ng.from = k;
        ng.parity = par;
        ng.version = eng_.shape_version();
        c.insert(ng);
        gr = &c.graphs.back();
    }
    if (!dry) GPGMP_HIP_CHECK(hipGraphLaunch(gr->exec, eng_.stream()));
}

// Captures the graphs that run(steps, graph_chunk) will replay, without running anything:
// the host loop is planned on a copy of the runner's state, which is restored afterwards.
// A run of a few steps then replays its chunk instead of capturing it inside the run.
void GmpRunner::prepare(int steps, int graph_chunk) {
    GPGMP_REQUIRE(steps >= 0, GPGMP_E_INVALID, "steps must be non-negative");
    if (graph_chunk <= 0 || steps == 0) return;
    const Schedule sched = sched_;
    const std::vector<SeedStream> seeds = seeds_;
    const StepDecision prev = prev_;
    const bool has_prev = has_prev_;
    const std::vector<int32_t> prev_seed = prev_seed_;
    const int parity = parity_;
    const int64_t iterations = iterations_, plan_step = plan_step_, step_no = step_no_;
    struct Restore {
        GmpRunner &r;
        const Schedule &sc;
        const std::vector<SeedStream> &se;
        const StepDecision &pv;
        bool hp;
        const std::vector<int32_t> &ps;
        int pa;
        int64_t it, pl, sn;
        ~Restore() {
            r.sched_ = sc;
            r.seeds_ = se;
            r.prev_ = pv;
            r.has_prev_ = hp;
            r.prev_seed_ = ps;
            r.parity_ = pa;
            r.iterations_ = it;
            r.plan_step_ = pl;
            r.step_no_ = sn;
        }
    } restore{*this, sched, seeds, prev, has_prev, prev_seed, parity, iterations, plan_step, step_no};
    run_impl(steps, graph_chunk, true);
}

void GmpRunner::run(int steps, int graph_chunk) { run_impl(steps, graph_chunk, false); }

// Advances the loop by `steps` iterations without launching anything: the schedule, the
// seed streams and the iteration count move on as if they had run, and nothing is pending
// (the caller's state is the state after them, e.g. a snapshot of an earlier run).
void GmpRunner::skip(int steps) {
    GPGMP_REQUIRE(steps >= 0, GPGMP_E_INVALID, "steps must be non-negative");
    GPGMP_REQUIRE(!has_prev_, GPGMP_E_STATE, "outflow of a run is pending");
    for (int j = 0; j < steps && !sched_.finished(); ++j) {
        (void)sched_.next();
        for (size_t r = 0; r < seeds_.size(); ++r) prev_seed_[r] = seeds_[r].next();
        ++plan_step_;
        ++iterations_;
    }
}

void GmpRunner::run_impl(int steps, int graph_chunk, bool dry) {
    GPGMP_REQUIRE(steps >= 0, GPGMP_E_INVALID, "steps must be non-negative");
    const int nreal = eng_.num_real();
    const bool use_graph = graph_chunk > 0;
    int cap = use_graph ? graph_chunk : 256;
    if (!eng_.inh_ && use_graph && (cap & 1)) ++cap;  // keeps the outflow-buffer parity per chunk
    if (cap != chunk_cap_) {
        GPGMP_HIP_CHECK(hipStreamSynchronize(eng_.stream()));
        for (Chunk *c : chunks_) delete c;
        chunks_.clear();
        chunks_.push_back(new Chunk(cap, nreal));
        chunks_.push_back(new Chunk(cap, nreal));
        chunk_cap_ = cap;
    }
    parity_ = 0;  // nothing is pending at entry
    int done = 0, ci = 0;
    while (done < steps && !sched_.finished()) {
        const int nmax = std::min(cap, steps - done);
        int n = 0;
        bool dump = false;
        float dump_time = 0.f;
        Chunk &c = *chunks_[ci];
        GPGMP_HIP_CHECK(hipEventSynchronize(c.copied));  // host buffers free again
        for (int r = 0; r < nreal; ++r) c.hseed[r] = prev_seed_[r];
        for (int j = 0; j < nmax; ++j) {
            if (j > 0 && sched_.finished()) break;  // the loop's simTime < maxTime test
            const StepDecision d = sched_.next();
            n = j + 1;
            for (int r = 0; r < nreal; ++r) c.hseed[(size_t)(j + 1) * nreal + r] = seeds_[r].next();
            const float dts = (eng_.table_.R > 0) ? d.deltat : 0.f;
            const int32_t spread = eng_.spread() ? 1 : 0;
            if (!eng_.inh_) {
                LaunchPlan &p = c.hplan[j];
                std::memset(&p, 0, sizeof p);
                p.interleave = spread;
                p.deltat = dts;
                p.gather_mask = has_prev_ ? prev_.due_x : 0u;
                p.gather_off = has_prev_ ? (eng_.table_.R > 0 ? 1 : 0) : 0;
                p.out_mask = d.due_x;
                for (int s = 0; s < kMaxSpecies; ++s) p.out_off[s] = d.off_x[s];
            } else {
                const bool split = eng_.split_x();
                const int lps = split ? 3 : 2;
                LaunchPlan &px = c.hplan[lps * j];
                LaunchPlan &py = c.hplan[lps * j + lps - 1];
                std::memset(&px, 0, sizeof px);
                std::memset(&py, 0, sizeof py);
                px.interleave = spread;
                py.interleave = spread;
                // split steps predict busy cells from the previous step's hot list
                px.predict = eng_.predict() ? 1 : 0;
                px.hot_parity = (int32_t)(plan_step_ & 1);
                // the Y launch empties the hot list the X-sweep launch consumed (or, when
                // predicting, the one this step's SSA launch read)
                py.reset_hot = eng_.fused_hot() ? 1 : 0;
                py.hot_parity = px.predict ? (px.hot_parity ^ 1) : 0;
                px.deltat = dts;
                px.gather_mask = has_prev_ ? prev_.due_y : 0u;
                px.out_mask = d.due_x;
                py.deltat = 0.f;
                py.gather_mask = d.due_x;
                py.out_mask = d.due_y;
                for (int s = 0; s < kMaxSpecies; ++s) {
                    px.out_off[s] = d.off_x[s];
                    py.out_off[s] = d.off_y[s];
                }
                if (split) {  // the X sweeps move from the SSA launch to their own
                    LaunchPlan &pxs = c.hplan[lps * j + 1];
                    pxs = px;
                    pxs.deltat = 0.f;
                    pxs.gather_mask = 0u;
                    px.out_mask = 0u;
                }
            }
            prev_ = d;
            has_prev_ = true;
            ++plan_step_;
            if (d.dump && dump_) {  // end the chunk here: the copy follows its launches
                dump = true;
                dump_time = d.dump_time;
                break;
            }
        }
        for (int r = 0; r < nreal; ++r) prev_seed_[r] = c.hseed[(size_t)n * nreal + r];
        issue(c, n, use_graph, dry);
        if (!eng_.inh_) parity_ = (parity_ + n) & 1;
        done += n;
        iterations_ += n;
        ci ^= 1;
        if (dump && !dry) dump_(dump_time);
    }
    if (dry) return;
    // final gather of the last launch's outflow (and source boundaries)
    if (has_prev_) {
        Chunk &c = *chunks_[ci];
        GPGMP_HIP_CHECK(hipEventSynchronize(c.copied));
        LaunchPlan &p = c.hplan[0];
        std::memset(&p, 0, sizeof p);
        p.deltat = 0.f;
        p.out_mask = 0;
        for (int r = 0; r < nreal; ++r) {
            c.hseed[r] = prev_seed_[r];
            c.hseed[nreal + r] = prev_seed_[r];
        }
        if (!eng_.inh_) {
            p.gather_mask = prev_.due_x;
            p.gather_off = eng_.table_.R > 0 ? 1 : 0;
        } else {
            p.gather_mask = prev_.due_y;
        }
        GPGMP_HIP_CHECK(hipMemcpyAsync(c.dplan.get(), c.hplan, sizeof(LaunchPlan), hipMemcpyHostToDevice,
                                       eng_.stream()));
        GPGMP_HIP_CHECK(hipMemcpyAsync(c.dseed.get(), c.hseed, sizeof(int32_t) * 2 * nreal,
                                       hipMemcpyHostToDevice, eng_.stream()));
        GPGMP_HIP_CHECK(hipEventRecord(c.copied, eng_.stream()));
        const int32_t *scur = c.dseed.get() + nreal;
        const int32_t *sprev = c.dseed.get();
        if (!eng_.inh_)
            eng_.launch(0, c.dplan.get(), scur, sprev, eng_.lost(parity_ ^ 1), eng_.lost(parity_));
        else
            eng_.launch(1, c.dplan.get(), scur, sprev, eng_.lost(1), eng_.lost(0));
        has_prev_ = false;
    }
    if (timing_) collect_times();
}

}  // namespace gpgmp
