// Deterministic solver types behind the solver interface: the fp32 concentration state of
// Solver(..., detState, ...) with doReactions running the RK4 or alpha-QSS integrator
// (Solver.cpp:250-266) and diffuse the reference's stencils with its time steps
// (HomogeneousSolver.cpp:97-121,157-160; InhomogeneousSolver.cpp:282-286,374-378).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "det_solver.h"
#include "deterministic_launch.h"

namespace gpgmp {

DetEngine::DetEngine(const Model &m, uint32_t flags, hipStream_t stream) : m_(m), flags_(flags), stream_(stream) {
    const gpgmp_model_desc &d = m_.desc;
    W_ = d.width;
    H_ = d.height;
    S_ = d.num_species;
    A_ = (size_t)W_ * H_;
    x_.alloc((size_t)S_ * A_);
    tmp_.alloc(A_);
    m_.table.act = nullptr;
    m_.table.grid_w = W_;
    {
        const std::vector<uint32_t> code = build_law_programs(m_, m_.table);
        m_.table.prog = nullptr;
        if (!code.empty()) {
            prog_.alloc(code.size());
            GPGMP_HIP_CHECK(hipMemcpy(prog_.get(), code.data(), code.size() * 4, hipMemcpyHostToDevice));
            m_.table.prog = prog_.get();
        }
    }
    if (m_.table.localized) {
        std::vector<uint16_t> act(A_);
        reaction_act(m_, W_, H_, act.data());
        act_.alloc(A_);
        GPGMP_HIP_CHECK(hipMemcpy(act_.get(), act.data(), A_ * sizeof(uint16_t), hipMemcpyHostToDevice));
        m_.table.act = act_.get();
    }
    GPGMP_HIP_CHECK(hipMemsetAsync(x_.get(), 0, (size_t)S_ * A_ * sizeof(float), stream_));
    dtx_.assign(S_, 0.f);
    dty_.assign(S_, 0.f);
    if (!m_.inhomogeneous()) {
        // HomogeneousSolver.cpp:97-121: dt_s = 1e-2 h^2 / (4 D_s) in float (maxTime for a
        // species that does not diffuse), then the smallest for every species
        const float h = d.physical_length / (float)d.width;
        float common = 0.f;
        for (int s = 0; s < S_; ++s) {
            const float D = d.diffusion[s];
            const float dts = (D > 0) ? (float)1e-2f * h * h / (4.f * D) : d.max_time;
            common = (s == 0 || dts < common) ? dts : common;
        }
        dtx_.assign(S_, common);
        dty_ = dtx_;
    } else {
        // the stochastic solver's first time steps (InhomogeneousSolver.cpp:189-212), which
        // the deterministic inhomogeneous types use too
        for (int s = 0; s < S_; ++s) {
            float mx = INFINITY, my = INFINITY;
            for (size_t c = 0; c < A_; ++c) {
                mx = std::min(mx, inhomogeneous_cell_dt(m_.Dx[s * A_ + c], m_.rx[s * A_ + c], m_.cell_width));
                my = std::min(my, inhomogeneous_cell_dt(m_.Dy[s * A_ + c], m_.ry[s * A_ + c], m_.cell_width));
            }
            dtx_[s] = mx;
            dty_[s] = my;
        }
        const size_t n = (size_t)S_ * A_;
        for (DevBuf<float> *b : {&Dx_, &Dy_, &rx_, &ry_}) b->alloc(n);
        GPGMP_HIP_CHECK(hipMemcpyAsync(Dx_.get(), m_.Dx.data(), n * 4, hipMemcpyHostToDevice, stream_));
        GPGMP_HIP_CHECK(hipMemcpyAsync(Dy_.get(), m_.Dy.data(), n * 4, hipMemcpyHostToDevice, stream_));
        GPGMP_HIP_CHECK(hipMemcpyAsync(rx_.get(), m_.rx.data(), n * 4, hipMemcpyHostToDevice, stream_));
        GPGMP_HIP_CHECK(hipMemcpyAsync(ry_.get(), m_.ry.data(), n * 4, hipMemcpyHostToDevice, stream_));
    }
    nx_ = dtx_;
    ny_ = dty_;
    GPGMP_HIP_CHECK(hipStreamSynchronize(stream_));
}

bool DetEngine::aqss() const {
    return m_.desc.solver_type == GPGMP_DETERMINISTIC_HOMOGENEOUS_AQSS ||
           m_.desc.solver_type == GPGMP_DETERMINISTIC_INHOMOGENEOUS_AQSS;
}

bool DetEngine::reacts() const {
    if (m_.table.R == 0) return false;
    // the reference's doReactions reacts for the homogeneous deterministic types only
    return !(m_.inhomogeneous() && (flags_ & GPGMP_SOLVER_REFERENCE_QUIRKS));
}

void DetEngine::set_state(const float *host) {
    const size_t n = (size_t)S_ * A_;
    for (size_t i = 0; i < n; ++i)
        GPGMP_REQUIRE(std::isfinite(host[i]) && host[i] >= 0.f, GPGMP_E_INVALID,
                      "concentrations must be finite and non-negative");
    GPGMP_HIP_CHECK(hipMemcpyAsync(x_.get(), host, n * sizeof(float), hipMemcpyHostToDevice, stream_));
    GPGMP_HIP_CHECK(hipStreamSynchronize(stream_));
}

void DetEngine::get_state(float *host) {
    GPGMP_HIP_CHECK(hipMemcpyAsync(host, x_.get(), (size_t)S_ * A_ * sizeof(float), hipMemcpyDeviceToHost, stream_));
    GPGMP_HIP_CHECK(hipStreamSynchronize(stream_));
}

// Solver::doReactions for the deterministic types (Solver.cpp:250-266): the integrator over
// dt on every cell, in place; the seed is not used.
void DetEngine::do_reactions(float dt) {
    if (!reacts() || !(dt > 0.f)) return;
    if (aqss())
        GPGMP_HIP_CHECK(launch_det_aqss(x_.get(), (int)A_, m_.table, dt, (flags_ & GPGMP_SOLVER_AQSS_PADE_FIX) ? 1 : 0,
                                        stream_));
    else
        GPGMP_HIP_CHECK(launch_det_rk4(x_.get(), (int)A_, m_.table, dt, stream_));
}

// One Jacobi stencil step of species s into the scratch slice, copied back (the other
// species' slices stay as they are).
void DetEngine::hom_step(int s, float dt) {
    const float h = m_.desc.physical_length / (float)m_.desc.width;
    float *xs = x_.get() + (size_t)s * A_;
    GPGMP_HIP_CHECK(launch_det_hom_diffuse(xs, tmp_.get(), W_, H_, dt * m_.desc.diffusion[s] / (h * h), stream_));
    GPGMP_HIP_CHECK(hipMemcpyAsync(xs, tmp_.get(), A_ * sizeof(float), hipMemcpyDeviceToDevice, stream_));
}

void DetEngine::inh_step(int s, int axis, float dt) {
    const float h = axis ? m_.cell_height : m_.cell_width;
    float *xs = x_.get() + (size_t)s * A_;
    const float *D = (axis ? Dy_.get() : Dx_.get()) + (size_t)s * A_;
    const float *r = (axis ? ry_.get() : rx_.get()) + (size_t)s * A_;
    GPGMP_HIP_CHECK(launch_det_inhom_sweep(xs, tmp_.get(), D, r, W_, H_, axis, dt / (h * h), dt / (2.f * h), stream_));
    GPGMP_HIP_CHECK(hipMemcpyAsync(xs, tmp_.get(), A_ * sizeof(float), hipMemcpyDeviceToDevice, stream_));
}

// HomogeneousSolver::diffuse / InhomogeneousSolver::diffuse for the deterministic types.
void DetEngine::diffuse(int s, float t) {
    if (!m_.inhomogeneous()) {
        if (flags_ & GPGMP_SOLVER_REFERENCE_QUIRKS) {
            // homogeneous_deterministic_diffuse moves every species (Deterministic.cl:32)
            for (int k = 0; k < S_; ++k)
                if (m_.desc.diffusion[k] > 0) hom_step(k, dtx_[s]);
        } else if (m_.desc.diffusion[s] > 0) {
            hom_step(s, dtx_[s]);
        }
        nx_[s] += dtx_[s];
        ny_[s] = nx_[s];
        return;
    }
    if (nx_[s] <= t) {
        inh_step(s, 0, dtx_[s]);
        nx_[s] += dtx_[s];
    }
    if (ny_[s] <= t) {
        inh_step(s, 1, dty_[s]);
        ny_[s] += dty_[s];
    }
}

std::vector<float> DetEngine::next_diffusion_time() const {
    std::vector<float> out(S_);
    for (int s = 0; s < S_; ++s) out[s] = (!m_.inhomogeneous()) ? nx_[s] : std::min(nx_[s], ny_[s]);
    return out;
}

// DiffusionModel::oclSolve's loop (DiffusionModel.cpp:1479-1603) with a deterministic solver:
// the same multi-rate schedule as the stochastic loop, reactions over each iteration's dt,
// then the due species' diffusion; stream-ordered, one synchronisation at the end.
void DetEngine::run(int steps) {
    GPGMP_REQUIRE(steps >= 0, GPGMP_E_INVALID, "steps must be non-negative");
    for (int n = 0; n < steps; ++n) {
        const std::vector<float> next = next_diffusion_time();
        float mindt = next[0];
        int minspec = 0;
        for (int i = 1; i < S_; ++i)
            if (next[i] < mindt) {
                mindt = next[i];
                minspec = i;
            }
        do_reactions(mindt - sim_time_);
        for (int i = 0; i < S_; ++i) {
            if (!(i == minspec || next[i] <= mindt)) continue;
            if (!(next[i] > 0)) continue;
            diffuse(i, mindt);
        }
        sim_time_ = mindt;
        ++iterations_;
    }
    GPGMP_HIP_CHECK(hipStreamSynchronize(stream_));
}

}  // namespace gpgmp
