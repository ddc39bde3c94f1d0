// Deterministic solver types of the plain solver handle (fp32 concentrations).
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "host.h"

namespace gpgmp {

class DetEngine {
  public:
    DetEngine(const Model &m, uint32_t flags, hipStream_t stream);
    void set_state(const float *host);
    void get_state(float *host);
    float *state() const { return x_.get(); }
    void do_reactions(float dt);
    void diffuse(int species, float t);
    std::vector<float> next_diffusion_time() const;
    const std::vector<float> &dt_x() const { return dtx_; }
    const std::vector<float> &dt_y() const { return dty_; }
    // the host loop for `steps` more iterations (from time 0 after reset)
    void run(int steps);
    void reset() {
        nx_ = dtx_;
        ny_ = dty_;
        sim_time_ = 0.f;
        iterations_ = 0;
    }
    float sim_time() const { return sim_time_; }
    int64_t iterations() const { return iterations_; }
    hipStream_t stream() const { return stream_; }

  private:
    bool aqss() const;
    bool reacts() const;
    void hom_step(int s, float dt);
    void inh_step(int s, int axis, float dt);
    Model m_;
    uint32_t flags_;
    hipStream_t stream_;
    int W_ = 0, H_ = 0, S_ = 0;
    size_t A_ = 0;
    DevBuf<float> x_, tmp_, Dx_, Dy_, rx_, ry_;
    DevBuf<uint16_t> act_;  // active-reaction bits per cell (localized reactions only)
    DevBuf<uint32_t> prog_;  // kinetic-law programs (custom laws only)
    std::vector<float> dtx_, dty_, nx_, ny_;
    float sim_time_ = 0.f;
    int64_t iterations_ = 0;
};

}  // namespace gpgmp
