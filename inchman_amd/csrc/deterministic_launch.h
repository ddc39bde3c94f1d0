// Host-side launch interface of the deterministic kernels.
#pragma once

#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include "stochastic_types.h"

namespace gpgmp {

struct RddConsts {
    float dt, cu, cv;  // c_s = D0_s / h^2
    float k[4];
    const float *kdev = nullptr;  // non-null: the kernels read k1..k4 from this device array instead
};

hipError_t launch_det_hom_diffuse(const float *in, float *out, int W, int H, float coef, hipStream_t s);
hipError_t launch_det_inhom_sweep(const float *in, float *out, const float *D, const float *r, int W, int H,
                                  int axis, float dt_h2, float dt_2h, hipStream_t s);
hipError_t launch_det_rk4(float *state, int A, const RxTable &T, float dt, hipStream_t s);
// alpha-QSS reactions over dt per cell; fixed = 1: the predictor's Pade factor in the corrector
hipError_t launch_det_aqss(float *state, int A, const RxTable &T, float dt, int fixed, hipStream_t s);
// ns = 1 .. 4 steps per launch (sides >= ns + 1); umid/vmid (may be NULL) receive the
// intermediate states: the state after intermediate step j (1 .. ns - 1) at umid/vmid +
// (j - 1) * mid_stride (mid_stride may be 0 for ns = 2)
hipError_t launch_rdd_step(const float *u, const float *v, const float *D, float *umid, float *vmid, float *un,
                           float *vn, int W, int H, const RddConsts &c, int ns, hipStream_t s,
                           hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr, size_t mid_stride = 0);
// adjoint of ns steps: u1/v1 = state before the later step, u0/v0 = before the earlier one
hipError_t launch_rdd_adjoint(const float *lu, const float *lv, const float *u1, const float *v1, const float *u0,
                              const float *v0, const float *D, float *lun, float *lvn, float *gD, double *gk_tiles,
                              int W, int H, const RddConsts &c, int ns, hipStream_t s);
// adjoint of ns = 3 or 4 steps: us[k]/vs[k] = the state before the k-th step back (k = 0:
// the latest)
hipError_t launch_rdd_adjoint_n(const float *lu, const float *lv, const float *const *us, const float *const *vs,
                                const float *D, float *lun, float *lvn, float *gD, double *gk_tiles, int W, int H,
                                const RddConsts &c, int ns, hipStream_t s);
hipError_t launch_rdd_gk_reduce(const double *gk_tiles, int ntiles, float *gk, hipStream_t s);
hipError_t launch_fill_f32(float *p, size_t n, float v, hipStream_t s);
hipError_t launch_fill_f64(double *p, size_t n, double v, hipStream_t s);
int rdd_num_tiles(int W, int H);

}  // namespace gpgmp
