// The model's drift/diffusivity method on the device: the generated
// computeDiffusionConstants_forSpecies_<id> kernels (Solver.cpp:659-777), one per species in
// the reference, as programs of the method compiler (law.cpp) run for every cell and
// species by one kernel.
#include <hip/hip_runtime.h>

#include "fields_launch.h"
#include "law.h"

namespace gpgmp {

// out[(f*S + s)*A + c]: field f (DiffusivityX, DiffusivityY, DriftX, DriftY) of species s in
// cell c; prog[2*(f*S+s)] is the program's offset in code, prog[2*(f*S+s)+1] its type.
// PhysicalX/Y as getPhysicalX/Y compute them (cl_include/helpers.cl:29-36):
// (Real)PhysicalModelWidth / (Real)GridModelWidth * ((Real)x - (Real)GridModelWidth/2.).
__global__ __launch_bounds__(256) void k_drift_fields(const uint32_t *__restrict__ code,
                                                      const uint32_t *__restrict__ prog, LawExtra ex, float pmw,
                                                      float pmh, float *__restrict__ out) {
    const int A = ex.W * ex.H;
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= A) return;
    const int gx = c % ex.W, gy = c / ex.W;
    LawExtra e = ex;
    e.px = (float)((double)(pmw / (float)ex.W) * ((double)(float)gx - (double)(float)ex.W / 2.));
    e.py = (float)((double)(pmh / (float)ex.H) * ((double)(float)gy - (double)(float)ex.H / 2.));
    const LawCell cell{gx, gy, 0u, 0.f};
    const int32_t none[1] = {0};
    for (int k = 0; k < 4 * ex.S; ++k) {
        const double v = law_run<int32_t, 1>(code + prog[2 * k], none, cell, &e);
        out[(size_t)k * A + c] = law_to_float(v, prog[2 * k + 1]);
    }
}

hipError_t launch_drift_fields(const uint32_t *code, const uint32_t *prog, const LawExtra &ex, float pmw,
                               float pmh, float *out, hipStream_t stream) {
    const int A = ex.W * ex.H;
    if (A <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_drift_fields, dim3((A + 255) / 256), dim3(256), 0, stream, code, prog, ex, pmw, pmh, out);
    return hipGetLastError();
}

}  // namespace gpgmp
