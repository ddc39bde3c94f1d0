// Stochastic GMP: per-cell time steps, the inhomogeneous hop-probability tables and
// the dispatch of k_gmp_step launches to the per-mode translation units
// (stochastic_hom.hip, stochastic_inh_x.hip, stochastic_inh_y.hip, stochastic_inh_xs.hip, which instantiate
// the kernels of stochastic_kernels.h).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "gpgmp_numeric.h"
#include "stochastic_launch.h"
#include "stochastic_types.h"

namespace gpgmp {

// Exact integer form of the test u <= p, u = u01(w) = float(w ^ 2^31) * 2^-32: u01 is
// non-decreasing in m = w ^ 2^31, so {w : u01(w) <= p} = {m <= T} with T the largest such
// m (binary search over the same u01), i.e. (int32)w <= (int32)(T ^ 2^31) (signed order of
// w is the unsigned order of m).  p < 0 or NaN admits no draw: none = 1 (the threshold
// then admits only m = 0, and the count is zeroed per cell afterwards).
__device__ __forceinline__ int32_t hop_threshold(float p, int32_t &none) {
    none = 0;
    if (!(p >= 0.f)) {
        none = 1;
        return (int32_t)0x80000000u;
    }
    uint32_t lo = 0u, hi = 0xffffffffu;  // u01 at m = lo is 0 <= p
    if (u01(hi ^ 0x80000000u) <= p) return (int32_t)(hi ^ 0x80000000u);
    while (hi - lo > 1u) {
        const uint32_t mid = lo + ((hi - lo) >> 1);
        if (u01(mid ^ 0x80000000u) <= p) lo = mid;
        else hi = mid;
    }
    return (int32_t)(lo ^ 0x80000000u);
}

// Inhomogeneous hop probabilities (StochasticInhomogeneousDriftDiffusion.cl:518-520 /
// :648-650) with the reference's float/double promotions, stored as the thresholds of
// (min(p1,p2), p2) with their none-mask (bit 0, bit 1).  They depend only on the fields
// and the sweep's dt, so they are tabulated once.
// fpe != 0: the stochastic_inhomogeneous_fpe solver's Fokker-Planck transition rates
// (computeIndividualTimestepX/Y, StochasticInhomogeneousDriftDiffusion.cl:191-197,341-347:
// D/dx^2 +- r/(2. dx), float quotient plus a double term, stored as float) times the sweep's
// dt in float (inhomogeneous_diffuseX/Y, :491-500, :626-633).
__global__ void k_inh_probs(int4 *__restrict__ out, const float *__restrict__ D,
                            const float *__restrict__ r, float dt, float dx, int n, int fpe) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float Di = D[i], ri = r[i];
    float p1, p2;
    if (fpe) {
        const float dD = Di / (dx * dx);
        const double dr = (double)ri / (2. * (double)dx);
        const float up = (float)((double)dD + dr), down = (float)((double)dD - dr);
        p1 = up * dt;
        p2 = p1 + down * dt;
    } else {
        const float rrdt = (ri * ri) * dt;
        const float dx2 = dx * dx;
        const float dxr = dx * ri;
        const float ps = (float)(1. - ((double)dt * (2. * (double)Di + (double)rrdt)) / (double)dx2);
        const double den = 2. * (double)Di + (double)rrdt;
        p1 = (float)((1. - (double)ps) * (1. + (double)dxr / den) / 2.);
        p2 = (float)((double)p1 + (1. - (double)ps) * (1. - (double)dxr / den) / 2.);
    }
    int32_t n1, n2;
    const int32_t t1 = hop_threshold((p1 <= p2) ? p1 : p2, n1);
    const int32_t t2 = hop_threshold(p2, n2);
    out[i] = make_int4(t1, t2, n1 | (n2 << 1), 0);
}

// Per-cell admissible time step of the stochastic inhomogeneous solver
// (computeIndividualTimestepX/Y, StochasticInhomogeneousDriftDiffusion.cl:130-151), in
// double, and its minimum over the cells: block minima, then one block over the
// minima (the reference reduces in a single work-group, which caps grids at 512^2,
// InhomogeneousSolver.cpp:63-85).  Min is order-independent, so the result is exact.
__device__ __forceinline__ float cell_dt(float Df, float rf, float hf) {
    const double D = Df, r = rf, h = hf;
    double ts;
    if (r == 0.0) {
        ts = h * h / (2.0 * D) * (1.0 - 2.0 / 3.0);
    } else if (D == 0.0) {
        ts = h / fabs(r);
    } else {
        const double eps = h * r / (2.0 * D);
        const double th = tanh(eps), sh = sinh(eps);
        double temp = 1.0 + (-1.0 / (th * eps) + 1.0 / (sh * sh));
        if (temp <= 0.0) temp = 1.0 / 3.0;
        ts = h / r * th * temp;
    }
    return (float)ts;
}

__device__ __forceinline__ float block_min(float v, float *smem) {
    for (int off = 32; off > 0; off >>= 1) v = fminf(v, __shfl_down(v, off, 64));
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) smem[wave] = v;
    __syncthreads();
    float m = smem[0];
    for (int w = 1; w < (int)(blockDim.x >> 6); ++w) m = fminf(m, smem[w]);
    return m;
}

__global__ __launch_bounds__(256) void k_min_dt(const float *__restrict__ D, const float *__restrict__ r, int n,
                                                float h, float *__restrict__ block_out) {
    __shared__ float sm[4];
    float v = __builtin_inff();
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        v = fminf(v, cell_dt(D[i], r[i], h));
    const float m = block_min(v, sm);
    if (threadIdx.x == 0) block_out[blockIdx.x] = m;
}

__global__ __launch_bounds__(256) void k_min_reduce(const float *__restrict__ in, int n, float *__restrict__ out) {
    __shared__ float sm[4];
    float v = __builtin_inff();
    for (int i = threadIdx.x; i < n; i += blockDim.x) v = fminf(v, in[i]);
    const float m = block_min(v, sm);
    if (threadIdx.x == 0) *out = m;
}

hipError_t launch_gmp_step(const StepLaunch &a) {
    if (a.ntiles <= 0) return hipSuccess;
    if (a.block <= 0 || a.block > 256 || (a.block % 64) != 0 || a.cpt <= 0) return hipErrorInvalidValue;
    // localized reactions, float literals or law programs: the extended instances
    const bool cl = (a.T->custom | a.T->localized | a.T->lit_is_float) != 0;
    switch (a.mode) {
        case 0: return cl ? launch_gmp_hom_cl(a) : launch_gmp_hom(a);
        case 1: return cl ? launch_gmp_inh_x_cl(a) : launch_gmp_inh_x(a);
        case 2: return launch_gmp_inh_y(a);
        case 3: return cl ? launch_gmp_inh_xs_cl(a) : launch_gmp_inh_xs(a);
        default: return hipErrorInvalidValue;
    }
}

// blocks: scratch of at least 1024 floats; out: one device float
hipError_t launch_min_dt(const float *D, const float *r, int n, float h, float *blocks, float *out,
                         hipStream_t stream) {
    const int nb = n > 0 ? std::min(1024, (n + 255) / 256) : 1;
    hipLaunchKernelGGL(k_min_dt, dim3(nb), dim3(256), 0, stream, D, r, n, h, blocks);
    hipLaunchKernelGGL(k_min_reduce, dim3(1), dim3(256), 0, stream, blocks, nb, out);
    return hipGetLastError();
}

hipError_t launch_inh_probs(int4 *out, const float *D, const float *r, float dt, float dx, int n, int fpe,
                            hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_inh_probs, dim3((n + 255) / 256), dim3(256), 0, stream, out, D, r, dt, dx, n, fpe);
    return hipGetLastError();
}

}  // namespace gpgmp
