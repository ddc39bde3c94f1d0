// Deterministic kernels for gfx950 (fp32): the reference's diffusion stencils and RK4
// reaction integrator (Deterministic.cl), and the fused Schnakenberg
// reaction-diffusion step with its hand-written adjoint.
//
// All stencils are Jacobi (read one buffer, write another): the reference updates
// in place, so its results depend on work-item scheduling (Deterministic.cl:33-62).
// Boundaries are the reference's mirror rule: the missing neighbour of an edge cell
// is the cell on the other side (Deterministic.cl:40-53).
#include <hip/hip_runtime.h>

#include "gpgmp_numeric.h"
#include "stochastic_types.h"

namespace gpgmp {

__device__ __forceinline__ int mlo(int i) { return i > 0 ? i - 1 : i + 1; }
__device__ __forceinline__ int mhi(int i, int n) { return i < n - 1 ? i + 1 : i - 1; }

// homogeneous_deterministic_diffuse (Deterministic.cl:17-65) for one species:
// out = c + dt*D*(sum of 4 mirror neighbours - 4c)/h^2
__global__ void k_det_hom_diffuse(const float *__restrict__ in, float *__restrict__ out, int W, int H,
                                  float coef /* dt*D/h^2 */) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y;
    if (x >= W) return;
    const size_t yc = (size_t)y * W;
    const float l = in[yc + mlo(x)], r = in[yc + mhi(x, W)];
    const float t = in[(size_t)mlo(y) * W + x], b = in[(size_t)mhi(y, H) * W + x];
    const float c = in[yc + x];
    out[yc + x] = c + coef * (t + b + l + r - 4.f * c);
}

// inhomogeneous_deterministic_diffuseX/Y (Deterministic.cl:71-136) for one species and axis:
// out = c + dt*[(Dc)_+ + (Dc)_- - 2(Dc)]/h^2 - dt*[(rc)_+ - (rc)_-]/(2h)
__global__ void k_det_inhom_sweep(const float *__restrict__ in, float *__restrict__ out,
                                  const float *__restrict__ D, const float *__restrict__ r, int W, int H,
                                  int axis, float dt_h2, float dt_2h) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y;
    if (x >= W) return;
    const size_t i = (size_t)y * W + x;
    size_t lo, hi;
    if (axis == 0) {
        lo = (size_t)y * W + mlo(x);
        hi = (size_t)y * W + mhi(x, W);
    } else {
        lo = (size_t)mlo(y) * W + x;
        hi = (size_t)mhi(y, H) * W + x;
    }
    const float c = in[i];
    const float lap = in[lo] * D[lo] + in[hi] * D[hi] - 2.f * c * D[i];
    const float drift = r[hi] * in[hi] - r[lo] * in[lo];
    out[i] = c + dt_h2 * lap - dt_2h * drift;
}

// deterministic_performReaction_rk4 (Deterministic.cl:179-238): 20 RK4 substeps of
// dt/20 with the deterministic mass-action law k*X*X*Y (Reaction.cpp:145-162), each
// followed by a clamp at 0.  In place: each cell touches only itself.
template <int S>
__device__ __forceinline__ void rk4_rhs(const RxTable &T, const float (&x)[S], float (&d)[S]) {
#pragma unroll
    for (int s = 0; s < S; ++s) d[s] = 0.f;
    for (int i = 0; i < T.R; ++i) {
        float law = (float)T.lit_dbl[i];
        for (int f = 0; f < T.nf[i]; ++f) {
            float v = x[0];
#pragma unroll
            for (int k = 1; k < S; ++k) v = (T.fs[i][f] == k) ? x[k] : v;
            law *= v;
        }
#pragma unroll
        for (int s = 0; s < S; ++s) d[s] += (float)T.net[i][s] * law;
    }
}

template <int S>
__global__ void k_det_rk4(float *__restrict__ state, int A, RxTable T, float dt) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= A) return;
    float x[S], k1[S], k2[S], k3[S], k4[S], tmp[S];
#pragma unroll
    for (int s = 0; s < S; ++s) x[s] = state[(size_t)s * A + c];
    const float h = dt / 20.f;
    for (int n = 0; n < 20; ++n) {
        rk4_rhs<S>(T, x, k1);
#pragma unroll
        for (int s = 0; s < S; ++s) tmp[s] = x[s] + h / 2.f * k1[s];
        rk4_rhs<S>(T, tmp, k2);
#pragma unroll
        for (int s = 0; s < S; ++s) tmp[s] = x[s] + h / 2.f * k2[s];
        rk4_rhs<S>(T, tmp, k3);
#pragma unroll
        for (int s = 0; s < S; ++s) tmp[s] = x[s] + h * k3[s];
        rk4_rhs<S>(T, tmp, k4);
#pragma unroll
        for (int s = 0; s < S; ++s) {
            x[s] += h / 6.f * (k1[s] + 2.f * k2[s] + 2.f * k3[s] + k4[s]);
            if (x[s] < 0.f) x[s] = 0.f;
        }
    }
#pragma unroll
    for (int s = 0; s < S; ++s) state[(size_t)s * A + c] = x[s];
}

// ---------------------------------------------------------------------------------------
// Fused Schnakenberg reaction-diffusion step (no reference counterpart; the diffusion is
// the FPE stencil of Deterministic.cl:71-136 with zero drift, both axes in one Jacobi
// update; the reactions are the concentration-form mass-action laws):
//   u' = u + dt*( cu*(sum_nbr D_n u_n - 4 D u) + k1 - k4 u + k3 u^2 v )
//   v' = v + dt*( cv*(sum_nbr D_n v_n - 4 D v) + k2 - k3 u^2 v ),  c_s = D0_s/h^2.
// A 256-thread block owns a 64x16 tile; D*u and D*v of the tile and its one-cell halo
// are staged in LDS once, so every product is computed once per cell.
struct RddK {
    float dt, cu, cv, k1, k2, k3, k4;
};

constexpr int RT_X = 64, RT_Y = 16;
constexpr int RL_X = RT_X + 2, RL_Y = RT_Y + 2;

__global__ __launch_bounds__(256) void k_rdd_step(const float *__restrict__ u, const float *__restrict__ v,
                                                  const float *__restrict__ D, float *__restrict__ un,
                                                  float *__restrict__ vn, int W, int H, RddK p) {
    __shared__ float su[RL_Y][RL_X];
    __shared__ float sv[RL_Y][RL_X];
    const int x0 = blockIdx.x * RT_X, y0 = blockIdx.y * RT_Y;
    for (int idx = threadIdx.x; idx < RL_X * RL_Y; idx += blockDim.x) {
        const int ly = idx / RL_X, lx = idx - ly * RL_X;
        int gx = x0 + lx - 1, gy = y0 + ly - 1;
        gx = gx < 0 ? 1 : (gx >= W ? W - 2 : gx);
        gy = gy < 0 ? 1 : (gy >= H ? H - 2 : gy);
        const size_t g = (size_t)gy * W + gx;
        const float d = D[g];
        su[ly][lx] = d * u[g];
        sv[ly][lx] = d * v[g];
    }
    __syncthreads();
    const int tx = threadIdx.x & (RT_X - 1);
    const int ty0 = (threadIdx.x >> 6) * 4;
    const int gx = x0 + tx;
    if (gx >= W) return;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int ty = ty0 + k;
        const int gy = y0 + ty;
        if (gy >= H) break;
        const size_t g = (size_t)gy * W + gx;
        const float uc = u[g], vc = v[g];
        const int lx = tx + 1, ly = ty + 1;
        const float lu = su[ly][lx - 1] + su[ly][lx + 1] + su[ly - 1][lx] + su[ly + 1][lx] - 4.f * su[ly][lx];
        const float lv = sv[ly][lx - 1] + sv[ly][lx + 1] + sv[ly - 1][lx] + sv[ly + 1][lx] - 4.f * sv[ly][lx];
        const float uuv = uc * uc * vc;
        const float fu = p.k1 - p.k4 * uc + p.k3 * uuv;
        const float fv = p.k2 - p.k3 * uuv;
        un[g] = uc + p.dt * (p.cu * lu + fu);
        vn[g] = vc + p.dt * (p.cv * lv + fv);
    }
}

// Adjoint of k_rdd_step.  With lambda = dL/d(state after the step), c = state before it:
//   T(l)(j)  = sum over c in {j-1, j+1} (per axis, inside the grid) of w(c) l(c), where
//              w(c) = 2 for the first/last cell of the axis (both its mirror neighbours
//              are j) and 1 elsewhere -- the transpose of the mirror neighbour sum;
//   lu_prev  = lu + dt*cu*D*(T(lu) - 4 lu) + dt*((-k4 + 2 k3 u v) lu - 2 k3 u v lv)
//   lv_prev  = lv + dt*cv*D*(T(lv) - 4 lv) + dt*(k3 u^2 lu - k3 u^2 lv)
//   dL/dD   += dt*( cu*u*(T(lu) - 4 lu) + cv*v*(T(lv) - 4 lv) )
//   dL/dk   += dt*( sum lu, sum lv, sum u^2 v (lu - lv), -sum u lu )
// Block sums of dL/dk go to a per-tile double accumulator (same tile every step, so
// there is no race and the summation order is fixed).
__global__ __launch_bounds__(256) void k_rdd_adjoint(
    const float *__restrict__ lu, const float *__restrict__ lv, const float *__restrict__ u,
    const float *__restrict__ v, const float *__restrict__ D, float *__restrict__ lun, float *__restrict__ lvn,
    float *__restrict__ gD, double *__restrict__ gk_tiles, int W, int H, RddK p) {
    __shared__ float slu[RL_Y][RL_X];
    __shared__ float slv[RL_Y][RL_X];
    __shared__ double red[4][4];
    const int x0 = blockIdx.x * RT_X, y0 = blockIdx.y * RT_Y;
    for (int idx = threadIdx.x; idx < RL_X * RL_Y; idx += blockDim.x) {
        const int ly = idx / RL_X, lx = idx - ly * RL_X;
        const int gx = x0 + lx - 1, gy = y0 + ly - 1;
        float a = 0.f, b = 0.f;
        if (gx >= 0 && gx < W && gy >= 0 && gy < H) {
            const size_t g = (size_t)gy * W + gx;
            a = lu[g];
            b = lv[g];
        }
        slu[ly][lx] = a;
        slv[ly][lx] = b;
    }
    __syncthreads();
    const int tx = threadIdx.x & (RT_X - 1);
    const int ty0 = (threadIdx.x >> 6) * 4;
    const int gx = x0 + tx;
    float g0 = 0.f, g1 = 0.f, g2 = 0.f, g3 = 0.f;
    if (gx < W) {
        const float wl = (gx - 1 == 0 || gx - 1 == W - 1) ? 2.f : 1.f;
        const float wr = (gx + 1 == 0 || gx + 1 == W - 1) ? 2.f : 1.f;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int ty = ty0 + k;
            const int gy = y0 + ty;
            if (gy >= H) break;
            const size_t g = (size_t)gy * W + gx;
            const int lx = tx + 1, ly = ty + 1;
            const float wt = (gy - 1 == 0 || gy - 1 == H - 1) ? 2.f : 1.f;
            const float wb = (gy + 1 == 0 || gy + 1 == H - 1) ? 2.f : 1.f;
            const float l0 = slu[ly][lx], m0 = slv[ly][lx];
            const float tu = wl * slu[ly][lx - 1] + wr * slu[ly][lx + 1] + wt * slu[ly - 1][lx] + wb * slu[ly + 1][lx];
            const float tv = wl * slv[ly][lx - 1] + wr * slv[ly][lx + 1] + wt * slv[ly - 1][lx] + wb * slv[ly + 1][lx];
            const float au = tu - 4.f * l0, av = tv - 4.f * m0;
            const float uc = u[g], vc = v[g], dc = D[g];
            const float uv = uc * vc, uu = uc * uc;
            lun[g] = l0 + p.dt * p.cu * dc * au + p.dt * ((-p.k4 + 2.f * p.k3 * uv) * l0 - 2.f * p.k3 * uv * m0);
            lvn[g] = m0 + p.dt * p.cv * dc * av + p.dt * (p.k3 * uu * l0 - p.k3 * uu * m0);
            gD[g] += p.dt * (p.cu * uc * au + p.cv * vc * av);
            g0 += l0;
            g1 += m0;
            g2 += uu * vc * (l0 - m0);
            g3 -= uc * l0;
        }
    }
    // block reduction of the four dL/dk partial sums (wave64 shuffles, then LDS)
    double d0 = g0, d1 = g1, d2 = g2, d3 = g3;
    for (int off = 32; off > 0; off >>= 1) {
        d0 += __shfl_down(d0, off, 64);
        d1 += __shfl_down(d1, off, 64);
        d2 += __shfl_down(d2, off, 64);
        d3 += __shfl_down(d3, off, 64);
    }
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (lane == 0) {
        red[wave][0] = d0;
        red[wave][1] = d1;
        red[wave][2] = d2;
        red[wave][3] = d3;
    }
    __syncthreads();
    if (threadIdx.x < 4) {
        const double s = red[0][threadIdx.x] + red[1][threadIdx.x] + red[2][threadIdx.x] + red[3][threadIdx.x];
        const size_t tile = (size_t)blockIdx.y * gridDim.x + blockIdx.x;
        gk_tiles[tile * 4 + threadIdx.x] += (double)p.dt * s;
    }
}

// dL/dk = sum over tiles (fixed order) -> float[4]
__global__ void k_rdd_gk_reduce(const double *__restrict__ gk_tiles, int ntiles, float *__restrict__ gk) {
    __shared__ double red[256];
    const int j = blockIdx.x;  // parameter index 0..3
    double s = 0.0;
    for (int t = threadIdx.x; t < ntiles; t += blockDim.x) s += gk_tiles[(size_t)t * 4 + j];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int off = blockDim.x / 2; off > 0; off >>= 1) {
        if (threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) gk[j] = (float)red[0];
}

__global__ void k_fill_f32(float *__restrict__ p, size_t n, float v) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void k_fill_f64(double *__restrict__ p, size_t n, double v) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

}  // namespace gpgmp

// ---------------------------------------------------------------------------------------
#include "deterministic_launch.h"

namespace gpgmp {

hipError_t launch_det_hom_diffuse(const float *in, float *out, int W, int H, float coef, hipStream_t s) {
    hipLaunchKernelGGL(k_det_hom_diffuse, dim3((W + 255) / 256, H), dim3(256), 0, s, in, out, W, H, coef);
    return hipGetLastError();
}

hipError_t launch_det_inhom_sweep(const float *in, float *out, const float *D, const float *r, int W, int H,
                                  int axis, float dt_h2, float dt_2h, hipStream_t s) {
    hipLaunchKernelGGL(k_det_inhom_sweep, dim3((W + 255) / 256, H), dim3(256), 0, s, in, out, D, r, W, H, axis,
                       dt_h2, dt_2h);
    return hipGetLastError();
}

hipError_t launch_det_rk4(float *state, int A, const RxTable &T, float dt, hipStream_t s) {
    const dim3 g((A + 255) / 256), b(256);
    switch (T.S) {
        case 1: hipLaunchKernelGGL(k_det_rk4<1>, g, b, 0, s, state, A, T, dt); break;
        case 2: hipLaunchKernelGGL(k_det_rk4<2>, g, b, 0, s, state, A, T, dt); break;
        case 3: hipLaunchKernelGGL(k_det_rk4<3>, g, b, 0, s, state, A, T, dt); break;
        case 4: hipLaunchKernelGGL(k_det_rk4<4>, g, b, 0, s, state, A, T, dt); break;
        case 5: hipLaunchKernelGGL(k_det_rk4<5>, g, b, 0, s, state, A, T, dt); break;
        case 6: hipLaunchKernelGGL(k_det_rk4<6>, g, b, 0, s, state, A, T, dt); break;
        case 7: hipLaunchKernelGGL(k_det_rk4<7>, g, b, 0, s, state, A, T, dt); break;
        case 8: hipLaunchKernelGGL(k_det_rk4<8>, g, b, 0, s, state, A, T, dt); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

static RddK make_k(const RddConsts &c) {
    RddK k;
    k.dt = c.dt;
    k.cu = c.cu;
    k.cv = c.cv;
    k.k1 = c.k[0];
    k.k2 = c.k[1];
    k.k3 = c.k[2];
    k.k4 = c.k[3];
    return k;
}

dim3 rdd_grid(int W, int H) { return dim3((W + RT_X - 1) / RT_X, (H + RT_Y - 1) / RT_Y); }

hipError_t launch_rdd_step(const float *u, const float *v, const float *D, float *un, float *vn, int W, int H,
                           const RddConsts &c, hipStream_t s, hipEvent_t e0, hipEvent_t e1) {
    if (e0) {
        hipExtLaunchKernelGGL(k_rdd_step, rdd_grid(W, H), dim3(256), 0, s, e0, e1, 0, u, v, D, un, vn, W, H,
                              make_k(c));
    } else {
        hipLaunchKernelGGL(k_rdd_step, rdd_grid(W, H), dim3(256), 0, s, u, v, D, un, vn, W, H, make_k(c));
    }
    return hipGetLastError();
}

hipError_t launch_rdd_adjoint(const float *lu, const float *lv, const float *u, const float *v, const float *D,
                              float *lun, float *lvn, float *gD, double *gk_tiles, int W, int H,
                              const RddConsts &c, hipStream_t s, hipEvent_t e0, hipEvent_t e1) {
    if (e0) {
        hipExtLaunchKernelGGL(k_rdd_adjoint, rdd_grid(W, H), dim3(256), 0, s, e0, e1, 0, lu, lv, u, v, D, lun,
                              lvn, gD, gk_tiles, W, H, make_k(c));
    } else {
        hipLaunchKernelGGL(k_rdd_adjoint, rdd_grid(W, H), dim3(256), 0, s, lu, lv, u, v, D, lun, lvn, gD,
                           gk_tiles, W, H, make_k(c));
    }
    return hipGetLastError();
}

hipError_t launch_rdd_gk_reduce(const double *gk_tiles, int ntiles, float *gk, hipStream_t s) {
    hipLaunchKernelGGL(k_rdd_gk_reduce, dim3(4), dim3(256), 0, s, gk_tiles, ntiles, gk);
    return hipGetLastError();
}

hipError_t launch_fill_f32(float *p, size_t n, float v, hipStream_t s) {
    if (!n) return hipSuccess;
    hipLaunchKernelGGL(k_fill_f32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, n, v);
    return hipGetLastError();
}

hipError_t launch_fill_f64(double *p, size_t n, double v, hipStream_t s) {
    if (!n) return hipSuccess;
    hipLaunchKernelGGL(k_fill_f64, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, n, v);
    return hipGetLastError();
}

int rdd_num_tiles(int W, int H) {
    const dim3 g = rdd_grid(W, H);
    return (int)(g.x * g.y);
}

}  // namespace gpgmp
