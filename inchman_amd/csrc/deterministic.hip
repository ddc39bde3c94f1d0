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
#include "law.h"
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

// Streaming forms of the two stencils for W % 4 == 0 (the HBM-bound case): a thread owns
// 4 consecutive cells of a row (16-byte loads, non-temporal 16-byte stores: the output
// is not read again by this launch) and marches down TY rows keeping the rows above and
// below in registers; the left/right neighbours of its 4 cells come from the adjacent
// lanes (global loads only at the wave's ends).  Short strips keep the most loads in
// flight: measured on 8192^2 ping-pong steps, TY = 2 (homogeneous, Y sweep) and 1 (X
// sweep) run at 0.96-0.99x the device-to-device copy rate.  Same arithmetic, same bits,
// as the one-cell kernels above.

typedef float f4v __attribute__((ext_vector_type(4)));
__device__ __forceinline__ float4 ld4(const float *__restrict__ p) { return *reinterpret_cast<const float4 *>(p); }
__device__ __forceinline__ void st4nt(float *__restrict__ p, float4 v) {
    const f4v t = {v.x, v.y, v.z, v.w};
    __builtin_nontemporal_store(t, reinterpret_cast<f4v *>(p));
}

// left neighbour of the thread's first cell and right neighbour of its last, mirrored at
// the grid edges (lanes past the row end hold a copy of the last 4 cells and store nothing)
__device__ __forceinline__ void row_nbrs(const float *__restrict__ rowp, const float4 &c, int xs, bool first,
                                         bool last, float &l, float &r) {
    const int lane = (int)(threadIdx.x & 63);
    l = __shfl_up(c.w, 1, 64);
    r = __shfl_down(c.x, 1, 64);
    if (lane == 0 && !first) l = rowp[xs - 1];
    if (lane == 63 && !last) r = rowp[xs + 4];
    if (first) l = c.y;
    if (last) r = c.z;
}

template <int TY>
__global__ __launch_bounds__(256) void k_det_hom_diffuse4(const float *__restrict__ in, float *__restrict__ out,
                                                          int W, int H, float coef) {
    const int x0 = (blockIdx.x * blockDim.x + threadIdx.x) * 4;
    const bool ok = x0 < W;
    const int xs = ok ? x0 : W - 4;
    const int y0 = blockIdx.y * TY;
    const int y1 = min(y0 + TY, H);
    const bool first = xs == 0, last = xs + 4 == W;
    float4 up = ld4(in + (size_t)mlo(y0) * W + xs);
    float4 cu = ld4(in + (size_t)y0 * W + xs);
    for (int y = y0; y < y1; ++y) {
        const size_t row = (size_t)y * W;
        const float4 dn = ld4(in + (size_t)mhi(y, H) * W + xs);
        float l, r;
        row_nbrs(in + row, cu, xs, first, last, l, r);
        float4 o;
        o.x = cu.x + coef * (up.x + dn.x + l + cu.y - 4.f * cu.x);
        o.y = cu.y + coef * (up.y + dn.y + cu.x + cu.z - 4.f * cu.y);
        o.z = cu.z + coef * (up.z + dn.z + cu.y + cu.w - 4.f * cu.z);
        o.w = cu.w + coef * (up.w + dn.w + cu.z + r - 4.f * cu.w);
        if (ok) st4nt(out + row + xs, o);
        up = cu;
        cu = dn;
    }
}

__device__ __forceinline__ float inh_cell(float c, float dc, float clo, float dlo, float rlo, float chi, float dhi,
                                          float rhi, float dt_h2, float dt_2h) {
    const float lap = clo * dlo + chi * dhi - 2.f * c * dc;
    const float drift = rhi * chi - rlo * clo;
    return c + dt_h2 * lap - dt_2h * drift;
}

template <int TY>
__global__ __launch_bounds__(256) void k_det_inhom_sweep4(const float *__restrict__ in, float *__restrict__ out,
                                                          const float *__restrict__ D, const float *__restrict__ r,
                                                          int W, int H, int axis, float dt_h2, float dt_2h) {
    const int x0 = (blockIdx.x * blockDim.x + threadIdx.x) * 4;
    const bool ok = x0 < W;
    const int xs = ok ? x0 : W - 4;
    const int y0 = blockIdx.y * TY;
    const int y1 = min(y0 + TY, H);
    if (axis == 0) {
        const bool first = xs == 0, last = xs + 4 == W;
        for (int y = y0; y < y1; ++y) {
            const size_t row = (size_t)y * W;
            const float4 c = ld4(in + row + xs), d = ld4(D + row + xs), q = ld4(r + row + xs);
            float cl, cr, dl, dr, ql, qr;
            row_nbrs(in + row, c, xs, first, last, cl, cr);
            row_nbrs(D + row, d, xs, first, last, dl, dr);
            row_nbrs(r + row, q, xs, first, last, ql, qr);
            float4 o;
            o.x = inh_cell(c.x, d.x, cl, dl, ql, c.y, d.y, q.y, dt_h2, dt_2h);
            o.y = inh_cell(c.y, d.y, c.x, d.x, q.x, c.z, d.z, q.z, dt_h2, dt_2h);
            o.z = inh_cell(c.z, d.z, c.y, d.y, q.y, c.w, d.w, q.w, dt_h2, dt_2h);
            o.w = inh_cell(c.w, d.w, c.z, d.z, q.z, cr, dr, qr, dt_h2, dt_2h);
            if (ok) st4nt(out + row + xs, o);
        }
    } else {
        const size_t ru = (size_t)mlo(y0) * W + xs, rc = (size_t)y0 * W + xs;
        float4 cu = ld4(in + ru), du = ld4(D + ru), qu = ld4(r + ru);
        float4 cc = ld4(in + rc), dc = ld4(D + rc), qc = ld4(r + rc);
        for (int y = y0; y < y1; ++y) {
            const size_t rd = (size_t)mhi(y, H) * W + xs;
            const float4 cd = ld4(in + rd), dd = ld4(D + rd), qd = ld4(r + rd);
            float4 o;
            o.x = inh_cell(cc.x, dc.x, cu.x, du.x, qu.x, cd.x, dd.x, qd.x, dt_h2, dt_2h);
            o.y = inh_cell(cc.y, dc.y, cu.y, du.y, qu.y, cd.y, dd.y, qd.y, dt_h2, dt_2h);
            o.z = inh_cell(cc.z, dc.z, cu.z, du.z, qu.z, cd.z, dd.z, qd.z, dt_h2, dt_2h);
            o.w = inh_cell(cc.w, dc.w, cu.w, du.w, qu.w, cd.w, dd.w, qd.w, dt_h2, dt_2h);
            if (ok) st4nt(out + (size_t)y * W + xs, o);
            cu = cc; du = dc; qu = qc;
            cc = cd; dc = dd; qc = qd;
        }
    }
}

// deterministic_performReaction_rk4 (Deterministic.cl:179-238): 20 RK4 substeps of
// dt/20 with the deterministic mass-action law k*X*X*Y (Reaction.cpp:145-162), each
// followed by a clamp at 0.  In place: each cell touches only itself.
// A localized reaction's law is multiplied by the cell's mask value w (0 or 1,
// Solver.cpp:947-948): act holds the cell's active-reaction bits (RxTable::act).
// A model with custom laws takes each species' generated derivative expression as a program
// (law.h; RxTable::det_rk).
template <int S, bool CL>
__device__ __forceinline__ void rk4_rhs(const RxTable &T, const float (&x)[S], float (&d)[S], const LawCell &lc) {
    if constexpr (CL) {
#pragma unroll
        for (int s = 0; s < S; ++s) d[s] = law_call<S>(T.prog + T.det_rk[s], T.det_rk_t[s], x, lc);
        return;
    }
    const uint32_t act = lc.act;
#pragma unroll
    for (int s = 0; s < S; ++s) d[s] = 0.f;
    for (int i = 0; i < T.R; ++i) {
        float law = (float)T.lit_dbl[i];
        if ((T.localized >> i) & 1u) law = law * (((act >> i) & 1u) ? 1.f : 0.f);
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

// The same right-hand side for compact tables (at most kCompactReactions reactions of at
// most kCompactFactors factors) in straight-line code: an unused factor is 1 (law * 1 is
// law), and a reaction at or past R leaves d unchanged (selected, not added), so the float
// operations on d are the table walk's.
template <int S>
__device__ __forceinline__ void rk4_rhs_compact(const RxTable &T, const float (&x)[S], float (&d)[S]) {
#pragma unroll
    for (int s = 0; s < S; ++s) d[s] = 0.f;
#pragma unroll
    for (int i = 0; i < kCompactReactions; ++i) {
        float law = (float)T.lit_dbl[i];
#pragma unroll
        for (int f = 0; f < kCompactFactors; ++f) {
            float v = x[0];
#pragma unroll
            for (int k = 1; k < S; ++k) v = (T.fs[i][f] == k) ? x[k] : v;
            law *= (f < (int)T.nf[i]) ? v : 1.f;
        }
#pragma unroll
        for (int s = 0; s < S; ++s) {
            const float ds = d[s] + (float)T.net[i][s] * law;
            d[s] = (i < T.R) ? ds : d[s];
        }
    }
}

template <int S, bool COMPACT, bool CL>
__device__ __forceinline__ void rk4_rhs_t(const RxTable &T, const float (&x)[S], float (&d)[S], const LawCell &act) {
    if constexpr (COMPACT) rk4_rhs_compact<S>(T, x, d);
    else rk4_rhs<S, CL>(T, x, d, act);
}

// The 20 substep increments are accumulated with a compensation term per species (Kahan
// summation; the build keeps float operations unfused and unreordered): an increment of a
// few ulps of the state would otherwise be rounded 20 times per call (h k = 1.25e-4 on a
// state of 64, ulp 7.6e-6: 16.4 ulps rounded to 16 each substep, a 2.4% loss of the call's
// increment), and the call ends within one rounding of the exact sum of its increments.  The
// right-hand sides see the rounded state, as the reference's do.
template <int S, bool COMPACT, bool CL>
__device__ __forceinline__ void rk4_cell(float *__restrict__ state, int A, int c, const RxTable &T, float dt) {
    const LawCell act{c % T.grid_w, c / T.grid_w, T.localized ? (uint32_t)T.act[c] : 0xffffu, 0.f};
    float x[S], k1[S], k2[S], k3[S], k4[S], tmp[S], comp[S];
#pragma unroll
    for (int s = 0; s < S; ++s) {
        x[s] = state[(size_t)s * A + c];
        comp[s] = 0.f;
    }
    const float h = dt / 20.f;
    for (int n = 0; n < 20; ++n) {
        rk4_rhs_t<S, COMPACT, CL>(T, x, k1, act);
#pragma unroll
        for (int s = 0; s < S; ++s) tmp[s] = x[s] + h / 2.f * k1[s];
        rk4_rhs_t<S, COMPACT, CL>(T, tmp, k2, act);
#pragma unroll
        for (int s = 0; s < S; ++s) tmp[s] = x[s] + h / 2.f * k2[s];
        rk4_rhs_t<S, COMPACT, CL>(T, tmp, k3, act);
#pragma unroll
        for (int s = 0; s < S; ++s) tmp[s] = x[s] + h * k3[s];
        rk4_rhs_t<S, COMPACT, CL>(T, tmp, k4, act);
#pragma unroll
        for (int s = 0; s < S; ++s) {
            const float inc = h / 6.f * (k1[s] + 2.f * k2[s] + 2.f * k3[s] + k4[s]);
            const float y = inc - comp[s];
            const float t = x[s] + y;
            comp[s] = (t - x[s]) - y;
            x[s] = t;
            if (x[s] < 0.f) {
                x[s] = 0.f;
                comp[s] = 0.f;
            }
        }
    }
#pragma unroll
    for (int s = 0; s < S; ++s) state[(size_t)s * A + c] = x[s];
}

// CL: the instance with the law programs of custom laws (RxTable::det_custom).
template <int S, bool CL>
__global__ void k_det_rk4(float *__restrict__ state, int A, RxTable T, float dt) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= A) return;
    if (!CL && T.compact) rk4_cell<S, true, false>(state, A, c, T, dt);
    else rk4_cell<S, false, CL>(state, A, c, T, dt);
}

// deterministic_performReaction_aqss (Deterministic.cl:291-522): Mott & Oran's alpha-QSS
// predictor-corrector over dtg, one thread per cell, every per-cell array of the reference
// (d_ys, d_qs, d_rtaus, d_y1, d_scrarray and the state in LDS) in registers.  Production
// q and destruction d per species as deterministic_rhs_ode_aqss computes them
// (Solver.cpp:960-1010): 0. + |n|. * k * X * Y + ... in double, in reaction order, rounded
// to float.  Float variables with the reference's double-literal promotions statement by
// statement; fixed = 1 puts the predictor's Pade factor into the corrector, 0 keeps the
// reference's `180.*+rtaub*(...)` (Deterministic.cl:423).  itermax = 2, so the stability
// factor is the constant 0.01.  At most AQSS_MAX_ITER inner passes per cell; a cell that
// reaches the cap is set to NaN.
constexpr int AQSS_MAX_ITER = 1 << 16;

// A localized reaction's term is |n|. * w * <law> (Solver.cpp:987-1000): the double chain
// starts at |n| * w for the cell's mask value w (0 or 1).
template <int S, bool CL>
__device__ __forceinline__ void aqss_rhs(const RxTable &T, const float (&y)[S], float (&q)[S], float (&d)[S],
                                         const LawCell &lc) {
    if constexpr (CL) {  // the generated sums as programs (RxTable::det_q, det_d)
#pragma unroll
        for (int s = 0; s < S; ++s) {
            q[s] = law_call<S>(T.prog + T.det_q[s], T.det_q_t[s], y, lc);
            d[s] = law_call<S>(T.prog + T.det_d[s], T.det_d_t[s], y, lc);
        }
        return;
    }
    const uint32_t act = lc.act;
    double qs[S], ds[S];
#pragma unroll
    for (int s = 0; s < S; ++s) qs[s] = ds[s] = 0.;
    for (int i = 0; i < T.R; ++i) {
#pragma unroll
        for (int s = 0; s < S; ++s) {
            const int n = T.net[i][s];
            if (n == 0) continue;
            double term = (double)(n > 0 ? n : -n);
            if ((T.localized >> i) & 1u) term = term * (((act >> i) & 1u) ? 1.0 : 0.0);
            term = term * T.lit_dbl[i];
            for (int f = 0; f < T.nf[i]; ++f) {
                float v = y[0];
#pragma unroll
                for (int k = 1; k < S; ++k) v = (T.fs[i][f] == k) ? y[k] : v;
                term = term * (double)v;
            }
            if (n > 0)
                qs[s] = qs[s] + term;
            else
                ds[s] = ds[s] + term;
        }
    }
#pragma unroll
    for (int s = 0; s < S; ++s) {
        q[s] = (float)qs[s];
        d[s] = (float)ds[s];
    }
}

template <int S, bool CL>
__global__ __launch_bounds__(256) void k_det_aqss(float *__restrict__ state, int A, RxTable T, float dtg,
                                                  int fixed) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= A) return;
    const LawCell act{c % T.grid_w, c / T.grid_w, T.localized ? (uint32_t)T.act[c] : 0xffffu, 0.f};
    const float epsmin = 1e-2f, sqreps = 0.5f, epscl = 100.f, epsmax = 10.f, tfd = 1.000008f, dtmin = 1e-15f,
                ymin = 1e-20f;
    float y[S], q[S], d[S], ys[S], qsv[S], rtaus[S], y1[S], scr[S];
    float tn = 0.f;
#pragma unroll
    for (int i = 0; i < S; ++i) y[i] = fmaxf(state[(size_t)i * A + c], ymin);
    aqss_rhs<S, CL>(T, y, q, d, act);
    float scrtch = 1e-25f;
#pragma unroll
    for (int i = 0; i < S; ++i) {
        const float ascr = fabsf(q[i]);
        const float scr2 = copysignf(1.f / y[i], 0.1f * epsmin * ascr - d[i]);
        const float scr1 = scr2 * d[i];
        scrtch = fmaxf(fmaxf(scr1, -fabsf(ascr - d[i]) * scr2), scrtch);
    }
    float dt = fminf(sqreps / scrtch, dtg);
    float ts, dtcurrent, eps;
    int iters = 0;
    bool capped = false;
    while (dtg > tn * tfd && !capped) {
        ts = tn;
        dtcurrent = dt;
#pragma unroll
        for (int i = 0; i < S; ++i) {
            ys[i] = y[i];
            qsv[i] = q[i];
            rtaus[i] = dtcurrent * d[i] / y[i];
        }
        while (dtg > tn * tfd) {
            if (++iters > AQSS_MAX_ITER) {
                capped = true;
                break;
            }
#pragma unroll
            for (int i = 0; i < S; ++i) {
                const float rtaui = dtcurrent * d[i] / y[i];
                const float alpha = (float)((180. + rtaui * (60. + rtaui * (11. + rtaui))) /
                                            (360. + rtaui * (60. + rtaui * (12. + rtaui))));
                scr[i] = (float)((q[i] - d[i]) / (1. + alpha * rtaui));
            }
            for (int iteration = 0; iteration < 2; iteration++) {
#pragma unroll
                for (int i = 0; i < S; ++i) y[i] = fmaxf(ys[i] + dt * scr[i], ymin);
                if (iteration == 0) {
                    tn = ts + dt;
#pragma unroll
                    for (int i = 0; i < S; ++i) y1[i] = y[i];
                }
                aqss_rhs<S, CL>(T, y, q, d, act);
                eps = 1e-10f;
#pragma unroll
                for (int i = 0; i < S; ++i) {
                    const float rtaub = (float)(0.5 * (rtaus[i] + dt * d[i] / y[i]));
                    const double px = 60. + rtaub * (11. + rtaub);
                    const double pd = 360. + rtaub * (60. + rtaub * (12. + rtaub));
                    // `180.*+rtaub*(...)` parses as (180. * rtaub) * (...)
                    const float alpha = fixed ? (float)((180. + rtaub * px) / pd) : (float)((180. * rtaub) * px / pd);
                    const float qt = (float)(qsv[i] * (1. - alpha) + q[i] * alpha);
                    const float pb = rtaub / dt;
                    scr[i] = (float)((qt - ys[i] * pb) / (1. + alpha * rtaub));
                }
            }
#pragma unroll
            for (int i = 0; i < S; ++i) {
                const float scr2 = fmaxf(ys[i] + dt * scr[i], 0.f);
                float scr1 = fabsf(scr2 - y1[i]);
                y[i] = fmaxf(scr2, ymin);
                if (0.25 * (ys[i] + y[i]) > ymin) {
                    scr1 = scr1 / y[i];
                    eps = fmaxf(0.5f * (scr1 + fminf(fabsf(q[i] - d[i]) / (q[i] + d[i] + 1e-30f), scr1)), eps);
                }
            }
            eps = eps * epscl;
            if (dt <= dtmin + 1e-16 * tn) {
                dt = dtg - ts;
                dt = fminf(dtmin, fabsf(dt));
            }
            const float stab = 0.01f;
            if (eps <= epsmax && stab <= 1.) {
                if (dtg <= tn * tfd) break;
            } else {
                tn = ts;
            }
            const float rteps = sqrtf(eps);
            float dto = dt;
            dt = fminf(fminf(dt * (1.0f / rteps + .005f), tfd * (dtg - tn)), dto / (stab + 0.001f));
            if (eps <= epsmax && stab <= 1) break;
            dto = dt / dto;
#pragma unroll
            for (int i = 0; i < S; ++i) rtaus[i] *= dto;
        }
        if (!capped) aqss_rhs<S, CL>(T, y, q, d, act);
    }
#pragma unroll
    for (int i = 0; i < S; ++i) state[(size_t)i * A + c] = capped ? __builtin_nanf("") : y[i];
}

// ---------------------------------------------------------------------------------------
// Fused Schnakenberg reaction-diffusion step (no reference counterpart; the diffusion is
// the FPE stencil of Deterministic.cl:71-136 with zero drift, both axes in one Jacobi
// update; the reactions are the concentration-form mass-action laws):
//   u' = u + dt*( cu*(sum_nbr D_n u_n - 4 D u) + k1 - k4 u + k3 u^2 v )
//   v' = v + dt*( cv*(sum_nbr D_n v_n - 4 D v) + k2 - k3 u^2 v ),  c_s = D0_s/h^2.
// A 256-thread block owns a 64x16 tile and stages u, v, D of the tile plus an NS-cell
// halo in LDS; NS (1..4) steps per launch (temporal blocking) recompute the shrinking
// rings of the intermediate steps, so each cell's inputs are read from HBM once per NS
// steps.  The halo outside the grid is the reference's mirror image (g -> -g,
// g -> 2(W-1)-g), and the neighbour sum is taken as (l + r) + (t + b), which is invariant
// under both reflections, so a mirrored halo cell evolves bit-identically to the cell it
// mirrors and NS steps per launch equal NS one-step launches bit for bit (sides >= NS + 1).
struct RddK {
    float dt, cu, cv, k1, k2, k3, k4;
};

constexpr int RT_X = 64, RT_Y = 16;

// rate constants k1..k4 from a device array (rates set on the device by the caller, so the
// host never waits for them)
__device__ __forceinline__ void rdd_rates(RddK &p, const float *kdev) {
    p.k1 = kdev[0];
    p.k2 = kdev[1];
    p.k3 = kdev[2];
    p.k4 = kdev[3];
}

// mirror index for g in [-(n-1), 2(n-1)]; halo cells of a partial tile that lie further
// out feed no in-grid result and are only clamped to a valid address
__device__ __forceinline__ int reflect(int g, int n) {
    g = g < 0 ? -g : g;
    g = g >= n ? 2 * (n - 1) - g : g;
    return g < 0 ? 0 : (g >= n ? n - 1 : g);
}

__device__ __forceinline__ void rdd_cell(float uc, float vc, float dc, float dul, float dur, float dut, float dub,
                                         float dvl, float dvr, float dvt, float dvb, const RddK &p, float &un,
                                         float &vn) {
    const float lu = ((dul + dur) + (dut + dub)) - 4.f * (dc * uc);
    const float lv = ((dvl + dvr) + (dvt + dvb)) - 4.f * (dc * vc);
    const float uuv = uc * uc * vc;
    const float fu = p.k1 - p.k4 * uc + p.k3 * uuv;
    const float fv = p.k2 - p.k3 * uuv;
    un = uc + p.dt * (p.cu * lu + fu);
    vn = vc + p.dt * (p.cv * lv + fv);
}

template <int NS>
__global__ __launch_bounds__(256) void k_rdd_step(const float *__restrict__ u, const float *__restrict__ v,
                                                  const float *__restrict__ D, float *__restrict__ umid,
                                                  float *__restrict__ vmid, float *__restrict__ un,
                                                  float *__restrict__ vn, int W, int H, RddK p,
                                                  const float *__restrict__ kdev, size_t mid_stride) {
    if (kdev) rdd_rates(p, kdev);
    constexpr int LX = RT_X + 2 * NS, LY = RT_Y + 2 * NS;
    __shared__ float sd[LY][LX];
    __shared__ float su[2][LY][LX];
    __shared__ float sv[2][LY][LX];
    const int x0 = blockIdx.x * RT_X, y0 = blockIdx.y * RT_Y;
    for (int idx = threadIdx.x; idx < LX * LY; idx += blockDim.x) {
        const int ly = idx / LX, lx = idx - ly * LX;
        const int gx = reflect(x0 + lx - NS, W), gy = reflect(y0 + ly - NS, H);
        const size_t g = (size_t)gy * W + gx;
        sd[ly][lx] = D[g];
        su[0][ly][lx] = u[g];
        sv[0][ly][lx] = v[g];
    }
    __syncthreads();
    // intermediate steps on the shrinking region (ring r = s .. L-1-s)
#pragma unroll
    for (int s = 1; s < NS; ++s) {
        const int a = (s - 1) & 1, b = s & 1;
        const int nx = LX - 2 * s, ny = LY - 2 * s;
        for (int idx = threadIdx.x; idx < nx * ny; idx += blockDim.x) {
            const int ly = idx / nx + s, lx = idx % nx + s;
            float uo, vo;
            rdd_cell(su[a][ly][lx], sv[a][ly][lx], sd[ly][lx], sd[ly][lx - 1] * su[a][ly][lx - 1],
                     sd[ly][lx + 1] * su[a][ly][lx + 1], sd[ly - 1][lx] * su[a][ly - 1][lx],
                     sd[ly + 1][lx] * su[a][ly + 1][lx], sd[ly][lx - 1] * sv[a][ly][lx - 1],
                     sd[ly][lx + 1] * sv[a][ly][lx + 1], sd[ly - 1][lx] * sv[a][ly - 1][lx],
                     sd[ly + 1][lx] * sv[a][ly + 1][lx], p, uo, vo);
            su[b][ly][lx] = uo;
            sv[b][ly][lx] = vo;
            if (umid && lx >= NS && lx < NS + RT_X && ly >= NS && ly < NS + RT_Y) {
                const int gx = x0 + lx - NS, gy = y0 + ly - NS;
                if (gx < W && gy < H) {  // the state after intermediate step s
                    umid[(size_t)(s - 1) * mid_stride + (size_t)gy * W + gx] = uo;
                    vmid[(size_t)(s - 1) * mid_stride + (size_t)gy * W + gx] = vo;
                }
            }
        }
        __syncthreads();
    }
    constexpr int a = (NS - 1) & 1;
    const int tx = threadIdx.x & (RT_X - 1);
    const int ty0 = (threadIdx.x >> 6) * 4;
    const int gx = x0 + tx;
    if (gx >= W) return;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int ty = ty0 + k;
        const int gy = y0 + ty;
        if (gy >= H) break;
        const int lx = tx + NS, ly = ty + NS;
        float uo, vo;
        rdd_cell(su[a][ly][lx], sv[a][ly][lx], sd[ly][lx], sd[ly][lx - 1] * su[a][ly][lx - 1],
                 sd[ly][lx + 1] * su[a][ly][lx + 1], sd[ly - 1][lx] * su[a][ly - 1][lx],
                 sd[ly + 1][lx] * su[a][ly + 1][lx], sd[ly][lx - 1] * sv[a][ly][lx - 1],
                 sd[ly][lx + 1] * sv[a][ly][lx + 1], sd[ly - 1][lx] * sv[a][ly - 1][lx],
                 sd[ly + 1][lx] * sv[a][ly + 1][lx], p, uo, vo);
        const size_t g = (size_t)gy * W + gx;
        un[g] = uo;
        vn[g] = vo;
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
// NS = 2 walks two steps back per launch: lambda of the intermediate step is formed on
// the tile plus a one-cell ring (zero outside the grid, where lambda vanishes), from
// lambda staged with a two-cell halo.  Block sums of dL/dk go, step by step, to a
// per-tile double accumulator (same tile every launch: no race, fixed order), and
// dL/dD is accumulated in the same order as by NS = 1 launches.
__device__ __forceinline__ float tw(int c, int n) { return (c == 0 || c == n - 1) ? 2.f : 1.f; }

__device__ __forceinline__ void adj_cell(float l0, float m0, float ll, float lr, float lt, float lb, float ml,
                                         float mr, float mt, float mb, float wl, float wr, float wt, float wb,
                                         float uc, float vc, float dc, const RddK &p, float &lo, float &mo,
                                         float &gd, float (&gk)[4]) {
    const float tu = (wl * ll + wr * lr) + (wt * lt + wb * lb);
    const float tv = (wl * ml + wr * mr) + (wt * mt + wb * mb);
    const float au = tu - 4.f * l0, av = tv - 4.f * m0;
    const float uv = uc * vc, uu = uc * uc;
    lo = l0 + p.dt * p.cu * dc * au + p.dt * ((-p.k4 + 2.f * p.k3 * uv) * l0 - 2.f * p.k3 * uv * m0);
    mo = m0 + p.dt * p.cv * dc * av + p.dt * (p.k3 * uu * l0 - p.k3 * uu * m0);
    gd = p.dt * (p.cu * uc * au + p.cv * vc * av);
    gk[0] += l0;
    gk[1] += m0;
    gk[2] += uu * vc * (l0 - m0);
    gk[3] -= uc * l0;
}

__device__ __forceinline__ void tile_gk_add(float (&gk)[4], double (*red)[4], double *gk_tiles, float dt) {
    double d[4] = {gk[0], gk[1], gk[2], gk[3]};
#pragma unroll
    for (int j = 0; j < 4; ++j)
        for (int off = 32; off > 0; off >>= 1) d[j] += __shfl_down(d[j], off, 64);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (lane == 0)
#pragma unroll
        for (int j = 0; j < 4; ++j) red[wave][j] = d[j];
    __syncthreads();
    if (threadIdx.x < 4) {
        const double s = red[0][threadIdx.x] + red[1][threadIdx.x] + red[2][threadIdx.x] + red[3][threadIdx.x];
        const size_t tile = (size_t)blockIdx.y * gridDim.x + blockIdx.x;
        gk_tiles[tile * 4 + threadIdx.x] += (double)dt * s;
    }
    __syncthreads();
}

// states: u1/v1 = state before the later step (NS = 1: the only one), u0/v0 = state
// before the earlier step (NS = 2)
template <int NS>
__global__ __launch_bounds__(256) void k_rdd_adjoint(
    const float *__restrict__ lu, const float *__restrict__ lv, const float *__restrict__ u1,
    const float *__restrict__ v1, const float *__restrict__ u0, const float *__restrict__ v0,
    const float *__restrict__ D, float *__restrict__ lun, float *__restrict__ lvn, float *__restrict__ gD,
    double *__restrict__ gk_tiles, int W, int H, RddK p, const float *__restrict__ kdev) {
    if (kdev) rdd_rates(p, kdev);
    constexpr int LX = RT_X + 2 * NS, LY = RT_Y + 2 * NS;
    constexpr int MX = RT_X + 2, MY = RT_Y + 2;  // ring of the intermediate lambda
    __shared__ float slu[LY][LX];
    __shared__ float slv[LY][LX];
    __shared__ float mlu[NS == 2 ? MY : 1][NS == 2 ? MX : 1];
    __shared__ float mlv[NS == 2 ? MY : 1][NS == 2 ? MX : 1];
    __shared__ double red[4][4];
    const int x0 = blockIdx.x * RT_X, y0 = blockIdx.y * RT_Y;
    for (int idx = threadIdx.x; idx < LX * LY; idx += blockDim.x) {
        const int ly = idx / LX, lx = idx - ly * LX;
        const int gx = x0 + lx - NS, gy = y0 + ly - NS;
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
    float gd_late[4] = {0.f, 0.f, 0.f, 0.f};
    if (NS == 2) {
        // later step.  Core cells by their owner threads, exactly as an NS = 1 launch
        // would (same partial sums of dL/dk, same dL/dD term) ...
        float gk[4] = {0.f, 0.f, 0.f, 0.f};
        if (gx < W) {
            const float wl = tw(gx - 1, W), wr = tw(gx + 1, W);
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const int ty = ty0 + k;
                const int gy = y0 + ty;
                if (gy >= H) break;
                const size_t g = (size_t)gy * W + gx;
                const int lx = tx + NS, ly = ty + NS;
                float lo, mo;
                adj_cell(slu[ly][lx], slv[ly][lx], slu[ly][lx - 1], slu[ly][lx + 1], slu[ly - 1][lx], slu[ly + 1][lx],
                         slv[ly][lx - 1], slv[ly][lx + 1], slv[ly - 1][lx], slv[ly + 1][lx], wl, wr, tw(gy - 1, H),
                         tw(gy + 1, H), u1[g], v1[g], D[g], p, lo, mo, gd_late[k], gk);
                mlu[ty + 1][tx + 1] = lo;
                mlv[ty + 1][tx + 1] = mo;
            }
        }
        // ... and the one-cell ring around the tile (no dL/dD or dL/dk: those cells
        // belong to the neighbouring tiles); zero outside the grid.
        constexpr int RING = 2 * (RT_X + 2) + 2 * RT_Y;
        for (int idx = threadIdx.x; idx < RING; idx += blockDim.x) {
            int mx, my;
            if (idx < RT_X + 2) { mx = idx; my = 0; }
            else if (idx < 2 * (RT_X + 2)) { mx = idx - (RT_X + 2); my = RT_Y + 1; }
            else if (idx < 2 * (RT_X + 2) + RT_Y) { mx = 0; my = idx - 2 * (RT_X + 2) + 1; }
            else { mx = RT_X + 1; my = idx - 2 * (RT_X + 2) - RT_Y + 1; }
            const int cx = x0 + mx - 1, cy = y0 + my - 1;
            float lo = 0.f, mo = 0.f;
            if (cx >= 0 && cx < W && cy >= 0 && cy < H) {
                const size_t g = (size_t)cy * W + cx;
                const int lx = mx + 1, ly = my + 1;
                float gd;
                float gtmp[4] = {0.f, 0.f, 0.f, 0.f};
                adj_cell(slu[ly][lx], slv[ly][lx], slu[ly][lx - 1], slu[ly][lx + 1], slu[ly - 1][lx], slu[ly + 1][lx],
                         slv[ly][lx - 1], slv[ly][lx + 1], slv[ly - 1][lx], slv[ly + 1][lx], tw(cx - 1, W),
                         tw(cx + 1, W), tw(cy - 1, H), tw(cy + 1, H), u1[g], v1[g], D[g], p, lo, mo, gd, gtmp);
            }
            mlu[my][mx] = lo;
            mlv[my][mx] = mo;
        }
        // core cells beyond the grid edge of a partial tile hold no lambda
        for (int idx = threadIdx.x; idx < RT_X * RT_Y; idx += blockDim.x) {
            const int cy = idx / RT_X, cx = idx - cy * RT_X;
            if (x0 + cx >= W || y0 + cy >= H) {
                mlu[cy + 1][cx + 1] = 0.f;
                mlv[cy + 1][cx + 1] = 0.f;
            }
        }
        tile_gk_add(gk, red, gk_tiles, p.dt);  // contains the barrier that publishes mlu/mlv
    }
    float gk[4] = {0.f, 0.f, 0.f, 0.f};
    if (gx < W) {
        const float wl = tw(gx - 1, W), wr = tw(gx + 1, W);
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int ty = ty0 + k;
            const int gy = y0 + ty;
            if (gy >= H) break;
            const size_t g = (size_t)gy * W + gx;
            const float wt = tw(gy - 1, H), wb = tw(gy + 1, H);
            float lo, mo, gd;
            if (NS == 2) {
                const int lx = tx + 1, ly = ty + 1;
                adj_cell(mlu[ly][lx], mlv[ly][lx], mlu[ly][lx - 1], mlu[ly][lx + 1], mlu[ly - 1][lx], mlu[ly + 1][lx],
                         mlv[ly][lx - 1], mlv[ly][lx + 1], mlv[ly - 1][lx], mlv[ly + 1][lx], wl, wr, wt, wb, u0[g], v0[g],
                         D[g], p, lo, mo, gd, gk);
                gD[g] = (gD[g] + gd_late[k]) + gd;
            } else {
                const int lx = tx + 1, ly = ty + 1;
                adj_cell(slu[ly][lx], slv[ly][lx], slu[ly][lx - 1], slu[ly][lx + 1], slu[ly - 1][lx], slu[ly + 1][lx],
                         slv[ly][lx - 1], slv[ly][lx + 1], slv[ly - 1][lx], slv[ly + 1][lx], wl, wr, wt, wb, u1[g], v1[g],
                         D[g], p, lo, mo, gd, gk);
                gD[g] += gd;
            }
            lun[g] = lo;
            lvn[g] = mo;
        }
    }
    tile_gk_add(gk, red, gk_tiles, p.dt);
}

// NS (3, 4) steps back per launch, generalising the NS = 2 walk: lambda after the latest
// step is staged with an NS-cell halo; step k back (k = 0: the latest) forms lambda before
// it on the tile plus a ring of NS - 1 - k cells (ring cells: no dL/dD or dL/dk, zero
// outside the grid), with the state before that step, st.u[k], st.v[k].  The tile's own
// cells are computed by their owner threads exactly as a one-step launch would (the same
// partial sums of dL/dk, added to the tile's accumulator step by step; dL/dD summed in the
// steps' order), so the result equals NS one-step launches bit for bit.
struct RddStates {
    const float *u[4], *v[4];
};

template <int NS>
__global__ __launch_bounds__(256) void k_rdd_adjoint_n(const float *__restrict__ lu, const float *__restrict__ lv,
                                                       RddStates st, const float *__restrict__ D,
                                                       float *__restrict__ lun, float *__restrict__ lvn,
                                                       float *__restrict__ gD, double *__restrict__ gk_tiles, int W,
                                                       int H, RddK p, const float *__restrict__ kdev) {
    if (kdev) rdd_rates(p, kdev);
    constexpr int LX = RT_X + 2 * NS, LY = RT_Y + 2 * NS;
    __shared__ float sl[2][2][LY][LX];  // [level parity][lambda_u, lambda_v]
    __shared__ double red[4][4];
    const int x0 = blockIdx.x * RT_X, y0 = blockIdx.y * RT_Y;
    for (int idx = threadIdx.x; idx < LX * LY; idx += blockDim.x) {
        const int ly = idx / LX, lx = idx - ly * LX;
        const int gx = x0 + lx - NS, gy = y0 + ly - NS;
        float a = 0.f, b = 0.f;
        if (gx >= 0 && gx < W && gy >= 0 && gy < H) {
            const size_t g = (size_t)gy * W + gx;
            a = lu[g];
            b = lv[g];
        }
        sl[0][0][ly][lx] = a;
        sl[0][1][ly][lx] = b;
    }
    __syncthreads();
    const int tx = threadIdx.x & (RT_X - 1);
    const int ty0 = (threadIdx.x >> 6) * 4;
    const int gx = x0 + tx;
    float gdv[NS][4];
#pragma unroll
    for (int k = 0; k < NS; ++k) {
        const int c = k & 1, n = c ^ 1;
        const int r = NS - 1 - k;  // ring of this step's lambda
        const float *__restrict__ us = st.u[k];
        const float *__restrict__ vs = st.v[k];
        float gk[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int j = 0; j < 4; ++j) gdv[k][j] = 0.f;
        if (gx < W) {  // the tile's cells, by their owner threads
            const float wl = tw(gx - 1, W), wr = tw(gx + 1, W);
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int ty = ty0 + j;
                const int gy = y0 + ty;
                if (gy >= H) break;
                const size_t g = (size_t)gy * W + gx;
                const int lx = tx + NS, ly = ty + NS;
                float lo, mo;
                adj_cell(sl[c][0][ly][lx], sl[c][1][ly][lx], sl[c][0][ly][lx - 1], sl[c][0][ly][lx + 1],
                         sl[c][0][ly - 1][lx], sl[c][0][ly + 1][lx], sl[c][1][ly][lx - 1], sl[c][1][ly][lx + 1],
                         sl[c][1][ly - 1][lx], sl[c][1][ly + 1][lx], wl, wr, tw(gy - 1, H), tw(gy + 1, H), us[g], vs[g],
                         D[g], p, lo, mo, gdv[k][j], gk);
                if (k == NS - 1) {
                    float acc = gD[g];
#pragma unroll
                    for (int kk = 0; kk < NS; ++kk) acc = acc + gdv[kk][j];
                    gD[g] = acc;
                    lun[g] = lo;
                    lvn[g] = mo;
                } else {
                    sl[n][0][ly][lx] = lo;
                    sl[n][1][ly][lx] = mo;
                }
            }
        }
        if (r > 0) {
            const int RX = RT_X + 2 * r, RY = RT_Y + 2 * r;
            for (int idx = threadIdx.x; idx < RX * RY; idx += blockDim.x) {
                const int ry = idx / RX, rx = idx - ry * RX;
                const int cxl = rx - r, cyl = ry - r;  // relative to the tile's corner
                if (cxl >= 0 && cxl < RT_X && cyl >= 0 && cyl < RT_Y) {
                    // a tile cell beyond the grid edge of a partial tile holds no lambda
                    if (x0 + cxl >= W || y0 + cyl >= H) {
                        sl[n][0][cyl + NS][cxl + NS] = 0.f;
                        sl[n][1][cyl + NS][cxl + NS] = 0.f;
                    }
                    continue;
                }
                const int cx = x0 + cxl, cy = y0 + cyl;
                const int lx = cxl + NS, ly = cyl + NS;
                float lo = 0.f, mo = 0.f;
                if (cx >= 0 && cx < W && cy >= 0 && cy < H) {
                    const size_t g = (size_t)cy * W + cx;
                    float gd;
                    float gtmp[4] = {0.f, 0.f, 0.f, 0.f};
                    adj_cell(sl[c][0][ly][lx], sl[c][1][ly][lx], sl[c][0][ly][lx - 1], sl[c][0][ly][lx + 1],
                             sl[c][0][ly - 1][lx], sl[c][0][ly + 1][lx], sl[c][1][ly][lx - 1], sl[c][1][ly][lx + 1],
                             sl[c][1][ly - 1][lx], sl[c][1][ly + 1][lx], tw(cx - 1, W), tw(cx + 1, W), tw(cy - 1, H),
                             tw(cy + 1, H), us[g], vs[g], D[g], p, lo, mo, gd, gtmp);
                }
                sl[n][0][ly][lx] = lo;
                sl[n][1][ly][lx] = mo;
            }
        }
        tile_gk_add(gk, red, gk_tiles, p.dt);  // contains the barriers that publish this level
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
    const bool vec = (W % 4 == 0) && ((reinterpret_cast<uintptr_t>(in) | reinterpret_cast<uintptr_t>(out)) % 16 == 0);
    if (vec)
        hipLaunchKernelGGL(k_det_hom_diffuse4<2>, dim3((W / 4 + 255) / 256, (H + 1) / 2), dim3(256), 0, s, in, out, W,
                           H, coef);
    else
        hipLaunchKernelGGL(k_det_hom_diffuse, dim3((W + 255) / 256, H), dim3(256), 0, s, in, out, W, H, coef);
    return hipGetLastError();
}

hipError_t launch_det_inhom_sweep(const float *in, float *out, const float *D, const float *r, int W, int H,
                                  int axis, float dt_h2, float dt_2h, hipStream_t s) {
    const uintptr_t al = reinterpret_cast<uintptr_t>(in) | reinterpret_cast<uintptr_t>(out) |
                         reinterpret_cast<uintptr_t>(D) | reinterpret_cast<uintptr_t>(r);
    if (W % 4 == 0 && al % 16 == 0 && axis == 0)
        hipLaunchKernelGGL(k_det_inhom_sweep4<1>, dim3((W / 4 + 255) / 256, H), dim3(256), 0, s, in, out, D, r, W, H,
                           axis, dt_h2, dt_2h);
    else if (W % 4 == 0 && al % 16 == 0)
        hipLaunchKernelGGL(k_det_inhom_sweep4<2>, dim3((W / 4 + 255) / 256, (H + 1) / 2), dim3(256), 0, s, in, out, D,
                           r, W, H, axis, dt_h2, dt_2h);
    else
        hipLaunchKernelGGL(k_det_inhom_sweep, dim3((W + 255) / 256, H), dim3(256), 0, s, in, out, D, r, W, H, axis,
                           dt_h2, dt_2h);
    return hipGetLastError();
}

hipError_t launch_det_aqss(float *state, int A, const RxTable &T, float dt, int fixed, hipStream_t s) {
    const dim3 g((A + 255) / 256), b(256);
    const bool cl = T.det_custom != 0;
    switch (T.S) {
        case 1: hipLaunchKernelGGL((cl ? k_det_aqss<1, true> : k_det_aqss<1, false>), g, b, 0, s, state, A, T, dt, fixed); break;
        case 2: hipLaunchKernelGGL((cl ? k_det_aqss<2, true> : k_det_aqss<2, false>), g, b, 0, s, state, A, T, dt, fixed); break;
        case 3: hipLaunchKernelGGL((cl ? k_det_aqss<3, true> : k_det_aqss<3, false>), g, b, 0, s, state, A, T, dt, fixed); break;
        case 4: hipLaunchKernelGGL((cl ? k_det_aqss<4, true> : k_det_aqss<4, false>), g, b, 0, s, state, A, T, dt, fixed); break;
        case 5: hipLaunchKernelGGL((cl ? k_det_aqss<5, true> : k_det_aqss<5, false>), g, b, 0, s, state, A, T, dt, fixed); break;
        case 6: hipLaunchKernelGGL((cl ? k_det_aqss<6, true> : k_det_aqss<6, false>), g, b, 0, s, state, A, T, dt, fixed); break;
        case 7: hipLaunchKernelGGL((cl ? k_det_aqss<7, true> : k_det_aqss<7, false>), g, b, 0, s, state, A, T, dt, fixed); break;
        case 8: hipLaunchKernelGGL((cl ? k_det_aqss<8, true> : k_det_aqss<8, false>), g, b, 0, s, state, A, T, dt, fixed); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_det_rk4(float *state, int A, const RxTable &T, float dt, hipStream_t s) {
    const dim3 g((A + 255) / 256), b(256);
    const bool cl = T.det_custom != 0;
    switch (T.S) {
        case 1: hipLaunchKernelGGL((cl ? k_det_rk4<1, true> : k_det_rk4<1, false>), g, b, 0, s, state, A, T, dt); break;
        case 2: hipLaunchKernelGGL((cl ? k_det_rk4<2, true> : k_det_rk4<2, false>), g, b, 0, s, state, A, T, dt); break;
        case 3: hipLaunchKernelGGL((cl ? k_det_rk4<3, true> : k_det_rk4<3, false>), g, b, 0, s, state, A, T, dt); break;
        case 4: hipLaunchKernelGGL((cl ? k_det_rk4<4, true> : k_det_rk4<4, false>), g, b, 0, s, state, A, T, dt); break;
        case 5: hipLaunchKernelGGL((cl ? k_det_rk4<5, true> : k_det_rk4<5, false>), g, b, 0, s, state, A, T, dt); break;
        case 6: hipLaunchKernelGGL((cl ? k_det_rk4<6, true> : k_det_rk4<6, false>), g, b, 0, s, state, A, T, dt); break;
        case 7: hipLaunchKernelGGL((cl ? k_det_rk4<7, true> : k_det_rk4<7, false>), g, b, 0, s, state, A, T, dt); break;
        case 8: hipLaunchKernelGGL((cl ? k_det_rk4<8, true> : k_det_rk4<8, false>), g, b, 0, s, state, A, T, dt); break;
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

hipError_t launch_rdd_step(const float *u, const float *v, const float *D, float *umid, float *vmid, float *un,
                           float *vn, int W, int H, const RddConsts &c, int ns, hipStream_t s, hipEvent_t e0,
                           hipEvent_t e1, size_t mid_stride) {
    const RddK k = make_k(c);
    const dim3 g = rdd_grid(W, H), b(256);
    if (ns < 1 || ns > 4 || (ns > 2 && umid && mid_stride == 0)) return hipErrorInvalidValue;
    auto kern = (ns == 4) ? k_rdd_step<4> : (ns == 3) ? k_rdd_step<3> : (ns == 2) ? k_rdd_step<2> : k_rdd_step<1>;
    if (e0)
        hipExtLaunchKernelGGL(kern, g, b, 0, s, e0, e1, 0, u, v, D, umid, vmid, un, vn, W, H, k, c.kdev, mid_stride);
    else hipLaunchKernelGGL(kern, g, b, 0, s, u, v, D, umid, vmid, un, vn, W, H, k, c.kdev, mid_stride);
    return hipGetLastError();
}

hipError_t launch_rdd_adjoint(const float *lu, const float *lv, const float *u1, const float *v1, const float *u0,
                              const float *v0, const float *D, float *lun, float *lvn, float *gD, double *gk_tiles,
                              int W, int H, const RddConsts &c, int ns, hipStream_t s) {
    const RddK k = make_k(c);
    const dim3 g = rdd_grid(W, H), b(256);
    if (ns == 2)
        hipLaunchKernelGGL(k_rdd_adjoint<2>, g, b, 0, s, lu, lv, u1, v1, u0, v0, D, lun, lvn, gD, gk_tiles, W, H, k,
                           c.kdev);
    else
        hipLaunchKernelGGL(k_rdd_adjoint<1>, g, b, 0, s, lu, lv, u1, v1, u0, v0, D, lun, lvn, gD, gk_tiles, W, H, k,
                           c.kdev);
    return hipGetLastError();
}

hipError_t launch_rdd_adjoint_n(const float *lu, const float *lv, const float *const *us, const float *const *vs,
                                const float *D, float *lun, float *lvn, float *gD, double *gk_tiles, int W, int H,
                                const RddConsts &c, int ns, hipStream_t s) {
    if (ns < 3 || ns > 4) return hipErrorInvalidValue;
    const RddK k = make_k(c);
    const dim3 g = rdd_grid(W, H), b(256);
    RddStates st;
    for (int j = 0; j < 4; ++j) {
        st.u[j] = j < ns ? us[j] : nullptr;
        st.v[j] = j < ns ? vs[j] : nullptr;
    }
    if (ns == 4)
        hipLaunchKernelGGL(k_rdd_adjoint_n<4>, g, b, 0, s, lu, lv, st, D, lun, lvn, gD, gk_tiles, W, H, k, c.kdev);
    else
        hipLaunchKernelGGL(k_rdd_adjoint_n<3>, g, b, 0, s, lu, lv, st, D, lun, lvn, gD, gk_tiles, W, H, k, c.kdev);
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
