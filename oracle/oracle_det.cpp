// CPU oracle: deterministic stencils (FP64, Jacobi) and the reaction-drift-diffusion
// forward/adjoint.  TEST INFRASTRUCTURE ONLY (see gpgmp_oracle.h).
#include "gpgmp_oracle.h"

#include <algorithm>
#include <cmath>
#include <functional>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {
inline int mirror_lo(int i) { return i > 0 ? i - 1 : i + 1; }
inline int mirror_hi(int i, int n) { return i < n - 1 ? i + 1 : i - 1; }
int nthr(int n) {
#ifdef _OPENMP
    return n > 0 ? n : omp_get_max_threads();
#else
    (void)n;
    return 1;
#endif
}
}  // namespace

// homogeneous_deterministic_diffuse (Deterministic.cl:17-65): c += dt*D*(sum of the 4
// mirror-BC neighbours - 4c)/h^2 for every species.  The reference updates in place
// (neighbour reads race with writes); this oracle is Jacobi (double-buffered) on purpose.
// Pinned by: decay of the discrete-Laplacian eigenmode cos(pi k x / (W-1)) (closed form).
extern "C" void oracle_det_hom_diffuse(double *state, int W, int H, int S, const double *D,
                                       double dt, double h) {
    const size_t A = (size_t)W * H;
    std::vector<double> out(A);
    for (int s = 0; s < S; ++s) {
        double *c = state + s * A;
        for (int y = 0; y < H; ++y)
            for (int x = 0; x < W; ++x) {
                const double l = c[mirror_lo(x) + (size_t)y * W], r = c[mirror_hi(x, W) + (size_t)y * W];
                const double t = c[x + (size_t)mirror_lo(y) * W], b = c[x + (size_t)mirror_hi(y, H) * W];
                const double here = c[x + (size_t)y * W];
                out[x + (size_t)y * W] = here + dt * D[s] * (t + b + l + r - 4.0 * here) / (h * h);
            }
        for (size_t i = 0; i < A; ++i) c[i] = out[i];
    }
}

// inhomogeneous_deterministic_diffuseX/Y (Deterministic.cl:71-136): Ito/FPE form
// c += dt*[(Dc)_{+} + (Dc)_{-} - 2(Dc)]/h^2 - dt*[(rc)_{+} - (rc)_{-}]/(2h), mirror BC,
// Jacobi here.  Pinned by: conservation of sum(c) for zero-flux interiors, the drift
// mean r*t and the variance 2 D t of a spreading Gaussian.
extern "C" void oracle_det_inhom_sweep(double *state, int W, int H, int species, int axis,
                                       double dt, const double *Dax, const double *rax, double h) {
    const size_t A = (size_t)W * H;
    double *c = state + species * A;
    const double *D = Dax + species * A, *r = rax + species * A;
    std::vector<double> out(A);
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            size_t i = x + (size_t)y * W, lo, hi;
            if (axis == 0) {
                lo = mirror_lo(x) + (size_t)y * W;
                hi = mirror_hi(x, W) + (size_t)y * W;
            } else {
                lo = x + (size_t)mirror_lo(y) * W;
                hi = x + (size_t)mirror_hi(y, H) * W;
            }
            const double lap = c[lo] * D[lo] + c[hi] * D[hi] - 2.0 * c[i] * D[i];
            const double drift = r[hi] * c[hi] - r[lo] * c[lo];
            out[i] = c[i] + dt * lap / (h * h) - dt * drift / (2.0 * h);
        }
    for (size_t i = 0; i < A; ++i) c[i] = out[i];
}

// deterministic_performReaction_rk4 (Deterministic.cl:143-238): 20 classical RK4
// substeps of dt/20 per call with the deterministic mass-action law
// k*X*X*Y (Reaction.cpp:145-162) and a clamp at 0 after each substep.
// Pinned by: the closed form of 0->A (k), A->0 (g): A(t) = k/g + (A0-k/g) e^{-g t}.
extern "C" void oracle_det_rk4(double *state, int W, int H, const oracle_reaction_table *t,
                               double dt) {
    const size_t A = (size_t)W * H;
    const int S = t->num_species, R = t->num_reactions;
    size_t cell = 0;
    // a localized reaction's rate is multiplied by the cell's mask value (Solver.cpp:947-948)
    std::function<void(const double *, double *)> rhs = [&](const double *x, double *d) {
        for (int s = 0; s < S; ++s) d[s] = 0.0;
        for (int i = 0; i < R; ++i) {
            double law = t->lit_dbl[i];
            if ((t->localized >> i) & 1u) law *= ((t->act[cell] >> i) & 1u) ? 1.0 : 0.0;
            for (int f = 0; f < t->n_factors[i]; ++f) law *= x[t->factor_species[i][f]];
            for (int s = 0; s < S; ++s)
                if (t->net[i][s] != 0) d[s] += (double)t->net[i][s] * law;
        }
    };
    const int nmax = 20;
    const double h = dt / nmax;
    if (t->rk_fn) {
        rhs = [&](const double *x, double *d) {
            const int Wd = t->grid_w > 0 ? t->grid_w : W;
            t->rk_fn(x, (int)(cell % Wd), (int)(cell / Wd), t->localized ? (uint32_t)t->act[cell] : 0xffffffffu, d);
        };
    }
    for (size_t c = 0; c < A; ++c) {
        cell = c;
        double x[ORACLE_MAX_SPECIES], k1[ORACLE_MAX_SPECIES], k2[ORACLE_MAX_SPECIES],
            k3[ORACLE_MAX_SPECIES], k4[ORACLE_MAX_SPECIES], tmp[ORACLE_MAX_SPECIES];
        for (int s = 0; s < S; ++s) x[s] = state[c + s * A];
        for (int n = 0; n < nmax; ++n) {
            rhs(x, k1);
            for (int s = 0; s < S; ++s) tmp[s] = x[s] + h / 2 * k1[s];
            rhs(tmp, k2);
            for (int s = 0; s < S; ++s) tmp[s] = x[s] + h / 2 * k2[s];
            rhs(tmp, k3);
            for (int s = 0; s < S; ++s) tmp[s] = x[s] + h * k3[s];
            rhs(tmp, k4);
            for (int s = 0; s < S; ++s) {
                x[s] += h / 6.0 * (k1[s] + 2.0 * k2[s] + 2.0 * k3[s] + k4[s]);
                if (x[s] < 0.0) x[s] = 0.0;
            }
        }
        for (int s = 0; s < S; ++s) state[c + s * A] = x[s];
    }
}

// ---------------------------------------------------------------------------------------
// deterministic_performReaction_aqss (Deterministic.cl:291-522): the alpha-QSS integrator
// of Mott & Oran (CHEMEQ2) per cell over dtg, with the production / destruction split
// q_s, d_s of deterministic_rhs_ode_aqss (Deterministic.cl:241-287) as Solver.cpp:960-1010
// generates it: q_s = 0. + |n|. * k * X * Y + ... over the reactions that produce s (net
// stoichiometry n > 0), d_s the same over those that consume it, left to right in double
// (the "1." multipliers make the whole expression double), stored in the real type R.
//
// R = float reproduces the reference's promotions statement by statement: Real variables
// are float, every subexpression with an unsuffixed literal (180., 0.5*, 1e-16*, ...) is
// evaluated in double and rounded on assignment; no contraction, IEEE division and sqrt
// (OpenCL allows 2.5 / 3 ulp: parity unpinned for those roundings).  R = double is the
// same algorithm in double, for the closed-form checks.
//
// Kept from the reference as it computes (fixed = 0): the corrector's Pade factor is written
// `180.*+rtaub*(...)` (Deterministic.cl:423), i.e. 180 * rtaub * (60 + ...), not
// 180 + rtaub * (...) as in the predictor (:388); fixed = 1 uses the predictor's form
// (Mott & Oran's alpha) in the corrector too.  itermax = 2, so the stability factor stays
// 0.01 and the ym1 / ym2 history (written only for it) is not kept.
// Not in the reference: at most kAqssMaxIter passes of the inner loop per cell (a cell
// whose step size underflows would otherwise never finish); such a cell is set to NaN.
// The GPU build has the same cap.
// Pinned by: the birth-death closed form A(t) = k/g + (A0 - k/g) e^{-g t}, the equilibrium
// of a stiff reversible pair, and the fp64 RK4 oracle on the Schnakenberg network.
namespace {
constexpr int kAqssMaxIter = 1 << 16;

// deterministic_rhs_ode_aqss as generated (Solver.cpp:960-1010): per species, production and
// destruction sums `0. + |n|. * [w *] <law> + ...` in double, w the cell's mask value of a
// localized reaction (act: the cell's active bits).
template <typename R>
void aqss_custom(const oracle_reaction_table *t, const R *y, R *q, R *d, uint32_t act, size_t cell);
template <>
void aqss_custom<float>(const oracle_reaction_table *t, const float *y, float *q, float *d, uint32_t act, size_t c) {
    t->aqss_f32_fn(y, (int)(c % t->grid_w), (int)(c / t->grid_w), act, q, d);
}
template <>
void aqss_custom<double>(const oracle_reaction_table *t, const double *y, double *q, double *d, uint32_t act,
                         size_t c) {
    t->aqss_f64_fn(y, (int)(c % t->grid_w), (int)(c / t->grid_w), act, q, d);
}

template <typename R>
void aqss_rhs(const oracle_reaction_table *t, const R *y, R *q, R *d, uint32_t act, size_t cell) {
    if ((sizeof(R) == 4 && t->aqss_f32_fn) || (sizeof(R) == 8 && t->aqss_f64_fn)) {
        aqss_custom<R>(t, y, q, d, act, cell);
        return;
    }
    const int S = t->num_species, NR = t->num_reactions;
    for (int s = 0; s < S; ++s) {
        double qs = 0., ds = 0.;
        for (int i = 0; i < NR; ++i) {
            const int n = t->net[i][s];
            if (n == 0) continue;
            double term = (double)std::abs(n);
            if ((t->localized >> i) & 1u) term = term * (((act >> i) & 1u) ? 1.0 : 0.0);
            term = term * t->lit_dbl[i];
            for (int f = 0; f < t->n_factors[i]; ++f) term = term * (double)y[t->factor_species[i][f]];
            if (n > 0)
                qs = qs + term;
            else
                ds = ds + term;
        }
        q[s] = (R)qs;
        d[s] = (R)ds;
    }
}

template <typename R>
int aqss_cell(const oracle_reaction_table *t, R *state, R dtg, int fixed, uint32_t act, size_t cell) {
    const int S = t->num_species;
    const R epsmin = (R)1e-2, sqreps = (R)0.5, epscl = (R)100., epsmax = (R)10., tfd = (R)1.000008,
            dtmin = (R)1e-15, ymin = (R)1e-20;
    const int itermax = 2;
    R y[ORACLE_MAX_SPECIES], q[ORACLE_MAX_SPECIES], d[ORACLE_MAX_SPECIES], ys[ORACLE_MAX_SPECIES],
        qs[ORACLE_MAX_SPECIES], rtaus[ORACLE_MAX_SPECIES], y1[ORACLE_MAX_SPECIES], scr[ORACLE_MAX_SPECIES];
    R tn = (R)0.;
    for (int i = 0; i < S; ++i) y[i] = std::fmax(state[i], ymin);
    aqss_rhs<R>(t, y, q, d, act, cell);
    R scrtch = (R)1e-25;
    for (int i = 0; i < S; ++i) {
        const R ascr = std::fabs(q[i]);
        const R scr2 = std::copysign((R)1. / y[i], (R)0.1 * epsmin * ascr - d[i]);
        const R scr1 = scr2 * d[i];
        scrtch = std::fmax(std::fmax(scr1, -std::fabs(ascr - d[i]) * scr2), scrtch);
    }
    R dt = std::fmin(sqreps / scrtch, dtg);
    R ts, dtcurrent, eps;
    int iters = 0;
    while (dtg > tn * tfd) {
        ts = tn;
        dtcurrent = dt;
        for (int i = 0; i < S; ++i) {
            ys[i] = y[i];
            qs[i] = q[i];
            rtaus[i] = dtcurrent * d[i] / y[i];
        }
        while (dtg > tn * tfd) {
            if (++iters > kAqssMaxIter) {
                for (int i = 0; i < S; ++i) state[i] = (R)NAN;
                return -1;
            }
            for (int i = 0; i < S; ++i) {  // predictor
                const R rtaui = dtcurrent * d[i] / y[i];
                const R alpha = (R)((180. + rtaui * (60. + rtaui * (11. + rtaui))) /
                                    (360. + rtaui * (60. + rtaui * (12. + rtaui))));
                scr[i] = (R)((q[i] - d[i]) / (1. + alpha * rtaui));
            }
            for (int iteration = 0; iteration < itermax; iteration++) {
                for (int i = 0; i < S; ++i) y[i] = std::fmax(ys[i] + dt * scr[i], ymin);
                if (iteration == 0) {
                    tn = ts + dt;
                    for (int i = 0; i < S; ++i) y1[i] = y[i];
                }
                aqss_rhs<R>(t, y, q, d, act, cell);
                eps = (R)1e-10;
                for (int i = 0; i < S; ++i) {  // corrector
                    const R rtaub = (R)(0.5 * (rtaus[i] + dt * d[i] / y[i]));
                    const R alpha = fixed ? (R)((180. + rtaub * (60. + rtaub * (11. + rtaub))) /
                                                (360. + rtaub * (60. + rtaub * (12. + rtaub))))
                                          : (R)((180. * +rtaub * (60. + rtaub * (11. + rtaub))) /
                                                (360. + rtaub * (60. + rtaub * (12. + rtaub))));
                    const R qt = (R)(qs[i] * (1. - alpha) + q[i] * alpha);
                    const R pb = rtaub / dt;
                    scr[i] = (R)((qt - ys[i] * pb) / (1. + alpha * rtaub));
                }
            }
            for (int i = 0; i < S; ++i) {  // convergence
                const R scr2 = std::fmax(ys[i] + dt * scr[i], (R)0.);
                R scr1 = std::fabs(scr2 - y1[i]);
                y[i] = std::fmax(scr2, ymin);
                if (0.25 * (ys[i] + y[i]) > ymin) {
                    scr1 = scr1 / y[i];
                    eps = std::fmax((R)0.5 * (scr1 + std::fmin(std::fabs(q[i] - d[i]) / (q[i] + d[i] + (R)1e-30), scr1)),
                                    eps);
                }
            }
            eps = eps * epscl;
            if (dt <= dtmin + 1e-16 * tn) {
                dt = dtg - ts;
                dt = std::fmin(dtmin, std::fabs(dt));
            }
            const R stab = (R)0.01;
            if (eps <= epsmax && stab <= 1.) {
                if (dtg <= tn * tfd) break;
            } else {
                tn = ts;
            }
            const R rteps = std::sqrt(eps);
            R dto = dt;
            dt = std::fmin(std::fmin(dt * ((R)1.0 / rteps + (R).005), tfd * (dtg - tn)), dto / (stab + (R)0.001));
            if (eps <= epsmax && stab <= 1) break;
            dto = dt / dto;
            for (int i = 0; i < S; ++i) rtaus[i] *= dto;
        }
        aqss_rhs<R>(t, y, q, d, act, cell);
    }
    for (int i = 0; i < S; ++i) state[i] = y[i];
    return iters;
}

template <typename R>
int aqss_grid(R *state, int W, int H, const oracle_reaction_table *t, R dtg, int fixed) {
    const size_t A = (size_t)W * H;
    const int S = t->num_species;
    int worst = 0;
    for (size_t c = 0; c < A; ++c) {
        R x[ORACLE_MAX_SPECIES];
        for (int s = 0; s < S; ++s) x[s] = state[c + s * A];
        const uint32_t act = t->localized ? (uint32_t)t->act[c] : 0xffffffffu;
        const int it = aqss_cell<R>(t, x, dtg, fixed, act, c);
        worst = (it < 0 || worst < 0) ? -1 : std::max(worst, it);
        for (int s = 0; s < S; ++s) state[c + s * A] = x[s];
    }
    return worst;
}
}  // namespace

// Returns the largest number of inner passes of any cell, or -1 if a cell hit the cap.
extern "C" int oracle_det_aqss_f32(float *state, int W, int H, const oracle_reaction_table *t, float dtg,
                                   int fixed) {
    return aqss_grid<float>(state, W, H, t, dtg, fixed);
}
extern "C" int oracle_det_aqss_f64(double *state, int W, int H, const oracle_reaction_table *t, double dtg,
                                   int fixed) {
    return aqss_grid<double>(state, W, H, t, dtg, fixed);
}

// ---------------------------------------------------------------------------------------
// Fused Schnakenberg reaction-drift-diffusion step (no reference counterpart: the
// reference computes no reactions in deterministic-inhomogeneous mode,
// Solver.cpp:107-171,250-266).  Diffusion follows the FPE stencil of
// Deterministic.cl:71-136 with r = 0 and mirror BC, both axes at once (Jacobi):
//   u' = u + dt*( D0u*(sum_nbr D_n u_n - 4 D u)/h^2 + k1 - k4 u + k3 u^2 v )
//   v' = v + dt*( D0v*(sum_nbr D_n v_n - 4 D v)/h^2 + k2 - k3 u^2 v )
// Pinned by: eigenmode decay with k = 0 (pure diffusion), the homogeneous ODE
// (D uniform) against a scalar integrator, and central finite differences for the
// gradient.
namespace {
void rdd_step(const oracle_rdd_params *p, const double *D, const double *u, const double *v,
              double *un, double *vn, int nt) {
    const int W = p->W, H = p->H;
    const double cu = p->D0u / (p->h * p->h), cv = p->D0v / (p->h * p->h);
#pragma omp parallel for schedule(static) num_threads(nthr(nt))
    for (int y = 0; y < H; ++y) {
        const size_t yl = (size_t)mirror_lo(y) * W, yh = (size_t)mirror_hi(y, H) * W, yc = (size_t)y * W;
        for (int x = 0; x < W; ++x) {
            const size_t i = yc + x, l = yc + mirror_lo(x), r = yc + mirror_hi(x, W), t = yl + x,
                         b = yh + x;
            const double su = D[l] * u[l] + D[r] * u[r] + D[t] * u[t] + D[b] * u[b] - 4.0 * D[i] * u[i];
            const double sv = D[l] * v[l] + D[r] * v[r] + D[t] * v[t] + D[b] * v[b] - 4.0 * D[i] * v[i];
            const double uuv = u[i] * u[i] * v[i];
            const double fu = p->k[0] - p->k[3] * u[i] + p->k[2] * uuv;
            const double fv = p->k[1] - p->k[2] * uuv;
            un[i] = u[i] + p->dt * (cu * su + fu);
            vn[i] = v[i] + p->dt * (cv * sv + fv);
        }
    }
}

// transpose of the mirror-BC neighbour sum: T(l)(j) = sum over c in {j-1, j+1} of
// w(c) l(c), w(c) = 2 at the first/last index (both its neighbours are j), else 1.
inline double tw(int c, int n) { return (c == 0 || c == n - 1) ? 2.0 : 1.0; }

void rdd_adjoint_step(const oracle_rdd_params *p, const double *D, const double *u,
                      const double *v, const double *lu, const double *lv, double *lun,
                      double *lvn, double *gD, double *gk_partial, int nt) {
    const int W = p->W, H = p->H;
    const double cu = p->D0u / (p->h * p->h), cv = p->D0v / (p->h * p->h), dt = p->dt;
    const double k3 = p->k[2], k4 = p->k[3];
    double g0 = 0, g1 = 0, g2 = 0, g3 = 0;
#pragma omp parallel for schedule(static) num_threads(nthr(nt)) reduction(+ : g0, g1, g2, g3)
    for (int y = 0; y < H; ++y) {
        for (int x = 0; x < W; ++x) {
            const size_t i = (size_t)y * W + x;
            double tu = 0, tv = 0;
            if (x > 0) { const size_t c = i - 1; tu += tw(x - 1, W) * lu[c]; tv += tw(x - 1, W) * lv[c]; }
            if (x < W - 1) { const size_t c = i + 1; tu += tw(x + 1, W) * lu[c]; tv += tw(x + 1, W) * lv[c]; }
            if (y > 0) { const size_t c = i - W; tu += tw(y - 1, H) * lu[c]; tv += tw(y - 1, H) * lv[c]; }
            if (y < H - 1) { const size_t c = i + W; tu += tw(y + 1, H) * lu[c]; tv += tw(y + 1, H) * lv[c]; }
            const double au = tu - 4.0 * lu[i], av = tv - 4.0 * lv[i];
            const double ui = u[i], vi = v[i];
            const double uv = ui * vi, uu = ui * ui;
            lun[i] = lu[i] + dt * cu * D[i] * au + dt * ((-k4 + 2.0 * k3 * uv) * lu[i] - 2.0 * k3 * uv * lv[i]);
            lvn[i] = lv[i] + dt * cv * D[i] * av + dt * (k3 * uu * lu[i] - k3 * uu * lv[i]);
            gD[i] += dt * (cu * ui * au + cv * vi * av);
            g0 += lu[i];
            g1 += lv[i];
            g2 += uu * vi * (lu[i] - lv[i]);
            g3 += -ui * lu[i];
        }
    }
    gk_partial[0] += dt * g0;
    gk_partial[1] += dt * g1;
    gk_partial[2] += dt * g2;
    gk_partial[3] += dt * g3;
}
}  // namespace

extern "C" void oracle_rdd_forward(const oracle_rdd_params *p, const double *D, double *u,
                                   double *v, int steps, int nt) {
    const size_t A = (size_t)p->W * p->H;
    std::vector<double> un(A), vn(A);
    for (int n = 0; n < steps; ++n) {
        rdd_step(p, D, u, v, un.data(), vn.data(), nt);
        std::copy(un.begin(), un.end(), u);
        std::copy(vn.begin(), vn.end(), v);
    }
}

// L = mean over cells of (u-u*)^2 + (v-v*)^2
extern "C" double oracle_rdd_loss(int n, const double *u, const double *v, const double *ut,
                                  const double *vt) {
    double s = 0;
    for (int i = 0; i < n; ++i) {
        const double a = u[i] - ut[i], b = v[i] - vt[i];
        s += a * a + b * b;
    }
    return s / n;
}

// Loss and gradient (dL/dD per cell, dL/dk1..k4) by reverse-mode through the explicit
// steps, with state checkpoints every ckpt_every steps and segment recomputation.
extern "C" double oracle_rdd_grad(const oracle_rdd_params *p, const double *D, const double *u0,
                                  const double *v0, const double *ut, const double *vt, int steps,
                                  int ckpt_every, double *gD, double *gk, int nt) {
    const size_t A = (size_t)p->W * p->H;
    if (ckpt_every <= 0) ckpt_every = steps > 0 ? steps : 1;
    const int nck = (steps + ckpt_every - 1) / ckpt_every;
    std::vector<std::vector<double>> cku(nck), ckv(nck);
    std::vector<double> u(u0, u0 + A), v(v0, v0 + A), un(A), vn(A);
    for (int n = 0; n < steps; ++n) {
        if (n % ckpt_every == 0) {
            cku[n / ckpt_every] = u;
            ckv[n / ckpt_every] = v;
        }
        rdd_step(p, D, u.data(), v.data(), un.data(), vn.data(), nt);
        u.swap(un);
        v.swap(vn);
    }
    const double loss = oracle_rdd_loss((int)A, u.data(), v.data(), ut, vt);
    std::vector<double> lu(A), lv(A), lun(A), lvn(A);
    for (size_t i = 0; i < A; ++i) {
        lu[i] = 2.0 * (u[i] - ut[i]) / (double)A;
        lv[i] = 2.0 * (v[i] - vt[i]) / (double)A;
        gD[i] = 0.0;
    }
    for (int j = 0; j < 4; ++j) gk[j] = 0.0;
    std::vector<std::vector<double>> segu, segv;
    for (int c = nck - 1; c >= 0; --c) {
        const int n0 = c * ckpt_every, n1 = std::min(steps, n0 + ckpt_every);
        segu.assign(n1 - n0, std::vector<double>());
        segv.assign(n1 - n0, std::vector<double>());
        segu[0] = cku[c];
        segv[0] = ckv[c];
        for (int n = n0 + 1; n < n1; ++n) {
            segu[n - n0].resize(A);
            segv[n - n0].resize(A);
            rdd_step(p, D, segu[n - n0 - 1].data(), segv[n - n0 - 1].data(), segu[n - n0].data(),
                     segv[n - n0].data(), nt);
        }
        for (int n = n1 - 1; n >= n0; --n) {
            rdd_adjoint_step(p, D, segu[n - n0].data(), segv[n - n0].data(), lu.data(), lv.data(),
                             lun.data(), lvn.data(), gD, gk, nt);
            lu.swap(lun);
            lv.swap(lvn);
        }
    }
    return loss;
}
