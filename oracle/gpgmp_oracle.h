/*
 * CPU oracle for the GPGMP hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library restates the reference algorithms
 * (Inchman exec/src/gpgmp/src/<kernel>.cl and the C++ host that drives them) on the
 * CPU so that the HIP path can be checked against it.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it, and only
 * as the checker: the product library never links or calls it.
 *
 * Every function cites the reference file:line it follows.  Parity pinning
 * (what each function is checked against) is listed per function; functions
 * whose only check is agreement with the HIP path are marked
 * "parity unpinned".
 *
 * Numeric contract shared with the HIP path (the reference leaves these
 * implementation-defined):
 *   - Threefry-2x32-20 (Salmon et al., SC'11), KAT-pinned.
 *   - uniform map u = (float)((double)(int32)w * 2^-32 + 0.5)
 *     (Gillespie.cl:72-73, constant Solver.cpp:469-471).
 *   - log: (float) of a double-precision log evaluated by a fixed sequence of
 *     IEEE double operations (oracle_logf below); pinned against libm log().
 *   - rate literals: the decimal the reference prints with `ss << rate`
 *     (%g, 6 significant digits, Reaction.cpp:122) parsed back; a literal
 *     without '.', 'e' or "inf"/"nan" is an int literal (int arithmetic),
 *     anything else a double literal (double arithmetic).
 *   - header constants printed "%.6e" then parsed as float (Solver.cpp:494-505).
 *   - no floating-point contraction (compiled with -ffp-contract=off).
 */
#ifndef GPGMP_ORACLE_H
#define GPGMP_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORACLE_MAX_SPECIES 16
#define ORACLE_MAX_REACTIONS 32
#define ORACLE_MAX_FACTORS 8

/* Compiled reaction table: one entry per reaction, species in index order. */
typedef struct {
    int num_species;
    int num_reactions;
    /* literal of the rate as printed by the reference */
    int lit_is_int[ORACLE_MAX_REACTIONS];
    int32_t lit_int[ORACLE_MAX_REACTIONS];
    double lit_dbl[ORACLE_MAX_REACTIONS];
    /* falling-factorial factors: species index and offset m of (X - m) */
    int n_factors[ORACLE_MAX_REACTIONS];
    int factor_species[ORACLE_MAX_REACTIONS][ORACLE_MAX_FACTORS];
    int factor_offset[ORACLE_MAX_REACTIONS][ORACLE_MAX_FACTORS];
    /* net stoichiometric change (product - reactant) */
    int net[ORACLE_MAX_REACTIONS][ORACLE_MAX_SPECIES];
    /* localized reactions (Reaction::addCompartment, Reaction.cpp:209-235): bit i set when
     * reaction i is localized; act[cell] (x + y*W) has bit i set where its mask w(x,y) is 1.
     * act is read only when localized != 0. */
    uint32_t localized;
    const uint16_t *act;
    /* custom kinetic laws (Reaction::setCustomLaw): when set, the reference's generated hazard
     * and right-hand-side code for the model, compiled by a C++ compiler from the laws' formula
     * text (oracle/lawgen.py); act = the cell's active-reaction bits */
    void (*hazards_fn)(const int32_t *x, int gx, int gy, uint32_t act, float *h);
    void (*rk_fn)(const double *x, int gx, int gy, uint32_t act, double *deriv);
    void (*aqss_f32_fn)(const float *y, int gx, int gy, uint32_t act, float *q, float *d);
    void (*aqss_f64_fn)(const double *y, int gx, int gy, uint32_t act, double *q, double *d);
    int grid_w; /* cells per row (GridX / GridY of the custom laws) */
} oracle_reaction_table;

/* the laws' math functions (exp/log/log10 = fn 0/1/2 in double; pow), the same bits as the
 * HIP library's */
double oracle_law_fn_d(int fn, double x);
double oracle_law_pow_d(double x, double y);

/* ---- primitives ---- */
void oracle_threefry2x32_20(uint32_t c0, uint32_t c1, uint32_t k0, uint32_t k1,
                            uint32_t *o0, uint32_t *o1);
float oracle_u01(uint32_t w);
float oracle_logf(float y);
/* the particle step's Gaussian pair (Box-Muller on Threefry key {id, seed}, counter {step, 0}) */
void oracle_gauss2(uint32_t id, uint32_t seed, uint32_t step, float *gx, float *gy);
double oracle_logd(double y);
void oracle_mt19937_seeds(uint32_t seed0, int n, int32_t *out);

/* ---- model compile ---- */
int oracle_compile_reactions(int num_species, int num_reactions, const float *rates,
                             const int *reactant_stoich, const int *product_stoich,
                             oracle_reaction_table *out);
float oracle_stay_probability(float p0);
float oracle_header_float(float v);
float oracle_homogeneous_dt(float p0, float h, float D);
float oracle_inhomogeneous_cell_dt(float D, float r, float h);
float oracle_inhomogeneous_dt(const float *D, const float *r, int n, float h);
/* stochastic_inhomogeneous_fpe: per-cell step and its minimum (stay: StayProbability) */
float oracle_fpe_cell_dt(float D, float r, float h, float stay);
float oracle_fpe_dt(const float *D, const float *r, int n, float h, float stay);

/* ---- stochastic kernels ---- */
/* cumulative hazards h[0..R] (h[0] = total) of state x[S] in cell `cell` */
void oracle_hazards(const oracle_reaction_table *tab, const int32_t *x, int cell, float *h);
void oracle_ssa(int32_t *state, int W, int H, const oracle_reaction_table *tab,
                int32_t seed, float deltat, int nthreads);
void oracle_hom_diffuse(int32_t *state, int W, int H, int species, int32_t seed,
                        float stay, const int *boundary_mask, const int *source_mask,
                        int nthreads);
void oracle_inhom_sweep(int32_t *state, int W, int H, int species, int axis, int32_t seed,
                        float dt, const float *D_axis, const float *r_axis, float h,
                        const int *boundary_mask, int nthreads);
/* the same sweep with the Fokker-Planck transition rates (stochastic_inhomogeneous_fpe) */
void oracle_fpe_sweep(int32_t *state, int W, int H, int species, int axis, int32_t seed,
                      float dt, const float *D_axis, const float *r_axis, float h,
                      const int *boundary_mask, int nthreads);

/* ---- host schedule (DiffusionModel::oclSolve) ---- */
typedef struct {
    int W, H, S;
    float physical_length;
    float p0;
    int inhomogeneous;          /* 0: stochastic_homogeneous, 1: stochastic_inhomogeneous,
                                   2: stochastic_inhomogeneous_fpe */
    const float *diffusion;     /* [S] (homogeneous) */
    const float *Dx, *Dy, *rx, *ry; /* [S][H][W] (inhomogeneous) */
    int boundary_mask[4];
    int source_mask[4];
    float max_time;             /* dt of non-diffusing species; loop bound */
} oracle_model;

int oracle_run(const oracle_model *m, const oracle_reaction_table *tab, int32_t *state,
               int steps, uint32_t seed0, int nthreads, float *sim_time_out);
/* the same loop with output events every `interval` and the stop at max_time (infinity:
 * off); snapshots of the state [S][A] at t = 0 and each write into dumps[max_dumps];
 * returns the number of writes */
int oracle_run_output(const oracle_model *m, const oracle_reaction_table *t, int32_t *state, int steps,
                      uint32_t seed0, int nthreads, float interval, float max_time, int32_t *dumps,
                      float *dump_times, int max_dumps, int64_t *iterations, float *sim_time_out);
/* the loop from iteration `skip` on: the first skip iterations advance only the schedule and
 * the seed stream, `state` being the state after them; then `steps` iterations run */
int oracle_run_from(const oracle_model *m, const oracle_reaction_table *t, int32_t *state, int skip, int steps,
                    uint32_t seed0, int nthreads, float *sim_time_out);

/* ---- deterministic (FP64 Jacobi) ---- */
void oracle_det_hom_diffuse(double *state, int W, int H, int S, const double *D, double dt,
                            double h);
void oracle_det_inhom_sweep(double *state, int W, int H, int species, int axis, double dt,
                            const double *D_axis, const double *r_axis, double h);
void oracle_det_rk4(double *state, int W, int H, const oracle_reaction_table *tab, double dt);

/* ---- reaction-drift-diffusion (Schnakenberg, FP64) ---- */
typedef struct {
    int W, H;
    double h, dt;
    double D0u, D0v;
    double k[4];
} oracle_rdd_params;
void oracle_rdd_forward(const oracle_rdd_params *p, const double *D, double *u, double *v,
                        int steps, int nthreads);
double oracle_rdd_loss(int n, const double *u, const double *v, const double *ut,
                       const double *vt);
double oracle_rdd_grad(const oracle_rdd_params *p, const double *D, const double *u0,
                       const double *v0, const double *ut, const double *vt, int steps,
                       int ckpt_every, double *gD, double *gk, int nthreads);

/* alpha-QSS reactions over dtg per cell (deterministic_performReaction_aqss); state [S][H][W];
 * fixed = 1: the predictor's Pade factor in the corrector too (0: the reference's).
 * Returns the largest inner-pass count of a cell, -1 if a cell hit the pass cap. */
int oracle_det_aqss_f32(float *state, int W, int H, const oracle_reaction_table *t, float dtg, int fixed);
int oracle_det_aqss_f64(double *state, int W, int H, const oracle_reaction_table *t, double dtg, int fixed);

/* ---- individual particles ---- */
void oracle_stable_sort_pairs(uint32_t *keys, uint32_t *vals, int n);
void oracle_cell_offsets(const uint32_t *sorted_keys, int n, int ncells, uint32_t *off);
void oracle_cell_left_right(const uint32_t *sorted_keys, int n, int ncells, uint32_t *left,
                            uint32_t *right);
typedef struct {
    int side;          /* cells per side (domain [0,side)^2, 1 um cells) */
    float D, dt, vortex_amp;
    float p_annihilate;
    uint32_t seed;
} oracle_particle_params;
int oracle_particle_step(const oracle_particle_params *p, int step, int n, float *x, float *y,
                         uint32_t *id, uint32_t *keys, uint32_t *off, int nthreads);

/* ---- individual-species lattice hops (IndividualSolver) ---- */
typedef struct {
    int W, H;
    float physical_length;
    int32_t boundary_mask[4];  /* -x, +x, -y, +y: 0 reflective, 1 absorbing */
    int species;               /* RNG key offset species * W * H */
} oracle_ind_params;
/* props: [4][cap] = DiffusivityX, DiffusivityY, DriftX, DriftY per particle index */
void oracle_ind_sweep(const oracle_ind_params *p, uint32_t *keys, const uint32_t *vals, const uint32_t *left,
                      const uint32_t *right, int n, const float *props, int cap, int axis, int32_t seed,
                      float dt);
void oracle_ind_constants(const oracle_ind_params *p, const uint32_t *keys, const uint32_t *vals,
                          const uint32_t *left, const uint32_t *right, int n, const float *props, int cap,
                          float *Dx, float *Dy, float *rx, float *ry, int32_t *count);
void oracle_ind_update(const oracle_ind_params *p, uint32_t *keys, uint32_t *vals, int n, const float *props,
                       int cap, uint32_t *left, uint32_t *right, float *Dx, float *Dy, float *rx, float *ry,
                       int32_t *count, float *dtx, float *dty);
double oracle_expd(double x);
double oracle_expm1d(double x);
float oracle_ind_cell_dt(float D, float r, float h);
int oracle_ind_run(const oracle_ind_params *p, uint32_t *keys, uint32_t *vals, int n, const float *props, int cap,
                   int steps, uint32_t seed0, int32_t *count, float *sim_time_out);

/* ---- population maintenance of an individual species (IndividualSolver::maintainPopulation) ---- */
typedef struct {
    int32_t mode[4]; /* per property: 0 keep the recycled id's value, 1 constant a, 2 a + b * RANDOM */
    double a[4], b[4];
} oracle_new_rule;
/* keys/vals: the full list [cap] (live sorted in [0, n), dead ids after); props [4][cap] by id;
 * goal[W*H]; returns 0, -1 (pool too small: nothing changed) or -2 (negative goal). */
int oracle_ind_maintain(const oracle_ind_params *p, uint32_t *keys, uint32_t *vals, int n, int cap, float *props,
                        const int32_t *goal, int32_t seed, const oracle_new_rule *rule, int *n_out);

#ifdef __cplusplus
}
#endif
#endif
