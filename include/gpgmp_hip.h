/*
 * gpgmp_hip.h -- C ABI of the MI355X-native GPGMP solver (libgpgmp_hip.so).
 *
 * This boundary replaces the reference's OpenCL solver plugin interface
 * (Inchman exec/src/gpgmp/src):
 *   gpgmp::Solver                 Solver.h:15-239
 *   gpgmp::HomogeneousSolver      HomogeneousSolver.h:15-81
 *   gpgmp::InhomogeneousSolver    InhomogeneousSolver.h:15-162
 *   gpgmp::IndividualSolver       IndividualSolver.h:36-297 (sort / cell indices)
 *   DiffusionModel::oclSolve      DiffusionModel.cpp:1414-1620 (the host loop)
 * Each entry point names the reference member it replaces.
 *
 * Conventions
 *   - Every call returns int: GPGMP_OK (0) or a negative GPGMP_E_* code; no C++
 *     exception crosses the ABI (the reference throws or exit(1)s, DiffusionModel.h:68-105).
 *     gpgmp_last_error() returns the message of the calling thread's last failure.
 *   - Host buffers are caller-owned and copied; device buffers are library-owned
 *     unless a function takes a device pointer ("_device" suffix or documented).
 *   - Streams are passed as void* (a hipStream_t; NULL = the default stream).
 *   - A handle is used by one thread at a time; different handles may be used
 *     concurrently on different streams.
 *   - Counts are int32 in layout [species][height][width] with x fastest, the
 *     reference's buffer order (index x + y*W + s*W*H, Solver.cpp:476-482).
 */
#ifndef GPGMP_HIP_H
#define GPGMP_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GPGMP_OK 0
#define GPGMP_E_INVALID (-1)     /* bad argument or model */
#define GPGMP_E_HIP (-2)         /* HIP runtime error */
#define GPGMP_E_CAPACITY (-3)    /* a size exceeds a capacity */
#define GPGMP_E_STATE (-4)       /* call not valid in the handle's state */
#define GPGMP_E_UNSUPPORTED (-5) /* feature outside this build */

#define GPGMP_MAX_SPECIES 8
#define GPGMP_MAX_REACTIONS 16

/* Solver types (DiffusionModel.cpp:250-266 string map; Solver.h SolverType). */
#define GPGMP_STOCHASTIC_HOMOGENEOUS 0
#define GPGMP_STOCHASTIC_INHOMOGENEOUS 1
#define GPGMP_DETERMINISTIC_HOMOGENEOUS_RK4 2
#define GPGMP_DETERMINISTIC_INHOMOGENEOUS_RK4 3
#define GPGMP_DETERMINISTIC_HOMOGENEOUS_AQSS 4
#define GPGMP_DETERMINISTIC_INHOMOGENEOUS_AQSS 5
/* stochastic_inhomogeneous_fpe (DiffusionModel.h:49, DiffusionModel.cpp:257): the inhomogeneous
 * solver with the Fokker-Planck discretisation -- per-cell step StayProbability/6/(D/h^2 +
 * |r|/(2h)) and hop rates D/h^2 +- r/(2h) times the sweep's dt
 * (StochasticInhomogeneousDriftDiffusion.cl:123-126,191-197,268-271,341-347,491-500,626-633) */
#define GPGMP_STOCHASTIC_INHOMOGENEOUS_FPE 6

/* One reaction: stoichiometry per species and the rate constant
 * (Reaction(id, Real rate, ...) Reaction.cpp:39-56; law generated at :119-138). */
typedef struct {
    float rate;
    int32_t reactant_stoich[GPGMP_MAX_SPECIES];
    int32_t product_stoich[GPGMP_MAX_SPECIES];
} gpgmp_reaction_desc;

/* Model description (the DiffusionModel setters used by the solvers,
 * DiffusionModel.h:149-773). */
typedef struct {
    int32_t width, height;        /* grid cells; multiples of 16 (Globals.h:27) */
    float physical_length;        /* domain side in um (physicalLength) */
    float p0;                     /* homogeneous jump probability (setP0) */
    int32_t num_species;
    float diffusion[GPGMP_MAX_SPECIES]; /* um^2/s, homogeneous solvers (Species::diffusionConstant) */
    int32_t num_reactions;
    const gpgmp_reaction_desc *reactions;
    int32_t boundary_mask[4];     /* -x,+x,-y,+y: 0 reflective, 1 absorbing (setBoundaryMasks) */
    int32_t source_mask[4];       /* source-boundary particle counts */
    float max_time;               /* time step of non-diffusing species (maxTime) */
    int32_t solver_type;          /* GPGMP_* solver type */
    /* inhomogeneous solvers: fields [S][H][W] on the host, copied at create
     * (the generated computeDiffusionConstants_forSpecies_* output, Solver.cpp:659-777) */
    const float *diffusivity_x, *diffusivity_y, *drift_x, *drift_y;
} gpgmp_model_desc;

typedef struct gpgmp_model gpgmp_model;
typedef struct gpgmp_solver gpgmp_solver;
typedef struct gpgmp_batch gpgmp_batch;
typedef struct gpgmp_rdd gpgmp_rdd;
typedef struct gpgmp_particles gpgmp_particles;
typedef struct gpgmp_individuals gpgmp_individuals;

/* The model's newIndividuals method (DiffusionModel::getNewIndividualsMethod, pasted into
 * updateNewParticles at IndividualSolver.cl:1284) as a rule per property (DiffusivityX,
 * DiffusivityY, DriftX, DriftY): mode 0 keeps the recycled particle's stored value (an
 * empty method), 1 sets the constant a, 2 sets a + b * RANDOM evaluated in double and stored
 * as float, RANDOM being the reference's macro (IndividualSolver.cl:1262-1282), drawn in
 * property order. */
typedef struct {
    int32_t mode[4];
    double a[4], b[4];
} gpgmp_new_individual_rule;

const char *gpgmp_last_error(void);
int gpgmp_version(int32_t *major, int32_t *minor);

/* ---- model ---------------------------------------------------------------------- */
/* Validates the description and compiles the reaction table; replaces the run-time
 * kernel source generation and build (Solver.cpp:449-1070). */
int gpgmp_model_create(const gpgmp_model_desc *desc, gpgmp_model **out);
int gpgmp_model_destroy(gpgmp_model *m);
/* Compiled per-reaction rate literal as the reference prints it (Reaction.cpp:122):
 * is_int[i], value[i]. */
int gpgmp_model_rate_literals(const gpgmp_model *m, int32_t *is_int, double *value);
/* Reaction::addCompartment (Reaction.cpp:209-212) with the compartment's bounds
 * (Compartment::setBounds, Compartment.cpp:27-32): reaction `reaction` becomes localized, and
 * its hazard and deterministic rate are multiplied in every cell by the reaction mask
 * w(x, y) = 1 inside any of its compartments (x0 <= x <= x1, y0 <= y <= y1,
 * Compartment::isInCompartment, Compartment.h:152-158), 0 elsewhere (Reaction::reactionMask,
 * Reaction.cpp:216-235; Solver::oclCreateReactionMask, Solver.cpp:358-392; hazards
 * Solver.cpp:853-854, RK4 Solver.cpp:947-948, alpha-QSS Solver.cpp:987-1000).  Applies to
 * solvers created from the model afterwards. */
int gpgmp_model_add_reaction_compartment(gpgmp_model *m, int32_t reaction, int32_t x0, int32_t y0, int32_t x1,
                                         int32_t y1);
/* Species::id of every species (ids[num_species]), the names custom laws use. */
int gpgmp_model_set_species_ids(gpgmp_model *m, const char *const *ids);
/* DiffusionModel::setParameter(id, value) (a Real): a name custom laws may use; it enters the
 * laws as the generated `const Real <id> = <value>;` does (Solver.cpp:834-841). */
int gpgmp_model_set_parameter(gpgmp_model *m, const char *name, float value);
/* Reaction::setCustomLaw (Reaction.cpp:108-111): the kinetic law of reaction `reaction` as the
 * C expression text of its formula (libSBML's formula syntax: species ids, parameters,
 * PhysicalCellWidth/Height, PhysicalModelWidth/Height, StayProbability, GridX/GridY, numbers,
 * + - * /, parentheses, pow exp log log10 sqrt fabs abs floor ceil).  It replaces the generated
 * law in the SSA hazards (`_h[0] + [mask *] <law>`, Solver.cpp:843-857) and in the
 * deterministic right-hand sides (Solver.cpp:918-1010), typed as C types the pasted text: int
 * counts, float parameters, int and double literals.  Laws of the mass-action shape (a constant
 * times species factors X or (X - m)) are evaluated by the reaction table, the others by a
 * compiled stack program.  Set the species ids and parameters first; the formula is checked
 * here (GPGMP_E_INVALID with the reason). */
int gpgmp_model_set_kinetic_law(gpgmp_model *m, int32_t reaction, const char *formula);
/* The value the kernels give reaction `reaction`'s hazard assignment `_h[0] + [mask *] <law>`
 * for counts x[S] in cell (gx, gy) with running total h0 and active-reaction bits act,
 * evaluated on the host: how 0 as the solver evaluates it, 1 as a law program, 2 by the
 * reaction table (mass-action shapes only).  No counterpart in the reference (its laws are
 * compiled kernel text); for checking the law compiler. */
int gpgmp_model_law_hazard(const gpgmp_model *m, int32_t reaction, const int32_t *x, int32_t gx, int32_t gy,
                           float h0, uint32_t act, int32_t how, float *out);
/* The same for species `species`'s deterministic expression (which 0: the RK4 derivative,
 * 1: the alpha-QSS production, 2: its destruction) of concentrations y[S]; models with custom
 * laws only. */
int gpgmp_model_law_rhs(const gpgmp_model *m, int32_t species, int32_t which, const float *y, int32_t gx,
                        int32_t gy, uint32_t act, float *out);
/* The model's drift/diffusivity method (DiffusionModel::computeDriftDiffusivityMethod; the C
 * statements the reference pastes into computeDiffusionConstants_forSpecies_<id>,
 * Solver.cpp:659-777, and InhomogeneousSolver::computeDiffusivityDrift runs,
 * InhomogeneousSolver.cpp:437-470) evaluated on the device for every cell and species:
 * `Real x = e;` locals, `X->DiffusivityX/DiffusivityY/DriftX/DriftY = e;` assignments (All =
 * the species being computed), X->State, getCentralDifferenceX/Y(X), PhysicalX/PhysicalY,
 * PhysicalSimTime, the header constants, the parameters, the field parameters
 * (field_names[nfields], values fields[nfields][H][W]), comparisons and ?:.  state: counts
 * [S][H][W]; current: the fields before the call [4][S][H][W] (DiffusivityX, DiffusivityY,
 * DriftX, DriftY; NULL: zeros); out: the same layout after it -- what
 * gpgmp_model_desc.diffusivity_x/_y and drift_x/_y take.  MeanX (computeMoments) is not
 * supported. */
int gpgmp_model_eval_drift_diffusivity(const gpgmp_model *m, const char *method, const char *const *field_names,
                                       int32_t nfields, const float *fields, const int32_t *state,
                                       const float *current, float sim_time, float *out, void *stream);
/* The same evaluated on the host, for checking the method compiler (current is required).
 * No counterpart in the reference. */
int gpgmp_model_method_host(const gpgmp_model *m, const char *method, const char *const *field_names,
                            int32_t nfields, const float *fields, const int32_t *state, const float *current,
                            float sim_time, float *out);
/* What the drift/diffusivity method reads besides the cell's position, the constants and the
 * field parameters: bit 0 the counts (X->State, getCentralDifferenceX/Y), bit 1
 * PhysicalSimTime, bit 2 the current fields (a field the method does not assign keeps its
 * value, which does not count).  With none of them, the re-evaluation a
 * non-linear model gets after every sweep (InhomogeneousSolver.cpp:304-336,384-417)
 * reproduces the fields and time steps it starts with.  No counterpart in the reference. */
int gpgmp_model_method_dependencies(const gpgmp_model *m, const char *method, const char *const *field_names,
                                    int32_t nfields, uint32_t *deps);
/* Bit i set when reaction i is localized (Reaction::isLocalized, Reaction.h:196). */
int gpgmp_model_localized(const gpgmp_model *m, uint32_t *mask);

/* ---- plain solver (gpgmp::Solver and subclasses) --------------------------------- */
/* Flags of gpgmp_solver_create. */
/* deterministic alpha-QSS types: the predictor's Pade factor in the corrector too
 * (see gpgmp_det_aqss_device) */
#define GPGMP_SOLVER_AQSS_PADE_FIX 1u
/* deterministic types: reproduce the reference where it departs from the model: the
 * inhomogeneous deterministic types run no reactions (Solver.cpp:250-266 reacts for the
 * homogeneous types only) and a homogeneous diffuse(species) call moves every species
 * (homogeneous_deterministic_diffuse, Deterministic.cl:32, called once per due species by
 * the host loop).  Without it, every deterministic type reacts and diffuse moves the one
 * species asked for. */
#define GPGMP_SOLVER_REFERENCE_QUIRKS 2u
/* Solver::Solver + HomogeneousSolver/InhomogeneousSolver constructors
 * (Solver.cpp:18-172, HomogeneousSolver.cpp:8-125, InhomogeneousSolver.cpp:10-213), for every
 * solver type.  Stochastic types hold int32 counts (gpgmp_solver_set_state/get_state), the
 * deterministic types fp32 concentrations (the reference's detState,
 * gpgmp_solver_set_state_f32/get_state_f32): doReactions runs RK4 or alpha-QSS over dt
 * (Solver.cpp:250-266), diffuse the reference's stencils (Jacobi) with its time steps --
 * homogeneous dt = min_s 1e-2 h^2 / (4 D_s) for every species (HomogeneousSolver.cpp:97-121),
 * inhomogeneous the per-cell minimum of the stochastic types per axis
 * (InhomogeneousSolver.cpp:189-212, 282-286, 374-378) -- and neither touches the seed.
 * flags: GPGMP_SOLVER_* (others: GPGMP_E_INVALID). */
int gpgmp_solver_create(const gpgmp_model *m, void *stream, uint32_t flags, gpgmp_solver **out);
int gpgmp_solver_destroy(gpgmp_solver *s);
/* Solver::synchronizeStateBufferToDevice / FromDevice (Solver.h:85-90). */
int gpgmp_solver_set_state(gpgmp_solver *s, const int32_t *host_counts);
int gpgmp_solver_get_state(gpgmp_solver *s, int32_t *host_counts);
/* The same for the deterministic types' fp32 concentrations [S][H][W] (Solver.h:85-90 with
 * the detState vector); GPGMP_E_STATE on a stochastic solver (and the int32 calls on a
 * deterministic one). */
int gpgmp_solver_set_state_f32(gpgmp_solver *s, const float *host_conc);
int gpgmp_solver_get_state_f32(gpgmp_solver *s, float *host_conc);
/* Device-resident state buffer (int32 or float, [S][H][W]); valid until destroy. */
int gpgmp_solver_state_device(gpgmp_solver *s, void **dev_ptr);
/* Solver::doReactions(Real dt, int &seed) (Solver.h:52, Solver.cpp:207-248): seed += 1. */
int gpgmp_solver_do_reactions(gpgmp_solver *s, float dt, int32_t *seed);
/* Solver::diffuse(int species, Real t, int &seed) (Solver.h:80; HomogeneousSolver.cpp:132-164,
 * InhomogeneousSolver.cpp:254-427): inhomogeneous sweeps add 1 to seed each. */
int gpgmp_solver_diffuse(gpgmp_solver *s, int32_t species, float t, int32_t *seed);
/* Solver::getNextDiffusionTime() (Solver.h:68; HomogeneousSolver.h:42,
 * InhomogeneousSolver.cpp:238-252): out[num_species]. */
int gpgmp_solver_next_diffusion_time(gpgmp_solver *s, float *out);
/* The seed_n of the host loop's first n iterations from seed0 (DiffusionModel.cpp:1467-1482:
 * boost::mt19937(seed0), uniform_int<>(0, INT_MAX), plus the iteration number), for a loop
 * written over the per-call API. */
int gpgmp_loop_seeds(uint32_t seed0, int32_t n, int32_t *out);
/* Non-linear diffusivity (DiffusionModel m_nonlinearDiffusivity; InhomogeneousSolver.cpp:304-336,
 * 384-417, the generated computeDiffusionConstants_forSpecies kernels of Solver.cpp:659-777):
 * the drift/diffusivity method (the language of gpgmp_model_eval_drift_diffusivity, with its
 * field parameters, nfields x [H][W] floats) is re-evaluated after every sweep that
 * gpgmp_solver_diffuse runs, on the swept counts at the call's time t: the species' four fields,
 * then the swept axis' time step and the sweep probabilities (both axes', the Fokker-Planck
 * type's swept axis' only, as its rates are stored per axis).  Stochastic inhomogeneous
 * solvers (both types); the model then runs through the per-call API (gpgmp_solver_run
 * refuses it). */
int gpgmp_solver_set_nonlinear_method(gpgmp_solver *s, const char *method, const char *const *field_names,
                                      int32_t nfields, const float *fields);
/* Makes `species` an individual (particle-based) species of this stochastic inhomogeneous
 * solver (Solver(..., hasIndividualSpecies), IndividualSolver.cpp:37-130): its counts in the
 * state become the particles' (gpgmp_individuals_bind_counts), gpgmp_solver_do_reactions
 * follows the SSA with gpgmp_individuals_maintain(seed++) for each individual species in
 * index order (Solver.cpp:241-247, `rule` = the model's newIndividuals method, NULL = keep),
 * and gpgmp_solver_diffuse hops its particles, each sweep followed by the constants and the
 * new time step (InhomogeneousSolver.cpp:254-427).  `g` must have the solver's grid, its
 * species index and its particles set; it stays caller-owned.  Such models run through this
 * per-call API (gpgmp_solver_run refuses them), synchronising once per call. */
int gpgmp_solver_attach_individuals(gpgmp_solver *s, int32_t species, gpgmp_individuals *g,
                                    const gpgmp_new_individual_rule *rule);
/* Solver::getIndividualSolver() (Solver.h:109, used by writeAllSpecies,
 * DiffusionModel.cpp:1186-1250): the handle attached to `species` by
 * gpgmp_solver_attach_individuals, or NULL. */
int gpgmp_solver_individuals(gpgmp_solver *s, int32_t species, gpgmp_individuals **out);
/* Per-species time steps: homogeneous dt[s] (both arrays), inhomogeneous dtX/dtY. */
int gpgmp_solver_time_steps(gpgmp_solver *s, float *dt_x, float *dt_y);
/* DiffusionModel::oclSolve main loop (DiffusionModel.cpp:1467-1603) for `steps`
 * iterations from seed0 (boost::mt19937(seed0), uniform_int<>(0,INT_MAX)), with the
 * output events and max_time stop set by gpgmp_solver_set_output (none by default; no
 * alert events), as fused launches without host synchronisation. */
int gpgmp_solver_run(gpgmp_solver *s, int32_t steps, uint32_t seed0, float *sim_time_out);
/* Output of the host loop (DiffusionModel.cpp:1442, 1519-1572, 1479-1480): with interval > 0
 * every run takes the reference's output iterations (a reaction step up to the output time
 * and, once the clock has reached it, a zero-length one that writes; both consume a seed)
 * and keeps a device snapshot of the state at t = 0 and at each write, up to max_dumps
 * of them (writeAllSpecies, DiffusionModel.cpp:1096-1296, without the file); with
 * max_time > 0 the loop stops once the simulation time reaches it.  <= 0 disables either. */
int gpgmp_solver_set_output(gpgmp_solver *s, float interval, float max_time, int32_t max_dumps);
/* The snapshots of the last run: counts[n_stored][S][H][W] and times[n_stored] (either may
 * be NULL), n_total = writes including those beyond max_dumps, and the iterations run. */
int gpgmp_solver_get_output(gpgmp_solver *s, int32_t *counts, float *times, int32_t *n_stored, int32_t *n_total,
                            int64_t *iterations);
/* Continues the loop of the last gpgmp_solver_run for `steps` more iterations;
 * graph_chunk > 0 replays chunks of that many iterations as a HIP graph. */
int gpgmp_solver_run_continue(gpgmp_solver *s, int32_t steps, int32_t graph_chunk, float *sim_time_out);
/* Advances the loop of the last gpgmp_solver_run by `steps` iterations without running
 * them: the schedule, the seed stream (seed_n of DiffusionModel.cpp:1482) and the iteration
 * count move on, and the state is left as the caller set it -- it must be the state after
 * those iterations (a snapshot of an earlier run), so that gpgmp_solver_run_continue resumes
 * the loop from there.  Skipped iterations write no output.  No counterpart in the
 * reference. */
int gpgmp_solver_skip(gpgmp_solver *s, int32_t steps, float *sim_time_out);
/* Captures the HIP graphs that gpgmp_solver_run_continue(s, steps, graph_chunk) will replay,
 * without running any step (the loop, state and seeds are unchanged), so that a following
 * run of a few steps replays graphs instead of capturing them while it runs.  Graphs are
 * kept per chunk buffer, chunk length and outflow-buffer parity.  No counterpart in the
 * reference (its loop launches every kernel eagerly, Kernel.h:248-272). */
int gpgmp_solver_prepare_run(gpgmp_solver *s, int32_t steps, int32_t graph_chunk);
/* launch shape of the solver's kernels: threads per block (64..256), cells per thread */
/* Flattened sweeps (n = 8 or 16 cells per wave): on (default) spreads each wave's cells over
 * the grid (stride W*H/n) so that waves over dense and sparse regions even out; off keeps n
 * consecutive cells.  Results are identical either way (no reference counterpart). */
int gpgmp_solver_set_sweep_spread(gpgmp_solver *s, int32_t on);
/* Split steps (one lane per cell, flattened sweeps): on runs the cells that were busy in
 * the previous step (more SSA events than the budget) whole in worker waves of the SSA
 * launch, beside the cheap cells; off (default, faster on the 256x256 Schnakenberg run)
 * defers them to the X-sweep launch as they turn up.  Results are identical either way
 * (no reference counterpart). */
int gpgmp_solver_set_predict_busy(gpgmp_solver *s, int32_t on);
/* Flattened launches without an SSA (X and Y sweeps): 8 (default) or 16 cells per wave.
 * 8 gives twice the waves (256x256 Schnakenberg: X sweeps + busy cells 24.2 -> 23.3 us).
 * Results are identical either way (no reference counterpart). */
int gpgmp_solver_set_sweep_wave_cells(gpgmp_solver *s, int32_t cells);
int gpgmp_solver_set_launch(gpgmp_solver *s, int32_t block, int32_t cells_per_thread);
/* Lanes per cell (1 or 4, default 1) of the launches with an SSA, and the sweep
 * scheme (default 0): 1 or 4 lanes per cell, or 0 = the particle loops of a wave's cells
 * flattened over all its lanes.  With 4 SSA lanes, four lanes share a cell's SSA (random words and logarithms of
 * 4 events drawn at once).  Results are identical for every choice.  No counterpart in
 * the reference (a launch shape of this build's fused kernel). */
int gpgmp_solver_set_lanes_per_cell(gpgmp_solver *s, int32_t ssa_lanes, int32_t sweep_lanes);
/* current launch shape */
int gpgmp_solver_get_launch(gpgmp_solver *s, int32_t *block, int32_t *cells_per_thread, int32_t *ssa_lanes,
                            int32_t *sweep_lanes);
/* Reads the accumulated per-launch HIP-event times of the current run's eager launches
 * (ms[5], n[5], algorithmic HBM bytes[5] by kernel: 0 homogeneous, 1 SSA (+ X sweeps
 * unless they run in their own launch), 2 Y sweeps, 3 hot-cell continuation in its own
 * launch, 4 X sweeps in their own launch, together with the hot-cell continuation; may be
 * NULL), then sets the timing of the following steps: on = 0
 * off, on = 1 every launch, on = k > 1 the launches of 4 consecutive host-loop steps out
 * of every k, each window led by one more step whose events are not counted (events on
 * launches right after untimed ones read skewed times).  Timing uses dispatch-tied HIP
 * events; each timed launch adds a few microseconds of dispatch overhead, so windows keep
 * the run's own speed. */
int gpgmp_solver_kernel_times(gpgmp_solver *s, int32_t on, double *ms, int64_t *n, double *bytes);
/* Deferral of busy cells: a cell whose SSA has drawn `budget` events (default 8) moves
 * to the one-wave-per-cell continuation kernel once fewer than `below_active` lanes of
 * its wave still run their SSA (default 24; 65 = always at the budget).  Results do not
 * depend on either.  No counterpart in the reference. */
int gpgmp_solver_set_event_budget(gpgmp_solver *s, uint32_t budget, int32_t below_active);
/* Workers of the deferred cells inside the X-sweep launch of a split step (the default
 * inhomogeneous launch shape): a deferred cell whose expected remaining events
 * (h0 * (dt - t) when it was deferred) are at most lane_events (default 16) is resumed by
 * one lane of one of lane_waves worker waves (default 1024; 64 cells per wave and round),
 * the others by one wave each of wave_waves (default 1024) with speculative event batches;
 * lane_events <= 0 sends every deferred cell to the wave workers.  Results do not depend
 * on any of them.  No counterpart in the reference. */
int gpgmp_solver_set_hot_workers(gpgmp_solver *s, float lane_events, int32_t lane_waves, int32_t wave_waves);
int gpgmp_solver_synchronize(gpgmp_solver *s);

/* ---- batched ensemble (no reference counterpart) ----------------------------------- */
/* n realizations of one network on grids of their own size, each with its own seed0;
 * the cell width h = physical_length/width of the template is kept for every grid.
 * Counts are bit-exact with a plain solver of the same grid and seed0. */
int gpgmp_batch_create(const gpgmp_model *tmpl, int32_t n, const int32_t *widths,
                       const int32_t *heights, const uint32_t *seed0s, void *stream,
                       gpgmp_batch **out);
int gpgmp_batch_destroy(gpgmp_batch *b);
int gpgmp_batch_set_state(gpgmp_batch *b, int32_t r, const int32_t *host_counts);
int gpgmp_batch_get_state(gpgmp_batch *b, int32_t r, int32_t *host_counts);
/* launch shape: threads per block (64..256) and cells per thread (1..16) */
int gpgmp_batch_set_launch(gpgmp_batch *b, int32_t block, int32_t cells_per_thread);
/* lanes per cell (1 or 4), as gpgmp_solver_set_lanes_per_cell */
int gpgmp_batch_set_lanes_per_cell(gpgmp_batch *b, int32_t ssa_lanes, int32_t sweep_lanes);
int gpgmp_batch_get_launch(gpgmp_batch *b, int32_t *block, int32_t *cells_per_thread, int32_t *ssa_lanes,
                           int32_t *sweep_lanes);
/* Times the candidate launch shapes (threads per block, cells per thread, lanes per
 * cell) on `trial_steps` steps from a copy of the current state (restored afterwards)
 * and keeps the fastest; reports block and cells per thread (gpgmp_batch_get_launch
 * gives the lanes too). */
int gpgmp_batch_autotune(gpgmp_batch *b, int32_t trial_steps, int32_t *block_out,
                         int32_t *cpt_out);
/* Runs `steps` host-loop iterations of every realization.  graph_chunk > 0 captures
 * chunks of that many iterations in a HIP graph that is replayed. */
int gpgmp_batch_run(gpgmp_batch *b, int32_t steps, int32_t graph_chunk);
/* As gpgmp_solver_prepare_run, for gpgmp_batch_run(b, steps, graph_chunk). */
int gpgmp_batch_prepare_run(gpgmp_batch *b, int32_t steps, int32_t graph_chunk);
/* Restarts the host loops at time 0 with the realizations' seed0 (state is kept). */
int gpgmp_batch_reset(gpgmp_batch *b);
int gpgmp_batch_kernel_times(gpgmp_batch *b, int32_t on, double *ms, int64_t *n, double *bytes);
int gpgmp_batch_set_event_budget(gpgmp_batch *b, uint32_t budget, int32_t below_active);
int gpgmp_batch_total_cells(gpgmp_batch *b, int64_t *cells);
int gpgmp_batch_synchronize(gpgmp_batch *b);

/* ---- inhomogeneous time step (computeIndividualTimestepX/Y + reduceBlocks) -------- */
/* Per-cell admissible time step (StochasticInhomogeneousDriftDiffusion.cl:130-151) and
 * its minimum over n cells, on the device in double; D, r device pointers. */
int gpgmp_min_timestep_device(const float *D, const float *r, int32_t n, float h, float *out_host,
                              void *stream);

/* ---- deterministic reaction-drift-diffusion (Schnakenberg, fp32) ------------------ */
typedef struct {
    int32_t width, height;
    float h, dt;
    float D0u, D0v;          /* per-species diffusivity scale of the field D */
    float k[4];              /* k1 (0->U), k2 (0->V), k3 (2U+V->3U), k4 (U->0) */
} gpgmp_rdd_params;
/* Workspace for `steps` steps with a checkpoint every ckpt_every steps. */
int gpgmp_rdd_create(const gpgmp_rdd_params *p, int32_t steps, int32_t ckpt_every, void *stream,
                     gpgmp_rdd **out);
int gpgmp_rdd_destroy(gpgmp_rdd *r);
/* Replaces the rate constants k1..k4 (host floats) for the following forward runs. */
int gpgmp_rdd_set_rates(gpgmp_rdd *r, const float *k);
/* The same from a device array k_dev[4], copied on the handle's stream: the host never waits
 * for the rates (a torch tensor of rates stays on the GPU); the kernels read them from the
 * handle's copy until gpgmp_rdd_set_rates is called again. */
int gpgmp_rdd_set_rates_device(gpgmp_rdd *r, const float *k_dev);
/* Generation of the saved forward run (1, 2, ... for each forward with save != 0; 0 when no
 * saved forward exists or the rates changed since). */
int gpgmp_rdd_generation(gpgmp_rdd *r, uint64_t *generation);
/* Euler steps per forward kernel launch, 1 .. 4 (sides >= ns + 1; default 4 when both sides
 * are >= 5, else 2 when >= 3): ns steps per launch keep the intermediate states in LDS
 * (ns-cell halo) and give results bit-identical to one step per launch; the adjoint and its
 * segment recompute take 4 steps per launch when ns = 4, else min(ns, 2).  No counterpart in
 * the reference (a launch shape of this build's forward/adjoint). */
int gpgmp_rdd_set_steps_per_launch(gpgmp_rdd *r, int32_t ns);
/* Forward: u,v (device, in/out) advanced `steps` fused Euler steps with field D; stores
 * checkpoints when save != 0. */
int gpgmp_rdd_forward(gpgmp_rdd *r, const float *D, float *u, float *v, int32_t save);
/* Adjoint of the last saved forward: lu/lv (device) carry dL/du_T, dL/dv_T in and
 * dL/du_0, dL/dv_0 out; outputs gD = dL/dD (device, [H][W]) and gk = dL/dk1..k4
 * (device, 4 floats).  generation (from gpgmp_rdd_generation after that forward) must be
 * the saved forward's, else GPGMP_E_STATE (a later forward replaced its checkpoints);
 * 0 skips the check. */
int gpgmp_rdd_backward(gpgmp_rdd *r, uint64_t generation, const float *D, float *lu, float *lv, float *gD,
                       float *gk);

/* Deterministic stencils and RK4 of Deterministic.cl on device fp32 buffers, Jacobi
 * (in != out): homogeneous_deterministic_diffuse (Deterministic.cl:17-65) for every
 * species, inhomogeneous_deterministic_diffuseX/Y (:71-136) for one species (only that
 * species' slice of `out` is written), deterministic_performReaction_rk4 (:179-238). */
int gpgmp_det_hom_diffuse_device(const float *in, float *out, int32_t W, int32_t H, int32_t S,
                                 const float *D_host, float dt, float h, void *stream);
int gpgmp_det_inhom_sweep_device(const float *in, float *out, int32_t W, int32_t H, int32_t species,
                                 int32_t axis, float dt, const float *D_axis, const float *r_axis,
                                 float h, void *stream);
int gpgmp_det_rk4_device(const gpgmp_model *m, float *state, float dt, void *stream);
/* deterministic_performReaction_aqss (Deterministic.cl:291-522, called by Solver::doReactions
 * for the alpha-QSS solver types, Solver.cpp:257-263): Mott & Oran's alpha-QSS
 * predictor-corrector over dt per cell on the fp32 state [S][H][W] in place, production and
 * destruction of deterministic_rhs_ode_aqss (Deterministic.cl:241-287).  flags 0 keeps the
 * reference's corrector factor `180.*+rtaub*(...)` (Deterministic.cl:423);
 * GPGMP_AQSS_PADE_FIX uses the predictor's Pade factor there too.  A cell that needs more
 * than 2^16 inner passes (a step size that underflows, e.g. a cell at zero under the
 * reference corrector) is set to NaN. */
#define GPGMP_AQSS_PADE_FIX 1
int gpgmp_det_aqss_device(const gpgmp_model *m, float *state, float dt, int32_t flags, void *stream);
/* inhomogeneous_computeMomentsX + inhomogeneous_reduceMomentsBlocks
 * (StochasticInhomogeneousDriftDiffusion.cl:5-82, InhomogeneousSolver.cpp:288-299): for
 * species `species` of the int32 state [S][H][W] on the device, sums[0] = sum of x * n(x, y)
 * (x = cell column) and sums[1] = sum of n(x, y), written to the device array sums[2]
 * (stream-ordered).  Exact 64-bit integer sums instead of the reference's float tree. */
int gpgmp_moments_x_device(const int32_t *state, int32_t W, int32_t H, int32_t species, int64_t *sums,
                           void *stream);

/* ---- individual particles (IndividualSolver sort / cell indices) ------------------- */
/* Stable LSD radix sort of n (key, value) pairs by the 32-bit key, device pointers;
 * keys_tmp/vals_tmp are n-element scratch (IndividualSolver::sortSpecies,
 * IndividualSolver.cpp:673-808).  key_bits limits the passes (32 = full key). */
int gpgmp_sort_pairs_device(uint32_t *keys, uint32_t *vals, uint32_t *keys_tmp, uint32_t *vals_tmp,
                            int32_t n, int32_t key_bits, void *stream);
/* gpgmp_sort_pairs_device without any device allocation: no counterpart in the reference;
 * for callers that cannot allocate (code under HIP graph capture, tight serving loops, hosts
 * that forbid hipMalloc).  The per-tile digit histograms live in a library-owned device array
 * of 4 MiB + 1 KiB per GPU (held until the process ends), so n is at most 2^24
 * (GPGMP_E_CAPACITY above); keys_tmp/vals_tmp are still the caller's n-element buffers.
 * Stream-ordered, no host synchronisation.  Every call on a GPU shares that array: calls
 * that overlap on one GPU (other streams or threads) race. */
int gpgmp_sort_pairs_device_noalloc(uint32_t *keys, uint32_t *vals, uint32_t *keys_tmp, uint32_t *vals_tmp,
                                    int32_t n, int32_t key_bits, void *stream);
/* CSR offsets off[c] = #(key < c), c = 0..ncells, of sorted keys. */
int gpgmp_cell_offsets_device(const uint32_t *sorted_keys, int32_t n, int32_t ncells, uint32_t *off,
                              void *stream);
/* computeCellIndices (IndividualSolver.cl:564-619): left/right per cell. */
int gpgmp_cell_left_right_device(const uint32_t *sorted_keys, int32_t n, int32_t ncells,
                                 uint32_t *left, uint32_t *right, void *stream);

typedef struct {
    int32_t side;            /* cells per side, 1 um cells, periodic domain [0, side)^2 */
    float D, dt, vortex_amp, p_annihilate;
    uint32_t seed;
} gpgmp_particle_params;
int gpgmp_particles_create(const gpgmp_particle_params *p, int32_t capacity, void *stream,
                           gpgmp_particles **out);
int gpgmp_particles_destroy(gpgmp_particles *ps);
int gpgmp_particles_set(gpgmp_particles *ps, const float *x, const float *y, const uint32_t *id,
                        int32_t n);
/* n_out holds the live count; buffers sized by the caller to that count */
int gpgmp_particles_get(gpgmp_particles *ps, float *x, float *y, uint32_t *id, uint32_t *keys,
                        uint32_t *off, int32_t *n_out);
int gpgmp_particles_count(gpgmp_particles *ps, int32_t *n_out);
/* `steps` steps starting at step index `step0`: move, stable sort, offsets, annihilate. */
int gpgmp_particles_step(gpgmp_particles *ps, int32_t step0, int32_t steps);
/* Which step runs once the list is sorted by cell (after the first step): 1 (default) the
 * step on the sorted list -- per-cell counts of the moves to the 3x3 neighbour cells and one
 * scatter of every particle to its stable-order place, with moves beyond the neighbours on an
 * exact but slow far list -- when 6 sigma + |v| dt < 1 cell, else the radix step; 0 always the
 * radix step (move, stable LSD radix sort, offsets, annihilation, compaction); 2 always the
 * sorted-list step.  Results are identical (no reference counterpart). */
int gpgmp_particles_set_sorted_step(gpgmp_particles *ps, int32_t mode);
/* Sum over the steps run so far of the particles each step processed. */
int gpgmp_particles_processed(gpgmp_particles *ps, uint64_t *particle_steps);

/* ---- individual (particle-based) species on the lattice (gpgmp::IndividualSolver) -- */
/* One particle-based species of an inhomogeneous stochastic model: particles are cell
 * keys with four properties each (DiffusivityX, DiffusivityY, DriftX, DriftY; the
 * computeDiffusivityDrift hook is the identity, IndividualSolver.cl:690-694). */
typedef struct {
    int32_t width, height;       /* grid cells */
    float physical_length;       /* domain side in um (cell width = header float of L/width) */
    int32_t boundary_mask[4];    /* -x,+x,-y,+y: 0 reflective, 1 absorbing (the key leaves or wraps) */
    int32_t species;             /* species index: RNG key = cell + species*width*height */
} gpgmp_individuals_desc;
/* IndividualSolver(...) buffers for `capacity` particles (IndividualSolver.cpp:37-395). */
int gpgmp_individuals_create(const gpgmp_individuals_desc *d, int32_t capacity, void *stream,
                             gpgmp_individuals **out);
int gpgmp_individuals_destroy(gpgmp_individuals *g);
/* Sets n particles (host): cells[n] in [0, W*H), props[4][n]; particle index i carries
 * props[k][i].  Then sorts, computes the per-cell constants and both first time steps
 * (InhomogeneousSolver.cpp:205-225 for an individual species). */
int gpgmp_individuals_set(gpgmp_individuals *g, const uint32_t *cells, const float *props, int32_t n);
/* Current time steps of the X and Y sweeps (m_diffusionDtX/Y). */
int gpgmp_individuals_time_steps(gpgmp_individuals *g, float *dt_x, float *dt_y);
/* IndividualSolver::diffuseX / diffuseY (IndividualSolver.cpp:416-456, kernels
 * IndividualSolver.cl:750-1014) with time step dt and *seed, then *seed += 1 as
 * InhomogeneousSolver::diffuse does (InhomogeneousSolver.cpp:268-269, 361-362). */
int gpgmp_individuals_diffuse(gpgmp_individuals *g, int32_t axis, float dt, int32_t *seed);
/* IndividualSolver::computeDiffusionConstants (IndividualSolver.cpp:1100-1153: sort,
 * cell indices, constants kernel IndividualSolver.cl:654-747) and the minimum per-cell
 * time step of the axes in `axes` (bit 0: X, bit 1: Y; InhomogeneousSolver.cpp:304-336). */
int gpgmp_individuals_update(gpgmp_individuals *g, int32_t axes, float *dt_x, float *dt_y);
/* DiffusionModel::oclSolve's loop (DiffusionModel.cpp:1467-1603) for this species alone
 * (no reactions): `steps` iterations from seed0, X then Y sweep when due, each followed by
 * the update of its time step (InhomogeneousSolver::diffuse, InhomogeneousSolver.cpp:254-427). */
int gpgmp_individuals_run(gpgmp_individuals *g, int32_t steps, uint32_t seed0, float *sim_time_out);
/* Sorted keys[n] and particle indices vals[n], per-cell counts[W*H] (any may be NULL),
 * and n (IndividualSolver::getPositionList, IndividualSolver.h:36-297). */
int gpgmp_individuals_get(gpgmp_individuals *g, uint32_t *keys, uint32_t *vals, int32_t *counts, int32_t *n);
/* Per-cell max diffusivities and max-magnitude drifts of the last update (host, [W*H]). */
int gpgmp_individuals_constants(gpgmp_individuals *g, float *dx, float *dy, float *rx, float *ry);
/* Hop sweeps (X and Y) done since the particles were set. */
int gpgmp_individuals_sweeps(gpgmp_individuals *g, int64_t *sweeps);
/* Device pointers of the sorted keys, particle indices and CSR cell offsets [W*H+1]. */
int gpgmp_individuals_keys_device(gpgmp_individuals *g, uint32_t **keys, uint32_t **vals, uint32_t **off);

/* IndividualSolver::maintainPopulation(speciesIndex, seed, d_state) (IndividualSolver.cpp:458-671,
 * kernels IndividualSolver.cl:1017-1599), called by Solver::doReactions after the SSA for each
 * individual species with seed++ (Solver.cpp:241-247): afterwards cell c holds exactly goal[c]
 * particles (goal = the species' counts after the reactions, W*H int32, on the device if
 * goal_on_device, e.g. the solver's state from gpgmp_solver_state_device, else host).
 * Kills pick uniformly with the reference's loop (RNG key {cell, 0xdecafbad + seed}); new
 * particles recycle dead ids from the front of the pool in cell order and go before the
 * cell's survivors; `rule` (NULL = keep) sets their properties with RANDOM keyed
 * {cell, 0xdecafbad + seed + 1}.  The list stays sorted; the per-cell counts and constants
 * are refreshed, the time steps are not (as in the reference).  GPGMP_E_CAPACITY when the
 * pool holds fewer ids than needed (nothing changes); GPGMP_E_INVALID for a negative goal. */
int gpgmp_individuals_maintain(gpgmp_individuals *g, const int32_t *goal, int32_t goal_on_device, int32_t seed,
                               const gpgmp_new_individual_rule *rule);
/* Mirror the per-cell counts into dev_counts[W*H] (device; NULL stops) after every update and
 * maintenance, as computeDiffusionConstants writes the species' slice of d_state
 * (IndividualSolver.cl:700-747); copies the current counts now. */
int gpgmp_individuals_bind_counts(gpgmp_individuals *g, int32_t *dev_counts);
/* Waits for the handle's queued work. */
int gpgmp_individuals_synchronize(gpgmp_individuals *g);
/* Properties of the live particles in sorted order, props[4][n]
 * (IndividualSolver::getPropertiesList, IndividualSolver.h:36-297). */
int gpgmp_individuals_get_props(gpgmp_individuals *g, float *props);
/* The dead-particle pool front first, i.e. the reference's key list after the live
 * particles (positions n .. cap-1, key UINT_MAX): ids[cap-n] and props[4][cap-n] (either
 * may be NULL), and its size. */
int gpgmp_individuals_get_pool(gpgmp_individuals *g, uint32_t *ids, float *props, int32_t *np);

#ifdef __cplusplus
}
#endif
#endif
