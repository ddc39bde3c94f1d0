"""Python mirror of the GPGMP solver interface over the C ABI.

Class and method names follow the reference's C++ solver plugin
(exec/src/gpgmp/src/Solver.h:15-239 and its subclasses) so that code driving a
reference solver reads the same here:

    model = DiffusionModel(width=32, height=32, physical_length=32.0, p0=0.15)
    a = model.addSpecies("A", 1.0) ...
    model.addReaction(rate=1.0, products={a: 1})
    solver = model.createSolver()           # HomogeneousSolver / InhomogeneousSolver
    seed = 7
    seed = solver.doReactions(dt, seed)     # the reference mutates `int &seed`
    seed = solver.diffuse(species, t, seed)
    solver.getNextDiffusionTime()

Seeds are returned instead of mutated in place (Python has no `int&`).
All compute runs in libgpgmp_hip.so; nothing here falls back to the CPU.
"""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import call, lib

STOCHASTIC_HOMOGENEOUS = _lib.STOCHASTIC_HOMOGENEOUS
STOCHASTIC_INHOMOGENEOUS = _lib.STOCHASTIC_INHOMOGENEOUS
STOCHASTIC_INHOMOGENEOUS_FPE = _lib.STOCHASTIC_INHOMOGENEOUS_FPE
DETERMINISTIC_HOMOGENEOUS_RK4 = _lib.DETERMINISTIC_HOMOGENEOUS_RK4
DETERMINISTIC_INHOMOGENEOUS_RK4 = _lib.DETERMINISTIC_INHOMOGENEOUS_RK4
DETERMINISTIC_HOMOGENEOUS_AQSS = _lib.DETERMINISTIC_HOMOGENEOUS_AQSS
DETERMINISTIC_INHOMOGENEOUS_AQSS = _lib.DETERMINISTIC_INHOMOGENEOUS_AQSS
DETERMINISTIC_TYPES = (DETERMINISTIC_HOMOGENEOUS_RK4, DETERMINISTIC_INHOMOGENEOUS_RK4,
                       DETERMINISTIC_HOMOGENEOUS_AQSS, DETERMINISTIC_INHOMOGENEOUS_AQSS)
# gpgmp_solver_create flags
SOLVER_AQSS_PADE_FIX = 1
SOLVER_REFERENCE_QUIRKS = 2


def _fptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _iptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_int32))


def _uptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_uint32))


class Species:
    """Species (Species.h): identifier, index and homogeneous diffusion constant."""

    def __init__(self, id, index, diffusion_constant):
        self.id = id
        self.index = index
        self.diffusionConstant = float(diffusion_constant)


class Compartment:
    """Rectangular compartment (Compartment.h:50-158): inclusive cell bounds
    x0 <= x <= x1, y0 <= y <= y1 (setBounds / isInCompartment)."""

    def __init__(self, id, x0=0, y0=0, x1=0, y1=0):
        self.id = id
        self.setBounds(x0, y0, x1, y1)

    def setBounds(self, x0, y0, x1, y1):
        self.bounds = (int(x0), int(y0), int(x1), int(y1))

    def getBounds(self):
        return self.bounds

    def isInCompartment(self, x, y, z=0):
        x0, y0, x1, y1 = self.bounds
        return x0 <= x <= x1 and y0 <= y <= y1


class Reaction:
    """Mass-action reaction (Reaction.h): rate and stoichiometry maps by species, and the
    compartments of a localized reaction (Reaction::addCompartment)."""

    def __init__(self, rate, reactants=None, products=None, id=None):
        self.rate = float(rate)
        self.reactants = dict(reactants or {})
        self.products = dict(products or {})
        self.id = id
        self.compartments = []
        self.customLaw = None

    # Reaction::setCustomLaw (Reaction.cpp:108-111): the law as C expression text
    def setCustomLaw(self, formula):
        self.customLaw = str(formula) if formula else None

    # Reaction::addCompartment (Reaction.cpp:209-212)
    def addCompartment(self, compartment):
        self.compartments.append(compartment)

    def isLocalized(self):
        return bool(self.compartments)

    # Reaction::reactionMask (Reaction.cpp:216-235)
    def reactionMask(self, x, y, z=0):
        if not self.compartments:
            return 1.0
        return 1.0 if any(c.isInCompartment(x, y, z) for c in self.compartments) else 0.0


class DiffusionModel:
    """The model description the solvers read (DiffusionModel.h:149-773 setters)."""

    def __init__(self, width, height, physical_length, p0=0.15,
                 solver=STOCHASTIC_HOMOGENEOUS, boundary_mask=(0, 0, 0, 0),
                 source_mask=(0, 0, 0, 0), max_time=1e30):
        self.width = int(width)
        self.height = int(height)
        self.physical_length = float(physical_length)
        self.p0 = float(p0)
        self.solver = int(solver)
        self.boundary_mask = tuple(int(b) for b in boundary_mask)
        self.source_mask = tuple(int(s) for s in source_mask)
        self.max_time = float(max_time)
        self.species = []
        self.reactions = []
        self.parameters = {}  # DiffusionModel::setParameter: name -> value (a Real)
        self.fields = None  # (Dx, Dy, rx, ry) each [S][H][W] float32

    # DiffusionModel::setParameter (a name custom laws may use)
    def setParameter(self, name, value):
        self.parameters[str(name)] = float(value)

    # DiffusionModel::addSpecies
    def addSpecies(self, id, diffusion_constant=0.0):
        sp = Species(id, len(self.species), diffusion_constant)
        self.species.append(sp)
        return sp

    # DiffusionModel::addReaction (stoichiometry given as {species: count})
    def addReaction(self, rate, reactants=None, products=None, id=None):
        r = Reaction(rate, reactants, products, id)
        self.reactions.append(r)
        return r

    def setDriftDiffusivityFields(self, Dx, Dy, rx, ry):
        S, H, W = len(self.species), self.height, self.width
        arrs = [np.ascontiguousarray(a, dtype=np.float32).reshape(S, H, W) for a in (Dx, Dy, rx, ry)]
        self.fields = tuple(arrs)

    def numSpecies(self):
        return len(self.species)

    def gridArea(self):
        return self.width * self.height

    def _desc(self):
        S, R = len(self.species), len(self.reactions)
        if R > _lib.MAX_REACTIONS or S > _lib.MAX_SPECIES:
            raise ValueError("model exceeds %d species / %d reactions" % (_lib.MAX_SPECIES, _lib.MAX_REACTIONS))
        rx = (_lib.ReactionDesc * max(R, 1))()
        for i, r in enumerate(self.reactions):
            rx[i].rate = r.rate
            for sp, n in r.reactants.items():
                rx[i].reactant_stoich[sp.index if isinstance(sp, Species) else int(sp)] += int(n)
            for sp, n in r.products.items():
                rx[i].product_stoich[sp.index if isinstance(sp, Species) else int(sp)] += int(n)
        d = _lib.ModelDesc()
        d.width, d.height = self.width, self.height
        d.physical_length, d.p0 = self.physical_length, self.p0
        d.num_species, d.num_reactions = S, R
        for i, sp in enumerate(self.species):
            d.diffusion[i] = sp.diffusionConstant
        d.reactions = C.cast(rx, C.POINTER(_lib.ReactionDesc))
        for i in range(4):
            d.boundary_mask[i] = self.boundary_mask[i]
            d.source_mask[i] = self.source_mask[i]
        d.max_time = self.max_time
        d.solver_type = self.solver
        keep = [rx]
        if self.fields is not None:
            for name, arr in zip(("diffusivity_x", "diffusivity_y", "drift_x", "drift_y"), self.fields):
                setattr(d, name, _fptr(arr))
                keep.append(arr)
        return d, keep

    def compile(self):
        d, keep = self._desc()
        h = C.c_void_p()
        call("gpgmp_model_create", C.byref(d), C.byref(h))
        cm = CompiledModel(h, self)
        ids = (C.c_char_p * max(len(self.species), 1))(*[sp.id.encode() for sp in self.species])
        call("gpgmp_model_set_species_ids", h, ids)
        for name, value in self.parameters.items():
            call("gpgmp_model_set_parameter", h, name.encode(), C.c_float(value))
        for i, r in enumerate(self.reactions):
            for cp in r.compartments:
                call("gpgmp_model_add_reaction_compartment", h, i, *cp.getBounds())
        for i, r in enumerate(self.reactions):
            if r.customLaw:
                call("gpgmp_model_set_kinetic_law", h, i, r.customLaw.encode())
        return cm

    def hasLocalizedReactions(self):
        return any(r.isLocalized() for r in self.reactions)

    def reactionActiveBits(self, width=None, height=None):
        """Per cell [H][W] uint16: bit i set where reaction i's mask is 1 (the reaction
        mask Solver::oclCreateReactionMask uploads, Solver.cpp:358-392)."""
        W = self.width if width is None else width
        H = self.height if height is None else height
        act = np.zeros((H, W), np.uint16)
        for i, r in enumerate(self.reactions):
            if not r.compartments:
                act |= np.uint16(1 << i)
                continue
            for cp in r.compartments:
                x0, y0, x1, y1 = cp.getBounds()
                x0, y0, x1, y1 = max(x0, 0), max(y0, 0), min(x1, W - 1), min(y1, H - 1)
                if x0 <= x1 and y0 <= y1:
                    act[y0:y1 + 1, x0:x1 + 1] |= np.uint16(1 << i)
        return act

    def createSolver(self, stream=None, flags=0):
        """Factory of DiffusionModel::oclSolve (DiffusionModel.cpp:1422-1440).  A model given a
        non-linear drift/diffusivity method (setNonlinearMethod) gets it attached."""
        cm = self.compile()
        if self.solver in DETERMINISTIC_TYPES:
            return DeterministicSolver(cm, stream, flags)
        cls = HomogeneousSolver if self.solver == STOCHASTIC_HOMOGENEOUS else InhomogeneousSolver
        s = cls(cm, stream)
        nl = getattr(self, "nonlinear_method", None)
        if nl is not None:
            s.setNonlinearMethod(*nl)
        return s

    def setNonlinearMethod(self, method, field_names=(), fields=None):
        """m_nonlinearDiffusivity: the drift/diffusivity method re-evaluated after every sweep
        (InhomogeneousSolver.cpp:304-336,384-417) by the solvers createSolver makes from now
        on; fields: [len(field_names)][H][W] values of the field parameters."""
        self.nonlinear_method = (method, tuple(field_names), fields)


class CompiledModel:
    def __init__(self, handle, model):
        self.handle = handle
        self.model = model

    def rateLiterals(self):
        R = len(self.model.reactions)
        is_int = np.zeros(max(R, 1), np.int32)
        val = np.zeros(max(R, 1), np.float64)
        call("gpgmp_model_rate_literals", self.handle, _iptr(is_int),
             val.ctypes.data_as(C.POINTER(C.c_double)))
        return is_int[:R].astype(bool), val[:R]

    def __del__(self):
        try:
            if self.handle:
                lib().gpgmp_model_destroy(self.handle)
                self.handle = None
        except Exception:
            pass


def _stream_ptr(stream):
    if stream is None:
        return None
    if isinstance(stream, int):
        return stream
    return int(stream.cuda_stream)  # torch.cuda.Stream


def loopSeeds(seed0, n):
    """seed_n of the host loop's first n iterations from seed0 (DiffusionModel.cpp:1467-1482)."""
    out = np.zeros(max(int(n), 1), np.int32)
    call("gpgmp_loop_seeds", C.c_uint32(seed0), int(n), _iptr(out))
    return out[: int(n)]


class Solver:
    """gpgmp::Solver (Solver.h:15-239) over the HIP library."""

    def __init__(self, compiled, stream=None, flags=0):
        self._cm = compiled
        self.model = compiled.model
        h = C.c_void_p()
        call("gpgmp_solver_create", compiled.handle, _stream_ptr(stream), int(flags), C.byref(h))
        self.handle = h

    def close(self):
        if getattr(self, "handle", None):
            lib().gpgmp_solver_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def shape(self):
        m = self.model
        return (m.numSpecies(), m.height, m.width)

    # Solver::synchronizeStateBufferToDevice (host state -> device)
    def synchronizeStateBufferToDevice(self, state):
        a = np.ascontiguousarray(state, dtype=np.int32).reshape(self.shape)
        call("gpgmp_solver_set_state", self.handle, _iptr(a))

    # Solver::synchronizeStateBufferFromDevice (device -> host state)
    def synchronizeStateBufferFromDevice(self):
        a = np.empty(self.shape, np.int32)
        call("gpgmp_solver_get_state", self.handle, _iptr(a))
        return a

    def stateDevicePointer(self):
        p = C.c_void_p()
        call("gpgmp_solver_state_device", self.handle, C.byref(p))
        return p.value

    # Solver::doReactions(Real dt, int &seed) -> returns the updated seed
    def doReactions(self, dt, seed):
        s = C.c_int32(seed)
        call("gpgmp_solver_do_reactions", self.handle, C.c_float(dt), C.byref(s))
        return s.value

    # Solver::diffuse(int species, Real currentTime, int &seed) -> returns the updated seed
    def diffuse(self, species, current_time, seed):
        s = C.c_int32(seed)
        call("gpgmp_solver_diffuse", self.handle, int(species), C.c_float(current_time), C.byref(s))
        return s.value

    def setNonlinearMethod(self, method, field_names=(), fields=None):
        """Non-linear diffusivity (gpgmp_solver_set_nonlinear_method): after every sweep that
        diffuse runs, the method on the swept counts at the call's time gives the species' fields,
        the swept axis' time step and both axes' probabilities.  The model then runs through
        runPerCall (run refuses it)."""
        H, W = self.model.height, self.model.width
        names = list(field_names)
        fv = np.zeros((max(len(names), 1), H, W), np.float32)
        if names:
            fv[: len(names)] = np.asarray(fields, np.float32).reshape(len(names), H, W)
        cnames = (C.c_char_p * max(len(names), 1))(*[n.encode() for n in names])
        call("gpgmp_solver_set_nonlinear_method", self.handle, method.encode(), cnames, len(names), _fptr(fv))

    def runPerCall(self, steps, seed0, sim_time=0.0, output_interval=float("inf"), max_time=float("inf"),
                   on_output=None):
        """DiffusionModel::oclSolve's loop (DiffusionModel.cpp:1479-1603) through doReactions,
        diffuse and getNextDiffusionTime, one call each per step, for what the device loop does
        not run (non-linear diffusivity, individual species): from the clock sim_time, seeds
        seed_n of seed0 (loopSeeds).  Output events every output_interval (:1519-1527,
        1553-1572): the initial state is written first (Dump_0), an output time before the next
        diffusion makes the iteration reactions only, and once the clock has reached it the
        iteration after writes -- each write calls on_output(time, counts); the loop stops once
        the clock reaches max_time (:1479-1480).  Returns (clock, iterations run)."""
        seeds = loopSeeds(seed0, steps)
        sim = np.float32(sim_time)
        interval, out_last = np.float32(output_interval), np.float32(0)
        react = len(self.model.reactions) > 0
        if on_output is not None and (np.isfinite(interval) or np.isfinite(max_time)):
            on_output(float(sim), self.synchronizeStateBufferFromDevice())
        it = 0
        while it < steps and sim < np.float32(max_time):
            nxt = self.getNextDiffusionTime()
            minspec = int(np.argmin(nxt))  # the first minimum
            mindt = nxt[minspec]
            out_time = np.float32(out_last + interval)
            output_event = out_time < mindt
            if output_event:
                mindt = out_time
            seed = int(seeds[it])
            if react:
                seed = self.doReactions(np.float32(mindt - sim), seed)
            if output_event:
                if sim >= np.float32(out_last + interval):
                    if on_output is not None:
                        on_output(float(mindt), self.synchronizeStateBufferFromDevice())
                    out_last = np.float32(out_last + interval)
            else:
                for i in range(len(nxt)):
                    if (i == minspec or nxt[i] <= mindt) and nxt[i] > 0:
                        seed = self.diffuse(i, mindt, seed)
            sim = mindt
            it += 1
        return float(sim), it

    # Solver::getNextDiffusionTime()
    def getNextDiffusionTime(self):
        out = np.zeros(self.model.numSpecies(), np.float32)
        call("gpgmp_solver_next_diffusion_time", self.handle, _fptr(out))
        return out

    def timeSteps(self):
        S = self.model.numSpecies()
        x = np.zeros(S, np.float32)
        y = np.zeros(S, np.float32)
        call("gpgmp_solver_time_steps", self.handle, _fptr(x), _fptr(y))
        return x, y

    def run(self, steps, seed0):
        """DiffusionModel::oclSolve's loop (DiffusionModel.cpp:1467-1603) for `steps`
        iterations from seed0, on the device without per-step synchronisation."""
        t = C.c_float()
        call("gpgmp_solver_run", self.handle, int(steps), C.c_uint32(seed0), C.byref(t))
        return t.value

    def runContinue(self, steps, graph_chunk=0):
        t = C.c_float()
        call("gpgmp_solver_run_continue", self.handle, int(steps), int(graph_chunk), C.byref(t))
        return t.value

    def prepareRun(self, steps, graph_chunk):
        """Captures the HIP graphs runContinue(steps, graph_chunk) will replay (runs nothing)."""
        call("gpgmp_solver_prepare_run", self.handle, int(steps), int(graph_chunk))

    def skip(self, steps):
        """Advances the loop by `steps` iterations without running them (the state set on the
        device must be the state after them); returns the simulation time."""
        t = C.c_float()
        call("gpgmp_solver_skip", self.handle, int(steps), C.byref(t))
        return t.value

    def attachIndividualSpecies(self, species, individuals, rule=None):
        """Makes `species` particle-based: `individuals` (an IndividualSpecies on this grid with
        that species index) holds its particles; doReactions then maintains their population
        and diffuse hops them.  rule: None or ((mode0..3), (a0..3), (b0..3))."""
        r = None
        if rule is not None:
            r = _lib.NewIndividualRule()
            for k in range(4):
                r.mode[k], r.a[k], r.b[k] = int(rule[0][k]), float(rule[1][k]), float(rule[2][k])
        idx = species.index if isinstance(species, Species) else int(species)
        call("gpgmp_solver_attach_individuals", self.handle, idx, individuals.handle, C.byref(r) if r else None)
        self._individuals = individuals  # keeps it alive as long as the solver

    def setOutput(self, interval, max_time=0.0, max_dumps=64):
        """Output events every `interval` of simulation time (DiffusionModel::outputInterval)
        and the stop at max_time (DiffusionModel::maxTime); <= 0 disables either.  Runs then
        keep up to max_dumps device snapshots (t = 0 and each output write)."""
        call("gpgmp_solver_set_output", self.handle, C.c_float(interval), C.c_float(max_time), int(max_dumps))
        self._max_dumps = int(max_dumps)

    def getOutput(self):
        """(counts [n][S][H][W], times [n], writes including dropped ones, iterations run)."""
        ns, nt, it = C.c_int32(), C.c_int32(), C.c_int64()
        call("gpgmp_solver_get_output", self.handle, None, None, C.byref(ns), C.byref(nt), C.byref(it))
        counts = np.zeros((max(ns.value, 1),) + self.shape, np.int32)
        times = np.zeros(max(ns.value, 1), np.float32)
        call("gpgmp_solver_get_output", self.handle, _iptr(counts), _fptr(times), C.byref(ns), C.byref(nt),
             C.byref(it))
        return counts[: ns.value], times[: ns.value], nt.value, it.value

    def setPredictBusy(self, on=True):
        """Split steps: run last step's busy cells in worker waves of the SSA launch, or
        defer them to the X-sweep launch as they turn up (default); results are identical."""
        call("gpgmp_solver_set_predict_busy", self.handle, int(bool(on)))

    def setSweepWaveCells(self, cells=8):
        """Flattened X/Y sweep launches: 8 (default) or 16 cells per wave; results are
        identical."""
        call("gpgmp_solver_set_sweep_wave_cells", self.handle, int(cells))

    def setSweepSpread(self, on=True):
        """Flattened sweeps: spread each wave's 16 cells over the grid (default) or keep them
        consecutive; results are identical."""
        call("gpgmp_solver_set_sweep_spread", self.handle, int(bool(on)))

    def setLaunch(self, block, cells_per_thread):
        call("gpgmp_solver_set_launch", self.handle, int(block), int(cells_per_thread))

    def setLanesPerCell(self, ssa_lanes, sweep_lanes=None):
        """1 or 4 lanes per cell for the launches with an SSA and the sweep-only launches
        (results identical; speed differs)."""
        sweep_lanes = ssa_lanes if sweep_lanes is None else sweep_lanes
        call("gpgmp_solver_set_lanes_per_cell", self.handle, int(ssa_lanes), int(sweep_lanes))

    def getLaunch(self):
        """(threads per block, cells per thread, SSA lanes per cell, sweep lanes per cell)."""
        v = [C.c_int32() for _ in range(4)]
        call("gpgmp_solver_get_launch", self.handle, *[C.byref(x) for x in v])
        return tuple(x.value for x in v)

    def setEventBudget(self, budget, below_active=24):
        """A cell past `budget` SSA events moves to the one-wave-per-cell kernel once
        fewer than `below_active` lanes of its wave still run (65: always)."""
        call("gpgmp_solver_set_event_budget", self.handle, int(budget), int(below_active))

    def setHotWorkers(self, lane_events=16.0, lane_waves=1024, wave_waves=1024):
        """Deferred-cell workers of split steps: tails of at most lane_events expected events
        resume one per lane, busier cells one per wave (results do not depend on these)."""
        call("gpgmp_solver_set_hot_workers", self.handle, float(lane_events), int(lane_waves), int(wave_waves))

    def kernelTimes(self, on=True):
        """Accumulated HIP-event times (ms, launches timed, algorithmic bytes) per kernel
        of the current run, then sets the timing of the following launches: False/0 off,
        True/1 every launch, k > 1 the launches of 4 consecutive steps out of every k (after
        one more step whose events are not counted)."""
        ms = np.zeros(5, np.float64)
        n = np.zeros(5, np.int64)
        b = np.zeros(5, np.float64)
        call("gpgmp_solver_kernel_times", self.handle, int(on), ms.ctypes.data_as(C.POINTER(C.c_double)),
             n.ctypes.data_as(C.POINTER(C.c_int64)), b.ctypes.data_as(C.POINTER(C.c_double)))
        return ms, n, b

    def synchronize(self):
        call("gpgmp_solver_synchronize", self.handle)


class HomogeneousSolver(Solver):
    """HomogeneousSolver (HomogeneousSolver.h:15-81)."""


class InhomogeneousSolver(Solver):
    """InhomogeneousSolver (InhomogeneousSolver.h:15-162)."""


class DeterministicSolver(Solver):
    """The deterministic solver types (Solver(..., detState, ..., deterministic=true)): fp32
    concentrations [S][H][W]; doReactions runs RK4 or alpha-QSS, diffuse the reference's
    stencils; neither uses the seed (returned unchanged)."""

    def synchronizeStateBufferToDevice(self, state):
        a = np.ascontiguousarray(state, dtype=np.float32).reshape(self.shape)
        call("gpgmp_solver_set_state_f32", self.handle, _fptr(a))

    def synchronizeStateBufferFromDevice(self):
        a = np.empty(self.shape, np.float32)
        call("gpgmp_solver_get_state_f32", self.handle, _fptr(a))
        return a


class Ensemble:
    """Batched ensemble: n realizations of one homogeneous network on grids of their
    own size, each with its own seed0, advanced together (no reference counterpart)."""

    def __init__(self, model, widths, heights, seed0s, stream=None):
        self._cm = model.compile()
        self.model = model
        self.widths = np.ascontiguousarray(widths, dtype=np.int32)
        self.heights = np.ascontiguousarray(heights, dtype=np.int32)
        self.seed0s = np.ascontiguousarray(seed0s, dtype=np.uint32)
        n = len(self.widths)
        h = C.c_void_p()
        call("gpgmp_batch_create", self._cm.handle, n, _iptr(self.widths), _iptr(self.heights),
             _uptr(self.seed0s), _stream_ptr(stream), C.byref(h))
        self.handle = h

    def close(self):
        if getattr(self, "handle", None):
            lib().gpgmp_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __len__(self):
        return len(self.widths)

    def setState(self, r, state):
        S = self.model.numSpecies()
        a = np.ascontiguousarray(state, dtype=np.int32).reshape(S, self.heights[r], self.widths[r])
        call("gpgmp_batch_set_state", self.handle, int(r), _iptr(a))

    def getState(self, r):
        S = self.model.numSpecies()
        a = np.empty((S, self.heights[r], self.widths[r]), np.int32)
        call("gpgmp_batch_get_state", self.handle, int(r), _iptr(a))
        return a

    def setLaunch(self, block, cells_per_thread):
        call("gpgmp_batch_set_launch", self.handle, int(block), int(cells_per_thread))

    def setEventBudget(self, budget, below_active=24):
        call("gpgmp_batch_set_event_budget", self.handle, int(budget), int(below_active))

    def setLanesPerCell(self, ssa_lanes, sweep_lanes=None):
        sweep_lanes = ssa_lanes if sweep_lanes is None else sweep_lanes
        call("gpgmp_batch_set_lanes_per_cell", self.handle, int(ssa_lanes), int(sweep_lanes))

    def getLaunch(self):
        """(threads per block, cells per thread, SSA lanes per cell, sweep lanes per cell)."""
        v = [C.c_int32() for _ in range(4)]
        call("gpgmp_batch_get_launch", self.handle, *[C.byref(x) for x in v])
        return tuple(x.value for x in v)

    def autotune(self, trial_steps=20):
        b = C.c_int32()
        c = C.c_int32()
        call("gpgmp_batch_autotune", self.handle, int(trial_steps), C.byref(b), C.byref(c))
        return b.value, c.value

    def run(self, steps, graph_chunk=0):
        call("gpgmp_batch_run", self.handle, int(steps), int(graph_chunk))

    def prepareRun(self, steps, graph_chunk):
        """Captures the HIP graphs run(steps, graph_chunk) will replay (runs nothing)."""
        call("gpgmp_batch_prepare_run", self.handle, int(steps), int(graph_chunk))

    def reset(self):
        call("gpgmp_batch_reset", self.handle)

    def kernelTimes(self, on=True):
        ms = np.zeros(5, np.float64)
        n = np.zeros(5, np.int64)
        b = np.zeros(5, np.float64)
        call("gpgmp_batch_kernel_times", self.handle, int(on), ms.ctypes.data_as(C.POINTER(C.c_double)),
             n.ctypes.data_as(C.POINTER(C.c_int64)), b.ctypes.data_as(C.POINTER(C.c_double)))
        return ms, n, b

    def totalCells(self):
        n = C.c_int64()
        call("gpgmp_batch_total_cells", self.handle, C.byref(n))
        return n.value

    def synchronize(self):
        call("gpgmp_batch_synchronize", self.handle)
