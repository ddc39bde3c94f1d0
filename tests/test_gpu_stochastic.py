"""Parity of the HIP stochastic kernels with the CPU oracle: bit-exact counts on the
same inputs and seeds (SSA, homogeneous and inhomogeneous diffusion, the host loop,
batched ensemble vs plain solver)."""
import numpy as np
import pytest

from inchman_amd import configs

pytestmark = pytest.mark.gpu


def _solver(model):
    return model.createSolver()


@pytest.mark.parametrize("seed,dt", [(7, 0.025), (123456789, 0.5), (-5, 3.0)])
def test_ssa_plain_gmp(gpu, oracle, seed, dt):
    m = configs.plain_gmp_model(32)
    st = configs.plain_gmp_state(32, seed=abs(seed) + 1)
    s = _solver(m)
    s.synchronizeStateBufferToDevice(st)
    new_seed = s.doReactions(dt, seed)
    assert new_seed == seed + 1
    got = s.synchronizeStateBufferFromDevice()
    want = oracle.ssa(st, oracle.table_from_model(m), seed, dt)
    assert np.array_equal(got, want)
    assert not np.array_equal(got, st)


def test_ssa_schnakenberg(gpu, oracle):
    m = configs.schnakenberg_model(64)
    st = configs.schnakenberg_state(64, seed=3)
    s = _solver(m)
    s.synchronizeStateBufferToDevice(st)
    s.doReactions(0.11, 99)
    got = s.synchronizeStateBufferFromDevice()
    want = oracle.ssa(st, oracle.table_from_model(m), 99, 0.11)
    assert np.array_equal(got, want)


@pytest.mark.parametrize("bm,sm", [((0, 0, 0, 0), (0, 0, 0, 0)), ((1, 0, 1, 0), (0, 0, 0, 0)),
                                   ((0, 0, 0, 0), (0, 5, 3, 2))])
def test_hom_diffuse(gpu, oracle, bm, sm):
    m = configs.plain_gmp_model(48, 32)
    m.boundary_mask = bm
    m.source_mask = sm
    st = configs.plain_gmp_state(48, 32, seed=11)
    st[2] = np.random.default_rng(5).poisson(3, st[2].shape)
    s = _solver(m)
    s.synchronizeStateBufferToDevice(st)
    seed = 4242
    want = st.copy()
    stay = oracle.stay_probability(m.p0)
    for sp in range(3):
        seed2 = s.diffuse(sp, 1.0, seed)
        assert seed2 == seed  # homogeneous diffusion keeps the seed
        want = oracle.hom_diffuse(want, sp, seed, stay, bm, sm)
    got = s.synchronizeStateBufferFromDevice()
    assert np.array_equal(got, want)
    if bm == (0, 0, 0, 0) and sm == (0, 0, 0, 0):
        assert np.array_equal(got.sum(axis=(1, 2)), st.sum(axis=(1, 2)))


def test_inhom_sweeps(gpu, oracle):
    m = configs.schnakenberg_model(64, 48)
    st = configs.schnakenberg_state(64, 48, seed=8)
    s = _solver(m)
    s.synchronizeStateBufferToDevice(st)
    dtx, dty = s.timeSteps()
    Dx, Dy, rx, ry = m.fields
    h = oracle.header_float(m.physical_length / m.width)
    hy = oracle.header_float(m.physical_length / m.height)
    for sp in range(2):
        assert dtx[sp] == oracle.inhomogeneous_dt(Dx[sp], rx[sp], h)
        assert dty[sp] == oracle.inhomogeneous_dt(Dy[sp], ry[sp], h)
    seed = 17
    want = st.copy()
    t = float(max(dtx.max(), dty.max()))
    for sp in range(2):
        want = oracle.inhom_sweep(want, sp, 0, seed, dtx[sp], Dx, rx, h)
        want = oracle.inhom_sweep(want, sp, 1, seed + 1, dty[sp], Dy, ry, hy)
        seed2 = s.diffuse(sp, t, seed)
        assert seed2 == seed + 2
        seed = seed2
    got = s.synchronizeStateBufferFromDevice()
    assert np.array_equal(got, want)
    assert np.array_equal(got.sum(axis=(1, 2)), st.sum(axis=(1, 2)))


def test_run_plain_gmp(gpu, oracle):
    m = configs.plain_gmp_model(32)
    st = configs.plain_gmp_state(32)
    s = _solver(m)
    s.synchronizeStateBufferToDevice(st)
    t_gpu = s.run(1000, 1)
    got = s.synchronizeStateBufferFromDevice()
    want, t_cpu = oracle.run(m, st, 1000, 1)
    assert t_gpu == t_cpu
    assert np.array_equal(got, want)


@pytest.mark.parametrize("side,steps", [(64, 300), (256, 40)])
def test_run_schnakenberg(gpu, oracle, side, steps):
    m = configs.schnakenberg_model(side)
    st = configs.schnakenberg_state(side)
    s = _solver(m)
    s.synchronizeStateBufferToDevice(st)
    t_gpu = s.run(steps, 5)
    got = s.synchronizeStateBufferFromDevice()
    want, t_cpu = oracle.run(m, st, steps, 5)
    assert t_gpu == t_cpu
    assert np.array_equal(got, want)


def test_run_continue_graph(gpu):
    m = configs.schnakenberg_model(64)
    st = configs.schnakenberg_state(64)
    a = _solver(m)
    a.synchronizeStateBufferToDevice(st)
    a.run(100, 9)
    ref = a.synchronizeStateBufferFromDevice()
    b = _solver(m)
    b.synchronizeStateBufferToDevice(st)
    b.run(30, 9)
    b.runContinue(50, graph_chunk=16)
    b.runContinue(20, graph_chunk=0)
    assert np.array_equal(b.synchronizeStateBufferFromDevice(), ref)


def test_graph_chunks_with_timing_windows(gpu):
    """Timing windows (4 counted steps of every 32) inside graph chunks of 32: the windows
    run eagerly at each chunk's start and the rest replays a graph; counts are the eager
    run's and every window is timed."""
    m = configs.schnakenberg_model(64)
    st = configs.schnakenberg_state(64)
    a = _solver(m)
    a.synchronizeStateBufferToDevice(st)
    a.run(200, 4)
    ref = a.synchronizeStateBufferFromDevice()
    b = _solver(m)
    b.synchronizeStateBufferToDevice(st)
    b.run(0, 4)
    b.kernelTimes(32)
    b.runContinue(200, graph_chunk=32)
    ms, n, _ = b.kernelTimes(0)
    assert np.array_equal(b.synchronizeStateBufferFromDevice(), ref)
    # 200 steps = 6 chunks of 32 + 8 eager steps; each window of 5 counts 4 steps
    assert n[1] == n[2] == n[4] == 7 * 4
    assert (ms[[1, 2, 4]] > 0).all()


def test_ensemble_matches_plain(gpu):
    m = configs.plain_gmp_model(32)
    sides = np.array([16, 32, 48, 64, 80, 128, 96], np.int32)
    seeds = configs.ensemble_seeds(len(sides))
    states = [configs.ensemble_state(sd, r) for r, sd in enumerate(sides)]
    from inchman_amd import Ensemble

    ens = Ensemble(m, sides, sides, seeds)
    for r, st in enumerate(states):
        ens.setState(r, st)
    ens.setLaunch(128, 2)
    ens.setLanesPerCell(4, 1)  # the plain solvers below run the default (1 lane, flattened sweeps)
    ens.run(40, graph_chunk=0)
    ens.run(60, graph_chunk=20)
    for r, sd in enumerate(sides):
        pm = configs.plain_gmp_model(int(sd))
        p = pm.createSolver()
        assert p.getLaunch()[2:] == (1, 0)
        p.synchronizeStateBufferToDevice(states[r])
        p.run(100, int(seeds[r]))
        assert np.array_equal(ens.getState(r), p.synchronizeStateBufferFromDevice()), r


@pytest.mark.parametrize("network", ["plain", "schnakenberg"])
def test_launch_shape_invariance(gpu, network):
    """Threads per block, cells per thread and lanes per cell (1: one lane runs a cell;
    4: a lane group draws 4 SSA events at once and splits the particle loops; sweep
    lanes 0: the particle loops of a wave's cells flattened over all its lanes) change
    the speed only, never a count."""
    if network == "plain":
        m, st, steps = configs.plain_gmp_model(64), configs.plain_gmp_state(64), 50
    else:
        m, st, steps = configs.schnakenberg_model(64), configs.schnakenberg_state(64, seed=3), 80
    outs = []
    for block, cpt, lanes in [(256, 1, (1, 1)), (64, 4, (4, 4)), (128, 3, (1, 4)), (192, 2, (4, 1)),
                              (256, 1, (4, 4)), (256, 1, (1, 0)), (128, 2, (4, 0)), (64, 3, (1, 0))]:
        s = m.createSolver()
        s.setLaunch(block, cpt)
        s.setLanesPerCell(*lanes)
        assert s.getLaunch() == (block, cpt) + lanes
        s.synchronizeStateBufferToDevice(st)
        s.run(steps, 3)
        outs.append(s.synchronizeStateBufferFromDevice())
    assert all(np.array_equal(outs[0], o) for o in outs[1:])


def test_min_timestep_device(gpu, oracle):
    """computeIndividualTimestepX/Y + reduceBlocks on the device (double, multi-level
    min) against the host evaluation the solvers use."""
    import ctypes as C
    import torch

    from inchman_amd._lib import call

    for side in (64, 256, 1024):
        m = configs.schnakenberg_model(side)
        Dx, Dy, rx, ry = m.fields
        h = oracle.header_float(m.physical_length / m.width)
        for s in range(2):
            for D, r in ((Dx[s], rx[s]), (Dy[s], ry[s])):
                tD = torch.from_numpy(np.ascontiguousarray(D)).cuda()
                tr = torch.from_numpy(np.ascontiguousarray(r)).cuda()
                out = C.c_float()
                call("gpgmp_min_timestep_device", tD.data_ptr(), tr.data_ptr(), D.size, h, C.byref(out), None)
                want = np.float32(oracle.inhomogeneous_dt(D, r, h))
                assert abs(np.float32(out.value) - want) <= np.spacing(want), (side, s)


@pytest.mark.parametrize("lanes", [(1, 1), (4, 4), (1, 0)])
@pytest.mark.parametrize("budget,below", [(1, 65), (5, 65), (1, 8), (64, 65), (1 << 30, 8)])
def test_event_budget_hot_cells(gpu, oracle, budget, below, lanes):
    """Cells that exceed the SSA event budget continue in the one-wave-per-cell kernel
    (speculative 64-event batches, lane-split particle loops): counts stay bit-exact for
    any budget and deferral threshold, including budget 1 with below_active 65 (every
    cell with an event is deferred), with 1 or 4 lanes per cell in the main kernel."""
    m = configs.schnakenberg_model(64)
    st = configs.schnakenberg_state(64, seed=21)
    s = m.createSolver()
    s.setLanesPerCell(*lanes)
    s.setEventBudget(budget, below)
    s.synchronizeStateBufferToDevice(st)
    t = s.run(120, 4)
    want, tw = oracle.run(m, st, 120, 4)
    assert t == tw
    assert np.array_equal(s.synchronizeStateBufferFromDevice(), want)
    # homogeneous network with sources through the same path
    mh = configs.plain_gmp_model(48)
    mh.source_mask = (0, 3, 0, 2)
    sh = mh.createSolver()
    sh.setLanesPerCell(*lanes)
    sh.setEventBudget(budget, below)
    st2 = configs.plain_gmp_state(48)
    sh.synchronizeStateBufferToDevice(st2)
    sh.run(60, 8)
    want2, _ = oracle.run(mh, st2, 60, 8)
    assert np.array_equal(sh.synchronizeStateBufferFromDevice(), want2)


def test_moments_x(gpu):
    """inhomogeneous_computeMomentsX + reduceMomentsBlocks: sum of x * n and of n for one
    species, exact, against numpy on odd and even grids."""
    import torch

    from inchman_amd._lib import call

    rng = np.random.default_rng(17)
    for S, H, W, sp in ((2, 256, 256, 1), (3, 37, 50, 0), (1, 1, 1, 0), (2, 300, 1024, 1)):
        st = rng.integers(0, 5000, (S, H, W)).astype(np.int32)
        x = torch.from_numpy(st).cuda()
        sums = torch.zeros(2, dtype=torch.int64, device="cuda")
        call("gpgmp_moments_x_device", x.data_ptr(), W, H, sp, sums.data_ptr(), None)
        torch.cuda.synchronize()
        n = st[sp].astype(np.int64)
        want = [int((n * np.arange(W)[None, :]).sum()), int(n.sum())]
        assert sums.cpu().tolist() == want


@pytest.mark.parametrize("case", ["homogeneous", "inhomogeneous", "max_time", "overflow"])
def test_output_events(gpu, oracle, case):
    """The host loop with the reference's output events (a reaction step up to each output
    time, then a zero-length iteration that writes) and its max_time stop, on the device,
    against the oracle: final counts, clock, iterations, every snapshot and its time; a
    run split by runContinue gives the same snapshots."""
    inf = float("inf")
    if case == "inhomogeneous":
        m, st = configs.schnakenberg_model(64), configs.schnakenberg_state(64, seed=3)
        steps, interval, max_time, md = 200, 0.5, 0.0, 64
    else:
        m, st = configs.plain_gmp_model(32), configs.plain_gmp_state(32)
        steps, interval, max_time, md = {"homogeneous": (300, 0.1, 0.0, 64), "max_time": (2000, 0.3, 3.0, 64),
                                         "overflow": (300, 0.05, 0.0, 4)}[case]
    s = m.createSolver()
    s.synchronizeStateBufferToDevice(st)
    s.setOutput(interval, max_time, md)
    if case == "homogeneous":
        s.run(steps // 2, 21)
        t = s.runContinue(steps - steps // 2)
    else:
        t = s.run(steps, 21)
    counts, times, total, iters = s.getOutput()
    fin = s.synchronizeStateBufferFromDevice()
    want, wt, wit, dumps, wtimes, writes = oracle.run_output(m, st, steps, 21, interval,
                                                             max_time if max_time > 0 else inf, max_dumps=md)
    assert t == wt and iters == wit and total == writes
    assert np.array_equal(fin, want)
    assert np.array_equal(times, wtimes) and np.array_equal(counts, dumps)
    if case == "max_time":
        assert iters < steps and t >= max_time
    if case == "overflow":
        assert total > md and counts.shape[0] == md


@pytest.mark.parametrize("lanes", [(1, 0), (4, 4), (1, 1)])
@pytest.mark.parametrize("kind", ["mixed", "pure_drift"])
def test_extreme_hop_probabilities(gpu, oracle, kind, lanes):
    """Fields whose hop probabilities leave [0, 1): cells with strong negative drift under a
    time step set by fast-diffusing cells (p1 < 0: no draw goes +), pure-drift cells
    (p2 = 1: every particle hops), zero fields; the whole host loop bit-exact against the
    oracle (the device compares draws with exact integer thresholds and a none-mask)."""
    from inchman_amd.gpgmp import STOCHASTIC_INHOMOGENEOUS, DiffusionModel

    W = 32
    rng = np.random.default_rng(8)
    m = DiffusionModel(W, W, physical_length=float(W), p0=0.15, solver=STOCHASTIC_INHOMOGENEOUS)
    a = m.addSpecies("A", 0.0)
    m.addReaction(5.0, products={a: 1})
    m.addReaction(0.1, reactants={a: 1})
    D = np.zeros((1, W, W), np.float32)
    r = np.zeros((1, W, W), np.float32)
    if kind == "mixed":
        D[0, :, : W // 2] = 3.33
        D[0, :, W // 2:] = 0.001
        r[0, :, W // 2:] = rng.choice([-0.5, 0.5, -2.0], (W, W // 2))
        D[0, ::5, ::3] = 0.0
    else:
        r[0] = rng.choice([-0.5, 0.5], (W, W))
    m.setDriftDiffusivityFields(D, D.copy(), r, -r)
    st = rng.poisson(30, (1, W, W)).astype(np.int32)
    s = m.createSolver()
    s.setLanesPerCell(*lanes)
    s.synchronizeStateBufferToDevice(st)
    t = s.run(40, 17)
    got = s.synchronizeStateBufferFromDevice()
    want, tw = oracle.run(m, st, 40, 17)
    assert t == tw
    assert np.array_equal(got, want)
    assert not np.array_equal(want, st)


@pytest.mark.parametrize("spread", [False, True])
@pytest.mark.parametrize("lanes", [(1, 0), (4, 0)])
def test_sweep_spread(gpu, oracle, spread, lanes):
    """Flattened sweeps with each wave's 16 cells consecutive or spread over the grid
    (stride W*H/16), on a non-square grid, the inhomogeneous and the homogeneous solver:
    counts and clock bit-exact against the oracle either way."""
    m = configs.schnakenberg_model(96, 48)
    st = configs.schnakenberg_state(96, 48, seed=4)
    s = m.createSolver()
    s.setLanesPerCell(*lanes)
    s.setSweepSpread(spread)
    s.synchronizeStateBufferToDevice(st)
    t = s.run(150, 6)
    want, tw = oracle.run(m, st, 150, 6)
    assert t == tw
    assert np.array_equal(s.synchronizeStateBufferFromDevice(), want)
    mh = configs.plain_gmp_model(64, 32)
    sh = mh.createSolver()
    sh.setLanesPerCell(*lanes)
    sh.setSweepSpread(spread)
    st2 = configs.plain_gmp_state(64, 32)
    sh.synchronizeStateBufferToDevice(st2)
    sh.run(80, 3)
    want2, _ = oracle.run(mh, st2, 80, 3)
    assert np.array_equal(sh.synchronizeStateBufferFromDevice(), want2)


@pytest.mark.parametrize("predict", [True, False])
@pytest.mark.parametrize("budget,below", [(1, 65), (8, 24), (3, 40)])
def test_predict_busy_cells(gpu, oracle, predict, budget, below):
    """Split steps with the previous step's busy cells run whole in worker waves of the SSA
    launch (hot lists alternating by step parity) or deferred as they turn up: counts and
    clock bit-exact against the oracle, for eager launches and for HIP-graph chunks
    continued across runs (the parity lives in the uploaded plans)."""
    m = configs.schnakenberg_model(64)
    st = configs.schnakenberg_state(64, seed=23)
    s = m.createSolver()
    s.setPredictBusy(predict)
    s.setEventBudget(budget, below)
    s.synchronizeStateBufferToDevice(st)
    s.run(61, 9)
    s.runContinue(40, 10)
    t = s.runContinue(19)
    want, tw = oracle.run(m, st, 120, 9)
    assert t == tw
    assert np.array_equal(s.synchronizeStateBufferFromDevice(), want)


def _law_model(kind, side=32):
    """Reaction tables on either side of the compact form (at most 4 reactions of at most
    3 factors): five reactions, a fourth-order law, int and double rate literals mixed,
    and a single reaction (three unused compact slots)."""
    m = configs.DiffusionModel(side, side, physical_length=side, p0=0.15, solver=configs.STOCHASTIC_HOMOGENEOUS)
    a, b, c = m.addSpecies("A", 1.0), m.addSpecies("B", 0.5), m.addSpecies("C", 0.1)
    if kind == "five":
        m.addReaction(1.0, products={a: 1})
        m.addReaction(1.0, products={b: 1})
        m.addReaction(0.01, reactants={a: 1, b: 1}, products={c: 1})
        m.addReaction(0.1, reactants={c: 1})
        m.addReaction(2e-4, reactants={a: 2, c: 1}, products={b: 1})
    elif kind == "order4":
        m.addReaction(20, products={a: 1})
        m.addReaction(1e-6, reactants={a: 4}, products={b: 1})
        m.addReaction(0.5, reactants={b: 1}, products={c: 2})
    elif kind == "mixed":
        m.addReaction(5, products={a: 1})
        m.addReaction(2, reactants={a: 1})
        m.addReaction(1e-3, reactants={a: 2, b: 1}, products={c: 1})
        m.addReaction(3, reactants={c: 1}, products={b: 1})
    else:
        m.addReaction(0.3, reactants={a: 1}, products={b: 1})
    return m


@pytest.mark.parametrize("kind", ["five", "order4", "mixed", "single"])
@pytest.mark.parametrize("budget", [None, (2, 65)])
def test_ssa_law_forms(gpu, oracle, kind, budget):
    """The general table walk (more than 4 reactions, or a law of more than 3 factors) and
    the compact straight-line form give the oracle's counts, in the one-lane SSA and in
    the one-wave-per-cell continuation of busy cells (budget 2: nearly every cell)."""
    m = _law_model(kind)
    st = configs.plain_gmp_state(32, seed=11)
    s = _solver(m)
    if budget:
        s.setEventBudget(*budget)
    s.synchronizeStateBufferToDevice(st)
    s.doReactions(0.2, 41)
    got = s.synchronizeStateBufferFromDevice()
    want = oracle.ssa(st, oracle.table_from_model(m), 41, 0.2)
    assert np.array_equal(got, want)
    assert not np.array_equal(got, st)


def test_run_schnakenberg_five_reactions(gpu, oracle):
    """The split inhomogeneous step (SSA launch, X sweeps with busy cells, Y sweeps) with a
    table outside the compact form."""
    m = configs.schnakenberg_model(64)
    m.addReaction(0.05, reactants={m.species[1]: 1}, id="V_out")
    st = configs.schnakenberg_state(64, seed=4)
    s = _solver(m)
    s.synchronizeStateBufferToDevice(st)
    t_gpu = s.run(60, 9)
    got = s.synchronizeStateBufferFromDevice()
    want, t_cpu = oracle.run(m, st, 60, 9)
    assert t_gpu == t_cpu
    assert np.array_equal(got, want)


@pytest.mark.parametrize("W,H", [(48, 32), (64, 48)])
@pytest.mark.parametrize("cells", [16, 8])
def test_sweep_wave_cells(gpu, oracle, W, H, cells):
    """Flattened X/Y sweeps over 16 or 8 cells per wave (spread over the grid); the host
    loop matches the oracle either way."""
    m = configs.schnakenberg_model(W, H)
    st = configs.schnakenberg_state(W, H, seed=6)
    s = _solver(m)
    s.setSweepWaveCells(cells)
    s.synchronizeStateBufferToDevice(st)
    t_gpu = s.run(120, 13)
    got = s.synchronizeStateBufferFromDevice()
    want, t_cpu = oracle.run(m, st, 120, 13)
    assert t_gpu == t_cpu
    assert np.array_equal(got, want)
