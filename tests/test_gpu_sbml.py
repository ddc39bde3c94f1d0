"""Models read by the front end (InchmanDiffusionModel::load restated in inchman_amd.sbml) run
on the GPU: the drift/diffusivity method evaluated by the device kernel equals the method text
compiled by g++ (oracle/lawgen.py) bit for bit, and whole runs of the models equal the oracle."""
import os

import numpy as np
import pytest

from inchman_amd import GpgmpError, sbml

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_homogeneous_model_file_run(gpu, oracle):
    im = sbml.load(os.path.join(GOLDEN, "inchman_homogeneous.xml"))
    st, real, fields = sbml.build(im)
    m = im.model
    s = m.createSolver()
    s.synchronizeStateBufferToDevice(st)
    t = s.run(200, 17)
    want, tw = oracle.run(m, st, 200, 17)
    assert t == tw
    got = s.synchronizeStateBufferFromDevice()
    assert np.array_equal(got, want)
    assert got[2].sum() > st[2].sum()  # the catalysis made product


def test_drift_method_fields_and_run(gpu, oracle):
    from oracle import lawgen

    im = sbml.load(os.path.join(GOLDEN, "inchman_drift.xml"))
    st, real, fields = sbml.initial_state(im)
    m = im.model
    cur = np.zeros((4, 1, 64, 64), np.float32)
    cur[0] = cur[1] = 1.0
    got = np.stack(sbml.drift_diffusivity_fields(im, st, fields, cur))
    want = lawgen.method_fields(m, im.drift_diffusivity_method, im.field_parameters,
                                fields["gradient"][: 64 * 64].reshape(1, 64, 64), st, cur)
    assert got.tobytes() == want.tobytes()
    st2, _, _ = sbml.build(im)
    assert np.array_equal(st2, st)
    s = m.createSolver()
    s.synchronizeStateBufferToDevice(st)
    t = s.run(60, 3)
    want_st, tw = oracle.run(m, st, 60, 3)
    assert t == tw
    assert np.array_equal(s.synchronizeStateBufferFromDevice(), want_st)


def test_nonlinear_model_with_fixed_fields_runs(gpu, oracle):
    """A non-linear model whose method reads only positions, parameters and field parameters
    (as the reference's OrnsteinUhlenbeck and MultiplicativeNoise examples): its fields after
    any sweep are the first ones, so it builds and runs as the oracle's loop over them; one that
    reads the counts gets the re-evaluating solver, which the device loop refuses."""
    im = sbml.load(os.path.join(GOLDEN, "inchman_drift.xml"))
    im.nonlinear_diffusivity = True
    im.drift_diffusivity_method = ("All->DiffusivityX = 0.8 + 0.2 * gradient;\nAll->DiffusivityY = 0.6;\n"
                                   "All->DriftX = -(0.01 * PhysicalX);\nAll->DriftY = -(0.005 * PhysicalY);")
    st, real, fields = sbml.build(im)
    m = im.model
    again = np.stack(sbml.drift_diffusivity_fields(im, st * 3 + 1, fields, np.ones((4, 1, 64, 64), np.float32),
                                                   sim_time=5.0))
    first = np.stack(sbml.drift_diffusivity_fields(im, st, fields))
    assert again.tobytes() == first.tobytes()  # other counts, clock and current fields: same fields
    s = m.createSolver()
    s.synchronizeStateBufferToDevice(st)
    t = s.run(40, 9)
    want, tw = oracle.run(m, st, 40, 9)
    assert t == tw
    assert np.array_equal(s.synchronizeStateBufferFromDevice(), want)
    im.drift_diffusivity_method = "All->DriftX = 0.001 * All->State;"
    st, real, fields = sbml.build(im)  # reads the counts: re-evaluated after every sweep
    s2 = im.model.createSolver()
    with pytest.raises(GpgmpError):
        s2.run(5, 9)  # the device loop does not re-evaluate it: the per-call loop does


def _inc(seed):
    """seed + 1 with int32 wrap-around, as the reference's int seed."""
    return (seed + 1 + 2 ** 31) % 2 ** 32 - 2 ** 31


@pytest.mark.parametrize("fpe", [False, True])
def test_nonlinear_recompute_matches_oracle_loop(gpu, oracle, fpe):
    """Non-linear diffusivity whose method reads the counts (the drift fixture's: a central
    difference of the counts over a crowding term): DiffusionModel::oclSolve's loop through the
    solver calls, with the method re-evaluated after every sweep on the swept counts at the
    current time (InhomogeneousSolver.cpp:304-336,384-417), against the same loop over the
    oracle with the g++-compiled method: counts, clock, next diffusion times and time steps
    bit-exact after every iteration; then runPerCall over the same steps.  The Fokker-Planck
    type's sweeps use the rates its swept axis' time-step kernel last stored (computed from the
    fields of that moment) and its own time-step formula."""
    from inchman_amd.gpgmp import STOCHASTIC_INHOMOGENEOUS_FPE
    from oracle import lawgen

    im = sbml.load(os.path.join(GOLDEN, "inchman_drift.xml"))
    im.nonlinear_diffusivity = True
    if fpe:
        im.model.solver = STOCHASTIC_INHOMOGENEOUS_FPE
    st, real, fields = sbml.build(im)
    m = im.model
    S, H, W = m.numSpecies(), m.height, m.width
    names = list(im.field_parameters)
    fp = np.stack([np.asarray(fields[n], np.float32).reshape(-1)[: H * W].reshape(H, W) for n in names])
    sol = m.createSolver()
    sol.synchronizeStateBufferToDevice(st)
    # oracle: the build's fields, the time steps as set up, the reference's loop
    Dx, Dy, rx, ry = [np.array(a, np.float32).reshape(S, H, W) for a in m.fields]
    h = oracle.header_float(m.physical_length / W)
    stay = oracle.stay_probability(m.p0)

    def step(D, r):
        if fpe:
            return np.float32(min(oracle.fpe_cell_dt(d, v, h, stay) for d, v in zip(D.ravel(), r.ravel())))
        return np.float32(oracle.inhomogeneous_dt(D, r, h))

    dtx = [step(Dx[i], rx[i]) for i in range(S)]
    dty = [step(Dy[i], ry[i]) for i in range(S)]
    rates = [[Dx.copy(), rx.copy()], [Dy.copy(), ry.copy()]]  # the fields the FPE rates were stored from
    gx, gy = sol.timeSteps()
    assert list(gx) == dtx and list(gy) == dty
    nx, ny = list(dtx), list(dty)
    table = oracle.table_from_model(m)
    steps, seed0 = 30, 11
    seeds = oracle.seeds(seed0, steps)
    ost = st.copy()
    sim = np.float32(0)
    changed = False
    for it in range(steps):
        seed = int(seeds[it])
        nxt = sol.getNextDiffusionTime()
        assert list(nxt) == [np.float32(min(nx[i], ny[i])) for i in range(S)], it
        minspec = int(np.argmin(nxt))
        mindt = nxt[minspec]
        deltat = np.float32(mindt - sim)
        g_seed = sol.doReactions(deltat, seed)
        ost = oracle.ssa(ost, table, seed, deltat)
        seed = _inc(seed)
        assert g_seed == seed
        for i in range(S):
            if not ((i == minspec or nxt[i] <= mindt) and nxt[i] > 0):
                continue
            g_seed = sol.diffuse(i, mindt, g_seed)
            for axis, nn, dd in ((0, nx, dtx), (1, ny, dty)):
                if not (nn[i] <= mindt):
                    continue
                if fpe:
                    ost = oracle.fpe_sweep(ost, i, axis, seed, dd[i], rates[axis][0], rates[axis][1], h)
                else:
                    ost = oracle.inhom_sweep(ost, i, axis, seed, dd[i], Dx if axis == 0 else Dy,
                                             rx if axis == 0 else ry, h)
                seed = _inc(seed)
                new = lawgen.method_fields(m, im.drift_diffusivity_method, names, fp, ost,
                                           np.stack([Dx, Dy, rx, ry]), sim_time=mindt)
                before = rx[i].copy()
                Dx[i], Dy[i], rx[i], ry[i] = new[0][i], new[1][i], new[2][i], new[3][i]
                changed |= not np.array_equal(before, rx[i])  # the drift follows the counts
                dd[i] = step(Dx[i] if axis == 0 else Dy[i], rx[i] if axis == 0 else ry[i])
                rates[axis][0][i] = (Dx if axis == 0 else Dy)[i]
                rates[axis][1][i] = (rx if axis == 0 else ry)[i]
                nn[i] = np.float32(nn[i] + dd[i])
            assert g_seed == seed
        sim = mindt
        np.testing.assert_array_equal(sol.synchronizeStateBufferFromDevice(), ost, err_msg="iteration %d" % it)
        gx, gy = sol.timeSteps()
        assert list(gx) == dtx and list(gy) == dty, it
    assert changed
    # the same steps by runPerCall on a fresh solver
    sol2 = m.createSolver()
    sol2.synchronizeStateBufferToDevice(st)
    t, n = sol2.runPerCall(steps, seed0)
    assert t == sim and n == steps
    np.testing.assert_array_equal(sol2.synchronizeStateBufferFromDevice(), ost)


@pytest.mark.parametrize("case", ["inhomogeneous", "max_time"])
def test_run_per_call_output_events_match_oracle(gpu, oracle, case):
    """runPerCall's output events (a reaction step up to each output time, then a zero-length
    iteration that writes) and its max_time stop, on models the device loop also runs, against
    the oracle's loop with output events: final counts, clock, iterations, every dump and its
    time."""
    from inchman_amd import configs

    if case == "inhomogeneous":
        m, st = configs.schnakenberg_model(64), configs.schnakenberg_state(64, seed=3)
        steps, interval, max_time = 120, 0.5, float("inf")
    else:
        m, st = configs.plain_gmp_model(32), configs.plain_gmp_state(32)
        steps, interval, max_time = 2000, 0.3, 3.0
    s = m.createSolver()
    s.synchronizeStateBufferToDevice(st)
    dumps, times = [], []
    t, n = s.runPerCall(steps, 21, output_interval=interval, max_time=max_time,
                        on_output=lambda tm, c: (times.append(np.float32(tm)), dumps.append(c)))
    want, wt, wit, wdumps, wtimes, writes = oracle.run_output(m, st, steps, 21, interval, max_time, max_dumps=64)
    assert np.float32(t) == wt and n == wit and len(dumps) == writes
    np.testing.assert_array_equal(s.synchronizeStateBufferFromDevice(), want)
    k = min(len(dumps), 64)
    assert np.array_equal(np.array(times[:k], np.float32), wtimes)
    assert np.array_equal(np.stack(dumps[:k]), wdumps)
    if case == "max_time":
        assert n < steps  # the clock reached max_time first
