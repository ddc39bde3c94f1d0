"""Deferred busy cells resumed one per lane or one per wave, HIP-graph chunks keyed by the
outflow-buffer parity, and graphs captured ahead of a run: counts, clock and snapshots
bit-exact against the oracle."""
import numpy as np
import pytest

from inchman_amd import configs

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dense256(oracle):
    """The first 30 host-loop iterations of the 256x256 Schnakenberg run (about 55 SSA
    events per cell in every other iteration), on the oracle."""
    m = configs.schnakenberg_model(256)
    st = configs.schnakenberg_state(256)
    want, t = oracle.run(m, st, 30, 5)
    return m, st, want, t


@pytest.mark.parametrize("workers", [(16.0, 1024, 1024), (0.0, 1024, 1024), (1e9, 1024, 1024),
                                     (1e9, 8, 8), (4.0, 64, 16)])
def test_lane_and_wave_workers_dense(gpu, dense256, workers):
    """Every deferred cell to the wave workers (lane_events 0), every one to the lane
    workers (1e9), the default split, and few workers (each wave takes several rounds of
    64 lane entries): the dense transient of the 256x256 Schnakenberg run is bit-exact."""
    m, st, want, tw = dense256
    s = m.createSolver()
    s.setHotWorkers(*workers)
    s.synchronizeStateBufferToDevice(st)
    t = s.run(30, 5)
    assert t == tw
    assert np.array_equal(s.synchronizeStateBufferFromDevice(), want)


@pytest.mark.parametrize("budget,below", [(1, 65), (4, 40), (8, 24)])
@pytest.mark.parametrize("lane_events", [0.0, 3.0, 1e9])
def test_lane_workers_budgets(gpu, oracle, budget, below, lane_events):
    """Deferral settings that send nearly every cell to the lists (budget 1, below 65)
    combined with lane lists of every size, on the 64x64 Schnakenberg run."""
    m = configs.schnakenberg_model(64)
    st = configs.schnakenberg_state(64, seed=31)
    s = m.createSolver()
    s.setEventBudget(budget, below)
    s.setHotWorkers(lane_events, 32, 32)
    s.synchronizeStateBufferToDevice(st)
    t = s.run(150, 2)
    want, tw = oracle.run(m, st, 150, 2)
    assert t == tw
    assert np.array_equal(s.synchronizeStateBufferFromDevice(), want)


@pytest.mark.parametrize("case", ["homogeneous", "inhomogeneous"])
@pytest.mark.parametrize("chunk", [16, 7])
def test_output_events_in_graph_chunks(gpu, oracle, case, chunk):
    """Output events end a chunk early, often after an odd number of steps, which flips
    the outflow-buffer parity of homogeneous steps: graphs are only replayed at the parity
    they were captured at, so every snapshot, the clock and the final counts stay the
    oracle's."""
    if case == "homogeneous":
        m, st, steps, interval = configs.plain_gmp_model(32), configs.plain_gmp_state(32), 400, 0.1
    else:
        m, st, steps, interval = configs.schnakenberg_model(64), configs.schnakenberg_state(64, seed=3), 200, 0.5
    s = m.createSolver()
    s.synchronizeStateBufferToDevice(st)
    s.setOutput(interval, 0.0, 128)
    s.run(0, 21)
    t = s.runContinue(steps // 2, graph_chunk=chunk)
    t = s.runContinue(steps - steps // 2, graph_chunk=chunk)
    counts, times, total, iters = s.getOutput()
    fin = s.synchronizeStateBufferFromDevice()
    want, wt, wit, dumps, wtimes, writes = oracle.run_output(m, st, steps, 21, interval, float("inf"),
                                                             max_dumps=128)
    assert t == wt and iters == wit and total == writes and total > 4
    assert np.array_equal(fin, want)
    assert np.array_equal(times, wtimes) and np.array_equal(counts, dumps)


@pytest.mark.parametrize("timing", [0, 128])
@pytest.mark.parametrize("network", ["plain", "schnakenberg"])
def test_prepared_run(gpu, network, timing):
    """Graphs captured ahead of a run (prepareRun runs nothing and leaves the loop where it
    was) replay a run shorter than the chunk, with or without a timing window: the same
    counts and clock as an eager run, and the prepared run times its window."""
    if network == "plain":
        m, st = configs.plain_gmp_model(64), configs.plain_gmp_state(64)
    else:
        m, st = configs.schnakenberg_model(64), configs.schnakenberg_state(64, seed=5)
    a = m.createSolver()
    a.synchronizeStateBufferToDevice(st)
    a.run(25, 5)
    t_ref = a.runContinue(0)
    ref = a.synchronizeStateBufferFromDevice()
    b = m.createSolver()
    b.synchronizeStateBufferToDevice(st)
    b.run(0, 5)
    if timing:
        b.kernelTimes(timing)
    b.prepareRun(25, 128)
    before = b.synchronizeStateBufferFromDevice()
    assert np.array_equal(before, st)  # nothing ran
    t = b.runContinue(25, graph_chunk=128)
    assert t == t_ref
    assert np.array_equal(b.synchronizeStateBufferFromDevice(), ref)
    if timing:
        ms, n, _ = b.kernelTimes(0)
        assert n.sum() > 0 and (ms[n > 0] > 0).all()
