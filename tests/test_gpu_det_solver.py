"""The deterministic solver types behind the solver interface (Solver(..., detState, ...)):
the reference's host loop through doReactions / diffuse / getNextDiffusionTime on fp32
concentrations, and the same loop on the device (run), against the FP64 oracle of that loop
within 1e-5 relative error."""
import numpy as np
import pytest

from inchman_amd import configs
from inchman_amd.gpgmp import (DETERMINISTIC_HOMOGENEOUS_AQSS, DETERMINISTIC_HOMOGENEOUS_RK4,
                               DETERMINISTIC_INHOMOGENEOUS_AQSS, DETERMINISTIC_INHOMOGENEOUS_RK4,
                               SOLVER_REFERENCE_QUIRKS)

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-6)))


def _model(kind):
    if kind in (DETERMINISTIC_HOMOGENEOUS_RK4, DETERMINISTIC_HOMOGENEOUS_AQSS):
        m = configs.plain_gmp_model(48, 32)
        st = configs.plain_gmp_state(48, 32).astype(np.float32)
        st[2] = 5.0
    else:
        m = configs.schnakenberg_model(64, 48)
        st = configs.schnakenberg_state(64, 48, seed=5).astype(np.float32)
    m.solver = kind
    return m, st


def _reference_loop(solver, steps, S):
    """DiffusionModel::oclSolve's loop (DiffusionModel.cpp:1479-1603) through the solver calls."""
    dt = solver.getNextDiffusionTime()
    sim = np.float32(0.0)
    seed = 1234
    for n in range(steps):
        mindt, minspec = dt[0], 0
        for i in range(1, S):
            if dt[i] < mindt:
                mindt, minspec = dt[i], i
        seed2 = solver.doReactions(np.float32(mindt - sim), seed)
        assert seed2 == seed  # the deterministic types leave the seed alone
        for i in range(S):
            if (i == minspec or dt[i] <= mindt) and dt[i] > 0:
                assert solver.diffuse(i, mindt, seed) == seed
        dt = solver.getNextDiffusionTime()
        sim = mindt
    return float(sim)


KINDS = [DETERMINISTIC_HOMOGENEOUS_RK4, DETERMINISTIC_HOMOGENEOUS_AQSS, DETERMINISTIC_INHOMOGENEOUS_RK4,
         DETERMINISTIC_INHOMOGENEOUS_AQSS]


@pytest.mark.parametrize("flags", [0, SOLVER_REFERENCE_QUIRKS])
@pytest.mark.parametrize("kind", KINDS)
def test_deterministic_types_through_the_solver_calls(gpu, oracle, kind, flags):
    m, st = _model(kind)
    steps = 60
    want, t_want = oracle.det_run(m, st, steps, flags)
    s = m.createSolver(flags=flags)
    s.synchronizeStateBufferToDevice(st)
    t = _reference_loop(s, steps, m.numSpecies())
    got = s.synchronizeStateBufferFromDevice()
    assert got.dtype == np.float32
    assert t == t_want
    assert _rel(got, want) <= 1e-5
    assert _rel(got, st.astype(np.float64)) > 1e-3  # it moved
    # the same loop on the device
    s2 = m.createSolver(flags=flags)
    s2.synchronizeStateBufferToDevice(st)
    t2 = s2.run(steps // 2, 99)
    t2 = s2.runContinue(steps - steps // 2)
    assert t2 == t_want
    assert _rel(s2.synchronizeStateBufferFromDevice(), want) <= 1e-5


def test_deterministic_time_steps_and_state_types(gpu):
    """Homogeneous types share min_s 1e-2 h^2/(4 D_s) (HomogeneousSolver.cpp:97-121); the
    int32 state calls are refused on a deterministic solver, the fp32 ones on a stochastic
    one; stochastic-only calls answer with a status."""
    from inchman_amd import GpgmpError

    m, st = _model(DETERMINISTIC_HOMOGENEOUS_RK4)
    s = m.createSolver()
    dx, dy = s.timeSteps()
    h = np.float32(m.physical_length) / np.float32(m.width)
    want = min(np.float32(np.float32(np.float32(1e-2) * h) * h) / np.float32(4.0 * sp.diffusionConstant)
               for sp in m.species)
    assert (dx == want).all() and (dy == want).all()
    with pytest.raises(GpgmpError):
        s.synchronizeStateBufferToDevice(st.astype(np.int32).astype(np.float32))  # fine: fp32
        from inchman_amd._lib import call
        import ctypes as C
        a = np.zeros(st.shape, np.int32)
        call("gpgmp_solver_set_state", s.handle, a.ctypes.data_as(C.POINTER(C.c_int32)))
    with pytest.raises(GpgmpError):
        s.setLaunch(128, 1)
    sto = configs.plain_gmp_model(32).createSolver()
    from inchman_amd._lib import call
    import ctypes as C
    f = np.zeros((3, 32, 32), np.float32)
    with pytest.raises(GpgmpError):
        call("gpgmp_solver_set_state_f32", sto.handle, f.ctypes.data_as(C.POINTER(C.c_float)))
