"""Custom kinetic laws (Reaction::setCustomLaw) on the GPU against the oracle, whose hazards and
right-hand sides are the reference's generated code for the same formula text compiled by
g++ (oracle/lawgen.py): counts bit-exact through every SSA path (one lane per cell, lane
groups, busy cells deferred to the lane and wave workers, predicted busy cells, batches),
alpha-QSS bit-exact against the oracle's float instance, RK4 within 1e-5 of fp64."""
import numpy as np
import pytest

from inchman_amd import configs
from inchman_amd.gpgmp import Compartment, DETERMINISTIC_HOMOGENEOUS_AQSS, DETERMINISTIC_HOMOGENEOUS_RK4

pytestmark = pytest.mark.gpu


def _plain_custom(side=32, height=None, localized=True):
    """The plain GMP network with editor-style laws: float parameters, an int-arithmetic
    quotient, a Michaelis-Menten term and a cell-dependent law."""
    m = configs.plain_gmp_model(side, height or side)
    for k, v in {"kA": 1.0, "kB": 1.0, "kAB": 0.01, "kC": 0.1, "Vm": 2.5, "Km": 30.0}.items():
        m.setParameter(k, v)
    a, b, c = m.species
    m.reactions[0].setCustomLaw("kA + 0.01 * GridX")                 # 0 -> A, program (cell-dependent)
    m.reactions[1].setCustomLaw("kB")                                # 0 -> B, table (float constant)
    m.reactions[2].setCustomLaw("kAB * SpeciesA * SpeciesB".replace("SpeciesA", a.id).replace("SpeciesB", b.id))
    m.reactions[3].setCustomLaw("Vm * %s / (Km + %s)" % (c.id, c.id))  # C -> 0, program
    m.addReaction(1.0, reactants={a: 2}, products={c: 1}, id="AA").setCustomLaw("%s * (%s - 1) / 2 / 50" % (a.id, a.id))
    if localized:
        m.reactions[3].addCompartment(Compartment("west", 0, 0, side // 2, 1000))
        m.reactions[2].addCompartment(Compartment("core", 4, 4, 20, 20))
    return m


@pytest.mark.parametrize("lanes", [(1, 1), (4, 4), (1, 0)])
@pytest.mark.parametrize("budget", [None, (2, 65)])
def test_custom_laws_ssa_and_run(gpu, oracle, lanes, budget):
    m = _plain_custom(32)
    st = configs.plain_gmp_state(32, seed=4)
    st[2] = 30
    s = m.createSolver()
    s.setLanesPerCell(*lanes)
    if budget:
        s.setEventBudget(*budget)
    s.synchronizeStateBufferToDevice(st)
    s.doReactions(0.6, 123)
    want = oracle.ssa(st, oracle.table_from_model(m), 123, 0.6)
    assert np.array_equal(s.synchronizeStateBufferFromDevice(), want)
    s2 = m.createSolver()
    s2.setLanesPerCell(*lanes)
    if budget:
        s2.setEventBudget(*budget)
    s2.synchronizeStateBufferToDevice(st)
    t = s2.run(150, 8)
    want2, tw = oracle.run(m, st, 150, 8)
    assert t == tw
    assert np.array_equal(s2.synchronizeStateBufferFromDevice(), want2)


@pytest.mark.parametrize("predict,budget", [(False, None), (False, (1, 65)), (True, (3, 40))])
def test_custom_laws_split_steps(gpu, oracle, predict, budget):
    """The Schnakenberg network with its laws as formulas (the autocatalysis as a quotient
    program), through the split inhomogeneous step."""
    m = configs.schnakenberg_model(64)
    u, v = m.species
    m.setParameter("k3", 1e-4)
    m.reactions[2].setCustomLaw("k3 * %s * (%s - 1) * %s / 2 * 2" % (u.id, u.id, v.id))
    m.reactions[0].setCustomLaw("1.0 + 0 * GridY")
    st = configs.schnakenberg_state(64, seed=9)
    s = m.createSolver()
    s.setPredictBusy(predict)
    if budget:
        s.setEventBudget(*budget)
    s.synchronizeStateBufferToDevice(st)
    s.run(40, 2)
    t = s.runContinue(40, 16)
    want, tw = oracle.run(m, st, 80, 2)
    assert t == tw
    assert np.array_equal(s.synchronizeStateBufferFromDevice(), want)


def test_custom_laws_ensemble_matches_plain(gpu):
    from inchman_amd import Ensemble

    sides = np.array([32, 48, 16], np.int32)
    seeds = np.array([3, 5, 7], np.uint32)
    m = _plain_custom(32)
    states = [configs.plain_gmp_state(int(sd), seed=int(r) + 1) for r, sd in enumerate(sides)]
    ens = Ensemble(m, sides, sides, seeds)
    for r, st in enumerate(states):
        ens.setState(r, st)
    ens.run(50, 16)
    for r, sd in enumerate(sides):
        p = _plain_custom(32)
        p.width = p.height = int(sd)
        p.physical_length = float(sd)
        s = p.createSolver()
        s.synchronizeStateBufferToDevice(states[r])
        s.run(50, int(seeds[r]))
        assert np.array_equal(ens.getState(r), s.synchronizeStateBufferFromDevice()), r


def test_custom_laws_aqss_bit_exact(gpu, oracle):
    m = _plain_custom(48, 32)
    m.solver = DETERMINISTIC_HOMOGENEOUS_AQSS
    st = configs.plain_gmp_state(48, 32).astype(np.float32)
    st[2] = 5.0
    s = m.createSolver()
    s.synchronizeStateBufferToDevice(st)
    s.doReactions(np.float32(0.05), 0)
    want, _ = oracle.det_aqss(st, oracle.table_from_model(m), np.float32(0.05))
    assert np.array_equal(s.synchronizeStateBufferFromDevice(), want)


def test_custom_laws_rk4_run(gpu, oracle):
    m = _plain_custom(48, 32)
    m.solver = DETERMINISTIC_HOMOGENEOUS_RK4
    st = configs.plain_gmp_state(48, 32).astype(np.float32)
    st[2] = 5.0
    want, t_want = oracle.det_run(m, st, 40)
    s = m.createSolver()
    s.synchronizeStateBufferToDevice(st)
    t = s.run(40, 1)
    got = s.synchronizeStateBufferFromDevice()
    assert t == t_want
    rel = float(np.max(np.abs(got - want) / np.maximum(np.abs(want), 1e-6)))
    assert rel <= 1e-5, rel
