"""Localized reactions in the CPU oracle (Reaction::addCompartment, Reaction.cpp:209-235;
the mask in the generated hazards, Solver.cpp:853-854, and ODE right-hand sides,
Solver.cpp:947-948,987-1000), pinned by what the generated C fixes:

* a mask of 1 in front of a double-literal law leaves the double chain unchanged, so a
  reaction localized to the whole grid gives the unlocalized counts bit for bit;
* a mask of 0 removes the law: the reaction never fires outside its compartments, and its
  births inside follow Poisson(k dt) per cell;
* a mask of 1 in front of an int-literal law turns the int product into a float chain
  (w * 1 * A * B * C evaluated left to right), which rounds differently from the int
  product for A = B = 4097, C = 3 (50356224 vs 50356228);
* compartment bounds are inclusive (Compartment::isInCompartment, Compartment.h:152-158).
"""
import numpy as np

from inchman_amd import configs
from inchman_amd import gpgmp
from inchman_amd.gpgmp import Compartment


def _birth_model(side=32, rate=0.8):
    m = configs.DiffusionModel(side, side, physical_length=side, p0=0.15, solver=configs.STOCHASTIC_HOMOGENEOUS)
    a = m.addSpecies("A", 1.0)
    b = m.addSpecies("B", 1.0)
    m.addReaction(rate, products={a: 1})
    m.addReaction(0.5, products={b: 1})
    return m


def test_whole_grid_mask_is_identity_for_double_literals(oracle):
    m = configs.plain_gmp_model(32)
    st = configs.plain_gmp_state(32, seed=3)
    plain = oracle.ssa(st, oracle.table_from_model(m), 17, 0.4)
    # localize the double-literal reactions (0.01 and 0.1) to compartments covering the grid
    m.reactions[2].addCompartment(Compartment("all", 0, 0, 31, 31))
    m.reactions[3].addCompartment(Compartment("left", 0, 0, 15, 31))
    m.reactions[3].addCompartment(Compartment("right", 16, 0, 40, 40))  # clipped to the grid
    t = oracle.table_from_model(m)
    assert t.localized == 0b1100
    assert np.array_equal(oracle.ssa(st, t, 17, 0.4), plain)


def test_zero_mask_removes_the_reaction(oracle):
    m = _birth_model(32, rate=0.8)
    box = (4, 6, 19, 13)  # x0, y0, x1, y1 inclusive: 16 x 8 cells
    m.reactions[0].addCompartment(Compartment("box", *box))
    st = np.zeros((2, 32, 32), np.int32)
    dt = 2.0
    got = oracle.ssa(st, oracle.table_from_model(m), 5, dt)
    inside = np.zeros((32, 32), bool)
    inside[box[1]:box[3] + 1, box[0]:box[2] + 1] = True
    assert got[0][~inside].sum() == 0
    # births inside: Poisson(k dt) per cell (the B births leave A's hazard alone)
    n = got[0][inside]
    assert n.size == 128
    lam = 0.8 * dt
    assert abs(n.mean() - lam) < 4 * np.sqrt(lam / n.size)
    # the unlocalized reaction still fires everywhere
    assert (got[1] > 0).mean() > 0.5


def test_int_literal_law_becomes_a_float_chain(oracle):
    m = configs.DiffusionModel(16, 16, physical_length=16, p0=0.15, solver=configs.STOCHASTIC_HOMOGENEOUS)
    a, b, c = m.addSpecies("A", 1.0), m.addSpecies("B", 1.0), m.addSpecies("C", 1.0)
    m.addReaction(1, reactants={a: 1, b: 1, c: 1})   # law "1 * A * B * C", int literal
    m.addReaction(0.25, reactants={c: 1})
    x = np.array([4097, 4097, 3], np.int32)
    h_plain = oracle.hazards(oracle.table_from_model(m), x)
    assert h_plain[1] == np.float32(50356228.0)   # (float)(int product 50356227)
    m.reactions[0].addCompartment(Compartment("c", 0, 0, 7, 7))
    t = oracle.table_from_model(m)
    h_in = oracle.hazards(t, x, cell=3 + 5 * 16)   # (3, 5): inside
    assert h_in[1] == np.float32(50356224.0)       # ((1.f * 4097) * 4097) * 3 in float
    assert h_in[2] == np.float32(h_in[1] + np.float32(0.25 * 3))
    h_out = oracle.hazards(t, x, cell=8 + 5 * 16)  # (8, 5): outside
    assert h_out[1] == 0.0
    assert h_out[2] == np.float32(0.75)


def test_bounds_are_inclusive(oracle):
    m = _birth_model(16, rate=50.0)
    m.reactions[0].addCompartment(Compartment("one", 3, 9, 3, 9))
    st = np.zeros((2, 16, 16), np.int32)
    got = oracle.ssa(st, oracle.table_from_model(m), 1, 1.0)
    nz = np.argwhere(got[0] > 0)
    assert nz.tolist() == [[9, 3]]
    act = m.reactionActiveBits()
    assert act[9, 3] == 0b11 and act[9, 4] == 0b10 and act[0, 0] == 0b10


def test_localized_deterministic_rates(oracle):
    """RK4 (fp64 oracle): a localized birth adds k dt inside and nothing outside."""
    m = configs.DiffusionModel(16, 16, physical_length=16, p0=0.15,
                               solver=gpgmp.DETERMINISTIC_HOMOGENEOUS_RK4)
    a = m.addSpecies("A", 1.0)
    m.addReaction(2.0, products={a: 1})
    m.reactions[0].addCompartment(Compartment("box", 2, 2, 5, 9))
    st = np.zeros((1, 16, 16), np.float64)
    got = oracle.det_rk4(st, oracle.table_from_model(m), 0.3)
    want = np.zeros((16, 16))
    want[2:10, 2:6] = 0.6
    assert np.allclose(got[0], want, rtol=1e-12, atol=1e-15)
