"""Particle-based (individual) species on the lattice against the CPU oracle: sorted
keys, particle indices, per-cell counts, constants and time steps bit-exact after
sweeps and whole host loops, reflective and absorbing boundaries."""
import numpy as np
import pytest

from oracle import oracle

pytestmark = pytest.mark.gpu


def _population(W, H, n, seed, uniform=False):
    rng = np.random.default_rng(seed)
    cells = rng.integers(0, W * H, n).astype(np.uint32)
    if uniform:
        props = np.stack([np.full(n, 0.4), np.full(n, 0.7), np.full(n, 0.05), np.full(n, -0.03)])
    else:
        props = np.stack([rng.uniform(0.05, 1.0, n), rng.uniform(0.05, 1.0, n), rng.uniform(-0.3, 0.3, n),
                          rng.uniform(-0.3, 0.3, n)])
    return cells, props.astype(np.float32)


def _species(W, H, n, bm=(0, 0, 0, 0), species=0, L=None):
    from inchman_amd.individuals import IndividualSpecies

    return IndividualSpecies(W, H, float(L if L else W), n, boundary_mask=bm, species=species)


@pytest.mark.parametrize("bm,species", [((0, 0, 0, 0), 0), ((1, 0, 1, 1), 1), ((1, 1, 1, 1), 2)])
def test_sweeps_and_updates(gpu, bm, species):
    W, H, n = 48, 40, 6000
    cells, props = _population(W, H, n, 5)
    g = _species(W, H, n, bm, species)
    g.setParticles(cells, props)
    p = oracle.ind_params(W, H, float(W), bm=bm, species=species)
    k, v, le, ri, (Dx, Dy, rx, ry, cnt), dtx, dty = oracle.ind_update(p, cells, np.arange(n, dtype=np.uint32), props)
    assert g.timeSteps() == (dtx, dty)
    gk, gv, gc = g.getPositionList()
    assert np.array_equal(gk, k) and np.array_equal(gv, v) and np.array_equal(gc.ravel(), cnt)
    seed = 1234567
    for axis in (0, 1, 0, 1, 1):
        dt = dtx if axis == 0 else dty
        seed_next = (g.diffuseX if axis == 0 else g.diffuseY)(seed, dt)
        assert seed_next == seed + 1
        gdx, gdy = g.computeDiffusionConstants(1 << axis)
        k = oracle.ind_sweep(p, k, v, le, ri, props, axis, seed, dt)
        k, v, le, ri, (Dx, Dy, rx, ry, cnt), ox, oy = oracle.ind_update(p, k, v, props)
        if axis == 0:
            dtx = ox
            assert gdx == ox
        else:
            dty = oy
            assert gdy == oy
        seed = seed_next
        gk, gv, gc = g.getPositionList()
        assert np.array_equal(gk, k) and np.array_equal(gv, v) and np.array_equal(gc.ravel(), cnt)
        for a, b in zip(g.constants(), (Dx, Dy, rx, ry)):
            assert np.array_equal(a, b)
    if any(bm):
        assert (k >= W * H).any() or cnt.sum() < n  # some particles left through absorbing edges


@pytest.mark.parametrize("W,H,n,steps,uniform", [(32, 32, 3000, 60, False), (64, 48, 20000, 25, True),
                                                 (256, 256, 200000, 8, False)])
def test_run_matches_oracle(gpu, W, H, n, steps, uniform):
    cells, props = _population(W, H, n, 9, uniform)
    g = _species(W, H, n)
    g.setParticles(cells, props)
    t = g.run(steps, 77)
    p = oracle.ind_params(W, H, float(W))
    k, v, cnt, tw = oracle.ind_run(p, cells, np.arange(n, dtype=np.uint32), props, steps, 77)
    assert t == tw
    gk, gv, gc = g.getPositionList()
    assert np.array_equal(gk, k) and np.array_equal(gv, v) and np.array_equal(gc.ravel(), cnt)


def test_empty_population(gpu):
    g = _species(16, 16, 8)
    g.setParticles(np.zeros(0, np.uint32), np.zeros((4, 0), np.float32))
    dx, dy = g.timeSteps()
    assert np.isinf(dx) and np.isinf(dy)
    assert g.run(5, 1) == 0.0
    k, v, c = g.getPositionList()
    assert k.size == 0 and c.sum() == 0
