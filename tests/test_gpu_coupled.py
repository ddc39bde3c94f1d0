"""A stochastic inhomogeneous model with an individual (particle-based) species coupled to
reactions, driven through the solver's per-call API as DiffusionModel::oclSolve drives the
reference (DiffusionModel.cpp:1479-1603; Solver::doReactions, Solver.cpp:207-248, with
maintainPopulation; InhomogeneousSolver::diffuse, InhomogeneousSolver.cpp:254-427, with the
individual hops and their time steps), against the same loop over the CPU oracle: counts
of every species, the particle list (keys, ids, properties), the pool and the clock,
bit-exact after every iteration."""
import numpy as np
import pytest

from oracle import oracle

pytestmark = pytest.mark.gpu

DEAD = 0xFFFFFFFF
RULE = ((2, 2, 2, 2), (0.1, 0.1, -0.05, -0.05), (0.4, 0.4, 0.1, 0.1))


def _inc(seed):
    """seed + 1 with int32 wrap-around, as the reference's int seed."""
    return (seed + 1 + 2 ** 31) % 2 ** 32 - 2 ** 31


def _model(W):
    from inchman_amd.gpgmp import STOCHASTIC_INHOMOGENEOUS, DiffusionModel

    m = DiffusionModel(W, W, physical_length=float(W), p0=0.15, solver=STOCHASTIC_INHOMOGENEOUS)
    a = m.addSpecies("A", 0.0)  # particle-based: its lattice fields are never used
    b = m.addSpecies("B", 0.0)
    m.addReaction(0.6, products={a: 1})
    m.addReaction(0.3, reactants={a: 1})
    m.addReaction(2.0, products={b: 1})
    m.addReaction(0.05, reactants={a: 1, b: 1}, products={a: 1})
    x = np.arange(W) - W / 2
    Dx = np.zeros((2, W, W), np.float32)
    Dx[1] = (0.4 * (1 + 0.5 * np.sin(2 * np.pi * x / W)))[None, :]
    rx = np.zeros((2, W, W), np.float32)
    rx[1] = 0.03
    m.setDriftDiffusivityFields(Dx, Dx.copy(), rx, np.zeros_like(rx))
    return m


def test_coupled_individual_species_loop(gpu):
    from inchman_amd.individuals import IndividualSpecies

    W, n0, cap, steps, seed0 = 32, 3000, 20000, 60, 99
    A = W * W
    rng = np.random.default_rng(5)
    cells = rng.integers(0, A, n0).astype(np.uint32)
    props = np.stack([rng.uniform(0.1, 0.5, n0), rng.uniform(0.1, 0.5, n0), rng.uniform(-0.05, 0.05, n0),
                      rng.uniform(-0.05, 0.05, n0)]).astype(np.float32)
    m = _model(W)
    st = np.zeros((2, W, W), np.int32)
    st[1] = rng.poisson(20, (W, W))
    # GPU: the solver with species 0 attached to its particles
    ind = IndividualSpecies(W, W, float(W), cap, species=0)
    ind.setParticles(cells, props)
    sol = m.createSolver()
    sol.synchronizeStateBufferToDevice(st)
    sol.attachIndividualSpecies(0, ind, RULE)
    # oracle: the reference's list layout and loop
    p = oracle.ind_params(W, W, float(W), species=0)
    pr = np.zeros((4, cap), np.float32)
    pr[:2, n0:] = 1.0
    pr[:, :n0] = props
    k, v, le, ri, (dxm, dym, rxm, rym, cnt), dtx0, dty0 = oracle.ind_update(p, cells, np.arange(n0, dtype=np.uint32),
                                                                             pr)
    keys = np.concatenate([k, np.full(cap - n0, DEAD, np.uint32)])
    vals = np.concatenate([v, np.arange(n0, cap, dtype=np.uint32)])
    n = n0
    ost = st.copy()
    ost[0] = cnt.reshape(W, W)
    table = oracle.table_from_model(m)
    Dx, Dy, rx, ry = m.fields
    h = oracle.header_float(m.physical_length / W)
    dtx = [dtx0, oracle.inhomogeneous_dt(Dx[1], rx[1], h)]
    dty = [dty0, oracle.inhomogeneous_dt(Dy[1], ry[1], h)]
    nx, ny = list(dtx), list(dty)
    np.testing.assert_array_equal(sol.synchronizeStateBufferFromDevice(), ost)
    seeds = oracle.seeds(seed0, steps)
    sim = np.float32(0)
    rule = oracle.new_rule(*RULE)
    for it in range(steps):
        seed = int(seeds[it])
        nxt = sol.getNextDiffusionTime()
        want_next = [min(nx[s], ny[s]) for s in range(2)]
        assert list(nxt) == [np.float32(x) for x in want_next]
        minspec = int(np.argmin(nxt))
        mindt = nxt[minspec]
        deltat = np.float32(mindt - sim)
        g_seed = sol.doReactions(deltat, seed)
        # oracle: SSA (seed++), then maintainPopulation(0, seed++)
        ost = oracle.ssa(ost, table, seed, deltat)
        seed = _inc(seed)
        stt, keys, vals, pr, n = oracle.ind_maintain(p, keys, vals, n, pr, ost[0].ravel(), seed, rule)
        assert stt == 0
        seed = _inc(seed)
        assert g_seed == seed
        for i in range(2):
            if not (i == minspec or nxt[i] <= mindt) or not (nxt[i] > 0):
                continue
            g_seed = sol.diffuse(i, mindt, g_seed)
            for axis, nn, dd in ((0, nx, dtx), (1, ny, dty)):
                if not (nn[i] <= mindt):
                    continue
                if i == 0:
                    lk, lv = keys[:n], vals[:n]
                    le, ri = oracle.cell_left_right(lk, A)
                    lk = oracle.ind_sweep(p, lk, lv, le, ri, pr, axis, seed, dd[i])
                    lk, lv, _, _, (_, _, _, _, cnt), ux, uy = oracle.ind_update(p, lk, lv, pr)
                    keys[:n], vals[:n] = lk, lv
                    ost[0] = cnt.reshape(W, W)
                    dd[i] = ux if axis == 0 else uy
                else:
                    ost = oracle.inhom_sweep(ost, i, axis, seed, dd[i], Dx if axis == 0 else Dy,
                                             rx if axis == 0 else ry, h)
                seed = _inc(seed)
                nn[i] = np.float32(nn[i] + dd[i])
            assert g_seed == seed
        sim = mindt
        np.testing.assert_array_equal(sol.synchronizeStateBufferFromDevice(), ost)
        gk, gv, gc = ind.getPositionList()
        np.testing.assert_array_equal(gk, keys[:n])
        np.testing.assert_array_equal(gv, vals[:n])
        np.testing.assert_array_equal(ind.getPropertiesList(), pr[:, vals[:n]])
    assert n != n0  # the population changed
