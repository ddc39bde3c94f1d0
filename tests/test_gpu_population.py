"""Population maintenance of an individual species (IndividualSolver::maintainPopulation)
on the GPU against the CPU oracle, bit-exact: live keys, particle ids and properties in
sorted order, the dead-particle pool (ids and properties, front first), per-cell counts;
interleaved with hop sweeps, with and without a newIndividuals rule, reflective and
absorbing boundaries, goals on the host and on the device, and the error paths."""
import numpy as np
import pytest

from oracle import oracle

pytestmark = pytest.mark.gpu

DEAD = 0xFFFFFFFF
RULE = ((2, 1, 2, 0), (0.25, 1.5, -0.1, 0.0), (1.0, 0.0, 0.2, 0.0))


def _oracle_rule(rule):
    return None if rule is None else oracle.new_rule(*rule)


class RefList:
    """The reference's layout: keys/vals [cap] (live sorted, dead after), props [4][cap] by id."""

    def __init__(self, p, cells, props_live, cap):
        n = cells.size
        self.p, self.cap, self.n = p, cap, n
        self.props = np.zeros((4, cap), np.float32)
        self.props[:2, n:] = 1.0  # pool defaults (IndividualSolver.cpp:84-98)
        self.props[:, :n] = props_live
        k, v, *_ = oracle.ind_update(p, cells, np.arange(n, dtype=np.uint32), self.props)
        self.keys = np.concatenate([k, np.full(cap - n, DEAD, np.uint32)])
        self.vals = np.concatenate([v, np.arange(n, cap, dtype=np.uint32)])

    def maintain(self, goal, seed, rule):
        st, k, v, pr, n_new = oracle.ind_maintain(self.p, self.keys, self.vals, self.n, self.props, goal, seed,
                                                  _oracle_rule(rule))
        assert st == 0
        self.keys, self.vals, self.props, self.n = k, v, pr, n_new

    def sweep_update(self, axis, seed, dt):
        A = self.p.W * self.p.H
        k, v = self.keys[: self.n], self.vals[: self.n]
        le, ri = oracle.cell_left_right(k, A)
        k = oracle.ind_sweep(self.p, k, v, le, ri, self.props, axis, seed, dt)
        k, v, _, _, _, dtx, dty = oracle.ind_update(self.p, k, v, self.props)
        self.keys[: self.n], self.vals[: self.n] = k, v
        return dtx if axis == 0 else dty


def _check(g, ref, W, H):
    A = W * H
    gk, gv, gc = g.getPositionList()
    n = ref.n
    assert gk.size == n
    np.testing.assert_array_equal(gk, ref.keys[:n])
    np.testing.assert_array_equal(gv, ref.vals[:n])
    np.testing.assert_array_equal(g.getPropertiesList(), ref.props[:, ref.vals[:n]])
    pid, pprops = g.pool()
    np.testing.assert_array_equal(pid, ref.vals[n:])
    assert np.all(ref.keys[n:] == DEAD)
    np.testing.assert_array_equal(pprops, ref.props[:, ref.vals[n:]])
    live = ref.keys[:n]
    np.testing.assert_array_equal(gc.ravel(), np.bincount(live[live < A], minlength=A)[:A])


def _goal(rng, counts, trailing=True):
    A = counts.size
    goal = np.maximum(counts + rng.integers(-3, 4, A), 0)
    goal[rng.random(A) < 0.1] = 0
    goal[rng.random(A) < 0.02] += 25
    if trailing:
        goal[-3:] += 5
    return goal.astype(np.int32)


def _species(W, H, cap, bm=(0, 0, 0, 0)):
    from inchman_amd.individuals import IndividualSpecies

    return IndividualSpecies(W, H, float(W), cap, boundary_mask=bm)


@pytest.mark.parametrize("W,H,n,extra,bm", [(7, 5, 60, 400, (0, 0, 0, 0)), (48, 40, 6000, 30000, (0, 0, 0, 0)),
                                            (48, 40, 6000, 30000, (1, 1, 1, 1)),
                                            (256, 256, 150000, 400000, (0, 0, 0, 0))])
def test_maintain_matches_oracle(gpu, W, H, n, extra, bm):
    rng = np.random.default_rng(W * 1000 + n)
    A = W * H
    cap = n + extra
    cells = rng.integers(0, A, n).astype(np.uint32)
    props = np.stack([rng.uniform(0.05, 1.0, n), rng.uniform(0.05, 1.0, n), rng.uniform(-0.3, 0.3, n),
                      rng.uniform(-0.3, 0.3, n)]).astype(np.float32)
    g = _species(W, H, cap, bm)
    g.setParticles(cells, props)
    p = oracle.ind_params(W, H, float(W), bm=bm)
    ref = RefList(p, cells, props, cap)
    _check(g, ref, W, H)
    dtx, dty = g.timeSteps()
    seed = 987654
    for it in range(6):
        counts = g.getPositionList()[2].ravel().astype(np.int64)
        goal = _goal(rng, counts)
        rule = RULE if it % 2 else None
        g.maintainPopulation(goal, seed, rule)
        ref.maintain(goal, seed, rule)
        _check(g, ref, W, H)
        seed += 2
        axis = it % 2
        (g.diffuseX if axis == 0 else g.diffuseY)(seed, dtx if axis == 0 else dty)
        gx, gy = g.computeDiffusionConstants(1 << axis)
        want = ref.sweep_update(axis, seed, dtx if axis == 0 else dty)
        assert (gx if axis == 0 else gy) == want
        if axis == 0:
            dtx = gx
        else:
            dty = gy
        seed += 1
        _check(g, ref, W, H)


def test_goal_on_device_and_errors(gpu):
    import torch

    from inchman_amd._lib import E_CAPACITY, E_INVALID, GpgmpError

    W, H, n, cap = 32, 24, 3000, 5000
    A = W * H
    rng = np.random.default_rng(5)
    cells = rng.integers(0, A, n).astype(np.uint32)
    props = rng.uniform(0.1, 1.0, (4, n)).astype(np.float32)
    g, h = _species(W, H, cap), _species(W, H, cap)
    g.setParticles(cells, props)
    h.setParticles(cells, props)
    counts = g.getPositionList()[2].ravel()
    goal = _goal(rng, counts.astype(np.int64))
    g.maintainPopulation(goal, 11, RULE)
    h.maintainPopulation(torch.from_numpy(goal).cuda(), 11, RULE)
    for a, b in zip(g.getPositionList(), h.getPositionList()):
        np.testing.assert_array_equal(a, b)
    np.testing.assert_array_equal(g.getPropertiesList(), h.getPropertiesList())
    before = (g.getPositionList(), g.getPropertiesList(), g.pool())
    # more new particles than the pool holds: refused, nothing changes
    big = goal + 10
    with pytest.raises(GpgmpError) as e:
        g.maintainPopulation(big, 12)
    assert e.value.code == E_CAPACITY
    bad = goal.copy()
    bad[7] = -1
    with pytest.raises(GpgmpError) as e:
        g.maintainPopulation(bad, 12)
    assert e.value.code == E_INVALID
    after = (g.getPositionList(), g.getPropertiesList(), g.pool())
    for a, b in zip(before[0], after[0]):
        np.testing.assert_array_equal(a, b)
    np.testing.assert_array_equal(before[1], after[1])
    np.testing.assert_array_equal(before[2][0], after[2][0])
    # a valid call afterwards still works (the error flag was reset)
    goal2 = np.zeros(A, np.int32)
    goal2[::5] = 2
    g.maintainPopulation(goal2, 13)
    np.testing.assert_array_equal(g.getPositionList()[2].ravel(), goal2)
    # everything dies, then the grid is refilled from the pool
    g.maintainPopulation(np.zeros(A, np.int32), 14)
    assert g.getPositionList()[0].size == 0 and g.pool()[0].size == cap
    g.maintainPopulation(np.full(A, 3, np.int32), 15)
    np.testing.assert_array_equal(g.getPositionList()[2].ravel(), np.full(A, 3))
