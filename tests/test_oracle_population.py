"""The oracle's population maintenance of an individual species (oracle_ind_maintain,
IndividualSolver::maintainPopulation) pinned three ways:

* against a literal restatement of the reference's multi-kernel procedure
  (IndividualSolver.cpp:458-671): the maintainPopulation kernel's kill loop, counts and
  segment flags (IndividualSolver.cl:1017-1131), the exclusive segmented scan (checked the
  way python/gpgmp/helpers.py:26-94 checks it), computeReductionOffsets (:1299-1334), the
  inclusive scans of scanIndividuals (:1344-1499), scatterIndividuals' index arithmetic
  (:1501-1581), updateNewParticles' indexNew / indexPool arithmetic (:1135-1293),
  writeLostParticles and copyTempToKeys (:1584-1599) -- a different formulation from the
  oracle's "list it leaves behind", on random small lists;
* by the reference's own consistency check (IndividualSolver.cpp:1320-1387): ids unique,
  keys sorted, per-cell count == goal;
* by statistics: kill picks uniform over a cell's particles, RANDOM uniform on [0, 1).
"""
import math

import numpy as np
import pytest

from oracle import oracle

DEAD = 0xFFFFFFFF
M32 = 0xFFFFFFFF


def _u(w):
    """(int)w * M_RAN_INVM32 + 0.5 in double (Solver.cpp:471, IndividualSolver.cl:1066)."""
    return float(np.int32(np.uint32(w).view(np.int32))) * 2.0 ** -32 + 0.5


class _Stream:
    """x then y words of Threefry calls with counter 1, 2, ... (the kernels' ri / c.v[0])."""

    def __init__(self, cell, seed):
        self.k0, self.k1, self.ri, self.ctr, self.w = cell, (0xDECAFBAD + seed) & M32, 0, 0, (0, 0)

    def next(self):
        if self.ri % 2 == 0:
            self.ctr += 1
            self.w = oracle.threefry(self.ctr, 0xF00DCAFE, self.k0, self.k1)
            r = self.w[0]
        else:
            r = self.w[1]
        self.ri += 1
        return r


def reference_procedure(keys, vals, n, A, props, goal, seed, rule):
    """The reference's kernels in order, as plain loops.  Returns (keys, vals, props,
    n_new, bug) where bug marks the case the reference gets wrong (the last listed particle
    killed while trailing empty cells need new particles)."""
    cap = len(keys)
    keys = [int(k) for k in keys]
    vals = [int(v) for v in vals]
    props = props.copy()
    temp = list(zip(keys, vals))  # d_tempKeys: copy of the list before the kernels
    left, right = oracle.cell_left_right(np.array(keys[:n], np.uint32), A)
    red = [0] * cap  # d_reductionOffsets
    newp, dro, flags = [0] * A, [0] * A, [0] * A
    # maintainPopulation kernel, one work-item per cell
    for c in range(A):
        goal_state, is_state = int(goal[c]), 0
        i = int(left[c])
        while i <= int(right[c]):
            if keys[i] == c:
                red[i] = 0
                is_state += 1
            i += 1
        if is_state > goal_state:
            rs = _Stream(c, seed)
            for _ in range(is_state - goal_state):
                while True:
                    rand = int(_u(rs.next()) * is_state)
                    kp = int(left[c]) + rand
                    if keys[kp] != DEAD:
                        break
                keys[kp] = DEAD  # KILL_PARTICLE: key UINT_MAX, id kept
                red[kp] = 1
        if c < A - 1:
            flags[c + 1] = 1 if is_state > 0 else 0
        if is_state < goal_state:
            newp[c] = dro[c] = goal_state - is_state
    # segmented scan of d_domainReductionOffsets, exclusive, a flag starts a segment
    run = 0
    for c in range(A):
        if flags[c]:
            run = 0
        dro[c], run = run, run + dro[c]
    # computeReductionOffsets
    adding = [0] * (cap + 1)
    for c in range(A):
        if (c < A - 1 and flags[c + 1] == 1) or c == A - 1:
            adding[int(left[c])] = newp[c] + dro[c]
    # scanIndividuals + scanIndividualsBlocks: inclusive scans over the particles
    red_incl, add_incl = [0] * n, [0] * n
    sr = sa = 0
    for i in range(n):
        sr += red[i]
        sa += adding[i]
        red_incl[i], add_incl[i] = sr, sa
    # scatterIndividuals
    lost = [None] * n
    for i in range(n):
        rl, ra = red_incl[i], add_incl[i]
        if keys[i] < DEAD:
            temp[i - rl + ra] = (keys[i], vals[i])
        else:
            lost[rl - 1] = (keys[i], vals[i])
    total = n - red_incl[n - 1]
    total_added = add_incl[n - 1]
    trailing = sum(newp[c] for c in range(A) if int(left[c]) >= n)
    bug = keys[n - 1] == DEAD and trailing > 0
    if keys[n - 1] < A - 1:
        total_added += dro[A - 1] + newp[A - 1]
    eol_new = total + total_added
    assert eol_new <= cap
    # updateNewParticles, one work-item per cell
    for c in range(A):
        rs = _Stream(c, seed + 1)
        for i in range(newp[c]):
            index_old = int(left[c])
            if index_old < n:
                index_new = index_old + add_incl[index_old]
                index_new -= red_incl[index_old]
                if keys[index_old] == DEAD:
                    index_new += 1
                nnp = add_incl[index_old] - (add_incl[index_old - 1] if index_old > 0 else 0)
                index_new -= nnp
                index_new += dro[c] + i
                index_pool = n + add_incl[index_old] - nnp + dro[c] + i
            else:
                index_new = index_old + add_incl[index_old - 1] - red_incl[index_old - 1] + dro[c] + i
                index_pool = n + add_incl[index_old - 1] + dro[c] + i
            pid = vals[index_pool]
            temp[index_new] = (c, pid)
            if rule is not None:
                for q in range(4):
                    if rule.mode[q] == 1:
                        props[q, pid] = np.float32(rule.a[q])
                    elif rule.mode[q] == 2:
                        props[q, pid] = np.float32(rule.a[q] + rule.b[q] * _u(rs.next()))
    # writeLostParticles, copyTempToKeys
    for t in range(n - total):
        temp[eol_new + t] = lost[t]
    return (np.array([k for k, _ in temp], np.uint32), np.array([v for _, v in temp], np.uint32), props,
            eol_new, bug)


def _random_case(rng, A, n, pool):
    cap = n + pool
    keys = np.full(cap, DEAD, np.uint32)
    # cluster particles into a few cells so cells with many particles (kills) occur
    keys[:n] = np.sort(rng.integers(0, A, n) if rng.random() < 0.5 else
                       rng.choice(rng.integers(0, A, max(1, A // 3)), n))
    vals = rng.permutation(cap).astype(np.uint32)
    props = rng.uniform(0.1, 2.0, (4, cap)).astype(np.float32)
    counts = np.bincount(keys[:n], minlength=A)
    goal = np.maximum(counts + rng.integers(-4, 5, A), 0)
    goal[rng.random(A) < 0.2] = 0
    return keys, vals, props, goal.astype(np.int32)


def test_against_reference_procedure():
    rng = np.random.default_rng(7)
    checked = 0
    for t in range(300):
        W, H = int(rng.integers(1, 7)), int(rng.integers(1, 7))
        A = W * H
        n = int(rng.integers(1, 40))
        keys, vals, props, goal = _random_case(rng, A, n, pool=int(rng.integers(60, 120)))
        rule = None if t % 3 == 0 else oracle.new_rule((2, 1, 2, 0), (0.25, 1.5, -0.1, 0), (1.0, 0, 0.2, 0))
        p = oracle.ind_params(W, H, float(W))
        seed = int(rng.integers(-2 ** 31, 2 ** 31 - 2))
        st, k, v, pr, n_new = oracle.ind_maintain(p, keys, vals, n, props, goal, seed, rule)
        assert st == 0
        rk, rv, rp, rn, bug = reference_procedure(keys, vals, n, A, props, goal, seed, rule)
        if bug:
            continue
        checked += 1
        assert n_new == rn == int(goal.sum())
        np.testing.assert_array_equal(k, rk)
        np.testing.assert_array_equal(v, rv)
        np.testing.assert_array_equal(pr, rp)
    assert checked > 250


def _consistent(k, v, n_new, goal, A):
    """IndividualSolver::checkPositionListConsistency (IndividualSolver.cpp:1320-1387)."""
    assert np.unique(v).size == v.size  # no duplicate ids
    assert np.all(k[1:] >= k[:-1])  # keys in order
    live = k[k != DEAD]
    assert live.size == n_new
    np.testing.assert_array_equal(np.bincount(live, minlength=A), goal)


def test_consistency_and_edges():
    rng = np.random.default_rng(3)
    W, H = 8, 5
    A = W * H
    p = oracle.ind_params(W, H, float(W))
    for t in range(60):
        n = int(rng.integers(0, 120))
        keys, vals, props, goal = _random_case(rng, A, n, pool=400)
        if t % 10 == 0:
            goal[:] = 0  # every particle dies
        if t % 10 == 1:
            goal[-5:] += 7  # trailing cells gain particles
        st, k, v, pr, n_new = oracle.ind_maintain(p, keys, vals, n, props, goal, t, None)
        assert st == 0
        _consistent(k, v, n_new, goal, A)
        # survivors keep their order and properties (props are indexed by id)
        np.testing.assert_array_equal(pr, props)
    # the pool is too small: status -1, nothing changes
    keys, vals, props, goal = _random_case(rng, A, 50, pool=3)
    goal = goal + 10
    st, k, _, _, _ = oracle.ind_maintain(p, keys, vals, 50, props, goal, 1, None)
    assert st == -1
    np.testing.assert_array_equal(k, keys)
    goal[0] = -1
    assert oracle.ind_maintain(p, keys, vals, 50, props, goal, 1, None)[0] == -2


def test_kill_picks_uniform():
    """10 particles in one cell, goal 7: each is killed with probability 3/10."""
    A, n, cap = 4, 10, 20
    p = oracle.ind_params(2, 2, 2.0)
    keys = np.full(cap, DEAD, np.uint32)
    keys[:n] = 2
    vals = np.arange(cap, dtype=np.uint32)
    props = np.ones((4, cap), np.float32)
    goal = np.array([0, 0, 7, 0], np.int32)
    hits = np.zeros(n)
    trials = 3000
    for s in range(trials):
        _, k, v, _, n_new = oracle.ind_maintain(p, keys, vals, n, props, goal, s, None)
        assert n_new == 7
        dead = v[n_new:n_new + 3]
        assert np.all(k[n_new:n_new + 3] == DEAD) and np.all(np.diff(dead) > 0)  # lost in list order
        hits[dead] += 1
    expect = trials * 0.3
    chi2 = float(((hits - expect) ** 2 / (expect * 0.7)).sum())
    assert chi2 < 27.9  # 9 degrees of freedom, p = 0.001


def test_random_rule_uniform_and_pool_order():
    """mode 2 with a = 0, b = 1 gives RANDOM itself: uniform on [0, 1); new particles take
    the pool's ids front first, cell by cell."""
    W, H = 16, 16
    A = W * H
    cap = 20000
    p = oracle.ind_params(W, H, float(W))
    keys = np.full(cap, DEAD, np.uint32)
    vals = np.arange(cap, dtype=np.uint32)
    props = np.zeros((4, cap), np.float32)
    goal = np.full(A, 40, np.int32)
    st, k, v, pr, n_new = oracle.ind_maintain(p, keys, vals, 0, props, goal, 99,
                                              oracle.new_rule((2, 0, 0, 0), (0, 0, 0, 0), (1, 0, 0, 0)))
    assert st == 0 and n_new == A * 40
    np.testing.assert_array_equal(v[:n_new], np.arange(n_new))
    np.testing.assert_array_equal(k[:n_new], np.repeat(np.arange(A), 40))
    x = np.sort(pr[0, :n_new].astype(np.float64))
    ks = np.max(np.abs(x - np.arange(1, x.size + 1) / x.size))
    assert ks < 1.63 / math.sqrt(x.size)  # Kolmogorov-Smirnov, p = 0.01
    assert np.all(pr[1:, :] == 0)
