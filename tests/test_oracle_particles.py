"""Pins the oracle's sort, offsets and particle step: std::stable_sort semantics,
lower-bound offsets, the computeCellIndices left/right convention, and the step's
invariants (sortedness, conservation without annihilation, displacement moments)."""
import numpy as np


def test_stable_sort_pairs(oracle):
    rng = np.random.default_rng(0)
    k = rng.integers(0, 50, 5000).astype(np.uint32)
    v = np.arange(5000, dtype=np.uint32)
    ks, vs = oracle.stable_sort_pairs(k, v)
    order = np.argsort(k, kind="stable")
    assert np.array_equal(ks, k[order]) and np.array_equal(vs, v[order])
    # dead keys 0xffffffff sort last
    k[::7] = 0xFFFFFFFF
    ks, _ = oracle.stable_sort_pairs(k, v)
    assert np.all(ks[-len(k[::7]):] == 0xFFFFFFFF)


def test_offsets_and_left_right(oracle):
    keys = np.array([2, 2, 3, 7, 7, 7, 9], np.uint32)
    off = oracle.cell_offsets(keys, 12)
    assert list(off) == [int(np.searchsorted(keys, c)) for c in range(13)]
    left, right = oracle.cell_left_right(keys, 12)
    # computeCellIndices: occupied [first, last]; empty -> index of the next particle
    assert list(left) == [0, 0, 0, 2, 3, 3, 3, 3, 6, 6, 7, 7]
    assert list(right) == [0, 0, 1, 2, 3, 3, 3, 5, 6, 6, 7, 7]


def test_particle_step_invariants(oracle):
    side, n = 64, 20000
    rng = np.random.default_rng(1)
    x = rng.uniform(0, side, n).astype(np.float32)
    y = rng.uniform(0, side, n).astype(np.float32)
    ids = np.arange(n, dtype=np.uint32)
    p = oracle.particle_params(side, 1.0, 0.01, 0.1, 0.0, 77)
    x1, y1, i1, k1, off = oracle.particle_steps(p, 0, 1, x, y, ids)
    assert len(x1) == n  # no annihilation at p = 0
    assert np.all(np.diff(k1.astype(np.int64)) >= 0)
    assert np.array_equal(k1, (y1.astype(np.int64) * side + x1.astype(np.int64)).astype(np.uint32))
    assert sorted(i1.tolist()) == list(range(n))
    # stability: equal keys keep the order of the moved (unsorted) list, i.e. ascending ids here
    for c in np.unique(k1)[:200]:
        sel = i1[k1 == c]
        assert np.all(np.diff(sel.astype(np.int64)) > 0)
    assert list(off) == [int(np.searchsorted(k1, c)) for c in range(side * side + 1)]
    # displacement moments: mean v dt (~1e-3), variance 2 D dt = 0.02 per axis
    dx = (x1[np.argsort(i1)] - x + side / 2) % side - side / 2
    assert abs(dx.mean()) < 0.004 and abs(dx.var() - 0.02) < 0.001


def test_particle_annihilation_rate(oracle):
    side, n = 32, 40000
    rng = np.random.default_rng(2)
    x = rng.uniform(0, side, n).astype(np.float32)
    y = rng.uniform(0, side, n).astype(np.float32)
    p = oracle.particle_params(side, 1.0, 0.01, 0.1, 0.05, 5)
    x1, y1, i1, k1, off = oracle.particle_steps(p, 0, 1, x, y, np.arange(n, dtype=np.uint32))
    # removed pairs ~ 0.05 * sum over cells of floor(n_c / 2)
    removed = n - len(x1)
    assert removed % 2 == 0
    assert abs(removed / 2 - 0.05 * n / 2) < 5 * np.sqrt(0.05 * n / 2)


def test_gauss2_is_box_muller(oracle):
    """The particle Gaussian is the Box-Muller transform of one Threefry call, in float: against
    libm on the same words (a swapped sine/cosine, a missing factor or a wrong word would differ),
    and the pair is standard normal and uncorrelated."""
    import math

    rng = np.random.default_rng(3)
    ids = rng.integers(0, 2**32, 2000, dtype=np.uint64)
    g = []
    for k, i in enumerate(ids):
        gx, gy = oracle.gauss2(int(i), 1234567, k % 7)
        w0, w1 = oracle.threefry(k % 7, 0, int(i), 1234567)
        u1, u2 = ((w0 >> 9) + 0.5) * 2.0**-23, ((w1 >> 9) + 0.5) * 2.0**-23
        r = math.sqrt(-2.0 * math.log(u1))
        # float arithmetic: a few float ulps of r, and the float rounding of the phase
        assert abs(gx - r * math.cos(2 * math.pi * u2)) <= 2e-6 * max(1.0, r)
        assert abs(gy - r * math.sin(2 * math.pi * u2)) <= 2e-6 * max(1.0, r)
        g.append((gx, gy))
    g = np.array(g)
    assert np.all(np.abs(g.mean(axis=0)) < 0.1)
    assert np.all(np.abs(g.var(axis=0) - 1.0) < 0.1)
    assert abs(np.corrcoef(g[:, 0], g[:, 1])[0, 1]) < 0.1
