"""The individual-species oracle (particle-based lattice hops, IndividualSolver) pinned
by closed forms: sweep moments (mean r dt, variance 2 D dt per axis), boundary rules,
cell locality of the RNG stream, per-cell constants against numpy group reductions, and
its exp/expm1 and time step against libm."""
import math

import numpy as np
import pytest

from oracle import oracle


def _setup(W, H, n, cells, D=(0.5, 0.3), r=(0.3, -0.2), seed=0):
    props = np.stack([np.full(n, D[0]), np.full(n, D[1]), np.full(n, r[0]), np.full(n, r[1])]).astype(np.float32)
    keys = np.asarray(cells, np.uint32)
    vals = np.arange(n, dtype=np.uint32)
    return keys, vals, props


def test_exp_expm1_against_libm():
    L = oracle.lib()
    for x in np.concatenate([np.linspace(-700, 700, 4001), np.geomspace(1e-300, 60, 2000)]):
        assert abs(L.oracle_expd(float(x)) - math.exp(x)) <= 4e-16 * math.exp(x)
        if x >= 0:
            assert abs(L.oracle_expm1d(float(x)) - math.expm1(x)) <= 4e-16 * max(math.expm1(x), 1e-300)


def test_cell_dt_against_libm_formula():
    """Same formula as the libm-based stochastic time step: identical floats where the
    formula is well conditioned (|eps| >= 1e-3), within its conditioning below."""
    L = oracle.lib()
    rng = np.random.default_rng(0)
    h = np.float32(1.0)
    for _ in range(3000):
        D = np.float32(rng.uniform(0.01, 2.0))
        eps = rng.uniform(1e-3, 3.0) * rng.choice([-1, 1])
        r = np.float32(eps * 2 * D)
        assert L.oracle_ind_cell_dt(D, r, h) == L.oracle_inhomogeneous_cell_dt(D, r, h)
    for _ in range(1000):
        D = np.float32(rng.uniform(0.01, 2.0))
        eps = rng.uniform(1e-5, 1e-3)
        r = np.float32(eps * 2 * D)
        a, b = L.oracle_ind_cell_dt(D, r, h), L.oracle_inhomogeneous_cell_dt(D, r, h)
        assert abs(a - b) <= np.spacing(np.float32(b)) + 1e-14 / float(eps) ** 2 * b
    # r = 0: h^2 / (6 D); D = 0: h / |r|; empty cell: inf
    assert L.oracle_ind_cell_dt(np.float32(0.5), np.float32(0.0), h) == np.float32(1.0 / 3.0)
    assert L.oracle_ind_cell_dt(np.float32(0.0), np.float32(-0.25), h) == np.float32(4.0)
    assert math.isinf(L.oracle_ind_cell_dt(np.float32(0.0), np.float32(0.0), h))


@pytest.mark.parametrize("axis", [0, 1])
def test_sweep_moments(axis):
    """One sweep of particles in one interior cell: displacement mean r dt and variance
    2 D dt along the swept axis (cell width 1), nothing along the other."""
    W = H = 9
    n = 200000
    c0 = 4 + 4 * W
    D, r, dt = (0.5, 0.3), (0.3, -0.2), 0.2
    keys, vals, props = _setup(W, H, n, np.full(n, c0), D, r)
    p = oracle.ind_params(W, H, float(W))
    k, v, le, ri, _, _, _ = oracle.ind_update(p, keys, vals, props)
    k2 = oracle.ind_sweep(p, k, v, le, ri, props, axis, 12345, dt)
    dx = (k2.astype(np.int64) % W) - 4
    dy = (k2.astype(np.int64) // W) - 4
    moved, still = (dx, dy) if axis == 0 else (dy, dx)
    assert np.all(still == 0) and np.all(np.abs(moved) <= 1)
    mean, var = moved.mean(), moved.var()
    want_m, want_v = r[axis] * dt, 2 * D[axis] * dt
    assert abs(mean - want_m) < 5 * math.sqrt(want_v / n)
    assert abs(var - want_v) < 5 * math.sqrt(2 * want_v ** 2 / n) + 0.01 * want_v


def test_sweep_reflective_and_absorbing_boundaries():
    W, H = 8, 6
    n = 40000
    rng = np.random.default_rng(1)
    edge = np.concatenate([np.arange(H) * W, np.arange(H) * W + W - 1, np.arange(W), (H - 1) * W + np.arange(W)])
    cells = rng.choice(edge, n)
    keys, vals, props = _setup(W, H, n, cells, D=(0.6, 0.6), r=(0.0, 0.0))
    # reflective everywhere: nothing leaves, particles that drew an outward hop stay
    p = oracle.ind_params(W, H, float(W))
    k, v, le, ri, _, _, _ = oracle.ind_update(p, keys, vals, props)
    for axis in (0, 1):
        k2 = oracle.ind_sweep(p, k, v, le, ri, props, axis, 7, 0.3)
        assert np.all(k2 < W * H)
        step = 1 if axis == 0 else W
        assert np.all(np.isin(k2.astype(np.int64) - k.astype(np.int64), [-step, 0, step]))
    # absorbing +y: top-row particles hopping +y leave the cell range
    pa = oracle.ind_params(W, H, float(W), bm=(0, 0, 0, 1))
    k2 = oracle.ind_sweep(pa, k, v, le, ri, props, 1, 7, 0.3)
    out = k2 >= W * H
    assert out.any() and np.all(k[out] >= (H - 1) * W) and np.all(k2[out] - k[out] == W)


def test_sweep_cell_locality():
    """A cell's particles draw from the stream keyed by that cell: particles elsewhere
    do not change their hops."""
    W = H = 16
    rng = np.random.default_rng(2)
    c0 = 5 + 7 * W
    base = np.full(50, c0)
    extra = rng.integers(0, W * H, 300)
    extra = extra[extra != c0]
    p = oracle.ind_params(W, H, float(W))
    res = []
    for cells in (base, np.concatenate([extra[:150], base, extra[150:]])):
        n = cells.size
        keys, vals, props = _setup(W, H, n, cells)
        k, v, le, ri, _, _, _ = oracle.ind_update(p, keys, vals, props)
        k2 = oracle.ind_sweep(p, k, v, le, ri, props, 0, 99, 0.25)
        res.append(k2[k == c0])
    assert np.array_equal(res[0], res[1])


def test_constants_against_numpy():
    W, H = 12, 10
    n = 3000
    rng = np.random.default_rng(3)
    cells = rng.integers(0, W * H, n)
    props = np.stack([rng.uniform(0, 1, n), rng.uniform(0, 1, n), rng.uniform(-1, 1, n),
                      rng.uniform(-1, 1, n)]).astype(np.float32)
    p = oracle.ind_params(W, H, float(W))
    k, v, le, ri, (Dx, Dy, rx, ry, cnt), _, _ = oracle.ind_update(p, cells.astype(np.uint32),
                                                                   np.arange(n, dtype=np.uint32), props)
    assert np.array_equal(cnt, np.bincount(cells, minlength=W * H))
    for arr, row in ((Dx, 0), (Dy, 1)):
        want = np.zeros(W * H, np.float32)
        np.maximum.at(want, cells, props[row])
        assert np.array_equal(arr, want)
    for arr, row in ((rx, 2), (ry, 3)):
        want = np.zeros(W * H, np.float32)
        for c in range(W * H):
            vals_c = props[row][cells == c]
            if vals_c.size:
                want[c] = vals_c[np.argmax(np.abs(vals_c))]
        assert np.array_equal(arr, want)


def test_run_conserves_particles_and_advances_time():
    W, H = 20, 16
    n = 4000
    rng = np.random.default_rng(4)
    cells = rng.integers(0, W * H, n).astype(np.uint32)
    props = np.stack([rng.uniform(0.1, 1, n), rng.uniform(0.1, 1, n), rng.uniform(-0.2, 0.2, n),
                      rng.uniform(-0.2, 0.2, n)]).astype(np.float32)
    p = oracle.ind_params(W, H, float(W))
    keys, vals, cnt, t = oracle.ind_run(p, cells, np.arange(n, dtype=np.uint32), props, 40, 3)
    assert cnt.sum() == n and np.all(keys < W * H) and np.all(np.diff(keys.astype(np.int64)) >= 0)
    assert sorted(vals.tolist()) == list(range(n))
    assert t > 0
