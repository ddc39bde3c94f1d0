"""Pins the oracle's deterministic stencils and the reaction-diffusion forward/adjoint:
discrete eigenmode decay (closed form), the homogeneous ODE limit, RK4 against the
birth-death closed form, and central finite differences for the gradient."""
import numpy as np
import pytest


def _eig_mode(N, m):
    j = np.arange(N)
    return np.cos(np.pi * m * j / (N - 1))


def test_hom_diffusion_eigenmode(oracle):
    """Mirror-BC Laplacian: cos(pi m j/(N-1)) is an eigenvector with eigenvalue
    -4 sin^2(pi m/(2(N-1))); one Jacobi step scales it by 1 - dt D lam/h^2."""
    W, H, m = 40, 24, 3
    mode = _eig_mode(W, m)[None, :] * _eig_mode(H, 2)[:, None]
    st = np.stack([mode, 2 * mode])
    D = np.array([0.7, 0.3])
    dt, h = 0.05, 1.3
    out = oracle.det_hom_diffuse(st, D, dt, h)
    lam = 4 * np.sin(np.pi * m / (2 * (W - 1))) ** 2 + 4 * np.sin(np.pi * 2 / (2 * (H - 1))) ** 2
    for s in range(2):
        np.testing.assert_allclose(out[s], st[s] * (1 - dt * D[s] * lam / h ** 2), rtol=1e-12, atol=1e-12)


def test_inhom_sweep_conservation_and_drift(oracle):
    """With uniform D and r, the X sweep conserves mass away from the boundary and moves
    the centre of mass of a bump by r dt."""
    W, H = 64, 8
    x = np.arange(W)
    bump = np.exp(-((x - 32.0) ** 2) / 20.0)
    st = np.tile(bump, (1, H, 1)).astype(np.float64)
    D = np.full((1, H, W), 0.5)
    r = np.full((1, H, W), 0.3)
    out = oracle.det_inhom_sweep(st, 0, 0, 0.2, D, r, 1.0)
    assert abs(out.sum() - st.sum()) < 1e-9 * st.sum()
    com0 = (st[0, 0] * x).sum() / st[0, 0].sum()
    com1 = (out[0, 0] * x).sum() / out[0, 0].sum()
    assert abs((com1 - com0) - 0.3 * 0.2) < 1e-9


def test_rk4_birth_death(oracle):
    """0 -> A (k), A -> 0 (g): A(t) = k/g + (A0 - k/g) e^{-g t}."""
    t = oracle.compile_reactions(1, [2.0, 0.5], [[0], [1]], [[1], [0]])
    st = np.full((1, 4, 4), 10.0)
    out = oracle.det_rk4 if hasattr(oracle, "det_rk4") else None
    from oracle.oracle import lib, _p
    import ctypes as C
    lib().oracle_det_rk4(_p(st), 4, 4, C.byref(t), C.c_double(1.5))
    want = 4.0 + (10.0 - 4.0) * np.exp(-0.5 * 1.5)
    # classical RK4 with 20 substeps of 0.075: global error O(h^4) ~ 1e-8 relative
    np.testing.assert_allclose(st, want, rtol=2e-8)
    assert abs(st[0, 0, 0] - want) > 1e-12  # it is RK4, not the exact solution


def _small_problem(W=16, H=12, steps=30):
    rng = np.random.default_rng(3)
    D = np.exp(0.3 * rng.standard_normal((H, W)))
    u0 = 200 + 10 * rng.standard_normal((H, W))
    v0 = 75 + 5 * rng.standard_normal((H, W))
    ut = 1 + 0.1 * rng.standard_normal((H, W))
    vt = 60 + 5 * rng.standard_normal((H, W))
    return D, u0, v0, ut, vt


def test_rdd_uniform_field_is_the_ode(oracle):
    """Uniform D and initial state: diffusion vanishes and every cell follows the
    explicit Euler recurrence of the Schnakenberg ODE."""
    W = H = 8
    k = (1.0, 3.0, 1e-4, 1.0)
    p = oracle.rdd_params(W, H, 1.0, 0.01, 0.05, 1.0, k)
    D = np.full((H, W), 1.7)
    u, v = oracle.rdd_forward(p, D, np.full((H, W), 200.0), np.full((H, W), 75.0), 50)
    a, b = 200.0, 75.0
    for _ in range(50):
        fu = k[0] - k[3] * a + k[2] * a * a * b
        fv = k[1] - k[2] * a * a * b
        a, b = a + 0.01 * fu, b + 0.01 * fv
    np.testing.assert_allclose(u, a, rtol=1e-13)
    np.testing.assert_allclose(v, b, rtol=1e-13)


@pytest.mark.parametrize("ckpt", [7, 100])
def test_rdd_gradient_finite_differences(oracle, ckpt):
    W, H, steps = 16, 12, 30
    D, u0, v0, ut, vt = _small_problem(W, H, steps)
    k = [1.0, 3.0, 1e-4, 1.0]
    dt = 0.01

    def loss_of(Dv, kv):
        p = oracle.rdd_params(W, H, 1.0, dt, 0.05, 1.0, kv)
        u, v = oracle.rdd_forward(p, Dv, u0, v0, steps)
        return oracle.rdd_loss(u, v, ut, vt) if hasattr(oracle, "rdd_loss") else \
            float(np.mean((u - ut) ** 2 + (v - vt) ** 2))

    p = oracle.rdd_params(W, H, 1.0, dt, 0.05, 1.0, k)
    loss, gD, gk = oracle.rdd_grad(p, D, u0, v0, ut, vt, steps, ckpt)
    assert abs(loss - loss_of(D, k)) < 1e-9 * abs(loss)
    for j in range(4):
        e = 1e-3 * k[j]
        kp = list(k); kp[j] += e
        km = list(k); km[j] -= e
        fd = (loss_of(D, kp) - loss_of(D, km)) / (2 * e)
        assert abs(fd - gk[j]) <= 1e-6 * max(abs(fd), 1e-12) + 1e-12, (j, fd, gk[j])
    rng = np.random.default_rng(9)
    for _ in range(3):
        # random directions; the central difference converges as e^2 (1e-6 at e = 1e-3,
        # 1e-8 at e = 1e-4), so e = 1e-4 separates truncation from adjoint error
        d = rng.standard_normal(D.shape)
        e = 1e-4
        fd = (loss_of(D + e * d, k) - loss_of(D - e * d, k)) / (2 * e)
        an = float((gD * d).sum())
        assert abs(fd - an) <= 1e-7 * abs(fd), (fd, an)


# ---- alpha-QSS (deterministic_performReaction_aqss) ------------------------------------
def _aqss(oracle, t, st, dt, dtype, fixed):
    return oracle.det_aqss(st, t, dt, dtype, fixed)


def test_aqss_birth_death(oracle):
    """0 -> A (k), A -> 0 (g): A(t) = k/g + (A0 - k/g) e^{-g t}.  With the predictor's Pade
    factor in the corrector (Mott & Oran's alpha) the integrator is accurate to its error
    control; the reference's corrector factor (180.*+rtaub*..., Deterministic.cl:423) is
    kept by default and lands about 11.6 % off at dt = 1.5, in both precisions."""
    t = oracle.compile_reactions(1, [2.0, 0.5], [[0], [1]], [[1], [0]])
    st = np.full((1, 4, 4), 10.0)
    for dt, tol in ((0.1, 1e-7), (1.5, 2e-6), (5.0, 1e-4)):
        want = 4.0 + (10.0 - 4.0) * np.exp(-0.5 * dt)
        for dtype in (np.float64, np.float32):
            out, it = _aqss(oracle, t, st, dt, dtype, True)
            assert it > 0
            np.testing.assert_allclose(out, want, rtol=tol)
    want = 4.0 + (10.0 - 4.0) * np.exp(-0.5 * 1.5)
    for dtype in (np.float64, np.float32):
        out, _ = _aqss(oracle, t, st, 1.5, dtype, False)
        assert abs(out[0, 0, 0] / want - 1.0 - 0.11573) < 2e-4


def test_aqss_constant_production(oracle):
    """0 -> A only: q constant, d = 0, so both corrector forms integrate it exactly."""
    t = oracle.compile_reactions(1, [2.5], [[0]], [[1]])
    st = np.full((1, 2, 2), 3.0)
    for fixed in (True, False):
        out, _ = _aqss(oracle, t, st, 4.0, np.float64, fixed)
        np.testing.assert_allclose(out, 3.0 + 2.5 * 4.0, rtol=1e-12)


def test_aqss_stiff_equilibrium(oracle):
    """A <-> B with kf = 1000, kb = 10 relaxes to A = 100 kb / (kf + kb)."""
    t = oracle.compile_reactions(2, [1000.0, 10.0], [[1, 0], [0, 1]], [[0, 1], [1, 0]])
    st = np.zeros((2, 2, 2))
    st[0] = 100.0
    out, it = _aqss(oracle, t, st, 1.0, np.float64, True)
    np.testing.assert_allclose(out[0], 100 * 10 / 1010, rtol=1e-2)
    np.testing.assert_allclose(out[1], 100 * 1000 / 1010, rtol=1e-2)


def test_aqss_schnakenberg_against_rk4(oracle):
    """The Schnakenberg network (concentration form) against the fp64 RK4 oracle with 200
    calls of 20 substeps; the float instance follows the double one step for step."""
    from oracle.oracle import lib, _p
    import ctypes as C

    t = oracle.compile_reactions(2, [1.0, 3.0, 1e-4, 1.0], [[0, 0], [0, 0], [2, 1], [1, 0]],
                                 [[1, 0], [0, 1], [3, 0], [0, 0]])
    rng = np.random.default_rng(0)
    st = np.stack([rng.uniform(50, 250, (8, 8)), rng.uniform(20, 120, (8, 8))])
    for dt, tol in ((0.01, 2e-5), (0.1, 5e-3), (1.0, 2e-2)):
        ref = st.copy()
        for _ in range(200):
            lib().oracle_det_rk4(_p(ref), 8, 8, C.byref(t), C.c_double(dt / 200))
        o64, it64 = _aqss(oracle, t, st, dt, np.float64, True)
        o32, it32 = _aqss(oracle, t, st, dt, np.float32, True)
        assert it64 == it32
        assert np.max(np.abs(o64 - ref) / ref) <= tol
        assert np.max(np.abs(o32 - o64) / o64) <= 1e-5
        r64, _ = _aqss(oracle, t, st, dt, np.float64, False)
        assert np.max(np.abs(r64 - ref) / ref) > np.max(np.abs(o64 - ref) / ref)
