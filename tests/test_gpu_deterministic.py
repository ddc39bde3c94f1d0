"""FP32 HIP deterministic kernels against the FP64 oracle: the reference's stencils
and RK4 (relative error <= 1e-5), the fused Schnakenberg forward (<= 1e-5) and the
hand-written adjoint (gradients within 1e-4 of the FP64 gradient)."""
import ctypes as C

import numpy as np
import pytest
import torch

from inchman_amd import configs
from inchman_amd._lib import call

pytestmark = pytest.mark.gpu


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(a, np.float32)).cuda()


def _rel(a, b):
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-30)))


@pytest.mark.parametrize("H,W", [(40, 48), (37, 50), (300, 1024), (5, 4)])
def test_det_hom_diffuse(gpu, oracle, H, W):
    """Streaming 4-cell form (W % 4 == 0, strips of 16 rows, ragged last strip) and the
    one-cell form (other widths)."""
    rng = np.random.default_rng(1)
    st = rng.uniform(1, 100, (2, H, W))
    D = np.array([0.6, 0.2], np.float32)
    dt, h = 0.05, 1.1
    x = _dev(st)
    y = torch.empty_like(x)
    call("gpgmp_det_hom_diffuse_device", x.data_ptr(), y.data_ptr(), W, H, 2,
         D.ctypes.data_as(C.POINTER(C.c_float)), dt, h, None)
    torch.cuda.synchronize()
    want = oracle.det_hom_diffuse(st.astype(np.float32).astype(np.float64), D.astype(np.float64), dt,
                                  float(np.float32(h)))
    assert _rel(y.cpu().numpy(), want) <= 1e-5


@pytest.mark.parametrize("H,W", [(32, 48), (37, 50), (300, 1024), (5, 4)])
@pytest.mark.parametrize("axis", [0, 1])
def test_det_inhom_sweep(gpu, oracle, axis, H, W):
    rng = np.random.default_rng(2 + axis)
    st = rng.uniform(1, 100, (2, H, W)).astype(np.float32)
    Dax = rng.uniform(0.1, 1.0, (2, H, W)).astype(np.float32)
    rax = rng.uniform(-0.3, 0.3, (2, H, W)).astype(np.float32)
    x, D, r = _dev(st), _dev(Dax), _dev(rax)
    y = x.clone()
    call("gpgmp_det_inhom_sweep_device", x.data_ptr(), y.data_ptr(), W, H, 1, axis, 0.1, D.data_ptr(),
         r.data_ptr(), 1.0, None)
    torch.cuda.synchronize()
    want = oracle.det_inhom_sweep(st.astype(np.float64), 1, axis, 0.1, Dax.astype(np.float64),
                                  rax.astype(np.float64), 1.0)
    got = y.cpu().numpy()
    assert _rel(got[1], want[1]) <= 1e-5
    assert np.array_equal(got[0], st[0])


@pytest.mark.parametrize("table", ["compact", "general"])
def test_det_rk4(gpu, oracle, table):
    """RK4 reactions against the fp64 oracle: the plain GMP network (4 reactions, the
    straight-line right-hand side) and the same with a fifth reaction (the table walk)."""
    m = configs.plain_gmp_model(32)
    if table == "general":
        a, b, c = m.species
        m.addReaction(2e-4, reactants={a: 2, c: 1}, products={b: 1})
    cm = m.compile()
    rng = np.random.default_rng(4)
    st = rng.uniform(0, 60, (3, 32, 32)).astype(np.float32)
    x = _dev(st)
    call("gpgmp_det_rk4_device", cm.handle, x.data_ptr(), 0.5, None)
    torch.cuda.synchronize()
    want = st.astype(np.float64).copy()
    t = oracle.table_from_model(m)
    oracle.lib().oracle_det_rk4(oracle._p(want), 32, 32, C.byref(t), C.c_double(0.5))
    assert _rel(x.cpu().numpy(), want) <= 1e-5


@pytest.mark.parametrize("fixed", [0, 1])
def test_det_aqss(gpu, oracle, fixed):
    """alpha-QSS reactions (deterministic_performReaction_aqss) bit-exact against the
    oracle's float instance (the reference's promotions, no contraction): the plain GMP
    network (3 species) and the Schnakenberg network (2 species), states with zeros and
    tiny values, step sizes from one pass to many; both corrector forms, including cells
    whose step size underflows (NaN on both sides)."""
    from inchman_amd._lib import AQSS_PADE_FIX

    rng = np.random.default_rng(40 + fixed)
    five = configs.plain_gmp_model(32)  # a fifth reaction: the table walk, not the compact form
    a, b, c = five.species
    five.addReaction(2e-4, reactants={a: 2, c: 1}, products={b: 1})
    for m, scale, dts in ((configs.plain_gmp_model(32), 60.0, (0.05, 0.5, 2.0)),
                          (configs.schnakenberg_model(32), 250.0, (0.01, 0.1, 1.0)),
                          (five, 60.0, (0.05, 0.5))):
        S = m.numSpecies()
        st = rng.uniform(0, scale, (S, 32, 32)).astype(np.float32)
        st[:, ::7, ::5] = 0.0
        st[:, 3, :] = 1e-6
        t = oracle.table_from_model(m)
        cm = m.compile()
        for dt in dts:
            x = _dev(st)
            call("gpgmp_det_aqss_device", cm.handle, x.data_ptr(), dt, AQSS_PADE_FIX if fixed else 0, None)
            torch.cuda.synchronize()
            want, it = oracle.det_aqss(st, t, dt, np.float32, bool(fixed))
            # it == -1: some cells' step size underflowed (the reference corrector on cells
            # at zero, Schnakenberg at dt = 1); both sides set them to NaN
            assert it != 0
            got = x.cpu().numpy()
            assert np.array_equal(got, want, equal_nan=True), (dt, int(np.sum(got != want)))


def _rdd(prob, steps=None, ckpt=100):
    from inchman_amd.rdd import ReactionDiffusion

    H, W = prob["D"].shape
    return ReactionDiffusion(W, H, steps if steps is not None else prob["steps"], prob["dt"], prob["h"],
                             prob["D0"], prob["k"], ckpt)


def _oracle_params(oracle, prob):
    H, W = prob["D"].shape
    f = lambda v: float(np.float32(v))  # noqa: E731  the GPU runs on the fp32 values
    return oracle.rdd_params(W, H, f(prob["h"]), f(prob["dt"]), f(prob["D0"][0]), f(prob["D0"][1]),
                             [f(k) for k in prob["k"]])


@pytest.mark.parametrize("side,steps", [(256, 300), (1024, 2000)])
def test_rdd_forward(gpu, oracle, side, steps):
    prob = configs.rdd_problem(side, steps=steps)
    r = _rdd(prob)
    D, u, v = _dev(prob["D"]), _dev(prob["u0"]), _dev(prob["v0"])
    r.forward(D, u, v, save=False)
    torch.cuda.synchronize()
    p = _oracle_params(oracle, prob)
    uw, vw = oracle.rdd_forward(p, prob["D"].astype(np.float64), prob["u0"], prob["v0"], steps)
    assert _rel(u.cpu().numpy(), uw) <= 1e-5
    assert _rel(v.cpu().numpy(), vw) <= 1e-5


@pytest.mark.parametrize("side,steps,ckpt", [(128, 300, 100), (96, 250, 64), (1024, 2000, 100)])
def test_rdd_gradient(gpu, oracle, side, steps, ckpt):
    prob = configs.rdd_problem(side, steps=steps)
    r = _rdd(prob, ckpt=ckpt)
    D, u, v = _dev(prob["D"]), _dev(prob["u0"]), _dev(prob["v0"])
    r.forward(D, u, v, save=True)
    n = u.numel()
    ut, vt = _dev(prob["ut"]), _dev(prob["vt"])
    lu = (2.0 / n) * (u - ut)
    lv = (2.0 / n) * (v - vt)
    gD, gk, gu0, gv0 = r.backward(D, lu.contiguous(), lv.contiguous())
    torch.cuda.synchronize()
    p = _oracle_params(oracle, prob)
    loss, gDw, gkw = oracle.rdd_grad(p, prob["D"], prob["u0"], prob["v0"], prob["ut"], prob["vt"], steps, ckpt)
    g = gD.cpu().numpy().astype(np.float64)
    assert np.linalg.norm(g - gDw) / np.linalg.norm(gDw) <= 1e-4
    gkn = gk.cpu().numpy().astype(np.float64)
    assert np.all(np.abs(gkn - gkw) <= 1e-4 * np.abs(gkw)), (gkn, gkw)


def test_rdd_autograd_function(gpu):
    from inchman_amd.rdd import simulate

    prob = configs.rdd_problem(128, steps=200)
    r = _rdd(prob)
    D = _dev(prob["D"]).requires_grad_(True)
    k = torch.tensor(prob["k"], dtype=torch.float32, requires_grad=True)
    u0, v0 = _dev(prob["u0"]).requires_grad_(True), _dev(prob["v0"])
    u, v = simulate(r, D, k, u0, v0)
    loss = ((u - _dev(prob["ut"])) ** 2 + (v - _dev(prob["vt"])) ** 2).mean()
    loss.backward()
    # same numbers through the raw calls
    r2 = _rdd(prob)
    D2, uu, vv = _dev(prob["D"]), _dev(prob["u0"]), _dev(prob["v0"])
    r2.forward(D2, uu, vv, save=True)
    n = uu.numel()
    gD, gk, gu0, _ = r2.backward(D2, ((2.0 / n) * (uu - _dev(prob["ut"]))).contiguous(),
                                 ((2.0 / n) * (vv - _dev(prob["vt"]))).contiguous())
    torch.cuda.synchronize()
    assert torch.allclose(D.grad, gD, rtol=1e-6, atol=1e-12)
    assert torch.allclose(k.grad, gk.cpu(), rtol=1e-6)
    assert torch.allclose(u0.grad, gu0, rtol=1e-6, atol=1e-12)


def _rdd_run(prob, steps, ckpt, ns):
    r = _rdd(prob, steps=steps, ckpt=ckpt)
    r.set_steps_per_launch(ns)
    D, u, v = _dev(prob["D"]), _dev(prob["u0"]), _dev(prob["v0"])
    r.forward(D, u, v, save=True)
    n = u.numel()
    lu = ((2.0 / n) * (u - _dev(prob["ut"]))).contiguous()
    lv = ((2.0 / n) * (v - _dev(prob["vt"]))).contiguous()
    gD, gk, gu0, gv0 = r.backward(D, lu, lv)
    torch.cuda.synchronize()
    return [t.cpu().numpy() for t in (u, v, gD, gk, gu0, gv0)]


@pytest.mark.parametrize("W,H,steps,ckpt", [(100, 37, 37, 10), (3, 5, 9, 4), (64, 16, 20, 7), (130, 70, 41, 100),
                                            (257, 19, 3, 1)])
def test_rdd_two_steps_per_launch_bit_identical(gpu, oracle, W, H, steps, ckpt):
    """Temporal blocking (two or four Euler steps per launch, 2- or 4-cell LDS halo) gives the
    same bits as one step per launch, forward and adjoint, on ragged tiles, odd step counts and
    checkpoint intervals that split the launch pairs; both agree with the FP64 oracle."""
    prob = configs.rdd_problem(W, H, steps=steps)
    one = _rdd_run(prob, steps, ckpt, 1)
    two = _rdd_run(prob, steps, ckpt, 2)
    for a, b in zip(one, two):
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    if min(W, H) >= 5:  # four steps per launch (4-cell halo; 3 where a checkpoint splits a launch)
        four = _rdd_run(prob, steps, ckpt, 4)
        for a, b in zip(one, four):
            assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    p = _oracle_params(oracle, prob)
    uw, vw = oracle.rdd_forward(p, prob["D"].astype(np.float64), prob["u0"], prob["v0"], steps)
    assert _rel(two[0], uw) <= 1e-5 and _rel(two[1], vw) <= 1e-5
    _, gDw, gkw = oracle.rdd_grad(p, prob["D"], prob["u0"], prob["v0"], prob["ut"], prob["vt"], steps, ckpt)
    assert np.linalg.norm(two[2] - gDw) / np.linalg.norm(gDw) <= 1e-4
    assert np.all(np.abs(two[3] - gkw) <= 1e-4 * np.abs(gkw))


def test_rdd_two_cell_sides(gpu, oracle):
    """Sides of 2 cells run one step per launch (the 2-cell halo needs sides >= 3)."""
    prob = configs.rdd_problem(2, 7, steps=9)
    r = _rdd(prob, steps=9, ckpt=4)
    with pytest.raises(Exception):
        r.set_steps_per_launch(2)
    u, v, gD, gk, _, _ = _rdd_run(prob, 9, 4, 1)
    p = _oracle_params(oracle, prob)
    uw, vw = oracle.rdd_forward(p, prob["D"].astype(np.float64), prob["u0"], prob["v0"], 9)
    assert _rel(u, uw) <= 1e-5 and _rel(v, vw) <= 1e-5
    _, gDw, gkw = oracle.rdd_grad(p, prob["D"], prob["u0"], prob["v0"], prob["ut"], prob["vt"], 9, 4)
    assert np.linalg.norm(gD - gDw) / np.linalg.norm(gDw) <= 1e-4


def test_rdd_autograd_refuses_a_replaced_forward(gpu):
    """The handle keeps one forward's checkpoints: a second simulate() on the same handle
    before the first one's backward makes that backward raise instead of differentiating
    the second run; the rates of a CUDA tensor reach the solver without a host copy and
    give the same gradients as host rates."""
    from inchman_amd import GpgmpError
    from inchman_amd.rdd import simulate

    prob = configs.rdd_problem(64, steps=50)
    r = _rdd(prob, steps=50, ckpt=10)
    k1 = torch.tensor(prob["k"], dtype=torch.float32, device="cuda", requires_grad=True)
    D = _dev(prob["D"]).requires_grad_(True)
    u, v = simulate(r, D, k1, _dev(prob["u0"]), _dev(prob["v0"]))
    k2 = torch.tensor([1.5, 3.0, 1e-4, 1.0], dtype=torch.float32, device="cuda", requires_grad=True)
    u2, v2 = simulate(r, D, k2, _dev(prob["u0"]), _dev(prob["v0"]))
    with pytest.raises((GpgmpError, RuntimeError), match="saved forward"):
        (u.sum() + v.sum()).backward()
    (u2.sum() + v2.sum()).backward()
    assert k2.grad is not None and torch.isfinite(k2.grad).all()
    # device rates == host rates
    r3 = _rdd(prob, steps=50, ckpt=10)
    r3.set_rates([1.5, 3.0, 1e-4, 1.0])
    D3, uu, vv = _dev(prob["D"]), _dev(prob["u0"]), _dev(prob["v0"])
    r3.forward(D3, uu, vv, save=True)
    gD, gk, _, _ = r3.backward(D3, torch.ones_like(uu), torch.ones_like(vv), generation=r3.generation())
    torch.cuda.synchronize()
    assert torch.equal(u2.detach(), uu) and torch.equal(v2.detach(), vv)
    assert torch.allclose(k2.grad, gk, rtol=1e-6) and torch.allclose(D.grad, gD, rtol=1e-5, atol=1e-12)
