"""Parity at the configured sizes: the batched ensemble on its 512 uneven grids with
Threefry-derived seeds, 10^7 particles (sort, offsets, steps) and the 256x256 Schnakenberg run
deep in its Turing regime, each against the plain solver and/or the oracle, bit-exact."""
import os

import numpy as np
import pytest

from inchman_amd import configs

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_ensemble_config_grids(gpu, oracle):
    """512 realizations of the plain GMP network on the configured sides (log-uniform in
    [64, 512], multiples of 16; 3.04e7 cells) with seeds s_r = Threefry(r, 0xBA7C), 20
    host-loop iterations in one batch (autotuned launch shape, HIP-graph chunks): every
    realization equals the plain solver run alone with the same seed, and a stratified
    sample (smallest, largest and six sizes between) equals the oracle."""
    from inchman_amd import Ensemble

    n, steps = 512, 20
    sides = configs.ensemble_sides(n)
    seeds = configs.ensemble_seeds(n)
    assert sides.min() >= 64 and sides.max() <= 512 and (sides % 16 == 0).all()
    m = configs.plain_gmp_model(32)
    states = [configs.ensemble_state(sd, r) for r, sd in enumerate(sides)]
    ens = Ensemble(m, sides, sides, seeds)
    for r, st in enumerate(states):
        ens.setState(r, st)
    ens.autotune(3)
    for r, st in enumerate(states):
        ens.setState(r, st)
    ens.reset()
    ens.run(steps, 10)
    got = [ens.getState(r) for r in range(n)]
    for r, sd in enumerate(sides):
        p = configs.plain_gmp_model(int(sd)).createSolver()
        p.synchronizeStateBufferToDevice(states[r])
        p.run(steps, int(seeds[r]))
        assert np.array_equal(got[r], p.synchronizeStateBufferFromDevice()), r
    order = np.argsort(sides, kind="stable")
    sample = sorted(set(order[np.linspace(0, n - 1, 8).astype(int)].tolist()))
    assert sides[sample].max() == sides.max() and sides[sample].min() == sides.min()
    for r in sample:
        want, _ = oracle.run(configs.plain_gmp_model(int(sides[r])), states[r], steps, int(seeds[r]))
        assert np.array_equal(got[r], want), (r, int(sides[r]))


def test_turing_regime_256(gpu, oracle):
    """The 256x256 Schnakenberg run resumed at iteration 6000 (the oracle's state there,
    tests/golden/schnakenberg256_it6000.npz), where the spots draw hundreds of SSA events
    per cell and iteration and the busy cells go to the lane and wave workers: 60 more
    iterations (HIP-graph chunks) bit-exact against the oracle resumed from the same state,
    clock included."""
    z = np.load(os.path.join(GOLDEN, "schnakenberg256_it6000.npz"))
    snap, it0, seed0 = z["state"], int(z["iterations"]), int(z["seed0"])
    m = configs.schnakenberg_model(256)
    s = m.createSolver()
    s.synchronizeStateBufferToDevice(configs.schnakenberg_state(256))
    s.run(0, seed0)
    s.synchronizeStateBufferToDevice(snap)
    t_skip = s.skip(it0)
    assert t_skip == float(z["sim_time"])
    t = s.runContinue(60, 16)
    got = s.synchronizeStateBufferFromDevice()
    want, tw = oracle.run(m, snap, 60, seed0, skip=it0)
    assert t == tw
    assert np.array_equal(got, want)
    # the regime: the spots hold many times the mean count of U
    assert snap[0].max() > 8 * snap[0].mean()


@pytest.mark.parametrize("network", ["plain", "schnakenberg"])
def test_skip_then_continue(gpu, network):
    """Skipping iterations over a snapshot resumes the loop exactly: a run of 50 + 30
    iterations equals 50 iterations, a fresh solver given their state, 50 skipped and 30
    run (eagerly and in graph chunks)."""
    if network == "plain":
        m, st = configs.plain_gmp_model(48), configs.plain_gmp_state(48)
    else:
        m, st = configs.schnakenberg_model(64), configs.schnakenberg_state(64, seed=9)
    a = m.createSolver()
    a.synchronizeStateBufferToDevice(st)
    a.run(50, 7)
    mid = a.synchronizeStateBufferFromDevice()
    t_all = a.runContinue(30)
    ref = a.synchronizeStateBufferFromDevice()
    for chunk in (0, 8):
        b = m.createSolver()
        b.synchronizeStateBufferToDevice(st)
        b.run(0, 7)
        b.synchronizeStateBufferToDevice(mid)
        b.skip(50)
        assert b.runContinue(30, chunk) == t_all
        assert np.array_equal(b.synchronizeStateBufferFromDevice(), ref)


def test_sort_1e7_keys(gpu, oracle):
    """The stable radix sort on 10^7 (cell key, id) pairs of 20-bit keys (the particle
    configuration's key range) against std::stable_sort."""
    import ctypes as C

    import torch

    from inchman_amd._lib import call

    n = 10_000_000
    rng = np.random.default_rng(10)
    k = rng.integers(0, 1 << 20, n, dtype=np.uint64).astype(np.uint32)
    v = np.arange(n, dtype=np.uint32)
    tk = torch.from_numpy(k.view(np.int32)).cuda()
    tv = torch.from_numpy(v.view(np.int32)).cuda()
    t1, t2 = torch.empty_like(tk), torch.empty_like(tv)
    call("gpgmp_sort_pairs_device", tk.data_ptr(), tv.data_ptr(), t1.data_ptr(), t2.data_ptr(), n, 20, None)
    torch.cuda.synchronize()
    ks, vs = oracle.stable_sort_pairs(k, v)
    assert np.array_equal(tk.cpu().numpy().view(np.uint32), ks)
    assert np.array_equal(tv.cpu().numpy().view(np.uint32), vs)
    del C


def test_particles_1e7(gpu, oracle):
    """The particle configuration at full size: 10^7 particles on 1024x1024 periodic cells,
    D = 1, dt = 0.01, vortex drift 0.1, A+A->0 with p = 0.05, three steps (the first from an
    unsorted list, the others from the sorted one): positions, ids, keys and offsets
    bit-exact against the oracle."""
    from inchman_amd import particles

    n, side, steps = 10_000_000, 1024, 3
    rng = np.random.default_rng(4)
    x = rng.uniform(0, side, n).astype(np.float32)
    y = rng.uniform(0, side, n).astype(np.float32)
    ids = np.arange(n, dtype=np.uint32)
    ps = particles.ParticleSystem(side, D=1.0, dt=0.01, vortex_amp=0.1, p_annihilate=0.05, seed=17, capacity=n)
    ps.set(x, y, ids)
    ps.step(0, steps)
    gx, gy, gi, gk, goff = ps.get(offsets=True)
    po = oracle.particle_params(side, 1.0, 0.01, 0.1, 0.05, 17)
    ox, oy, oi, ok, ooff = oracle.particle_steps(po, 0, steps, x, y, ids)
    assert len(gx) == len(ox) < n
    assert np.array_equal(gi, oi) and np.array_equal(gk, ok)
    assert np.array_equal(gx, ox) and np.array_equal(gy, oy)
    assert np.array_equal(goff, ooff)
