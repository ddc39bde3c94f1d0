"""HIP radix sort, offsets and particle step against the oracle: bit-exact order,
keys, positions and offsets."""
import ctypes as C

import numpy as np
import pytest
import torch

from inchman_amd import particles
from inchman_amd._lib import call

pytestmark = pytest.mark.gpu


def _u32(a):
    return torch.from_numpy(np.ascontiguousarray(a, np.uint32).view(np.int32)).cuda()


def _np_u32(t):
    return t.cpu().numpy().view(np.uint32)


@pytest.mark.parametrize("n,bits,hi", [(1, 8, 200), (4095, 12, 4096), (4097, 20, 1 << 20), (100003, 20, 1 << 20),
                                       (250000, 32, 2**32 - 1), (2000000, 20, 5)])
def test_sort_pairs(gpu, oracle, n, bits, hi):
    rng = np.random.default_rng(n)
    k = rng.integers(0, hi, n, dtype=np.uint64).astype(np.uint32)
    if bits == 32:
        k[::11] = 0xFFFFFFFF
    v = rng.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32)
    tk, tv = _u32(k), _u32(v)
    t1, t2 = torch.empty_like(tk), torch.empty_like(tv)
    call("gpgmp_sort_pairs_device", tk.data_ptr(), tv.data_ptr(), t1.data_ptr(), t2.data_ptr(), n, bits, None)
    torch.cuda.synchronize()
    ks, vs = oracle.stable_sort_pairs(k, v)
    assert np.array_equal(_np_u32(tk), ks)
    assert np.array_equal(_np_u32(tv), vs)


@pytest.mark.parametrize("shift", [1, 2, 3])
def test_sort_pairs_unaligned(gpu, oracle, shift):
    """Keys that start 4, 8 or 12 bytes past a 16-byte boundary (a slice of a larger buffer):
    the histogram's 16-byte loads are skipped for them, the order is the same."""
    n = 3 * 4096 + 77
    rng = np.random.default_rng(shift)
    k = rng.integers(0, 1 << 20, n, dtype=np.uint64).astype(np.uint32)
    v = np.arange(n, dtype=np.uint32)
    big_k, big_v = _u32(np.zeros(n + 4, np.uint32)), _u32(np.zeros(n + 4, np.uint32))
    big_k[shift:shift + n] = _u32(k)
    big_v[shift:shift + n] = _u32(v)
    tk, tv = big_k[shift:shift + n], big_v[shift:shift + n]
    assert tk.data_ptr() % 16 == 4 * shift % 16
    t1, t2 = torch.empty_like(tk), torch.empty_like(tv)
    call("gpgmp_sort_pairs_device", tk.data_ptr(), tv.data_ptr(), t1.data_ptr(), t2.data_ptr(), n, 20, None)
    torch.cuda.synchronize()
    ks, vs = oracle.stable_sort_pairs(k, v)
    assert np.array_equal(_np_u32(tk), ks)
    assert np.array_equal(_np_u32(tv), vs)
    assert np.array_equal(_np_u32(big_k[:shift]), np.zeros(shift, np.uint32))  # nothing written before the slice


def test_offsets_left_right(gpu, oracle):
    rng = np.random.default_rng(3)
    ncells = 1 << 14
    k = np.sort(rng.integers(0, ncells, 30000).astype(np.uint32))
    k = k[(k < 100) | (k > 200)]  # a gap of empty cells
    tk = _u32(k)
    off = torch.empty(ncells + 1, dtype=torch.int32, device="cuda")
    left = torch.empty(ncells, dtype=torch.int32, device="cuda")
    right = torch.empty(ncells, dtype=torch.int32, device="cuda")
    call("gpgmp_cell_offsets_device", tk.data_ptr(), len(k), ncells, off.data_ptr(), None)
    call("gpgmp_cell_left_right_device", tk.data_ptr(), len(k), ncells, left.data_ptr(), right.data_ptr(), None)
    torch.cuda.synchronize()
    assert np.array_equal(_np_u32(off), oracle.cell_offsets(k, ncells))
    lw, rw = oracle.cell_left_right(k, ncells)
    assert np.array_equal(_np_u32(left), lw) and np.array_equal(_np_u32(right), rw)


@pytest.mark.parametrize("side,n,steps,p", [(64, 30000, 5, 0.05), (128, 200000, 3, 0.0), (1024, 1000000, 2, 0.05)])
def test_particle_steps(gpu, oracle, side, n, steps, p):
    rng = np.random.default_rng(side)
    x = rng.uniform(0, side, n).astype(np.float32)
    y = rng.uniform(0, side, n).astype(np.float32)
    ids = np.arange(n, dtype=np.uint32)
    sysm = particles.ParticleSystem(side, D=1.0, dt=0.01, vortex_amp=0.1, p_annihilate=p, seed=99, capacity=n)
    sysm.set(x, y, ids)
    sysm.step(0, steps)
    gx, gy, gi, gk, goff = sysm.get(offsets=True)
    po = oracle.particle_params(side, 1.0, 0.01, 0.1, p, 99)
    ox, oy, oi, ok, ooff = oracle.particle_steps(po, 0, steps, x, y, ids)
    assert len(gx) == len(ox)
    assert np.array_equal(gi, oi)
    assert np.array_equal(gk, ok)
    assert np.array_equal(gx, ox) and np.array_equal(gy, oy)
    assert np.array_equal(goff, ooff)


@pytest.mark.parametrize("mode", [0, 2])
@pytest.mark.parametrize("side,n,D,dt,amp,p", [(64, 40000, 1.0, 0.01, 0.1, 0.05), (3, 500, 0.05, 0.1, 0.3, 0.3),
                                                (16, 20000, 1.0, 0.2, 2.0, 0.1), (32, 60000, 0.02, 0.01, 0.0, 0.5),
                                                (128, 300, 1.0, 0.01, 0.1, 0.3), (96, 400, 1.0, 0.2, 2.0, 0.2)])
def test_sorted_list_step(gpu, oracle, mode, side, n, D, dt, amp, p):
    """The step on the sorted list (counts of the moves to the 3x3 neighbours and one scatter
    to the stable-order place, far moves on the far list) and the radix step give the
    oracle's particles, keys and offsets: ordinary moves, the smallest torus (side 3, every
    neighbour wraps), moves of several cells (sigma 0.63 plus drift: many far moves), crowded
    cells (pairs beyond the 32 of the kill mask), sparse lists where most cells and most
    64-cell groups are empty, with and without far moves."""
    rng = np.random.default_rng(side * 7 + n)
    if n > 50000:  # crowded: every particle in 4 cells
        x = (rng.integers(0, 2, n) + rng.uniform(0, 1, n)).astype(np.float32)
        y = (rng.integers(0, 2, n) + rng.uniform(0, 1, n)).astype(np.float32)
    else:
        x = rng.uniform(0, side, n).astype(np.float32)
        y = rng.uniform(0, side, n).astype(np.float32)
    ids = np.arange(n, dtype=np.uint32)
    sysm = particles.ParticleSystem(side, D=D, dt=dt, vortex_amp=amp, p_annihilate=p, seed=5, capacity=n)
    sysm.setSortedStep(mode)
    sysm.set(x, y, ids)
    sysm.step(0, 4)
    gx, gy, gi, gk, goff = sysm.get(offsets=True)
    po = oracle.particle_params(side, D, dt, amp, p, 5)
    ox, oy, oi, ok, ooff = oracle.particle_steps(po, 0, 4, x, y, ids)
    assert np.array_equal(gi, oi) and np.array_equal(gk, ok)
    assert np.array_equal(gx, ox) and np.array_equal(gy, oy)
    assert np.array_equal(goff, ooff)
