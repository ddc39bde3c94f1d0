"""The allocation-free radix sort entry point against the allocating one, through the C ABI:
identical keys and values on two calls in a row (the second smaller, so stale scratch from
the first call would show), an odd number of digit passes (result copied back), n = 0, and
the capacity limit refused with a status code."""
import ctypes as C

import numpy as np
import pytest
import torch

from inchman_amd._lib import E_CAPACITY, lib

pytestmark = pytest.mark.gpu


def _u32(a):
    return torch.from_numpy(np.ascontiguousarray(a, np.uint32).view(np.int32)).cuda()


def _sort(name, k, v, bits):
    tk, tv = _u32(k), _u32(v)
    t1, t2 = torch.empty_like(tk), torch.empty_like(tv)
    rc = getattr(lib(), name)(C.c_void_p(tk.data_ptr()), C.c_void_p(tv.data_ptr()), C.c_void_p(t1.data_ptr()),
                              C.c_void_p(t2.data_ptr()), len(k), bits, None)
    torch.cuda.synchronize()
    assert rc == 0, rc
    return tk.cpu().numpy().view(np.uint32), tv.cpu().numpy().view(np.uint32)


@pytest.mark.parametrize("n1,n2,bits", [(3_000_001, 70_001, 20), (1 << 20, 4097, 32), (250_000, 1, 12)])
def test_noalloc_sort_matches_allocating_sort(gpu, n1, n2, bits):
    rng = np.random.default_rng(n1)
    for n in (n1, n2):
        k = rng.integers(0, 1 << bits, n, dtype=np.uint64).astype(np.uint32)
        v = rng.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32)
        want = _sort("gpgmp_sort_pairs_device", k, v, bits)
        got = _sort("gpgmp_sort_pairs_device_noalloc", k, v, bits)
        assert np.array_equal(got[0], want[0]) and np.array_equal(got[1], want[1]), n
        assert np.all(np.diff(got[0].astype(np.int64)) >= 0)


def test_noalloc_sort_limits(gpu):
    L = lib()
    assert L.gpgmp_sort_pairs_device_noalloc(None, None, None, None, 0, 20, None) == 0
    assert L.gpgmp_sort_pairs_device_noalloc(None, None, None, None, (1 << 24) + 1, 20, None) == E_CAPACITY
