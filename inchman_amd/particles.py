"""Individual-particle system over the C ABI: positions in a periodic square of 1 um
cells, Brownian motion plus a vortex drift, a stable radix sort by cell key every
step, CSR cell offsets, and in-cell pair annihilation A+A->0."""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import call, lib


class ParticleSystem:
    def __init__(self, side, D=1.0, dt=0.01, vortex_amp=0.1, p_annihilate=0.05, seed=1, capacity=1 << 20,
                 stream=None):
        p = _lib.ParticleParams()
        p.side, p.D, p.dt, p.vortex_amp, p.p_annihilate, p.seed = int(side), D, dt, vortex_amp, p_annihilate, seed
        self.side = int(side)
        h = C.c_void_p()
        st = None if stream is None else (stream if isinstance(stream, int) else int(stream.cuda_stream))
        call("gpgmp_particles_create", C.byref(p), int(capacity), st, C.byref(h))
        self.handle = h

    def close(self):
        if getattr(self, "handle", None):
            lib().gpgmp_particles_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set(self, x, y, ids):
        x = np.ascontiguousarray(x, np.float32)
        y = np.ascontiguousarray(y, np.float32)
        ids = np.ascontiguousarray(ids, np.uint32)
        call("gpgmp_particles_set", self.handle, x.ctypes.data_as(C.POINTER(C.c_float)),
             y.ctypes.data_as(C.POINTER(C.c_float)), ids.ctypes.data_as(C.POINTER(C.c_uint32)), len(x))

    def count(self):
        n = C.c_int32()
        call("gpgmp_particles_count", self.handle, C.byref(n))
        return n.value

    def step(self, step0, steps):
        call("gpgmp_particles_step", self.handle, int(step0), int(steps))

    def setSortedStep(self, mode=1):
        """0: radix step every time; 1: the step on the sorted list when moves stay within the
        neighbour cells; 2: that step always (results are identical)."""
        call("gpgmp_particles_set_sorted_step", self.handle, int(mode))

    def processed(self):
        """Particle-steps done so far (sum of the live count at each step)."""
        v = C.c_uint64()
        call("gpgmp_particles_processed", self.handle, C.byref(v))
        return v.value

    def get(self, offsets=False):
        n = self.count()
        x = np.empty(n, np.float32)
        y = np.empty(n, np.float32)
        ids = np.empty(n, np.uint32)
        keys = np.empty(n, np.uint32)
        off = np.empty(self.side * self.side + 1, np.uint32) if offsets else None
        nn = C.c_int32()
        P = lambda a, t: a.ctypes.data_as(C.POINTER(t)) if a is not None else None  # noqa: E731
        call("gpgmp_particles_get", self.handle, P(x, C.c_float), P(y, C.c_float), P(ids, C.c_uint32),
             P(keys, C.c_uint32), P(off, C.c_uint32), C.byref(nn))
        return x, y, ids, keys, off
