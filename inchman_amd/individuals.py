"""Particle-based (individual) species on the lattice: the reference's
gpgmp::IndividualSolver as driven by InhomogeneousSolver::diffuse
(IndividualSolver.h:36-297, InhomogeneousSolver.cpp:254-427), over libgpgmp_hip.so.

Each particle is a cell key plus four properties (DiffusivityX, DiffusivityY, DriftX,
DriftY).  A sweep hops every particle along one axis with the inhomogeneous
drift-diffusion probabilities of its own properties; the constants update re-sorts the
particles by cell and recomputes the per-cell maxima, counts and the next time step.
After a reaction step, maintainPopulation makes every cell hold the species' new count
(IndividualSolver::maintainPopulation, IndividualSolver.cpp:458-671): it kills randomly
picked particles and creates new ones from the pool of dead ids.
"""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import call, lib


def _u32(a):
    return np.ascontiguousarray(a, np.uint32)


class IndividualSpecies:
    """Method names follow IndividualSolver / InhomogeneousSolver."""

    def __init__(self, width, height, physical_length, capacity, boundary_mask=(0, 0, 0, 0), species=0,
                 stream=None):
        d = _lib.IndividualsDesc()
        d.width, d.height, d.physical_length, d.species = int(width), int(height), float(physical_length), int(species)
        for i in range(4):
            d.boundary_mask[i] = int(boundary_mask[i])
        self.width, self.height, self.capacity = int(width), int(height), int(capacity)
        h = C.c_void_p()
        call("gpgmp_individuals_create", C.byref(d), self.capacity, stream, C.byref(h))
        self.handle = h
        self.n = 0

    def close(self):
        if getattr(self, "handle", None):
            lib().gpgmp_individuals_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def setParticles(self, cells, props):
        """cells[n] (cell index x + y*W), props[4][n]; sorts and computes the first steps."""
        cells = _u32(cells)
        props = np.ascontiguousarray(props, np.float32)
        assert props.shape == (4, cells.size)
        call("gpgmp_individuals_set", self.handle, cells.ctypes.data, props.ctypes.data, cells.size)
        self.n = cells.size

    def timeSteps(self):
        x, y = C.c_float(), C.c_float()
        call("gpgmp_individuals_time_steps", self.handle, C.byref(x), C.byref(y))
        return x.value, y.value

    def _diffuse(self, axis, dt, seed):
        s = C.c_int32(seed)
        call("gpgmp_individuals_diffuse", self.handle, axis, float(dt), C.byref(s))
        return s.value

    def diffuseX(self, seed, dt):
        """One X sweep; returns seed + 1 (the caller's seed++)."""
        return self._diffuse(0, dt, seed)

    def diffuseY(self, seed, dt):
        return self._diffuse(1, dt, seed)

    def computeDiffusionConstants(self, axes=3):
        """Sort, cell indices, per-cell constants and the time steps of `axes` (1 X, 2 Y, 3 both)."""
        x, y = C.c_float(), C.c_float()
        call("gpgmp_individuals_update", self.handle, int(axes), C.byref(x), C.byref(y))
        return x.value, y.value

    def run(self, steps, seed0):
        t = C.c_float()
        call("gpgmp_individuals_run", self.handle, int(steps), C.c_uint32(seed0), C.byref(t))
        return t.value

    def sweeps(self):
        """Hop sweeps done since the particles were set."""
        v = C.c_int64()
        call("gpgmp_individuals_sweeps", self.handle, C.byref(v))
        return v.value

    def _refresh_n(self):
        npool = C.c_int32()
        call("gpgmp_individuals_get_pool", self.handle, None, None, C.byref(npool))
        self.n = self.capacity - npool.value

    def getPositionList(self):
        """(sorted cell keys, particle indices, per-cell counts)."""
        self._refresh_n()
        keys = np.zeros(max(self.n, 1), np.uint32)
        vals = np.zeros(max(self.n, 1), np.uint32)
        cnt = np.zeros(self.width * self.height, np.int32)
        n = C.c_int32()
        call("gpgmp_individuals_get", self.handle, keys.ctypes.data, vals.ctypes.data, cnt.ctypes.data, C.byref(n))
        return keys[: n.value], vals[: n.value], cnt.reshape(self.height, self.width)

    def constants(self):
        out = [np.zeros(self.width * self.height, np.float32) for _ in range(4)]
        call("gpgmp_individuals_constants", self.handle, *[o.ctypes.data for o in out])
        return out

    def maintainPopulation(self, goal, seed, rule=None):
        """IndividualSolver::maintainPopulation(species, seed, d_state): afterwards cell c holds
        goal[c] particles.  goal: host int32 [H][W] / [W*H], or a device pointer (int) / torch
        tensor on the GPU (e.g. the species' slice of the solver state).  rule: None (keep the
        recycled particle's properties) or ((mode0..3), (a0..3), (b0..3)) per property with mode
        0 keep, 1 constant a, 2 a + b * RANDOM."""
        r = None
        if rule is not None:
            r = _lib.NewIndividualRule()
            modes, a, b = rule
            for k in range(4):
                r.mode[k], r.a[k], r.b[k] = int(modes[k]), float(a[k]), float(b[k])
        if isinstance(goal, int):
            ptr, on_dev = goal, 1
        elif hasattr(goal, "data_ptr"):
            assert goal.is_cuda and goal.is_contiguous() and goal.numel() == self.width * self.height
            ptr, on_dev = goal.data_ptr(), 1
        else:
            g = np.ascontiguousarray(goal, np.int32).reshape(-1)
            assert g.size == self.width * self.height
            ptr, on_dev, keep = g.ctypes.data, 0, g  # noqa: F841 (keeps the buffer alive)
        call("gpgmp_individuals_maintain", self.handle, ptr, on_dev, int(seed), C.byref(r) if r is not None else None)
        npool = C.c_int32()
        call("gpgmp_individuals_get_pool", self.handle, None, None, C.byref(npool))
        self.n = self.capacity - npool.value

    def getPropertiesList(self):
        """Properties [4][n] of the live particles in sorted order."""
        self._refresh_n()
        out = np.zeros((4, max(self.n, 1)), np.float32)
        call("gpgmp_individuals_get_props", self.handle, out.ctypes.data)
        return out[:, : self.n] if self.n else np.zeros((4, 0), np.float32)

    def pool(self):
        """(ids, props[4][np]) of the dead-particle pool, front first."""
        npool = C.c_int32()
        call("gpgmp_individuals_get_pool", self.handle, None, None, C.byref(npool))
        ids = np.zeros(max(npool.value, 1), np.uint32)
        props = np.zeros((4, max(npool.value, 1)), np.float32)
        call("gpgmp_individuals_get_pool", self.handle, ids.ctypes.data, props.ctypes.data, C.byref(npool))
        m = npool.value
        return ids[:m], props.reshape(-1)[: 4 * m].reshape(4, m) if m else np.zeros((4, 0), np.float32)
