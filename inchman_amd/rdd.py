"""Differentiable Schnakenberg reaction-diffusion on the GPU (torch.autograd.Function).

Forward: `steps` fused explicit Euler steps (reaction + mirror-BC diffusion with the
per-cell diffusivity field D scaled by D0_u, D0_v) in fp32, storing the state every
`ckpt_every` steps.  Backward: the hand-written adjoint kernel walks the steps in
reverse, recomputing each checkpoint segment, and returns dL/dD, dL/dk1..k4 and
dL/du0, dL/dv0.  Everything runs in libgpgmp_hip.so on torch's current stream.
"""
import ctypes as C

import torch

from . import _lib
from ._lib import call, lib


def _stream_handle():
    return int(torch.cuda.current_stream().cuda_stream)


class ReactionDiffusion:
    """Handle of one problem size: grid, step count, checkpoint interval, dt, D0."""

    def __init__(self, width, height, steps, dt, h=1.0, D0=(0.05, 1.0), k=(1.0, 3.0, 1e-4, 1.0),
                 ckpt_every=100):
        self.width, self.height, self.steps = int(width), int(height), int(steps)
        p = _lib.RddParams()
        p.width, p.height = self.width, self.height
        p.h, p.dt = float(h), float(dt)
        p.D0u, p.D0v = float(D0[0]), float(D0[1])
        for i in range(4):
            p.k[i] = float(k[i])
        self.params = p
        hnd = C.c_void_p()
        call("gpgmp_rdd_create", C.byref(p), self.steps, int(ckpt_every), _stream_handle(), C.byref(hnd))
        self.handle = hnd

    def close(self):
        if getattr(self, "handle", None):
            lib().gpgmp_rdd_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, *ts):
        for t in ts:
            assert t.is_cuda and t.dtype == torch.float32 and t.is_contiguous(), "fp32 contiguous CUDA tensors"
            assert t.numel() == self.width * self.height

    def set_rates(self, k):
        kk = (C.c_float * 4)(*[float(x) for x in k])
        call("gpgmp_rdd_set_rates", self.handle, kk)

    def set_rates_device(self, k):
        """Rates k1..k4 from a 4-element fp32 CUDA tensor, copied on the stream (no host wait)."""
        assert k.is_cuda and k.dtype == torch.float32 and k.is_contiguous() and k.numel() == 4
        call("gpgmp_rdd_set_rates_device", self.handle, k.data_ptr())

    def generation(self):
        """Generation of the saved forward run (0: none)."""
        g = C.c_uint64()
        call("gpgmp_rdd_generation", self.handle, C.byref(g))
        return g.value

    def set_steps_per_launch(self, ns):
        """1 or 2 Euler steps per kernel launch (results are bit-identical)."""
        call("gpgmp_rdd_set_steps_per_launch", self.handle, int(ns))

    def forward(self, D, u, v, save=True):
        """Advances u, v (in place) by `steps` steps."""
        self._check(D, u, v)
        call("gpgmp_rdd_forward", self.handle, D.data_ptr(), u.data_ptr(), v.data_ptr(), int(save))
        return u, v

    def backward(self, D, lu, lv, generation=0):
        """Adjoint of the last saved forward; lu, lv (in place): dL/du_T -> dL/du_0.  A
        non-zero generation must be that forward's (else GpgmpError: another forward ran)."""
        self._check(D, lu, lv)
        gD = torch.empty_like(D)
        gk = torch.empty(4, dtype=torch.float32, device=D.device)
        call("gpgmp_rdd_backward", self.handle, int(generation), D.data_ptr(), lu.data_ptr(), lv.data_ptr(),
             gD.data_ptr(), gk.data_ptr())
        return gD, gk, lu, lv


class RDDFunction(torch.autograd.Function):
    """(D, k, u0, v0) -> (u_T, v_T); k is a 4-element tensor of rate constants.

    The rates go to the solver as a device tensor (no host synchronisation).  The handle
    keeps the checkpoints of one forward only: backward checks that the handle's saved
    forward is still this one (another forward on the same handle before this backward
    raises instead of differentiating the wrong run)."""

    @staticmethod
    def forward(ctx, D, k, u0, v0, solver):
        kd = k.detach().to(device=D.device, dtype=torch.float32).contiguous()
        solver.set_rates_device(kd)
        Dc = D.detach().contiguous()
        u = u0.detach().clone().contiguous()
        v = v0.detach().clone().contiguous()
        solver.forward(Dc, u, v, save=True)
        ctx.solver = solver
        ctx.generation = solver.generation()
        ctx.save_for_backward(Dc, k, kd)
        return u, v

    @staticmethod
    def backward(ctx, gu, gv):
        Dc, k, _ = ctx.saved_tensors
        lu = (gu if gu is not None else torch.zeros_like(Dc)).detach().clone().contiguous()
        lv = (gv if gv is not None else torch.zeros_like(Dc)).detach().clone().contiguous()
        gD, gk, gu0, gv0 = ctx.solver.backward(Dc, lu, lv, generation=ctx.generation)
        return gD, gk.to(k.device, k.dtype), gu0, gv0, None


def simulate(solver, D, k, u0, v0):
    return RDDFunction.apply(D, k, u0, v0, solver)
