"""Adjoint of the BGK stream-collide map, exposed to PyTorch.

``Adjoint`` wraps plbm_adjoint_* (checkpointed forward, fused reverse kernel).
``VelocityMisfit`` is a torch.autograd.Function: loss = 1/2 sum |u(f^N) - u*|^2 after N
BGK steps from f^0 = feq(1, u0), differentiable with respect to the initial velocity
field (ux0, uy0) and the relaxation time tau (omega = 1/tau).

Fields are torch float64 CUDA tensors of shape (nx, ny), contiguous: element [x, y] at
x*ny + y, which is the engine's (ny, nx) Fortran-order layout.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _abi
from ._abi import check, dptr, lib


def band_limited_field(nx: int, ny: int, kmax: int = 8, amp: float = 0.02, seed: int = 1):
    """Divergence-free velocity from a streamfunction with modes 0 < |k| <= kmax
    (amplitude 1/|k|^2, random phases of `seed`), scaled to max |u| = amp.
    Returns (ux, uy) as (ny, nx) Fortran-ordered numpy arrays."""
    ux = np.zeros((ny, nx), order="F")
    uy = np.zeros((ny, nx), order="F")
    check(lib().plbm_field_band_limited(nx, ny, kmax, amp, seed, dptr(ux), dptr(uy)))
    return ux, uy


class Adjoint:
    def __init__(self, nx: int, ny: int, nsteps: int, checkpoint_every: int = 50,
                 storage: str | int = "f32s"):
        h = ctypes.c_void_p()
        check(lib().plbm_adjoint_create(nx, ny, _abi.STORAGE_IDS.get(storage, storage), nsteps,
                                        checkpoint_every, ctypes.byref(h)))
        self.h = h
        self.nx, self.ny, self.nsteps = nx, ny, nsteps
        self.generation = 0  # forward() calls so far; backward() differentiates the last

    def close(self):
        if getattr(self, "h", None) and self.h.value:
            check(lib().plbm_adjoint_destroy(self.h))
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, *ts):
        for t in ts:
            if t.dtype.itemsize != 8 or not t.is_cuda or not t.is_contiguous() \
                    or t.numel() != self.nx * self.ny:
                raise ValueError("fields must be contiguous float64 CUDA tensors of nx*ny elements")

    def forward(self, rho0, ux0, uy0, omega: float, uxt, uyt) -> float:
        """Loss of N steps from feq(rho0, u0); keeps the checkpoints for backward().
        The tensors must stay alive until backward()."""
        import torch
        self._check(rho0, ux0, uy0, uxt, uyt)
        torch.cuda.synchronize()
        loss = ctypes.c_double()
        check(lib().plbm_adjoint_forward_dev(self.h, rho0.data_ptr(), ux0.data_ptr(),
                                             uy0.data_ptr(), float(omega), uxt.data_ptr(),
                                             uyt.data_ptr(), ctypes.byref(loss)))
        self.generation += 1
        return loss.value

    def backward(self):
        """Returns (dL/drho0, dL/dux0, dL/duy0, dL/domega) of the last forward()."""
        import torch
        dev = torch.device("cuda")
        g = [torch.empty((self.nx, self.ny), dtype=torch.float64, device=dev) for _ in range(3)]
        # the engine writes the gradients on its own stream: blocks that the caching
        # allocator just handed out may still be read by queued work on torch's stream
        torch.cuda.current_stream().synchronize()
        gom = ctypes.c_double()
        check(lib().plbm_adjoint_backward_dev(self.h, g[0].data_ptr(), g[1].data_ptr(),
                                              g[2].data_ptr(), ctypes.byref(gom)))
        return g[0], g[1], g[2], gom.value


def _make_function():
    import torch

    class VelocityMisfit(torch.autograd.Function):
        """loss = VelocityMisfit.apply(ux0, uy0, tau, uxt, uyt, engine)"""

        @staticmethod
        def forward(ctx, ux0, uy0, tau, uxt, uyt, engine):
            rho0 = torch.ones_like(ux0)
            t = float(tau)
            ux0c, uy0c = ux0.detach().contiguous(), uy0.detach().contiguous()
            loss = engine.forward(rho0, ux0c, uy0c, 1.0 / t, uxt.contiguous(), uyt.contiguous())
            ctx.engine = engine
            ctx.generation = engine.generation
            ctx.tau = t
            ctx.keep = (rho0, ux0c, uy0c)  # the engine reads them again in backward
            return torch.tensor(loss, dtype=torch.float64, device=ux0.device)

        @staticmethod
        def backward(ctx, gl):
            if ctx.engine.generation != ctx.generation:
                raise RuntimeError(
                    "velocity_misfit: the Adjoint engine ran another forward since this loss "
                    "was computed; its checkpoints belong to the newer call (use one engine "
                    "per loss that is still to be differentiated)")
            _, gux, guy, gomega = ctx.engine.backward()
            gtau = -gomega / (ctx.tau * ctx.tau)
            g = gl.to(torch.float64)
            return (g * gux, g * guy, g * torch.tensor(gtau, dtype=torch.float64,
                                                        device=gux.device), None, None, None)

    return VelocityMisfit


_FN = None


def velocity_misfit(ux0, uy0, tau, uxt, uyt, engine: Adjoint):
    """Differentiable loss (see VelocityMisfit)."""
    global _FN
    if _FN is None:
        _FN = _make_function()
    return _FN.apply(ux0, uy0, tau, uxt, uyt, engine)
