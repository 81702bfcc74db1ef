"""Discrete adjoint of the RK2 hydro step behind torch.autograd (C ABI fx_adj_*).

The differentiable scheme is PLM with the smooth van Albada slope and HLLE
fluxes (Davis speeds), RK2 with a fixed dt, on a periodic box; the reverse
sweep is hand-written (fornaxamr_amd/csrc/fx_adjoint.hip) and checkpoints every
``checkpoint_every`` steps, recomputing one segment at a time.

    solver = AdjointSolver((64, 64, 64), xlim, gamma)
    W0 = prim(rho, p)                      # any differentiable torch expression
    WT = evolve(W0, solver, nsteps=40, dt=dt)
    J = (WT[0][mask] ** 2).sum()
    J.backward()                           # d J / d rho, d J / d p

Tensors are float64 on the GPU, W[var][X][Y][Z] with var = rho, v1, v2, v3,
eint, Ye.  All work runs on torch's current stream.
"""
from __future__ import annotations

import ctypes as C

from . import _native


def _stream():
    import torch
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def _check_tensor(t, shape, name):
    import torch
    if not (t.is_cuda and t.dtype == torch.float64 and t.is_contiguous()):
        raise ValueError(f"{name} must be a contiguous float64 CUDA tensor")
    if tuple(t.shape) != shape:
        raise ValueError(f"{name} has shape {tuple(t.shape)}, expected {shape}")


class AdjointSolver:
    """Device state of the forward/reverse sweeps for one box size."""

    def __init__(self, n, xlim=(0.0, 1.0, 0.0, 1.0, 0.0, 1.0), gamma=5.0 / 3.0, va_eps2=None,
                 checkpoint_every=8):
        self.lib = _native.load()
        self.n = tuple(int(v) for v in n)
        self.xlim = tuple(float(v) for v in xlim)
        self.gamma = float(gamma)
        if va_eps2 is None:
            # eps^2 = dx^3 (smallest cell width): no limiting of smooth data
            # whose slopes are below dx^1.5
            va_eps2 = min((self.xlim[2 * d + 1] - self.xlim[2 * d]) / self.n[d] for d in range(3)) ** 3
        self.va_eps2 = float(va_eps2)
        self.checkpoint_every = int(checkpoint_every)
        h = C.c_void_p()
        _native.check(self.lib.fx_adj_create((C.c_int * 3)(*self.n), (C.c_double * 6)(*self.xlim),
                                             self.gamma, self.va_eps2, self.checkpoint_every,
                                             C.byref(h)))
        self.h = h
        self.generation = 0
        self.shape = (6,) + self.n

    def close(self):
        if getattr(self, "h", None):
            self.lib.fx_adj_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def forward(self, W0, nsteps, dt):
        """W_T after nsteps steps (keeps the checkpoints for backward)."""
        import torch
        _check_tensor(W0, self.shape, "W0")
        WT = torch.empty_like(W0)
        _native.check(self.lib.fx_adj_forward(self.h, C.c_void_p(W0.data_ptr()),
                                              C.c_void_p(WT.data_ptr()), int(nsteps), float(dt),
                                              _stream()))
        self.generation += 1
        return WT

    def backward(self, gWT):
        """Cotangent of W_0 from that of W_T, for the last forward."""
        import torch
        gWT = gWT.contiguous()
        _check_tensor(gWT, self.shape, "gWT")
        gW0 = torch.empty_like(gWT)
        _native.check(self.lib.fx_adj_backward(self.h, C.c_void_p(gWT.data_ptr()),
                                               C.c_void_p(gW0.data_ptr()), _stream()))
        return gW0


def _function():
    import torch

    class RK2Evolve(torch.autograd.Function):
        @staticmethod
        def forward(ctx, W0, solver, nsteps, dt):
            WT = solver.forward(W0.detach().contiguous(), nsteps, dt)
            ctx.solver = solver
            ctx.generation = solver.generation
            return WT

        @staticmethod
        def backward(ctx, gWT):
            solver = ctx.solver
            if solver.generation != ctx.generation:
                raise RuntimeError("the solver ran another forward after this one; its "
                                   "checkpoints are gone (use one solver per pending graph)")
            return solver.backward(gWT), None, None, None

    return RK2Evolve


_RK2Evolve = None


def evolve(W0, solver: AdjointSolver, nsteps: int, dt: float):
    """Differentiable nsteps RK2 steps of the scheme (torch.autograd.Function)."""
    global _RK2Evolve
    if _RK2Evolve is None:
        _RK2Evolve = _function()
    return _RK2Evolve.apply(W0, solver, int(nsteps), float(dt))
