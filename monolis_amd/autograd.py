"""torch.autograd harness over the C-ABI: x = A(values)^{-1} b on the GPU, with
the backward pass done by the library's adjoint (monolis_hip_adjoint_R):

    A^T lambda = dL/dx,   dL/db = lambda,   dL/dvalues = -lambda x^T on the pattern.

torch provides the autograd graph; every solve and the gradient kernel run in
libmonolis_hip on its own device buffers.  The tensors crossing the boundary
are float64 host tensors: ``values`` in monolis's A layout (nz * n_dof^2,
row-major blocks), ``b`` and the returned ``x`` of length N * n_dof.  (The torch
wheel ships its own HIP runtime, separate from the system ROCm runtime the
library links, and device pointers do not cross between two runtimes in one
process, so the harness stages through host memory.)  The matrix handle takes
``values`` on every forward call.
"""
from __future__ import annotations

import numpy as np
import torch

from . import _lib
from .api import DeviceBuffer, DeviceMatrix, Prm


class _Buffers:
    """Device buffers of one handle, reused across calls."""

    def __init__(self, dm: DeviceMatrix, nvals: int):
        self.values = DeviceBuffer(nvals)
        self.b = DeviceBuffer(dm.vec)
        self.x = DeviceBuffer(dm.vec)
        self.g = DeviceBuffer(dm.vec)
        self.lam = DeviceBuffer(dm.vec)
        self.dA = DeviceBuffer(nvals)


def _buffers(dm: DeviceMatrix, nvals: int) -> _Buffers:
    bufs = getattr(dm, "_autograd_buffers", None)
    if bufs is None or bufs.values.count != nvals:
        bufs = _Buffers(dm, nvals)
        dm._autograd_buffers = bufs
    return bufs


def _host(t: torch.Tensor, n: int = None) -> np.ndarray:
    a = t.detach().to("cpu", torch.float64).contiguous().numpy()
    if n is not None and a.size < n:
        a = np.concatenate([a, np.zeros(n - a.size)])
    return a


class SparseSolve(torch.autograd.Function):
    """x = A(values)^{-1} b; CG / BiCGSTAB and the preconditioner from ``prm``."""

    @staticmethod
    def forward(ctx, values: torch.Tensor, b: torch.Tensor, dm: DeviceMatrix, prm: Prm, symmetric: bool):
        bufs = _buffers(dm, values.numel())
        bufs.values.upload(_host(values))
        dm.set_values(bufs.values)
        bufs.b.upload(_host(b, dm.vec))
        dm.solve(bufs.b, bufs.x, prm)
        if prm.Iarray[5] != 0:
            raise _lib.MonolisHipError(f"forward solve did not converge (status {prm.Iarray[5]})")
        ctx.dm, ctx.prm, ctx.symmetric, ctx.nvals = dm, prm, bool(symmetric), values.numel()
        ctx.out_device = values.device
        return torch.from_numpy(bufs.x.numpy()[: dm.m]).to(values.device)

    @staticmethod
    def backward(ctx, gx: torch.Tensor):
        dm, prm = ctx.dm, ctx.prm
        bufs = _buffers(dm, ctx.nvals)
        bufs.g.upload(_host(gx, dm.vec))
        dm.adjoint(bufs.x, bufs.g, bufs.lam, bufs.dA, prm, symmetric=ctx.symmetric)
        if prm.Iarray[5] != 0:
            raise _lib.MonolisHipError(f"adjoint solve did not converge (status {prm.Iarray[5]})")
        dA = torch.from_numpy(bufs.dA.numpy()).to(ctx.out_device)
        lam = torch.from_numpy(bufs.lam.numpy()[: dm.m]).to(ctx.out_device)
        return dA, lam, None, None, None


def sparse_solve(values: torch.Tensor, b: torch.Tensor, dm: DeviceMatrix, prm: Prm,
                 symmetric: bool = False) -> torch.Tensor:
    """Differentiable x = A(values)^{-1} b (float64 host tensors in and out)."""
    return SparseSolve.apply(values, b, dm, prm, symmetric)
