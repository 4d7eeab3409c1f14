"""Differentiable weighted projection for PyTorch.

``project(Z, w, design)`` returns the demeaned output ``Z - D x`` with
``x = argmin ||W^{1/2} (Z - D x)||`` (one column per right-hand side), computed on the GPU by
the library's batched LSMR, and is differentiable in ``Z`` and in the weights ``w``.

The backward pass is hand-written: for an upstream gradient ``G``,

    lambda = (D'WD)^+ D'G            (one adjoint solve with the forward's preconditioner)
    grad_Z = G - W (D lambda)
    grad_w = -sum_j R_j * (D lambda)_j     (R = the forward's output)

where the adjoint solve is the library's solve of ``G / w`` (its normal equations are exactly
``D'WD lambda = D'G`` for positive weights) and one fused gather kernel forms both gradients
(``wx_solver_projection_grad_device``).

``design`` may be a Solver of the same design: the forward then reweights it
(``Solver.reweight``: the device design is copied on the device rather than rebuilt from the
codes), which is what a training loop that re-solves one design under changing weights wants;
with a prebuilt Preconditioner as well, nothing is built on the host at all.  PyTorch is used only here, at the binding: tensors
are float64 CUDA tensors shaped ``(n, k)`` (or ``(n,)``), passed to the C ABI by pointer.
"""
from __future__ import annotations

import ctypes as C
from typing import Optional

import numpy as np
import torch

from . import _lib
from ._within import LsmrOptions, Solver

__all__ = ["WeightedProjection", "project"]


def _colmajor(t: torch.Tensor) -> torch.Tensor:
    """(n, k) row-major tensor -> (k, n) contiguous, i.e. column-major (n, k) memory."""
    return t.t().contiguous()


class WeightedProjection(torch.autograd.Function):
    """R = Z - D x(Z, w).  Forward builds a Solver for the weights (from the design, or by
    reweighting a given Solver) and runs the batched device solve; backward runs the adjoint
    solve and the fused gradient kernel with the same Solver."""

    @staticmethod
    def forward(ctx, Z: torch.Tensor, w: torch.Tensor, design, options: Optional[LsmrOptions], preconditioner):
        if Z.dtype != torch.float64 or w.dtype != torch.float64:
            raise TypeError("Z and w must be float64 tensors")
        if not Z.is_cuda or not w.is_cuda:
            raise ValueError("Z and w must be CUDA (HIP) tensors")
        one = Z.dim() == 1
        Z2 = Z.reshape(Z.shape[0], -1)
        n, k = Z2.shape
        if w.shape != (n,):
            raise ValueError(f"w must have shape ({n},), got {tuple(w.shape)}")
        opts = options or LsmrOptions()
        wh = w.detach().cpu().numpy()
        if isinstance(design, Solver):
            solver = design.reweight(wh, preconditioner)
        else:
            solver = Solver(design, weights=wh, preconditioner=preconditioner)
        L = _lib.lib()
        Zc = _colmajor(Z2.detach())
        Rc = torch.empty_like(Zc)
        stats = (_lib.RhsStats * max(1, k))()
        stream = torch.cuda.current_stream().cuda_stream
        _lib.check(L.wx_solver_solve_batch_device(solver._h, Zc.data_ptr(), n, k, C.byref(opts._native()), None,
                                                  Rc.data_ptr(), stats, C.c_void_p(stream)))
        ctx.solver, ctx.opts, ctx.one = solver, opts, one
        ctx.stats = [stats[j] for j in range(k)]
        ctx.save_for_backward(Rc)
        R = Rc.t()
        return R.reshape(-1) if one else R

    @staticmethod
    def backward(ctx, G: torch.Tensor):
        (Rc,) = ctx.saved_tensors
        k, n = Rc.shape
        Gc = _colmajor(G.reshape(n, k).to(torch.float64))
        gz = torch.empty_like(Gc)
        gw = torch.empty(n, dtype=torch.float64, device=Gc.device) if ctx.needs_input_grad[1] else None
        stats = (_lib.RhsStats * max(1, k))()
        stream = torch.cuda.current_stream().cuda_stream
        _lib.check(_lib.lib().wx_solver_projection_grad_device(
            ctx.solver._h, Gc.data_ptr(), n, Rc.data_ptr(), n, k, C.byref(ctx.opts._native()), gz.data_ptr(),
            None if gw is None else gw.data_ptr(), stats, C.c_void_p(stream)))
        grad_z = gz.t()
        if ctx.one:
            grad_z = grad_z.reshape(-1)
        return (grad_z if ctx.needs_input_grad[0] else None), gw, None, None, None


def project(Z: torch.Tensor, w: torch.Tensor, design, options: Optional[LsmrOptions] = None,
            preconditioner=None) -> torch.Tensor:
    """Differentiable demeaned output ``Z - D x`` of the weighted fixed-effects projection.
    ``design`` is a design (codes or effects) or a Solver of that design to reweight."""
    return WeightedProjection.apply(Z, w, design, options, preconditioner)
