"""Thin torch.autograd.Function wrapper of the RHS and its hand-written adjoint (abl_fv_rhs / abl_fv_rhs_vjp).

Used to check the adjoint against other gradients; the adjoint itself is the HIP kernel behind the C ABI.
``rhs(U, solver, theta=None, slot=0)``: U is the [n_comp, n_cells] float64 device state; when ``theta``
(a length-3 float64 tensor holding rho_in, |u_in|, T_in of inflow slot ``slot``) is given, the slot is set from it
before the forward evaluation and receives its gradient in the backward pass.
"""
from __future__ import annotations

import torch

from . import fv


class _RHS(torch.autograd.Function):
    @staticmethod
    def forward(ctx, U, theta, solver, slot, direction, Yi):
        if theta is not None:
            t = theta.detach().cpu().tolist()
            solver.set_boundary(slot, fv.InflowGhost([slot], t[0], [t[1] * d for d in direction], t[2], Yi))
        Uw = U.detach().clone()
        R = torch.zeros_like(Uw)
        solver.compute_rhs(0.0, Uw, R)
        ctx.solver, ctx.slot, ctx.has_theta = solver, slot, theta is not None
        ctx.save_for_backward(Uw)
        return R

    @staticmethod
    def backward(ctx, Rbar):
        (Uw,) = ctx.saved_tensors
        Ubar = torch.empty_like(Uw)
        theta_bar = ctx.solver.rhs_vjp(0.0, Uw.clone(), Rbar.contiguous(), Ubar)
        tb = None
        if ctx.has_theta:
            tb = torch.tensor(theta_bar[ctx.slot], dtype=torch.float64)
        return Ubar, tb, None, None, None, None


def rhs(U: torch.Tensor, solver: "fv.FiniteVolumeSolver", theta: torch.Tensor | None = None, slot: int = 0,
        direction=(1.0, 0.0, 0.0), Yi=None) -> torch.Tensor:
    return _RHS.apply(U, theta, solver, slot, list(direction)[: solver.dim], Yi)
