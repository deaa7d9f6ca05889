"""torch.autograd.Function wrappers of the transforms: a thin test binding used
for gradient checks.

    grid   = ZernikeBackward.apply(coeffs, plan, order)   # B c
    coeffs = ZernikeForward.apply(grid, plan, order)      # F g

Tensors are float64 on the GPU, batched along dim 0: coefficients (batch,
ncoeff, 2) in zest's {cos, sin} layout, grids (batch, *grid_shape). The
gradients are the hand-written transposes of the C ABI:

    d(B c)/dc  ->  B^T  (zgl_backward_adjoint)
    d(F g)/dg  ->  F^T  (zgl_forward_adjoint)

The same Functions serve spherical-harmonic plans (grids (batch, lon, lat) or
(batch, lat, lon), coefficients (batch, L(L+1)/2, 2)).
"""
from __future__ import annotations

import torch

from . import SH, Plan, st_num_coeffs, zt_num_coeffs


def _ncoeff(plan: Plan, order: int) -> int:
    return st_num_coeffs(order) if plan.kind == SH else zt_num_coeffs(order)


class ZernikeBackward(torch.autograd.Function):
    """grid = B(coeffs)."""

    @staticmethod
    def forward(ctx, coeffs: torch.Tensor, plan: Plan, order: int) -> torch.Tensor:
        coeffs = coeffs.contiguous()
        batch = coeffs.shape[0]
        grid = torch.empty((batch,) + tuple(plan.grid_shape(order)), dtype=torch.float64,
                           device=coeffs.device)
        plan.backward(coeffs, grid, order, batch=batch)
        ctx.plan, ctx.order, ctx.batch = plan, order, batch
        return grid

    @staticmethod
    def backward(ctx, grad_grid: torch.Tensor):
        plan, order, batch = ctx.plan, ctx.order, ctx.batch
        gc = torch.zeros((batch, _ncoeff(plan, order), 2), dtype=torch.float64,
                         device=grad_grid.device)
        plan.backward_adjoint(grad_grid.contiguous(), gc, order, batch=batch)
        return gc, None, None


class ZernikeForward(torch.autograd.Function):
    """coeffs = F(grid)."""

    @staticmethod
    def forward(ctx, grid: torch.Tensor, plan: Plan, order: int) -> torch.Tensor:
        grid = grid.contiguous()
        batch = grid.shape[0]
        coeffs = torch.zeros((batch, _ncoeff(plan, order), 2), dtype=torch.float64,
                             device=grid.device)
        plan.forward(grid, coeffs, order, batch=batch)
        ctx.plan, ctx.order, ctx.batch = plan, order, batch
        return coeffs

    @staticmethod
    def backward(ctx, grad_coeffs: torch.Tensor):
        plan, order, batch = ctx.plan, ctx.order, ctx.batch
        gg = torch.empty((batch,) + tuple(plan.grid_shape(order)), dtype=torch.float64,
                         device=grad_coeffs.device)
        plan.forward_adjoint(grad_coeffs.contiguous(), gg, order, batch=batch)
        return gg, None, None
