"""Differentiable pose energies: ``torch.autograd.Function`` over the pose kernels.

    E = pose_energy(plan, centers, axes, precision="mixed")   # (n,) float64
    E.sum().backward()                                        # centers.grad, axes.grad

``centers`` are Cartesian (Angstrom), ``axes`` the directors, both (n, 3)
float64 CUDA tensors; directors are used as given (normalize them in torch if
they are free parameters, and autograd carries the normalization).  Forward
runs the energy kernel; backward runs a separate kernel that recomputes every
pair term and returns dE/dc and dE/du scaled by the upstream gradient.  There
is no counterpart in the reference, whose String forces are forward
differences of the energy (native/string/source_code/global.h:1281-1324).
"""
from __future__ import annotations

import torch

from .field import Plan, PoseList


class PoseEnergy(torch.autograd.Function):
    @staticmethod
    def forward(ctx, centers, axes, plan: Plan, precision: str):
        if centers.shape != axes.shape or centers.ndim != 2 or centers.shape[1] != 3:
            raise ValueError("centers and axes must both be (n, 3)")
        device = centers.device if centers.is_cuda else torch.device("cuda")
        poses = PoseList(plan, centers.detach().cpu().double().numpy(),
                         axes.detach().cpu().double().numpy())
        energy = torch.empty(len(centers), dtype=torch.float64, device=device)
        torch.cuda.current_stream(device).synchronize()
        poses.eval(precision, energy=energy.data_ptr())
        ctx.poses = poses
        ctx.precision = precision
        ctx.device = device
        ctx.dtypes = (centers.dtype, axes.dtype)
        return energy

    @staticmethod
    def backward(ctx, grad_energy):
        n = ctx.poses.n
        g = grad_energy.detach().to(device=ctx.device, dtype=torch.float64).contiguous()
        grad_c = torch.empty((n, 3), dtype=torch.float64, device=ctx.device)
        grad_u = torch.empty((n, 3), dtype=torch.float64, device=ctx.device)
        torch.cuda.current_stream(ctx.device).synchronize()
        ctx.poses.eval(ctx.precision, grad_out=g.data_ptr(), grad_center=grad_c.data_ptr(),
                       grad_axis=grad_u.data_ptr())
        return grad_c.to(ctx.dtypes[0]), grad_u.to(ctx.dtypes[1]), None, None


def pose_energy(plan: Plan, centers: torch.Tensor, axes: torch.Tensor,
                precision: str = "fp64") -> torch.Tensor:
    """Differentiable energies of independent poses (center_i, director_i)."""
    return PoseEnergy.apply(centers, axes, plan, precision)
