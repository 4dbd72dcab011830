"""Differentiable pose energies: ``torch.autograd.Function`` over the pose kernels.

    E = pose_energy(plan, centers, axes, precision="mixed")   # (n,) float64
    E.sum().backward()                                        # centers.grad, axes.grad

``centers`` are Cartesian (Angstrom), ``axes`` the directors, both (n, 3)
tensors, normally on the GPU (host tensors are copied there, and their
energies and gradients come back to the host); directors are used as given (normalize them in torch if they are
free parameters, and autograd carries the normalization).  The poses stay on
the device: each call bins them into the plan's cached tile grid on the GPU,
enqueued on torch's current stream without a host synchronization
(sv_pose_grid_eval_async); invalid poses give NaN energies, like other torch
operations, and ``check_poses(plan)`` raises on them.

Two ways to the gradient:
  * fused (the default when gradients are being recorded, i.e. grad mode is on
    and centers or axes require a gradient): forward runs one kernel that sums
    each site's energy and gradient together and keeps the pose gradients
    dE/dc, dE/du (6 doubles per pose); backward only scales them by the
    upstream gradient.  One pass over the pair terms instead of two; a forward
    whose backward never runs has paid about a third more than the energy alone;
  * two-pass (``fused=False``, and whenever nothing is recorded): forward runs
    the energy kernel; backward runs a separate kernel that recomputes every
    pair term and returns dE/dc and dE/du scaled by the upstream gradient.
    When no other evaluation has used the grid since this forward, backward
    reuses the forward's binning (sv_pose_grid_reeval_async).
Both give the same energies and gradients bit for bit.  The kernels give
first derivatives only: differentiating the gradient again (create_graph with
a loss nonlinear in E) raises rather than dropping the Hessian term.

There is no counterpart in the reference, whose String forces are forward
differences of the energy (native/string/source_code/global.h:1281-1324).
"""
from __future__ import annotations

import torch
from torch.autograd.function import once_differentiable

from .field import Plan, PoseGrid


def _grid_for(plan: Plan, n: int) -> PoseGrid:
    """The plan's pose grid for batches of about n poses (one per power of 2)."""
    grids = plan.__dict__.setdefault("_pose_grids", {})
    key = max(1, int(n)).bit_length()
    if key not in grids:
        grids[key] = PoseGrid(plan, 1 << (key - 1))
    return grids[key]


class PoseEnergy(torch.autograd.Function):
    @staticmethod
    def forward(ctx, centers, axes, plan: Plan, precision: str, fused: bool = False):
        if centers.shape != axes.shape or centers.ndim != 2 or centers.shape[1] != 3:
            raise ValueError("centers and axes must both be (n, 3)")
        if len(centers) == 0:
            raise ValueError("at least one pose required")
        device = centers.device if centers.is_cuda else torch.device("cuda")
        c = centers.detach().to(device=device, dtype=torch.float64).contiguous()
        u = axes.detach().to(device=device, dtype=torch.float64).contiguous()
        energy = torch.empty(len(c), dtype=torch.float64, device=device)
        grid = _grid_for(plan, len(c))
        stream = torch.cuda.current_stream(device).cuda_stream
        token = object()
        ctx.fused = bool(fused) and any(ctx.needs_input_grad[:2])
        if ctx.fused:
            # dE/dc and dE/du side by side, so that the backward scales both
            # with one multiplication
            grads = torch.empty((2, len(c), 3), dtype=torch.float64, device=device)
            grid.eval_async(stream, len(c), c.data_ptr(), u.data_ptr(), precision,
                            energy=energy.data_ptr(), grad_center=grads[0].data_ptr(),
                            grad_axis=grads[1].data_ptr(), token=token)
            ctx.save_for_backward(grads)
        else:
            grid.eval_async(stream, len(c), c.data_ptr(), u.data_ptr(), precision,
                            energy=energy.data_ptr(), token=token)
            ctx.save_for_backward(c, u)
        ctx.grid = grid
        ctx.token = token
        ctx.precision = precision
        ctx.dtypes = (centers.dtype, axes.dtype)
        ctx.devices = (centers.device, axes.device)
        # host inputs get host energies (and host gradients in the backward)
        return energy if centers.is_cuda else energy.to(centers.device)

    @staticmethod
    @once_differentiable   # the kernels give first derivatives only: a second raises
    def backward(ctx, grad_energy):
        if ctx.fused:
            (grads,) = ctx.saved_tensors
            g = grad_energy.detach().to(device=grads.device, dtype=torch.float64)
            scaled = grads * g[None, :, None]
            return (scaled[0].to(device=ctx.devices[0], dtype=ctx.dtypes[0]),
                    scaled[1].to(device=ctx.devices[1], dtype=ctx.dtypes[1]), None, None, None)
        c, u = ctx.saved_tensors
        n = len(c)
        g = grad_energy.detach().to(device=c.device, dtype=torch.float64).contiguous()
        grad_c = torch.empty((n, 3), dtype=torch.float64, device=c.device)
        grad_u = torch.empty((n, 3), dtype=torch.float64, device=c.device)
        stream = torch.cuda.current_stream(c.device).cuda_stream
        if ctx.grid.binned_token is ctx.token:
            # the grid still holds this forward's binned poses (saved tensors
            # cannot have changed: autograd rejects in-place edits of them)
            ctx.grid.reeval_async(stream, ctx.precision, grad_out=g.data_ptr(),
                                  grad_center=grad_c.data_ptr(), grad_axis=grad_u.data_ptr())
        else:
            ctx.grid.eval_async(stream, n, c.data_ptr(), u.data_ptr(), ctx.precision,
                                grad_out=g.data_ptr(), grad_center=grad_c.data_ptr(),
                                grad_axis=grad_u.data_ptr())
        return (grad_c.to(device=ctx.devices[0], dtype=ctx.dtypes[0]),
                grad_u.to(device=ctx.devices[1], dtype=ctx.dtypes[1]), None, None, None)


def check_poses(plan: Plan) -> None:
    """Raise if any pose evaluation since the last check had a nonfinite center,
    a nonfinite or zero director, or a nonfinite result (waits for the GPU)."""
    for grid in plan.__dict__.get("_pose_grids", {}).values():
        grid.check()


def pose_energy(plan: Plan, centers: torch.Tensor, axes: torch.Tensor,
                precision: str = "fp64", fused: bool | None = None) -> torch.Tensor:
    """Differentiable energies of independent poses (center_i, director_i).

    ``fused`` None (default) takes the fused route exactly when autograd is
    recording the call; True or False forces one route."""
    if fused is None:
        fused = torch.is_grad_enabled() and (centers.requires_grad or axes.requires_grad)
    return PoseEnergy.apply(centers, axes, plan, precision, fused)
