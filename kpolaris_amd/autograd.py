"""torch.autograd over the imaging kernel's plasma-parameter derivatives.

``riaf_stokes(log_scales, params, model)`` returns the Stokes image
[4, npix] (I, Q, U, V in the observer basis, invariant units) of the analytic
RIAF at plasma log-scales ``log_scales = (q_density, q_temperature,
q_magnetic)`` (the PlasmaPerturbation families of the reference,
src/radiation/plasma_perturbation.hpp:10-33), differentiable with respect to
``log_scales``.

The forward pass runs kp_run_riaf_image_gradient, which carries the three
tangent directions through the transfer along the image's geodesics, so the
full Jacobian d S / d q ([3, 4, npix]) is known when forward returns; the
backward pass is the contraction of the incoming gradient with it (no second
sweep, no stored ray states).
"""
from __future__ import annotations

import torch

from kpolaris_amd import image


class RiafStokes(torch.autograd.Function):
    @staticmethod
    def forward(ctx, log_scales: torch.Tensor, params, model):
        q = [float(v) for v in log_scales.detach().cpu().reshape(3)]
        out, grad, _ = image.run_riaf_image_gradient(params, model, q, with_stats=False)
        stokes = torch.stack([out["image_i"], out["image_q"], out["image_u"], out["image_v"]])
        ctx.save_for_backward(grad)
        ctx.q_device = log_scales.device
        ctx.q_dtype = log_scales.dtype
        return stokes

    @staticmethod
    def backward(ctx, grad_stokes: torch.Tensor):
        (jac,) = ctx.saved_tensors
        g = (jac * grad_stokes.to(jac.dtype).unsqueeze(0)).sum(dim=(1, 2))
        return g.to(device=ctx.q_device, dtype=ctx.q_dtype), None, None


def riaf_stokes(log_scales: torch.Tensor, params, model) -> torch.Tensor:
    """Differentiable Stokes image [4, npix] at plasma log-scales (3,)."""
    return RiafStokes.apply(log_scales, params, model)


class SlowLightStokes(torch.autograd.Function):
    """Slow-light Stokes image [4, npix] at (q_density, R_high, q_magnetic).

    R_high enters through the snapshots' derived scalars: when it differs from
    the grids' current value the derived scalars of every snapshot are
    recomputed on the device (kp_grmhd_derive_scalars) before the image."""

    @staticmethod
    def forward(ctx, theta: torch.Tensor, params, grids, times, observation_time, step_mode, interpolation):
        q_density, r_high, q_magnetic = [float(v) for v in theta.detach().cpu().reshape(3)]
        for g in grids:
            if g.model.trat_large != r_high:
                g.model.trat_large = r_high
                g.derive()
        out, grad, _ = image.run_grmhd_slow_light_image_gradient(
            params, grids, times, observation_time, (q_density, q_magnetic), step_mode=step_mode,
            interpolation=interpolation, with_stats=False)
        stokes = torch.stack([out["image_i"], out["image_q"], out["image_u"], out["image_v"]])
        ctx.save_for_backward(grad)
        ctx.q_device = theta.device
        ctx.q_dtype = theta.dtype
        return stokes

    @staticmethod
    def backward(ctx, grad_stokes: torch.Tensor):
        (jac,) = ctx.saved_tensors
        g = (jac * grad_stokes.to(jac.dtype).unsqueeze(0)).sum(dim=(1, 2))
        return g.to(device=ctx.q_device, dtype=ctx.q_dtype), None, None, None, None, None, None


def slow_light_stokes(theta: torch.Tensor, params, grids, times, observation_time: float,
                      step_mode: int = 0, interpolation: int = 1) -> torch.Tensor:
    """Differentiable slow-light Stokes image [4, npix] at
    theta = (q_density, R_high, q_magnetic) (3,)."""
    return SlowLightStokes.apply(theta, params, grids, times, observation_time, step_mode, interpolation)


class GrmhdStokes(torch.autograd.Function):
    """Static GRMHD Stokes image [4, npix] at (q_density, R_high, q_magnetic);
    R_high re-derives the snapshot's scalars when it changes."""

    @staticmethod
    def forward(ctx, theta: torch.Tensor, params, grid):
        q_density, r_high, q_magnetic = [float(v) for v in theta.detach().cpu().reshape(3)]
        if grid.model.trat_large != r_high:
            grid.model.trat_large = r_high
            grid.derive()
        out, grad, _ = image.run_grmhd_image_gradient(params, grid, (q_density, q_magnetic), with_stats=False)
        stokes = torch.stack([out["image_i"], out["image_q"], out["image_u"], out["image_v"]])
        ctx.save_for_backward(grad)
        ctx.q_device = theta.device
        ctx.q_dtype = theta.dtype
        return stokes

    @staticmethod
    def backward(ctx, grad_stokes: torch.Tensor):
        (jac,) = ctx.saved_tensors
        g = (jac * grad_stokes.to(jac.dtype).unsqueeze(0)).sum(dim=(1, 2))
        return g.to(device=ctx.q_device, dtype=ctx.q_dtype), None, None


def grmhd_stokes(theta: torch.Tensor, params, grid) -> torch.Tensor:
    """Differentiable static GRMHD Stokes image [4, npix] at
    theta = (q_density, R_high, q_magnetic) (3,)."""
    return GrmhdStokes.apply(theta, params, grid)
