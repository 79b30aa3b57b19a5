"""Differentiable segmentation: a ``torch.autograd.Function`` over the C ABI.

The forward pass runs ``iters`` iterations of the engine from phi0 with
parameters theta = (T, eps, alpha) and records them (``lsseg_fwd_record``);
the backward pass is the engine's hand-written adjoint (``lsseg_bwd``), which
returns dL/dphi0 and dL/dtheta.  The active set and the branch decisions of
every update are those of the forward run and are held fixed, so the gradient
is exact for the recorded map (``Record.replay`` evaluates that map at other
inputs, e.g. for finite differences).  The record keeps the values every
update overwrote, so the backward pass steps back from the final phi without
recomputation.

Tensors stay on the GPU: the engine reads phi0 and dL/dphi from, and writes
phi and dL/dphi0 to, torch's device memory directly.  torch is the plumbing
here; all arithmetic runs in the engine's HIP kernels.

torch brings its own HIP runtime.  Import torch (or this module) before the
engine library is first loaded, so that one runtime serves both.
"""
from __future__ import annotations

import sys

from . import engine as _engine

if _engine._lib is not None and "torch" not in sys.modules:
    # torch ships its own HIP runtime; loaded after the engine's, it cannot open
    # the device.  Imported first, torch's runtime is the one both use.
    raise ImportError("import torch (or this module) before the engine library is loaded: torch must provide "
                      "the process's HIP runtime")

import torch  # noqa: E402

from .engine import MODE_REF, Context  # noqa: E402


class LevelSetFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, phi0, theta, engine: Context, dt, iters, band, mode):
        if not phi0.is_cuda:
            raise ValueError("phi0 must be a CUDA (HIP) tensor")
        if tuple(phi0.shape) != (engine.nz, engine.ny, engine.nx):
            raise ValueError(f"phi0 shape {tuple(phi0.shape)} does not match the volume "
                             f"{(engine.nz, engine.ny, engine.nx)}")
        p0 = phi0.detach().to(torch.float32).contiguous()
        T, eps, alpha = (float(v) for v in theta.detach().to("cpu", torch.float64).tolist())
        torch.cuda.current_stream(p0.device).synchronize()
        engine.set_phi_ptr(p0.data_ptr(), band)
        rec = engine.fwd_record(T, eps, alpha, dt, iters, mode)
        out = torch.empty_like(p0)
        engine.phi_into_ptr(out.data_ptr(), 0)
        ctx.rec = rec
        ctx.theta_meta = (theta.dtype, theta.device)
        return out

    @staticmethod
    def backward(ctx, grad_out):
        g = grad_out.detach().to(torch.float32).contiguous()
        torch.cuda.current_stream(g.device).synchronize()
        dphi0 = torch.empty_like(g)
        th = ctx.rec.bwd_ptr(g.data_ptr(), dphi0.data_ptr())
        dtype, device = ctx.theta_meta
        return dphi0, torch.tensor(th, dtype=dtype, device=device), None, None, None, None, None


def segment(engine: Context, phi0: torch.Tensor, theta: torch.Tensor, dt: float = 0.5, iters: int = 200,
            band: float = 1.0, mode: int = MODE_REF) -> torch.Tensor:
    """phi after `iters` iterations from phi0 with theta = (T, eps, alpha), differentiable in phi0 and theta."""
    return LevelSetFunction.apply(phi0, theta, engine, dt, iters, band, mode)


def soft_dice_loss(phi: torch.Tensor, mask: torch.Tensor, temperature: float = 0.5) -> torch.Tensor:
    """1 - soft Dice between sigmoid(-phi / temperature) (inside is phi < 0) and a binary mask."""
    s = torch.sigmoid(-phi / temperature)
    m = mask.to(s.dtype)
    return 1.0 - 2.0 * (s * m).sum() / (s.sum() + m.sum())
