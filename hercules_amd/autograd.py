"""Thin PyTorch binding of the discrete adjoint.

``wall_shear_loss(channel, u0, v0, w0, nsteps, ckpt_every)`` returns the
plane-averaged bottom wall shear after ``nsteps`` RK3 steps of ``channel``
started from the velocity fields ``u0, v0, w0`` (torch tensors of shape
(nz, ny, nx)) as a differentiable scalar.  Its backward is the library's
hand-written adjoint (hc_wall_shear_gradient: checkpointed forward run,
recomputation per segment, transposed substeps); the channel's pressure and
other fields are taken as they are and not differentiated.
"""
from __future__ import annotations

import torch


class _WallShear(torch.autograd.Function):
    @staticmethod
    def forward(ctx, u0, v0, w0, channel, nsteps, ckpt_every):
        for name, t in (("u", u0), ("v", v0), ("w", w0)):
            channel.set_field(name, t.detach().to("cpu", torch.float64).numpy())
        loss = channel.wall_shear_gradient(nsteps, ckpt_every)
        grads = [torch.from_numpy(channel.get_adjoint(n)).to(u0.device, u0.dtype) for n in ("u", "v", "w")]
        ctx.save_for_backward(*grads)
        return torch.tensor(loss, dtype=u0.dtype, device=u0.device)

    @staticmethod
    def backward(ctx, g):
        gu, gv, gw = ctx.saved_tensors
        return g * gu, g * gv, g * gw, None, None, None


def wall_shear_loss(channel, u0, v0, w0, nsteps: int, ckpt_every: int = 20):
    return _WallShear.apply(u0, v0, w0, channel, int(nsteps), int(ckpt_every))


def numpy_gradient(channel, nsteps: int, ckpt_every: int = 20):
    """(loss, {u, v, w, p: gradient}) from the channel's current state."""
    loss = channel.wall_shear_gradient(nsteps, ckpt_every)
    return loss, {n: channel.get_adjoint(n) for n in ("u", "v", "w", "p")}


__all__ = ["wall_shear_loss", "numpy_gradient"]
