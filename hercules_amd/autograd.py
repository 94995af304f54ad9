"""Thin PyTorch binding of the discrete adjoint.

``wall_shear_loss(channel, u0, v0, w0, nsteps, ckpt_every)`` returns the
plane-averaged bottom wall shear after ``nsteps`` RK3 steps of ``channel``
started from the velocity fields ``u0, v0, w0`` (torch tensors of shape
(nz, ny, nx)) as a differentiable scalar.  Its backward is the library's
hand-written adjoint (hc_wall_shear_gradient: checkpointed forward run,
recomputation per segment, transposed substeps); the channel's pressure and
other fields are taken as they are and not differentiated.  Tensors on the
GPU stay there: they are copied device-to-device into the channel and the
gradients are written into new device tensors.

torch ships its own HIP runtime under the soname the library links
(libamdhip64.so.7), and a process keeps the first one it loads: import torch
(or this module) before the first Channel is created, or torch may find no
GPU when it first touches one.
"""
from __future__ import annotations

import sys
import warnings

from . import channel as _channel

if "torch" not in sys.modules and _channel._lib is not None:
    warnings.warn("hercules_amd.autograd: torch is imported after the HIP library was loaded; torch's GPU "
                  "tensors may then find no device. Import torch before creating the first Channel.",
                  RuntimeWarning, stacklevel=2)

import torch  # noqa: E402


def _to_channel(channel, name, t):
    """A (nz, ny, nx) tensor into the channel's field: device tensors are
    copied device-to-device on the library's stream after torch's stream has
    produced them, host tensors through numpy."""
    t = t.detach().to(torch.float64).contiguous()
    if t.is_cuda:
        torch.cuda.current_stream(t.device).synchronize()
        channel.set_field_from(name, t.data_ptr())
    else:
        channel.set_field(name, t.numpy())


def _adjoint(channel, name, like):
    """The adjoint field `name` as a tensor on like's device (written in
    place by the library when that device is a GPU)."""
    if like.is_cuda:
        g = torch.empty(like.shape, dtype=torch.float64, device=like.device)
        channel.get_adjoint_into(name, g.data_ptr())  # returns after its stream finished
        return g.to(like.dtype)
    return torch.from_numpy(channel.get_adjoint(name)).to(like.dtype)


class _WallShear(torch.autograd.Function):
    """The forward runs the checkpointed gradient computation at once (the
    adjoint needs the forward's checkpoints), so backward only scales it."""

    @staticmethod
    def forward(ctx, u0, v0, w0, channel, nsteps, ckpt_every):
        for name, t in (("u", u0), ("v", v0), ("w", w0)):
            _to_channel(channel, name, t)
        loss = channel.wall_shear_gradient(nsteps, ckpt_every)
        grads = [_adjoint(channel, n, u0) for n in ("u", "v", "w")]
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
