"""The reference's autodiff_utils module (src/my_swampe/autodiff_utils.py): ``fwd_grad``, the
gradient of a scalar loss with respect to a 1-D parameter vector, with the reference's arguments,
checks and return shape.

The reference pushes p tangent directions forward (jax.jacfwd or chunked JVPs).  Here the
differentiable rollouts (``my_swampe_amd.autograd``) carry a hand-written reverse-mode adjoint,
so the same derivative comes from one reverse sweep (torch.autograd) whatever p is; ``chunk``,
the reference's memory knob for the tangent batches, is checked and has nothing to tune.
"""

from __future__ import annotations

from typing import Callable, Optional


def fwd_grad(loss_fn: Callable, theta, *, chunk: Optional[int] = None):
    """d loss_fn(theta) / d theta for a 1-D float tensor theta (autodiff_utils.py:20-79): a
    tensor of theta's shape, dtype and device.  ValueError for a theta that is not 1-D or a chunk
    that is not a positive int or None."""
    import torch

    theta = theta if isinstance(theta, torch.Tensor) else torch.as_tensor(theta, dtype=torch.float64)
    if theta.dim() != 1:
        raise ValueError(f"theta must be 1-D, got shape {tuple(theta.shape)}.")
    if theta.shape[0] == 0:
        return torch.zeros((0,), dtype=theta.dtype, device=theta.device)
    if chunk is not None and int(chunk) <= 0:
        raise ValueError(f"chunk must be a positive int or None, got {chunk!r}.")
    x = theta.detach().clone().requires_grad_(True)
    loss = loss_fn(x)
    if not isinstance(loss, torch.Tensor) or loss.numel() != 1:
        raise ValueError("loss_fn must return a scalar tensor")
    (g,) = torch.autograd.grad(loss.reshape(()), x)
    return g
