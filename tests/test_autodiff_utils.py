"""fwd_grad (autodiff_utils.py:20-79) on the host: the gradient of closed-form losses, the empty
vector, and the argument errors.  Its use over the differentiable rollout is covered by the GPU
adjoint tests."""

from __future__ import annotations

import numpy as np
import pytest
import torch

from my_swampe_amd import fwd_grad


def test_gradient_of_closed_form_losses():
    """L = 0.5 x^T A x + b^T x has gradient A_sym x + b; L = sum sin(x_i) x_(i+1) has
    cos(x_i) x_(i+1) + sin(x_(i-1)); chunk does not change the result."""
    rng = np.random.default_rng(3)
    A = torch.tensor(rng.standard_normal((7, 7)))
    b = torch.tensor(rng.standard_normal(7))
    x = torch.tensor(rng.standard_normal(7))
    want = 0.5 * (A + A.T) @ x + b
    for chunk in (None, 1, 3, 7, 50):
        g = fwd_grad(lambda t: 0.5 * t @ A @ t + b @ t, x, chunk=chunk)
        assert g.shape == x.shape and g.dtype == x.dtype
        torch.testing.assert_close(g, want, rtol=1e-14, atol=1e-14)
    g = fwd_grad(lambda t: (torch.sin(t[:-1]) * t[1:]).sum(), x)
    want = torch.zeros_like(x)
    want[:-1] += torch.cos(x[:-1]) * x[1:]
    want[1:] += torch.sin(x[:-1])
    torch.testing.assert_close(g, want, rtol=1e-14, atol=1e-14)
    assert not x.requires_grad


def test_empty_vector_and_argument_errors():
    z = fwd_grad(lambda t: t.sum(), torch.zeros(0, dtype=torch.float32), chunk=-1)
    assert z.shape == (0,) and z.dtype == torch.float32
    with pytest.raises(ValueError, match="theta must be 1-D"):
        fwd_grad(lambda t: t.sum(), torch.zeros((2, 2)))
    with pytest.raises(ValueError, match="chunk must be a positive int or None"):
        fwd_grad(lambda t: t.sum(), torch.ones(3), chunk=0)
    with pytest.raises(ValueError, match="scalar"):
        fwd_grad(lambda t: t * 2.0, torch.ones(3))
