"""Phase-curve observable of geopotential grids on the GPU.

``phase_curve(Phi, phase_angles)`` is the visibility-weighted disc integral of
unit_tests/test_parity_reference_regression.py:67-82 (the projection the retrieval fits to a
phase curve) for grids of shape (..., J, I): per phase angle phi_k,

    flux_k = sum_j sqrt(max(1 - mu_j^2, 0)) sum_i Phi[j, i] max(cos(lambda_i - phi_k), 0),

with lambda_i = linspace(-pi, pi, I, endpoint=False) and mu_j numpy's leggauss(J) nodes (the
reference function's choices), divided by mean_k |flux_k| when that is > 0 and ``normalize``
is set.  Forward and backward run in HIP kernels (swx_phase_curve, swx_phase_curve_vjp); the
function is differentiable in Phi through torch.autograd.
"""

from __future__ import annotations

import ctypes
from typing import Optional

import numpy as np

from . import _capi
from .plan import require_gpu


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


_MUS: dict = {}


def leggauss_nodes(J: int, device):
    """numpy leggauss(J) nodes (ascending) as a float64 device tensor (cached per J and device)."""
    torch = require_gpu()
    key = (int(J), str(device))
    if key not in _MUS:
        mus = np.polynomial.legendre.leggauss(int(J))[0]
        _MUS[key] = torch.as_tensor(np.ascontiguousarray(mus, dtype=np.float64), device=device)
    return _MUS[key]


def _phases(phase_angles, device):
    torch = require_gpu()
    ph = torch.as_tensor(phase_angles, dtype=torch.float64, device=device).reshape(-1).contiguous()
    if ph.numel() == 0:
        raise ValueError("phase_angles must hold at least one angle")
    return ph


def phase_curve_raw(Phi, phases, normalize: bool = True, mus=None):
    """(flux, flux_raw) of the grids Phi (n, J, I) (float64 device tensor, contiguous) at the phase
    angles ``phases`` (K,) (float64 device tensor): both (n, K)."""
    torch = require_gpu()
    n, J, I = Phi.shape
    K = phases.numel()
    mus = leggauss_nodes(J, Phi.device) if mus is None else mus
    flux = torch.empty((n, K), dtype=torch.float64, device=Phi.device)
    raw = torch.empty((n, K), dtype=torch.float64, device=Phi.device)
    stream = ctypes.c_void_p(torch.cuda.current_stream(Phi.device).cuda_stream)
    _capi.check(_capi.load().swx_phase_curve(_ptr(Phi), J * I, n, J, I, _ptr(mus), _ptr(phases), K,
                                             int(bool(normalize)), _ptr(flux), _ptr(raw), stream),
                "swx_phase_curve")
    return flux, raw


def phase_curve_vjp(g_flux, flux_raw, phases, J: int, I: int, normalize: bool = True, mus=None):
    """Cotangent (n, J, I) of the grids from the cotangent g_flux (n, K) of phase_curve's output
    and the un-normalized flux_raw (n, K) of the forward call."""
    torch = require_gpu()
    n, K = g_flux.shape
    dev = g_flux.device
    mus = leggauss_nodes(J, dev) if mus is None else mus
    g_flux = g_flux.to(torch.float64).contiguous()
    g_Phi = torch.empty((n, J, I), dtype=torch.float64, device=dev)
    stream = ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream)
    _capi.check(_capi.load().swx_phase_curve_vjp(_ptr(g_flux), _ptr(flux_raw), n, J, I, _ptr(mus), _ptr(phases), K,
                                                 int(bool(normalize)), _ptr(g_Phi), stream),
                "swx_phase_curve_vjp")
    return g_Phi


class _PhaseCurve:
    """torch.autograd.Function factory (torch imported lazily)."""

    _cls = None

    @classmethod
    def get(cls):
        if cls._cls is not None:
            return cls._cls
        torch = require_gpu()

        class Fn(torch.autograd.Function):
            @staticmethod
            def forward(ctx, Phi, phases, normalize, mus):
                flux, raw = phase_curve_raw(Phi, phases, normalize, mus)
                ctx.save_for_backward(raw, phases)
                ctx.normalize, ctx.mus, ctx.shape = normalize, mus, Phi.shape
                return flux

            @staticmethod
            def backward(ctx, g):
                raw, phases = ctx.saved_tensors
                _, J, I = ctx.shape
                return phase_curve_vjp(g, raw, phases, J, I, ctx.normalize, ctx.mus), None, None, None

        cls._cls = Fn
        return Fn


def phase_curve(Phi, phase_angles, normalize: bool = True, mus: Optional[object] = None):
    """Phase curve (..., K) of geopotential grids Phi (..., J, I) on the GPU (see module docstring).

    Phi: float64 CUDA tensor (a numpy array is moved to the current device); phase_angles: K
    angles; mus: optional (J,) Gauss nodes (default numpy leggauss(J), as the reference's
    projection).  Differentiable in Phi.
    """
    torch = require_gpu()
    if not torch.is_tensor(Phi):
        Phi = torch.as_tensor(np.asarray(Phi, dtype=np.float64), device=torch.device("cuda", torch.cuda.current_device()))
    if Phi.dtype != torch.float64:
        raise TypeError("phase_curve: Phi must be float64")
    if not Phi.is_cuda:
        raise ValueError("phase_curve: Phi must be a GPU tensor (there is no CPU path)")
    if Phi.dim() < 2:
        raise ValueError("phase_curve: Phi must have shape (..., J, I)")
    lead, (J, I) = Phi.shape[:-2], Phi.shape[-2:]
    ph = _phases(phase_angles, Phi.device)
    if mus is not None:
        mus = torch.as_tensor(mus, dtype=torch.float64, device=Phi.device).reshape(-1).contiguous()
        if mus.numel() != J:
            raise ValueError(f"phase_curve: mus must hold J={J} nodes")
    flat = Phi.reshape((-1, J, I)).contiguous()
    out = _PhaseCurve.get().apply(flat, ph, bool(normalize), mus)
    return out.reshape(tuple(lead) + (ph.numel(),))
