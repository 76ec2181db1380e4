"""Standalone spectral transforms on the GPU, with the reference's interface.

``diagnostic_eta_delta(Um, Vm, fmn, I, J, M, N, Pmn, Hmn, w, tstepcoeff, mJarray, dt)`` is
spectral_transform.py:599-650 of the reference: the vorticity (plus the Coriolis term fmn) and
divergence of Fourier winds, as spectra and as grids.  Same arguments and return tuple
``(neweta, newdelta, etamn, deltamn)``; the inputs may be numpy arrays or torch tensors, a
leading batch axis on Um / Vm runs several members in one launch, and the results are
complex128 torch tensors on the GPU.  It runs in FP64 HIP kernels (swx_diagnostic_eta_delta);
there is no CPU path.
"""

from __future__ import annotations

import ctypes

import numpy as np

from . import _capi
from .plan import require_gpu


def _dev(x, torch, dtype):
    t = x if isinstance(x, torch.Tensor) else torch.as_tensor(np.asarray(x))
    return t.to(device="cuda", dtype=dtype).contiguous()


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def diagnostic_eta_delta(Um, Vm, fmn, I, J, M, N, Pmn, Hmn, w, tstepcoeff, mJarray, dt):
    """(eta, delta) from Fourier winds (reference spectral_transform.py:599-650).

    Um, Vm: (J, M+1) or (B, J, M+1) complex; fmn: (M+1, N+1); Pmn, Hmn: (J, M+1, N+1); w: (J,);
    tstepcoeff, mJarray: (J, M+1); dt: scalar.  Returns neweta, newdelta ((B,) J, I) and etamn,
    deltamn ((B,) M+1, N+1), complex128 on the GPU."""
    torch = require_gpu()
    I, J, M, N = int(I), int(J), int(M), int(N)
    u = _dev(Um, torch, torch.complex128)
    v = _dev(Vm, torch, torch.complex128)
    if u.shape != v.shape or tuple(u.shape[-2:]) != (J, M + 1) or u.dim() not in (2, 3):
        raise ValueError(f"Um and Vm must be ({J}, {M + 1}) or (B, {J}, {M + 1}); got {tuple(u.shape)}, "
                         f"{tuple(v.shape)}")
    batched = u.dim() == 3
    B = u.shape[0] if batched else 1
    f = _dev(fmn, torch, torch.complex128)
    P = _dev(Pmn, torch, torch.float64)
    H = _dev(Hmn, torch, torch.float64)
    wj = _dev(w, torch, torch.float64).reshape(-1)
    tc = _dev(tstepcoeff, torch, torch.float64)
    mj = _dev(mJarray, torch, torch.float64)
    for name, t, shape in (("fmn", f, (M + 1, N + 1)), ("Pmn", P, (J, M + 1, N + 1)), ("Hmn", H, (J, M + 1, N + 1)),
                           ("w", wj, (J,)), ("tstepcoeff", tc, (J, M + 1)), ("mJarray", mj, (J, M + 1))):
        if tuple(t.shape) != shape:
            raise ValueError(f"{name} must have shape {shape}; got {tuple(t.shape)}")
    eta = torch.empty((B, J, I), dtype=torch.complex128, device=u.device)
    delta = torch.empty_like(eta)
    etamn = torch.empty((B, M + 1, N + 1), dtype=torch.complex128, device=u.device)
    deltamn = torch.empty_like(etamn)
    lib = _capi.load()
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    _capi.check(lib.swx_diagnostic_eta_delta(_ptr(u), _ptr(v), _ptr(f), B, I, J, M, N, _ptr(P), _ptr(H), _ptr(wj),
                                             _ptr(tc), _ptr(mj), float(dt), _ptr(eta), _ptr(delta), _ptr(etamn),
                                             _ptr(deltamn), stream), "swx_diagnostic_eta_delta")
    if not batched:
        return eta[0], delta[0], etamn[0], deltamn[0]
    return eta, delta, etamn, deltamn
