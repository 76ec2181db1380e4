"""The reference's spectral_transform module (src/my_swampe/spectral_transform.py), with the same
function names, arguments and return values.

Geometry (host FP64, computed once per truncation, as in the reference):
``build_lambdas(I)`` (:54-77) and ``PmnHmn(J, M, N, mus)`` (:219-287) return numpy arrays from the
native library's geometry code; ``gauss_legendre(J)`` (:80-114) returns the reference's own
quadrature (SciPy's) unless asked for the native one; ``weighted_legendre_basis`` (:351-367) is the
elementwise product ``w[:, None, None] * Pmn`` in the caller's array type.

Transforms (device FP64, HIP kernels, no CPU path): ``fwd_fft_trunc`` / ``fwd_fft_trunc_batch``
(:290-328), ``invrs_fft`` (:331-348), ``fwd_leg_w`` / ``fwd_leg_w_batch`` / ``fwd_leg``
(:370-435), ``invrs_leg`` (:438-481), ``invrsUV`` / ``invrsUV_with_coeffs`` (:484-596) and
``diagnostic_eta_delta`` (:599-650).  Inputs may be numpy arrays or torch tensors; results are
complex128 torch tensors on the GPU.  The Legendre transforms, ``invrsUV(_with_coeffs)`` and
``diagnostic_eta_delta`` also take a leading batch axis on their field arguments (one launch for
several members), where the reference's einsums accept only the unbatched shapes; the DFTs keep
the reference's axes (``fwd_fft_trunc`` and ``invrs_fft`` transform axis 1 of a 2-D array,
``fwd_fft_trunc_batch`` the last axis of any array).
"""

from __future__ import annotations

import ctypes
from typing import Tuple

import numpy as np

from . import _capi
from .plan import require_gpu


def _dev(x, torch, dtype):
    t = x if isinstance(x, torch.Tensor) else torch.as_tensor(np.asarray(x))
    return t.to(device="cuda", dtype=dtype).contiguous()


def _is_complex(x, torch) -> bool:
    return x.is_complex() if isinstance(x, torch.Tensor) else np.iscomplexobj(x)


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def _stream(torch):
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _host(x) -> np.ndarray:
    if hasattr(x, "detach"):
        x = x.detach().cpu().numpy()
    return np.ascontiguousarray(np.asarray(x, dtype=np.float64))


def _check_shape(name, t, shape):
    if tuple(t.shape) != tuple(shape):
        raise ValueError(f"{name} must have shape {tuple(shape)}; got {tuple(t.shape)}")


def _batched(name, t, tail):
    """Whether t is tail-shaped or (B,) + tail; ValueError otherwise."""
    if t.dim() == len(tail) and tuple(t.shape) == tuple(tail):
        return False
    if t.dim() == len(tail) + 1 and tuple(t.shape[1:]) == tuple(tail):
        return True
    raise ValueError(f"{name} must have shape {tuple(tail)} or (B, {', '.join(map(str, tail))}); "
                     f"got {tuple(t.shape)}")


# --------------------------------------------------------------------------------------------
# Geometry (host)
# --------------------------------------------------------------------------------------------

def build_lambdas(I: int, dtype=None) -> np.ndarray:
    """Longitudes linspace(-pi, pi, I, endpoint=False) (spectral_transform.py:54-77)."""
    lam = np.zeros(int(I))
    _capi.check(_capi.load().swx_lambdas(int(I), lam.ctypes.data_as(ctypes.c_void_p)), "build_lambdas")
    return lam if dtype is None else lam.astype(dtype)


def gauss_legendre(J: int, dtype=None, *, exact: bool = False) -> Tuple[np.ndarray, np.ndarray]:
    """Gauss-Legendre nodes (ascending) and weights of order J (spectral_transform.py:80-114).

    By default the reference's own values: scipy.special.roots_legendre (numpy's leggauss where
    SciPy is missing), which the plans also take so that their results match the reference's.
    ``exact=True`` returns the native library's Newton solution instead, whose weights are closer
    to the exact ones (scipy's are off by up to 2.9e-11 relative at J = 160)."""
    if not exact:
        mus, w = _capi.reference_quadrature(int(J))
        return (mus, w) if dtype is None else (mus.astype(dtype), w.astype(dtype))
    mus, w = np.zeros(int(J)), np.zeros(int(J))
    _capi.check(_capi.load().swx_gauss_legendre(int(J), mus.ctypes.data_as(ctypes.c_void_p),
                                                w.ctypes.data_as(ctypes.c_void_p)), "gauss_legendre")
    if dtype is not None:
        mus, w = mus.astype(dtype), w.astype(dtype)
    return mus, w


def PmnHmn(J: int, M: int, N: int, mus) -> Tuple[np.ndarray, np.ndarray]:
    """The normalized associated Legendre basis Pmn and Hmn = (1 - mu^2) dPmn/dmu at the nodes mus,
    (J, M+1, N+1) each, with the reference's factorial normalization (spectral_transform.py:
    219-287)."""
    J, M, N = int(J), int(M), int(N)
    x = _host(mus).reshape(-1)
    if x.shape != (J,):
        raise ValueError(f"mus must have {J} entries; got {x.shape[0]}")
    P = np.zeros((J, M + 1, N + 1))
    H = np.zeros_like(P)
    vp = ctypes.c_void_p
    _capi.check(_capi.load().swx_legendre_basis(J, M, N, x.ctypes.data_as(vp), 0, P.ctypes.data_as(vp),
                                                H.ctypes.data_as(vp)), "PmnHmn")
    return P, H


def weighted_legendre_basis(Pmn, w):
    """w[:, None, None] * Pmn (spectral_transform.py:351-367), in the type of Pmn."""
    if hasattr(Pmn, "detach"):
        wt = w if hasattr(w, "detach") else Pmn.new_tensor(np.asarray(w))
        return wt.to(Pmn.device).reshape(-1, 1, 1) * Pmn
    return np.asarray(w)[:, None, None] * np.asarray(Pmn)


# --------------------------------------------------------------------------------------------
# Longitude DFTs (device)
# --------------------------------------------------------------------------------------------

def _fft_trunc(data, I, M, torch):
    I, M = int(I), int(M)
    cplx = _is_complex(data, torch)
    x = _dev(data, torch, torch.complex128 if cplx else torch.float64)
    if x.dim() < 1:
        raise ValueError("fwd_fft_trunc: data must have a longitude axis")
    lead = tuple(x.shape[:-1])
    C = x.shape[-1]
    rows = int(np.prod(lead)) if lead else 1
    J = lead[-1] if lead else 1
    K = rows // J if J else 0
    out = torch.empty(lead + (min(M + 1, I),), dtype=torch.complex128, device=x.device)
    if rows == 0:
        return out
    _capi.check(_capi.load().swx_fwd_fft_trunc(_ptr(x), int(cplx), K, J, C, I, M, _ptr(out), _stream(torch)),
                "fwd_fft_trunc")
    return out


def fwd_fft_trunc(data, I: int, M: int):
    """fft(data / I, n=I, axis=1)[:, :M+1] of a (J, I) field, real or complex
    (spectral_transform.py:290-306): (J, M+1) complex."""
    torch = require_gpu()
    ndim = data.dim() if isinstance(data, torch.Tensor) else np.ndim(data)
    if ndim != 2:
        raise ValueError("fwd_fft_trunc transforms axis 1 of a 2-D (J, I) array; use fwd_fft_trunc_batch "
                         "for batched fields")
    return _fft_trunc(data, I, M, torch)


def fwd_fft_trunc_batch(data, I: int, M: int):
    """fft(data / I, n=I, axis=-1)[..., :M+1] of (..., J, I) fields (spectral_transform.py:309-328)."""
    torch = require_gpu()
    return _fft_trunc(data, I, M, torch)


def invrs_fft(approxXim, I: int):
    """ifft(I * approxXim, n=I, axis=1) of a (J, C) array of Fourier coefficients, zero-padded or
    truncated to I columns (spectral_transform.py:331-348): (J, I) complex."""
    torch = require_gpu()
    I = int(I)
    X = _dev(approxXim, torch, torch.complex128)
    if X.dim() != 2:
        raise ValueError("invrs_fft transforms axis 1 of a 2-D (J, C) array")
    J, C = X.shape
    out = torch.empty((J, I), dtype=torch.complex128, device=X.device)
    if J == 0:
        return out
    _capi.check(_capi.load().swx_invrs_fft(_ptr(X), 1, J, C, I, _ptr(out), _stream(torch)), "invrs_fft")
    return out


# --------------------------------------------------------------------------------------------
# Legendre transforms and the wind inversion (device)
# --------------------------------------------------------------------------------------------

def _fwd_leg(data, basis, w, torch):
    B = _dev(basis, torch, torch.float64)
    if B.dim() != 3:
        raise ValueError(f"the Legendre basis must be (J, M+1, N+1); got {tuple(B.shape)}")
    J, Mp, Np = B.shape
    x = _dev(data, torch, torch.complex128)
    batched = _batched("data", x, (J, Mp))
    K = x.shape[0] if batched else 1
    wj = None
    if w is not None:
        wj = _dev(w, torch, torch.float64).reshape(-1)
        _check_shape("w", wj, (J,))
    out = torch.empty((K, Mp, Np), dtype=torch.complex128, device=x.device)
    if K:
        _capi.check(_capi.load().swx_fwd_leg(_ptr(x), _ptr(B), _ptr(wj) if wj is not None else None, K, J, Mp - 1,
                                             Np - 1, _ptr(out), _stream(torch)), "fwd_leg")
    return out if batched else out[0]


def fwd_leg_w(data, Pmnw):
    """einsum('jm,jmn->mn', data, Pmnw) with a preweighted basis (spectral_transform.py:370-385)."""
    return _fwd_leg(data, Pmnw, None, require_gpu())


def fwd_leg_w_batch(data, Pmnw):
    """einsum('kjm,jmn->kmn', data, Pmnw) (spectral_transform.py:388-402)."""
    torch = require_gpu()
    x = data if isinstance(data, torch.Tensor) else np.asarray(data)
    if x.ndim != 3:
        raise ValueError(f"fwd_leg_w_batch: data must be (K, J, M+1); got {tuple(x.shape)}")
    return _fwd_leg(data, Pmnw, None, torch)


def fwd_leg(data, J: int, M: int, N: int, Pmn, w):
    """S[m, n] = sum_j w_j data[j, m] Pmn[j, m, n] (spectral_transform.py:405-435): (M+1, N+1)
    complex, or (B, M+1, N+1) for (B, J, M+1) data."""
    torch = require_gpu()
    B = _dev(Pmn, torch, torch.float64)
    _check_shape("Pmn", B, (int(J), int(M) + 1, int(N) + 1))
    return _fwd_leg(data, B, w, torch)


def invrs_leg(legcoeff, I: int, J: int, M: int, N: int, Pmn):
    """Fourier coefficients (J, I) of a spectrum (M+1, N+1): +m at column m, the conjugates at
    column I - m (spectral_transform.py:438-481).  A (B, M+1, N+1) spectrum gives (B, J, I)."""
    torch = require_gpu()
    I, J, M, N = int(I), int(J), int(M), int(N)
    B = _dev(Pmn, torch, torch.float64)
    _check_shape("Pmn", B, (J, M + 1, N + 1))
    S = _dev(legcoeff, torch, torch.complex128)
    batched = _batched("legcoeff", S, (M + 1, N + 1))
    K = S.shape[0] if batched else 1
    out = torch.empty((K, J, I), dtype=torch.complex128, device=S.device)
    if K:
        _capi.check(_capi.load().swx_invrs_leg(_ptr(S), _ptr(B), K, I, J, M, N, _ptr(out), _stream(torch)),
                    "invrs_leg")
    return out if batched else out[0]


def _invrs_uv(deltamn, etamn, fmn, I, J, M, N, Pmn, Hmn, tstepcoeffmn, marray, coeffs):
    torch = require_gpu()
    I, J, M, N = int(I), int(J), int(M), int(N)
    d = _dev(deltamn, torch, torch.complex128)
    e = _dev(etamn, torch, torch.complex128)
    batched = _batched("deltamn", d, (M + 1, N + 1))
    if tuple(e.shape) != tuple(d.shape):
        raise ValueError(f"etamn must have the shape of deltamn {tuple(d.shape)}; got {tuple(e.shape)}")
    K = d.shape[0] if batched else 1
    f = _dev(fmn, torch, torch.complex128)
    P = _dev(Pmn, torch, torch.float64)
    H = _dev(Hmn, torch, torch.float64)
    tc = _dev(tstepcoeffmn, torch, torch.float64)
    ma = _dev(marray, torch, torch.float64)
    for name, t, shape in (("fmn", f, (M + 1, N + 1)), ("Pmn", P, (J, M + 1, N + 1)), ("Hmn", H, (J, M + 1, N + 1)),
                           ("tstepcoeffmn", tc, (M + 1, N + 1)), ("marray", ma, (M + 1, N + 1))):
        _check_shape(name, t, shape)
    U = torch.empty((K, J, I), dtype=torch.complex128, device=d.device)
    V = torch.empty_like(U)
    Um = Vm = None
    if coeffs:
        Um = torch.empty((K, J, M + 1), dtype=torch.complex128, device=d.device)
        Vm = torch.empty_like(Um)
    if K:
        _capi.check(_capi.load().swx_invrs_uv(_ptr(d), _ptr(e), _ptr(f), K, I, J, M, N, _ptr(P), _ptr(H), _ptr(tc),
                                              _ptr(ma), _ptr(U), _ptr(V), _ptr(Um) if coeffs else None,
                                              _ptr(Vm) if coeffs else None, _stream(torch)), "invrsUV")
    outs = (U, V, Um, Vm) if coeffs else (U, V)
    return outs if batched else tuple(o[0] for o in outs)


def invrsUV(deltamn, etamn, fmn, I: int, J: int, M: int, N: int, Pmn, Hmn, tstepcoeffmn, marray):
    """Winds (U, V), (J, I) complex, from spectral divergence and vorticity
    (spectral_transform.py:484-532)."""
    return _invrs_uv(deltamn, etamn, fmn, I, J, M, N, Pmn, Hmn, tstepcoeffmn, marray, False)


def invrsUV_with_coeffs(deltamn, etamn, fmn, I: int, J: int, M: int, N: int, Pmn, Hmn, tstepcoeffmn, marray):
    """(U, V, Um, Vm): the winds and their positive-m Fourier coefficients (J, M+1)
    (spectral_transform.py:535-596)."""
    return _invrs_uv(deltamn, etamn, fmn, I, J, M, N, Pmn, Hmn, tstepcoeffmn, marray, True)


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
        _check_shape(name, t, shape)
    eta = torch.empty((B, J, I), dtype=torch.complex128, device=u.device)
    delta = torch.empty_like(eta)
    etamn = torch.empty((B, M + 1, N + 1), dtype=torch.complex128, device=u.device)
    deltamn = torch.empty_like(etamn)
    lib = _capi.load()
    _capi.check(lib.swx_diagnostic_eta_delta(_ptr(u), _ptr(v), _ptr(f), B, I, J, M, N, _ptr(P), _ptr(H), _ptr(wj),
                                             _ptr(tc), _ptr(mj), float(dt), _ptr(eta), _ptr(delta), _ptr(etamn),
                                             _ptr(deltamn), _stream(torch)), "swx_diagnostic_eta_delta")
    if not batched:
        return eta[0], delta[0], etamn[0], deltamn[0]
    return eta, delta, etamn, deltamn
