"""The reference's initial_conditions module (src/my_swampe/initial_conditions.py), with the same
function names, arguments and return values.

These build a run's inputs once, before any step: the resolution table and geometry
(``spectral_params``), the Williamson test-case fields (``test1_init``, ``state_var_init``,
``velocity_init``), the nonlinear products of the initial fields (``ABCDE_init``) and the
spectral Coriolis term (``coriolismn``).  The time step itself computes its products in the
grid kernel and its Coriolis term from the plan's tables.

Grid functions compute in the array type of their field arguments: numpy arrays on the host, or
torch tensors where they are (on the GPU).  Results are float64.
"""

from __future__ import annotations

import math
from typing import Optional, Tuple

import numpy as np

from . import _capi
from .spectral_transform import build_lambdas, gauss_legendre


def _lib(x):
    """numpy for host arrays, torch for tensors."""
    if hasattr(x, "detach"):
        import torch

        return torch
    return np


def _col(x, xp):
    """x as a float64 column (J, 1) in the array type xp."""
    if xp is np:
        return np.asarray(x, dtype=np.float64).reshape(-1, 1)
    return x.to(dtype=xp.float64).reshape(-1, 1)


def _row(x, xp, like):
    """x as a float64 row (1, I) in the array type xp (on like's device for tensors)."""
    if xp is np:
        return np.asarray(x, dtype=np.float64).reshape(1, -1)
    t = x if hasattr(x, "detach") else xp.as_tensor(np.asarray(x, dtype=np.float64))
    return t.to(device=like.device, dtype=xp.float64).reshape(1, -1)


def test1_init(a: float, omega: float, a1: float) -> Tuple[float, float, float, float, float]:
    """(SU0, sina, cosa, etaamp, Phiamp) of Williamson test 1, the cosine bell over the pole
    (initial_conditions.py:18-47): solid-body speed 2 pi a / 12 days, the rotation axis tilted
    by a1."""
    a, omega, a1 = float(a), float(omega), float(a1)
    SU0 = 2.0 * math.pi * a / (3600.0 * 24.0 * 12.0)
    etaamp = 2.0 * ((SU0 / a) + omega)
    Phiamp = SU0 * a * omega + 0.5 * SU0 ** 2
    return SU0, math.sin(a1), math.cos(a1), etaamp, Phiamp


def spectral_params(M: int):
    """(N, I, J, dt, lambdas, mus, w) of truncation M (initial_conditions.py:50-85): N = M, the
    grid and reference time step of Jakob et al. (1993) Tables 1 and 2, the longitudes and the
    reference's Gauss nodes and weights.  ValueError for M outside 42, 63, 106."""
    M = int(M)
    J, I = _capi.grid_size(M)  # ValueError with the reference's message for other M
    dt = {42: 1200, 63: 900, 106: 600}[M]
    mus, w = gauss_legendre(J)
    return M, I, J, dt, build_lambdas(I), mus, w


def state_var_init(I: int, J: int, mus, lambdas, test: Optional[int], etaamp: float, *args):
    """(etaic0, etaic1, deltaic0, deltaic1, Phiic0, Phiic1), each (J, I)
    (initial_conditions.py:88-174).  For test 1 or 2, args = (a, sina, cosa, Phibar, Phiamp):
    vorticity of the tilted solid-body rotation; Phi the cosine bell of radius a / 3 centred on
    the equator at lambda = 3 pi / 2 (test 1) or the balanced geopotential (test 2).  Otherwise
    eta = etaamp mu and Phi = 0.  delta is zero and both levels are equal."""
    I, J = int(I), int(J)
    xp = _lib(mus)
    mu = _col(mus, xp)
    lam = _row(lambdas, xp, mu)
    zeros = xp.zeros((J, I), dtype=xp.float64) if xp is np else xp.zeros((J, I), dtype=xp.float64, device=mu.device)
    delta = zeros
    if test is not None:
        if len(args) != 5:
            raise ValueError("For test!=None, expected args=(a,sina,cosa,Phibar,Phiamp).")
        a, sina, cosa, Phibar, Phiamp = (float(v) for v in args)
    if test in (1, 2):
        s1m = xp.sqrt(1.0 - mu ** 2)
        tilt = -xp.cos(lam) * s1m * sina + mu * cosa
        eta = float(etaamp) * tilt
        if test == 1:
            radius = a / 3.0
            # great-circle distance to (mu, lambda) = (0, 3 pi / 2): cos d = sqrt(1 - mu^2) cos(lam - 3 pi / 2)
            cosd = 0.0 * mu + xp.cos(xp.arcsin(mu)) * xp.cos(lam - 3.0 * math.pi / 2.0)
            dist = a * xp.arccos(xp.clip(cosd, -1.0, 1.0))
            bell = 0.5 * Phibar * (1.0 + xp.cos(math.pi * dist / radius))
            Phi = xp.where(dist < radius, bell, zeros)
        else:
            Phi = (Phibar - Phiamp) * tilt ** 2
    else:
        eta = float(etaamp) * (mu + 0.0 * lam)
        Phi = zeros
    return eta, eta, delta, delta, Phi, Phi


def velocity_init(I: int, J: int, SU0: float, cosa: float, sina: float, mus, lambdas, test: Optional[int]):
    """(Uic, Vic), each (J, I) (initial_conditions.py:177-239): the winds of the tilted
    solid-body rotation (test 1 carries the extra sqrt(1 - mu^2) factor of the reference), zero
    otherwise."""
    I, J = int(I), int(J)
    xp = _lib(mus)
    mu = _col(mus, xp)
    lam = _row(lambdas, xp, mu)
    SU0, cosa, sina = float(SU0), float(cosa), float(sina)
    if test in (1, 2):
        s1m = xp.sqrt(1.0 - mu ** 2)
        if test == 1:
            U = SU0 * (s1m * cosa + mu * xp.cos(lam) * sina) * s1m
            V = -SU0 * xp.sin(lam) * sina * s1m
        else:
            U = SU0 * (s1m * cosa + xp.cos(lam) * mu * sina)
            V = -SU0 * (xp.sin(lam) * sina) + 0.0 * mu
        return U, V
    if xp is np:
        return np.zeros((J, I)), np.zeros((J, I))
    return xp.zeros((J, I), dtype=xp.float64, device=mu.device), xp.zeros((J, I), dtype=xp.float64, device=mu.device)


def ABCDE_init(Uic, Vic, etaic0, Phiic0, mus, I: int, J: int):
    """(A, B, C, D, E) = (U eta, V eta, U Phi, V Phi, (U^2 + V^2) / (2 (1 - mu^2))), each (J, I)
    (initial_conditions.py:242-292)."""
    xp = _lib(Uic)
    mu = _col(mus, xp) if xp is np else _row(mus, xp, Uic).reshape(-1, 1)
    return (Uic * etaic0, Vic * etaic0, Uic * Phiic0, Vic * Phiic0,
            (Uic * Uic + Vic * Vic) / (2.0 * (1.0 - mu ** 2)))


def coriolismn(M: int, omega: float) -> np.ndarray:
    """Spectral Coriolis parameter (M+1, M+1) (initial_conditions.py:295-316): f = 2 omega mu is
    omega / sqrt(0.375) times the normalized P_1^0, so only (m, n) = (0, 1) is nonzero."""
    M = int(M)
    f = np.zeros((M + 1, M + 1))
    f[0, 1] = float(omega) / math.sqrt(0.375)
    return f
