"""The reference's time_stepping module (src/my_swampe/time_stepping.py): the one-member steps
``tstepping`` (:16-384) and ``tstepping_semi_implicit`` (:387-488), the coefficient arrays
(:491-534) and the RMS-wind diagnostic (:537-576),
with the same function names, arguments and return values.

The time step of the model runs in the fused HIP kernels: one plan step per call of
``my_swampe_amd.model._step_once`` or many per ``Plan.run`` / ``Plan.run_graph``, with these
coefficients in the plans' own FP64 tables.  ``tstepping`` here is the reference's standalone
step for one member's Fourier fields, composed from the standalone FP64 device updates
(``modEuler_tdiff`` / ``explicit_tdiff``) and the wind inversion (``invrsUV_with_coeffs``).  The
coefficient arrays are float64 numpy; ``tstepcoeff`` and ``RMS_winds`` compute in the array type of
their field arguments (numpy, or torch tensors where they are).
"""

from __future__ import annotations

import math

import numpy as np


def tstepcoeffmn(M: int, N: int, a: float) -> np.ndarray:
    """a / (n (n + 1)) broadcast to (M+1, N+1), zero at n = 0 (time_stepping.py:491-501)."""
    n = np.arange(int(N) + 1, dtype=np.float64)
    nn = n * (n + 1.0)
    nn[0] = 1.0
    c = float(a) / nn
    c[0] = 0.0
    return np.broadcast_to(c[None, :], (int(M) + 1, int(N) + 1)).copy()


def tstepcoeff2(J: int, M: int, dt: float, a: float) -> np.ndarray:
    """2 dt / a^2 everywhere in (J, M+1) (time_stepping.py:504-506)."""
    return np.full((int(J), int(M) + 1), 2.0 * float(dt) / (float(a) ** 2))


def narray(M: int, N: int) -> np.ndarray:
    """n (n + 1) broadcast to (M+1, N+1) (time_stepping.py:509-513)."""
    n = np.arange(int(N) + 1, dtype=np.float64)
    return np.broadcast_to((n * (n + 1.0))[None, :], (int(M) + 1, int(N) + 1)).copy()


def tstepcoeff(J: int, M: int, dt: float, mus, a: float):
    """2 dt / (a (1 - mu^2)) broadcast to (J, M+1) (time_stepping.py:516-522)."""
    J, M = int(J), int(M)
    if hasattr(mus, "detach"):
        mu = mus.reshape(-1, 1).double()
        return ((2.0 * float(dt)) / (float(a) * (1.0 - mu ** 2))).expand(J, M + 1).contiguous()
    mu = np.asarray(mus, dtype=np.float64).reshape(-1, 1)
    return np.broadcast_to((2.0 * float(dt)) / (float(a) * (1.0 - mu ** 2)), (J, M + 1)).copy()


def mJarray(J: int, M: int) -> np.ndarray:
    """m broadcast to (J, M+1) (time_stepping.py:525-528)."""
    return np.broadcast_to(np.arange(int(M) + 1, dtype=np.float64)[None, :], (int(J), int(M) + 1)).copy()


def marray(M: int, N: int) -> np.ndarray:
    """m broadcast to (M+1, N+1) (time_stepping.py:531-534)."""
    return np.broadcast_to(np.arange(int(M) + 1, dtype=np.float64)[:, None], (int(M) + 1, int(N) + 1)).copy()


def RMS_winds(a: float, I: int, J: int, lambdas, mus, U, V):
    """Area-weighted RMS wind speed (time_stepping.py:537-576): sqrt(sum(a^2 cos^2(phi) dphi
    dlambda ((U / cos phi)^2 + (V / cos phi)^2)) / (4 pi a^2)), with dphi and dlambda taken
    from nodes 1 and 2 as the reference does."""
    a = float(a)
    if hasattr(U, "detach"):
        import torch

        mu = (mus if hasattr(mus, "detach") else torch.as_tensor(np.asarray(mus, dtype=np.float64)))
        mu = mu.to(device=U.device, dtype=torch.float64).reshape(-1, 1)
        lam = lambdas if hasattr(lambdas, "detach") else torch.as_tensor(np.asarray(lambdas, dtype=np.float64))
        lam = lam.to(device=U.device, dtype=torch.float64)
        xp = torch
    else:
        mu = np.asarray(mus, dtype=np.float64).reshape(-1, 1)
        lam = np.asarray(lambdas, dtype=np.float64)
        xp = np
    phis = xp.arcsin(mu)
    dlam = lam[2] - lam[1]
    dphi = phis[2, 0] - phis[1, 0]
    area = (a ** 2) * (xp.sin(phis + math.pi / 2.0) ** 2) * dphi * dlam
    integrand = (U / xp.cos(phis)) ** 2 + (V / xp.cos(phis)) ** 2
    return xp.sqrt(xp.sum(area * integrand / (4.0 * math.pi * a ** 2)))


def tstepping(etam0, etam1, deltam0, deltam1, Phim0, Phim1, I, J, M, N, Am, Bm, Cm, Dm, Em, Fm, Gm, Um, Vm, fmn,
              Pmn, Hmn, Pmnw, Hmnw, tstepcoeff, tstepcoeff2, tstepcoeffmn, marray, mJarray, narray, PhiFm, dt, a,
              Phibar, taurad, taudrag, forcflag, diffflag, expflag, sigma, sigmaPhi, test, t):
    """One step of either scheme (time_stepping.py:16-384): the updates of Phi, delta and eta
    (explicit_tdiff when expflag, else modEuler_tdiff), then the winds of the new delta and eta.
    Returns (etamn, eta, etam, deltamn, delta, deltam, Phimn, Phi, Phim, U, V, Um, Vm): spectra
    (M+1, N+1), grids (J, I) and positive-m Fourier coefficients (J, M+1), complex128 on the GPU."""
    from . import explicit_tdiff, modEuler_tdiff
    from .spectral_transform import invrsUV_with_coeffs

    scheme = explicit_tdiff if bool(expflag) else modEuler_tdiff
    args = (etam0, etam1, deltam0, deltam1, Phim0, Phim1, I, J, M, N, Am, Bm, Cm, Dm, Em, Fm, Gm, Um, Vm, Pmn, Pmnw,
            Hmnw, tstepcoeff, tstepcoeff2, mJarray, narray, PhiFm, dt, a, Phibar, taurad, taudrag, forcflag, diffflag,
            sigma, sigmaPhi, test, t)
    Phimn, Phi, Phim = scheme.phi_timestep(*args)
    deltamn, delta, deltam = scheme.delta_timestep(*args)
    etamn, eta, etam = scheme.eta_timestep(*args)
    U, V, Um_new, Vm_new = invrsUV_with_coeffs(deltamn, etamn, fmn, I, J, M, N, Pmn, Hmn, tstepcoeffmn, marray)
    return etamn, eta, etam, deltamn, delta, deltam, Phimn, Phi, Phim, U, V, Um_new, Vm_new


def tstepping_semi_implicit(Rum_lag, Rvm_lag, etam0, etam1, deltam0, deltam1, Phim0, Phim1, I, J, M, N, Am, Bm, Cm,
                            Dm, Em, Fm, Gm, Um, Vm, fmn, Pmn, Hmn, Pmnw, Hmnw, tstepcoeff, tstepcoeff2, tstepcoeffmn,
                            marray, mJarray, narray, PhiFm, dt, a, Phibar, taurad, taudrag, forcflag, diffflag,
                            relax_implicit, sigma_exp, sigmaPhi_exp, si_alpha, test, t):
    """The semi-implicit leapfrog step with tstepping's 13-array return (time_stepping.py:387-488):
    semi_implicit_tdiff.si_timestep."""
    from .semi_implicit_tdiff import si_timestep

    return si_timestep(Rum_lag, Rvm_lag, etam0, etam1, deltam0, deltam1, Phim0, Phim1, I, J, M, N, Am, Bm, Cm, Dm, Em,
                       Fm, Gm, Um, Vm, fmn, Pmn, Hmn, Pmnw, Hmnw, tstepcoeff, tstepcoeff2, tstepcoeffmn, marray, mJarray,
                       narray, PhiFm, dt, a, Phibar, taurad, taudrag, forcflag, diffflag, relax_implicit, sigma_exp,
                       sigmaPhi_exp, si_alpha, test, t)
