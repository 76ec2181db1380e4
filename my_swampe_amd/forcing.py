"""The reference's forcing module (src/my_swampe/forcing.py): the equilibrium geopotential and the
radiative and velocity forcing of Perez-Becker and Showman (2013), with the same function names,
arguments and return values.

The time step evaluates Q and (F, G) pointwise inside the grid kernel (its FP32 or FP64 physics)
and takes Phieq from the plan's per-member table; these functions give the same fields for
callers of the standalone transforms, computing in the type of their field arguments (numpy
arrays on the host, torch tensors where they are).  ``Phieqfun`` returns float64 numpy.
"""

from __future__ import annotations

import math

import numpy as np


def Phieqfun(Phibar, DPhieq, lambdas, mus, I: int, J: int, g):
    """Equilibrium geopotential (J, I) (forcing.py:14-62): Phibar + DPhieq cos(lambda)
    sqrt(1 - mu^2) on the day side -pi/2 < lambda < pi/2 (both bounds strict), Phibar elsewhere;
    g is not used, as in the reference."""
    lam = np.asarray(lambdas, dtype=np.float64)
    mu = np.asarray(mus, dtype=np.float64)[:, None]
    base = np.full((int(J), int(I)), float(Phibar))
    day = ((lam > -math.pi / 2) & (lam < math.pi / 2))[None, :]
    return np.where(day, base + float(DPhieq) * np.cos(lam[None, :]) * np.sqrt(1 - mu ** 2), base)


def Qfun(Phieq, Phi, Phibar, taurad):
    """Radiative forcing Q = (Phieq - (Phi + Phibar)) / taurad (forcing.py:65-95)."""
    return (1 / taurad) * (Phieq - (Phi + Phibar))


def Rfun(U, V, Q, Phi, Phibar, taudrag):
    """Velocity forcing (F, G) (forcing.py:98-155): R = -(U, V) max(Q, 0) / (Phi + Phibar), the
    denominator replaced by the smallest normal double where it is zero, minus the Rayleigh drag
    (U, V) / taudrag unless taudrag == -1."""
    if hasattr(Q, "detach"):
        import torch

        Qc = torch.where(Q < 0, torch.zeros_like(Q), Q)
        tot = Phi + Phibar
        safe = torch.where(tot.abs() > 0, tot, torch.full_like(tot, torch.finfo(tot.dtype).tiny))
    else:
        Qc = np.where(Q < 0, 0.0, Q)
        tot = Phi + Phibar
        safe = np.where(np.abs(tot) > 0, tot, np.finfo(np.asarray(tot).dtype).tiny)
    Ru = -U * Qc / safe
    Rv = -V * Qc / safe
    if np.ndim(taudrag) == 0 and not hasattr(taudrag, "detach"):
        if float(taudrag) == -1:
            return Ru, Rv
        return Ru - (U / taudrag), Rv - (V / taudrag)
    # an array of drag times (per member, broadcast against the fields): -1 entries drag-free
    if hasattr(taudrag, "detach"):
        import torch as xp
    else:
        xp = np
    off = taudrag == -1
    td = xp.where(off, xp.ones_like(taudrag), taudrag)
    return xp.where(off, Ru, Ru - (U / td)), xp.where(off, Rv, Rv - (V / td))
