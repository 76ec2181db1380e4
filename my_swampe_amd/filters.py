"""The reference's filters module (src/my_swampe/filters.py): the hyperdiffusion factors per total
wavenumber and their application, with the same function names, arguments and return values.

The plans build the sixth-order factors into their FP64 coefficient tables and apply them inside
the Legendre kernel's spectral update; these functions give the same factors as (M+1, N+1)
float64 numpy arrays for callers of the standalone transforms, and ``diffusion`` multiplies in
the type of its arguments.  The arithmetic grouping of each factor is the reference's, so the
factors are bitwise the same.
"""

from __future__ import annotations

import numpy as np


def _spread(row: np.ndarray, M: int) -> np.ndarray:
    return np.broadcast_to(row[None, :], (int(M) + 1, row.size)).copy()


def _n6(N: int, a: float) -> np.ndarray:
    """(n^3 / a^3) ((n + 1)^3 / a^3), n = 0..N."""
    n = np.arange(int(N) + 1, dtype=np.float64)
    return ((n * n * n) / a ** 3) * (((n + 1) * (n + 1) * (n + 1)) / a ** 3)


def _n4(N: int, a: float) -> np.ndarray:
    """(n^2 / a^2) ((n + 1)^2 / a^2), n = 0..N."""
    n = np.arange(int(N) + 1, dtype=np.float64)
    return (n * n / a ** 2) * ((n + 1) * (n + 1) / a ** 2)


def diffusion(Ximn, sigma):
    """Ximn filtered by the factors sigma (filters.py:11-27)."""
    return Ximn * sigma


def sigma(M: int, N: int, K4: float, a: float, dt: float) -> np.ndarray:
    """Implicit fourth-order factor for vorticity and divergence, 1 / (1 + 2 dt K4 (n^2 (n+1)^2 /
    a^4 - 4 / a^4)) (filters.py:30-60)."""
    a = float(a)
    return _spread(1 / (1 + (2 * float(dt) * float(K4)) * (_n4(N, a) - 4 / a ** 4)), M)


def sigmaPhi(M: int, N: int, K4: float, a: float, dt: float) -> np.ndarray:
    """Implicit fourth-order factor for the geopotential, without the 4 / a^4 offset
    (filters.py:63-92)."""
    a = float(a)
    return _spread(1 / (1 + (2 * float(dt) * float(K4)) * _n4(N, a)), M)


def sigma6(M: int, N: int, K6: float, a: float, dt: float) -> np.ndarray:
    """Implicit sixth-order factor for vorticity and divergence, 1 / (1 + 2 dt K6 (n^3 (n+1)^3 /
    a^6 - 8 / a^6)) (filters.py:95-123)."""
    a = float(a)
    return _spread(1 / (1 + (2 * float(dt) * float(K6)) * (_n6(N, a) - 8 / a ** 6)), M)


def sigma6Phi(M: int, N: int, K6: float, a: float, dt: float) -> np.ndarray:
    """Implicit sixth-order factor for the geopotential, without the 8 / a^6 offset
    (filters.py:126-151)."""
    a = float(a)
    return _spread(1 / (1 + (2 * float(dt) * float(K6)) * _n6(N, a)), M)


def sigma6_exponential(M: int, N: int, K6: float, a: float, dt: float) -> np.ndarray:
    """Exponential sixth-order factor over a leapfrog step, exp(-2 dt K6 max(n^3 (n+1)^3 / a^6 -
    8 / a^6, 0)) (filters.py:154-182; the semi-implicit scheme's filter)."""
    a = float(a)
    return _spread(np.exp(-(2 * float(dt) * float(K6)) * np.maximum(_n6(N, a) - 8 / a ** 6, 0.0)), M)


def sigma6Phi_exponential(M: int, N: int, K6: float, a: float, dt: float) -> np.ndarray:
    """Exponential sixth-order factor for the geopotential, without the offset
    (filters.py:185-213)."""
    a = float(a)
    return _spread(np.exp(-(2 * float(dt) * float(K6)) * _n6(N, a)), M)
