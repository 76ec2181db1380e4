"""FP64 NumPy restatement of the SWAMPE-JAX spectral time-step core.

TEST INFRASTRUCTURE ONLY.  This module is the oracle: ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` use it
to check (or time beside) the HIP path.  Nothing in ``my_swampe_amd`` imports
it, and the HIP product path never falls back to it.

Every function restates the algorithm of the reference file:line it cites
(paths relative to the reference repository root).  The arithmetic follows the
reference step by step (complex128 Fourier fields, rectangular (M+1, N+1)
spectral arrays, einsum Legendre contractions, numpy FFTs), so that the oracle
reproduces the reference's rounding behaviour closely enough to meet its own
parity contract against upstream SWAMPE fixtures (eta/delta 1e-10, Phi 5e-8,
U/V 1e-9 atol; unit_tests/test_parity_reference_regression.py:58-64).

Pinning (each function is checked in tests/ by something other than a retyping
of its own formula):

* geometry (build_lambdas, gauss_legendre, scaling_table, pmn_hmn): numpy
  ``leggauss``, the closed forms Pmn[:,2,2]=(sqrt15/4)(1-mu^2) and
  Hmn[:,0,1]=(sqrt6/2)(1-mu^2), Gauss-quadrature orthonormality, the
  equatorial parity P(-mu)=(-1)^(n+m) P(mu), and an independent normalized
  recurrence written in C++ (the product's own basis builder).
* transforms (fwd_fft_trunc, invrs_fft, fwd_leg, invrs_leg, invrs_uv,
  diagnostic_eta_delta): a brute-force DFT sum on a small grid, the Coriolis
  closed form fmn[0,1]=omega/sqrt(0.375), the forward/inverse and wind
  inversion round trips of unit_tests/test_transform_stack.py.
* forcing / filters / tendencies / step / driver: the two upstream-SWAMPE
  fixtures (terminal fields, snapshots, spinup/geopot diagnostics and the
  phase-curve projection), the quirk formulas of
  unit_tests/test_parity_quirks.py, and member-independence.
* phase_curve / phase_curve_vjp: the fixtures' phase curves, the closed form of a
  uniform grid at grid-aligned phases (cot(pi/I) sum_j sqrt(1 - mu_j^2)), the
  longitude-roll shift property, and central differences of phase_curve for the VJP.
* explicit scheme (expflag=True; no reference fixture exists for it): the delta
  carry-only quirk (test_parity_quirks.py:85-135) and identities of the
  mathematics with delta_mn, zeta_mn from diagnostic_eta_delta: the flux terms of
  a uniform scalar (Phi from C, D; eta from A, B) equal -2 dt q0 delta_mn(U, V),
  the continuity term is -2 dt Phibar delta_mn, and drag-only forcing (quirk 5)
  moves delta by -(4 dt / taudrag) delta_mn(U, V) and eta by 0 (zonal winds) or
  -(4 dt / taudrag) zeta_mn(0, V) (meridional winds).
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field, replace
from typing import Dict, Optional, Tuple

import numpy as np
import scipy.special as sp

__all__ = [
    "build_lambdas", "gauss_legendre", "scaling_table", "pmn_hmn", "spectral_params",
    "fwd_fft_trunc", "invrs_fft", "fwd_leg", "invrs_leg", "invrs_uv", "diagnostic_eta_delta",
    "Static", "OracleState", "build_static", "init_state_from_fields", "analytic_ic",
    "diagnose_winds", "step_once", "run_model_scan", "run_model", "rms_winds", "phase_curve", "phase_curve_vjp",
]

# ----------------------------------------------------------------------------
# Geometry (spectral_transform.py:54-287, initial_conditions.py:50-85)
# ----------------------------------------------------------------------------


def build_lambdas(I: int) -> np.ndarray:
    """Longitudes linspace(-pi, pi, I, endpoint=False) (spectral_transform.py:54-77)."""
    return np.linspace(-np.pi, np.pi, num=int(I), endpoint=False)


def gauss_legendre(J: int) -> Tuple[np.ndarray, np.ndarray]:
    """Gauss-Legendre nodes (ascending) and weights via scipy (spectral_transform.py:80-114)."""
    mus, w = sp.roots_legendre(int(J))
    return np.asarray(mus, np.float64), np.asarray(w, np.float64)


def scaling_table(M: int, N: int) -> np.ndarray:
    """sqrt((2n+1)(n-m)!/(2(n+m)!)) with exact integer factorials (spectral_transform.py:117-146)."""
    s = np.zeros((M + 1, N + 1))
    for m in range(M + 1):
        for n in range(m, N + 1):
            s[m, n] = math.sqrt(((2 * n + 1) * math.factorial(n - m)) / (2 * math.factorial(n + m)))
    return s


def pmn_hmn(J: int, M: int, N: int, mus: np.ndarray) -> Tuple[np.ndarray, np.ndarray]:
    """Normalized associated Legendre basis Pmn and Hmn=(1-mu^2)dPmn/dmu, shape (J, M+1, N+1).

    scipy's assoc_legendre_p_all (unnormalized, Condon-Shortley phase), scaled by
    scaling_table and sign-flipped for odd m (spectral_transform.py:219-287).
    """
    P = np.zeros((J, M + 1, N + 1))
    H = np.zeros((J, M + 1, N + 1))
    for j, mu in enumerate(np.asarray(mus, np.float64)):
        a = np.asarray(sp.assoc_legendre_p_all(N, M, float(mu), diff_n=1), np.float64)
        P[j] = a[0, : N + 1, : M + 1].T
        H[j] = (1.0 - mu * mu) * a[1, : N + 1, : M + 1].T
    s = scaling_table(M, N)
    P *= s[None]
    H *= s[None]
    odd = (np.arange(M + 1) % 2) == 1
    P[:, odd, 1:] *= -1.0
    H[:, odd, 1:] *= -1.0
    return P, H


RESOLUTIONS = {42: (64, 128, 1200), 63: (96, 192, 900), 106: (160, 320, 600)}


def spectral_params(M: int):
    """(N, I, J, dt_default, lambdas, mus, w) for M in {42, 63, 106} (initial_conditions.py:50-85)."""
    if M not in RESOLUTIONS:
        raise ValueError(f"Unsupported value of M={M}. Only 42, 63, and 106 are supported.")
    J, I, dt = RESOLUTIONS[M]
    mus, w = gauss_legendre(J)
    return int(M), I, J, dt, build_lambdas(I), mus, w


# ----------------------------------------------------------------------------
# Transforms (spectral_transform.py:290-650)
# ----------------------------------------------------------------------------


def fwd_fft_trunc(data: np.ndarray, I: int, M: int) -> np.ndarray:
    """fft(data/I) along the last axis, truncated to m=0..M (spectral_transform.py:290-328)."""
    return np.fft.fft(np.asarray(data) / I, n=I, axis=-1)[..., : M + 1]


def invrs_fft(X: np.ndarray, I: int) -> np.ndarray:
    """ifft(I*X) along the longitude axis (spectral_transform.py:331-348)."""
    return np.fft.ifft(I * X, n=I, axis=-1)


def fwd_leg(data: np.ndarray, Bw: np.ndarray) -> np.ndarray:
    """S[..., m, n] = sum_j data[..., j, m] Bw[j, m, n] (spectral_transform.py:370-402)."""
    return np.einsum("...jm,jmn->...mn", data, Bw)


def invrs_leg(coeff: np.ndarray, I: int, J: int, M: int, N: int, B: np.ndarray) -> np.ndarray:
    """Hermitian (J, I) Fourier layout: +m at [:, :M+1], conj(-m) at [:, I-M:] (spectral_transform.py:438-481)."""
    out = np.zeros((J, I), dtype=np.complex128)
    out[:, : M + 1] = np.einsum("jmn,mn->jm", B, coeff)
    if M > 0:
        neg = np.einsum("jmn,mn->jm", B[:, 1:, :], np.conj(coeff[1:, :]))
        out[:, I - M:] = neg[:, ::-1]
    return out


def invrs_uv(deltamn, etamn, fmn, I, J, M, N, Pmn, Hmn, tcmn, marray):
    """Winds from spectral divergence/vorticity; also the truncated Fourier winds.

    spectral_transform.py:535-596 (invrsUV_with_coeffs; invrsUV :484-532 is the
    same without the Fourier outputs).
    """
    d = deltamn.copy()
    e = etamn.copy()
    d[:, 0] = 0.0
    e[:, 0] = 0.0
    ds = d * tcmn
    es = (e - fmn) * tcmn
    Um_full = -(invrs_leg(1j * (marray * ds), I, J, M, N, Pmn) - invrs_leg(es, I, J, M, N, Hmn))
    Vm_full = -(invrs_leg(1j * (marray * es), I, J, M, N, Pmn) + invrs_leg(ds, I, J, M, N, Hmn))
    return invrs_fft(Um_full, I), invrs_fft(Vm_full, I), Um_full[:, : M + 1], Vm_full[:, : M + 1]


def diagnostic_eta_delta(Um, Vm, fmn, I, J, M, N, Pmn, Hmn, w, tstepcoeff, mJarray, dt):
    """Inverse of invrs_uv: Fourier winds -> (eta, delta) grids and spectra (spectral_transform.py:599-650)."""
    coeff = tstepcoeff / (2.0 * dt)
    Pw = w[:, None, None] * Pmn
    Hw = w[:, None, None] * Hmn
    zetamn = fwd_leg(coeff * 1j * mJarray * Vm, Pw) + fwd_leg(coeff * Um, Hw)
    etamn = zetamn + fmn
    deltamn = fwd_leg(coeff * 1j * mJarray * Um, Pw) - fwd_leg(coeff * Vm, Hw)
    newdelta = invrs_fft(invrs_leg(deltamn, I, J, M, N, Pmn), I)
    neweta = invrs_fft(invrs_leg(etamn, I, J, M, N, Pmn), I)
    return neweta, newdelta, etamn, deltamn


# ----------------------------------------------------------------------------
# Coefficients, filters, forcing, initial conditions
# ----------------------------------------------------------------------------


def tstepcoeffmn(M, N, a):
    """a/(n(n+1)) with n=0 zeroed, broadcast (M+1, N+1) (time_stepping.py:491-501)."""
    n = np.arange(N + 1, dtype=np.float64)
    c = n * (n + 1)
    c[0] = 1.0
    t = a / c
    t[0] = 0.0
    return np.broadcast_to(t[None, :], (M + 1, N + 1)).copy()


def tstepcoeff(J, M, dt, mus, a):
    """2 dt / (a (1-mu^2)) broadcast (J, M+1) (time_stepping.py:516-522)."""
    base = (2.0 * dt) / (a * (1.0 - mus[:, None] ** 2))
    return np.broadcast_to(base, (J, M + 1)).copy()


def tstepcoeff2(J, M, dt, a):
    """2 dt / a^2 filled (J, M+1) (time_stepping.py:504-506)."""
    return np.full((J, M + 1), 2.0 * dt / (a ** 2))


def narray(M, N):
    """n(n+1) broadcast (M+1, N+1) (time_stepping.py:509-513)."""
    n = np.arange(N + 1, dtype=np.float64)
    return np.broadcast_to((n * (n + 1))[None, :], (M + 1, N + 1)).copy()


def mJarray(J, M):
    """m broadcast (J, M+1) (time_stepping.py:525-528)."""
    return np.broadcast_to(np.arange(M + 1, dtype=np.float64)[None, :], (J, M + 1)).copy()


def marray(M, N):
    """m broadcast (M+1, N+1) (time_stepping.py:531-534)."""
    return np.broadcast_to(np.arange(M + 1, dtype=np.float64)[:, None], (M + 1, N + 1)).copy()


def sigma6(M, N, K6, a, dt):
    """Implicit 6th-order hyperdiffusion for eta/delta (filters.py:95-123)."""
    n = np.arange(N + 1, dtype=np.float64)
    ncoeff = ((n * n * n) / a ** 3) * (((n + 1) * (n + 1) * (n + 1)) / a ** 3)
    s = 1.0 / (1.0 + (2.0 * dt * K6) * (ncoeff - 8.0 / a ** 6))
    return np.broadcast_to(s[None, :], (M + 1, N + 1)).copy()


def sigma6Phi(M, N, K6, a, dt):
    """Implicit 6th-order hyperdiffusion for Phi, no 8/a^6 offset (filters.py:126-151)."""
    n = np.arange(N + 1, dtype=np.float64)
    ncoeff = ((n * n * n) / a ** 3) * (((n + 1) * (n + 1) * (n + 1)) / a ** 3)
    s = 1.0 / (1.0 + (2.0 * dt * K6) * ncoeff)
    return np.broadcast_to(s[None, :], (M + 1, N + 1)).copy()


def sigma6_exponential(M, N, K6, a, dt):
    """Exponential (integrating-factor) 6th-order hyperdiffusion for eta/delta over a leapfrog
    step, offset clamped at zero (filters.py:154-182)."""
    n = np.arange(N + 1, dtype=np.float64)
    ncoeff = ((n * n * n) / a ** 3) * (((n + 1) * (n + 1) * (n + 1)) / a ** 3)
    s = np.exp(-(2.0 * dt * K6) * np.maximum(ncoeff - 8.0 / a ** 6, 0.0))
    return np.broadcast_to(s[None, :], (M + 1, N + 1)).copy()


def sigma6Phi_exponential(M, N, K6, a, dt):
    """Exponential hyperdiffusion for Phi, no offset (filters.py:185-213)."""
    n = np.arange(N + 1, dtype=np.float64)
    ncoeff = ((n * n * n) / a ** 3) * (((n + 1) * (n + 1) * (n + 1)) / a ** 3)
    s = np.exp(-(2.0 * dt * K6) * ncoeff)
    return np.broadcast_to(s[None, :], (M + 1, N + 1)).copy()


def coriolismn(M, omega):
    """f_mn with only [0, 1] = omega/sqrt(0.375) (initial_conditions.py:295-316)."""
    f = np.zeros((M + 1, M + 1))
    f[0, 1] = omega / np.sqrt(0.375)
    return f


def phieqfun(Phibar, DPhieq, lambdas, mus):
    """Dayside equilibrium geopotential, strict -pi/2 < lambda < pi/2 (forcing.py:14-62)."""
    lam = lambdas[None, :]
    mu = mus[:, None]
    day = ((lambdas > -np.pi / 2) & (lambdas < np.pi / 2))[None, :]
    base = np.full((mus.size, lambdas.size), float(Phibar))
    term = DPhieq * np.cos(lam) * np.sqrt(1 - mu ** 2)
    return np.where(day, base + term, base)


def qfun(Phieq, Phi, Phibar, taurad):
    """Q = (1/taurad)(Phieq - (Phi + Phibar)) (forcing.py:65-95)."""
    return (1 / taurad) * (Phieq - (Phi + Phibar))


def rfun(U, V, Q, Phi, Phibar, taudrag):
    """Velocity forcing with the Q<0 clamp and taudrag == -1 sentinel (forcing.py:98-155)."""
    Qc = np.where(Q < 0, 0.0, Q)
    tot = Phi + Phibar
    safe = np.where(np.abs(tot) > 0, tot, np.finfo(np.float64).tiny)
    Ru = -U * Qc / safe
    Rv = -V * Qc / safe
    if taudrag == -1:
        return Ru, Rv
    return Ru - (U / taudrag), Rv - (V / taudrag)


def test1_init(a, omega, a1):
    """Williamson test 1/2 amplitudes (initial_conditions.py:18-47)."""
    SU0 = 2.0 * np.pi * a / (3600.0 * 24.0 * 12.0)
    return SU0, np.sin(a1), np.cos(a1), 2.0 * ((SU0 / a) + omega), (SU0 * a * omega + 0.5 * SU0 ** 2)


def state_var_init(I, J, mus, lambdas, test, etaamp, *args):
    """Analytic eta/delta/Phi initial fields (initial_conditions.py:88-174)."""
    mu = mus[:, None]
    lam = lambdas[None, :]
    s1m = np.sqrt(1.0 - mu ** 2)
    delta0 = np.zeros((J, I))
    Phi0 = np.zeros((J, I))
    if test is not None:
        a, sina, cosa, Phibar, Phiamp = args
    if test == 1:
        arg = -np.cos(lam) * s1m * sina + mu * cosa
        eta0 = etaamp * arg
        bumpr = a / 3.0
        dist_arg = 0.0 * mu + np.cos(np.arcsin(0.0)) * np.cos(np.arcsin(mu)) * np.cos(lam - 3.0 * np.pi / 2.0)
        dist = a * np.arccos(np.clip(dist_arg, -1.0, 1.0))
        bump = (Phibar / 2.0) * (1.0 + np.cos(np.pi * dist / bumpr))
        Phi0 = np.where(dist < bumpr, bump, 0.0)
    elif test == 2:
        arg = -np.cos(lam) * s1m * sina + mu * cosa
        eta0 = etaamp * arg
        Phi0 = (Phibar - Phiamp) * (arg ** 2)
    else:
        eta0 = etaamp * np.broadcast_to(mu, (J, I)).copy()
    return eta0, delta0, Phi0


def velocity_init(I, J, SU0, cosa, sina, mus, lambdas, test):
    """Analytic initial winds (initial_conditions.py:177-239)."""
    mu = mus[:, None]
    lam = lambdas[None, :]
    s1m = np.sqrt(1.0 - mu ** 2)
    if test == 1:
        U = SU0 * (s1m * cosa + mu * np.cos(lam) * sina) * s1m
        V = -SU0 * np.sin(lam) * sina * s1m
    elif test == 2:
        U = SU0 * (s1m * cosa + np.cos(lam) * mu * sina)
        V = np.broadcast_to(-SU0 * (np.sin(lam) * sina), (J, I)).copy()
    else:
        U = np.zeros((J, I))
        V = np.zeros((J, I))
    return U, V


def abcde(U, V, eta, Phi, mus):
    """A=U eta, B=V eta, C=U Phi, D=V Phi, E=(U^2+V^2)/(2(1-mu^2)) (initial_conditions.py:242-292)."""
    denom = 2.0 * (1.0 - mus[:, None] ** 2)
    return U * eta, V * eta, U * Phi, V * Phi, (U * U + V * V) / denom


def rms_winds(a, I, J, lambdas, mus, U, V):
    """Area-weighted RMS wind (time_stepping.py:537-576)."""
    phis = np.arcsin(mus)[:, None]
    dl = lambdas[2] - lambdas[1]
    dp = phis[2, 0] - phis[1, 0]
    area_planet = 4.0 * np.pi * a ** 2
    area = (a ** 2) * (np.sin(phis + np.pi / 2.0) ** 2) * dp * dl
    integrand = (U / np.cos(phis)) ** 2 + (V / np.cos(phis)) ** 2
    return float(np.sqrt(np.sum(area * integrand / area_planet)))


# ----------------------------------------------------------------------------
# Static + state (model.py:62-86, :228-557, :694-772)
# ----------------------------------------------------------------------------

_GEOM_CACHE: Dict[int, tuple] = {}


def cached_geometry(M: int):
    """Quadrature + basis per truncation (model.py:62-86)."""
    if M not in _GEOM_CACHE:
        N, I, J, _, lambdas, mus, w = spectral_params(M)
        Pmn, Hmn = pmn_hmn(J, M, N, mus)
        _GEOM_CACHE[M] = (N, I, J, lambdas, mus, w, Pmn, Hmn)
    return _GEOM_CACHE[M]


@dataclass
class Static:
    """Geometry, operators and constants of one run (model.py:228-378)."""
    M: int
    N: int
    I: int
    J: int
    dt: float
    a: float
    omega: float
    g: float
    Phibar: float
    taurad: float
    taudrag: float
    lambdas: np.ndarray
    mus: np.ndarray
    w: np.ndarray
    Pmn: np.ndarray
    Hmn: np.ndarray
    Pmnw: np.ndarray
    Hmnw: np.ndarray
    fmn: np.ndarray
    tstepcoeff: np.ndarray
    tstepcoeff2: np.ndarray
    tstepcoeffmn: np.ndarray
    marray: np.ndarray
    mJarray: np.ndarray
    narray: np.ndarray
    sigma: np.ndarray
    sigmaPhi: np.ndarray
    Phieq: np.ndarray
    sigma_exp: np.ndarray = None
    sigmaPhi_exp: np.ndarray = None


def build_static(*, M, dt, a, omega, g, Phibar, taurad, taudrag, DPhieq, K6, K6Phi, test) -> Static:
    """Assemble the Static bundle (model.py:457-557)."""
    N, I, J, lambdas, mus, w, Pmn, Hmn = cached_geometry(int(M))
    K6Phi_eff = K6 if K6Phi is None else K6Phi
    Phieq = phieqfun(Phibar, DPhieq, lambdas, mus) if test is None else np.zeros((J, I))
    return Static(
        M=int(M), N=N, I=I, J=J, dt=float(dt), a=float(a), omega=float(omega), g=float(g),
        Phibar=float(Phibar), taurad=float(taurad), taudrag=float(taudrag),
        lambdas=lambdas, mus=mus, w=w, Pmn=Pmn, Hmn=Hmn,
        Pmnw=w[:, None, None] * Pmn, Hmnw=w[:, None, None] * Hmn,
        fmn=coriolismn(M, omega),
        tstepcoeff=tstepcoeff(J, M, dt, mus, a), tstepcoeff2=tstepcoeff2(J, M, dt, a),
        tstepcoeffmn=tstepcoeffmn(M, N, a), marray=marray(M, N), mJarray=mJarray(J, M), narray=narray(M, N),
        sigma=sigma6(M, N, K6, a, dt), sigmaPhi=sigma6Phi(M, N, K6Phi_eff, a, dt), Phieq=Phieq,
        sigma_exp=sigma6_exponential(M, N, K6, a, dt), sigmaPhi_exp=sigma6Phi_exponential(M, N, K6Phi_eff, a, dt),
    )


@dataclass
class OracleState:
    """The 27-field scan carry with the reference's field names (model.py:380-425)."""
    etam_prev: np.ndarray
    etam_curr: np.ndarray
    deltam_prev: np.ndarray
    deltam_curr: np.ndarray
    Phim_prev: np.ndarray
    Phim_curr: np.ndarray
    eta_prev: np.ndarray
    eta_curr: np.ndarray
    delta_prev: np.ndarray
    delta_curr: np.ndarray
    Phi_prev: np.ndarray
    Phi_curr: np.ndarray
    U_curr: np.ndarray
    V_curr: np.ndarray
    Um_curr: np.ndarray
    Vm_curr: np.ndarray
    Am_curr: np.ndarray
    Bm_curr: np.ndarray
    Cm_curr: np.ndarray
    Dm_curr: np.ndarray
    Em_curr: np.ndarray
    PhiFm_curr: np.ndarray
    Fm_curr: np.ndarray
    Gm_curr: np.ndarray
    Rum_lag: np.ndarray
    Rvm_lag: np.ndarray
    dead: bool = False
    # Spectral coefficients of the last update (not part of the reference carry;
    # kept for inspection by tests).
    spec: Optional[dict] = field(default=None, repr=False)


def forcing_phys(static: Static, test, Phi, U, V):
    """(PhiF, F, G); computed whenever test is None, regardless of forcflag (model.py:560-603)."""
    if test is None:
        Q = qfun(static.Phieq, Phi, static.Phibar, static.taurad)
        F, G = rfun(U, V, Q, Phi, static.Phibar, static.taudrag)
        return Q, F, G
    z = np.zeros((static.J, static.I))
    return z, z, z


def nonlinear_spectral(static: Static, eta, Phi, U, V):
    """Truncated Fourier of A..E (model.py:606-641)."""
    A, B, C, D, E = abcde(U, V, eta, Phi, static.mus)
    return fwd_fft_trunc(np.stack((A, B, C, D, E)), static.I, static.M)


def diagnose_winds(eta, delta, static: Static):
    """U, V from eta, delta by forward transforms + invrs_uv (model.py:644-666)."""
    etam, deltam = fwd_fft_trunc(np.stack((eta, delta)), static.I, static.M)
    etamn = fwd_leg(etam, static.Pmnw)
    deltamn = fwd_leg(deltam, static.Pmnw)
    U, V, _, _ = invrs_uv(deltamn, etamn, static.fmn, static.I, static.J, static.M, static.N,
                          static.Pmn, static.Hmn, static.tstepcoeffmn, static.marray)
    return np.real(U), np.real(V)


def analytic_ic(static: Static, test, a1):
    """Analytic (eta0, delta0, Phi0, U0, V0) (model.py:669-691)."""
    SU0, sina, cosa, etaamp, Phiamp = test1_init(static.a, static.omega, a1)
    if test in (1, 2):
        eta0, delta0, Phi0 = state_var_init(static.I, static.J, static.mus, static.lambdas, test, etaamp,
                                            static.a, sina, cosa, static.Phibar, Phiamp)
    else:
        eta0, delta0, Phi0 = state_var_init(static.I, static.J, static.mus, static.lambdas, test, etaamp)
    U0, V0 = velocity_init(static.I, static.J, SU0, cosa, sina, static.mus, static.lambdas, test)
    return eta0, delta0, Phi0, U0, V0


def init_state_from_fields(static: Static, test, eta0, delta0, Phi0, U0, V0) -> OracleState:
    """Two-level start prev == curr == initial (model.py:694-772)."""
    PhiF0, F0, G0 = forcing_phys(static, test, Phi0, U0, V0)
    etam0, deltam0, Phim0, Um0, Vm0 = fwd_fft_trunc(np.stack((eta0, delta0, Phi0, U0, V0)), static.I, static.M)
    PhiFm0, Fm0, Gm0 = fwd_fft_trunc(np.stack((PhiF0, F0, G0)), static.I, static.M)
    Am0, Bm0, Cm0, Dm0, Em0 = nonlinear_spectral(static, eta0, Phi0, U0, V0)
    inv_td = 0.0 if static.taudrag == -1.0 else 1.0 / static.taudrag
    return OracleState(
        etam_prev=etam0, etam_curr=etam0, deltam_prev=deltam0, deltam_curr=deltam0,
        Phim_prev=Phim0, Phim_curr=Phim0,
        eta_prev=eta0, eta_curr=eta0, delta_prev=delta0, delta_curr=delta0, Phi_prev=Phi0, Phi_curr=Phi0,
        U_curr=U0, V_curr=V0, Um_curr=Um0, Vm_curr=Vm0,
        Am_curr=Am0, Bm_curr=Bm0, Cm_curr=Cm0, Dm_curr=Dm0, Em_curr=Em0,
        PhiFm_curr=PhiFm0, Fm_curr=Fm0, Gm_curr=Gm0,
        Rum_lag=Fm0 + Um0 * inv_td, Rvm_lag=Gm0 + Vm0 * inv_td, dead=False,
    )


# ----------------------------------------------------------------------------
# Tendencies (modEuler_tdiff.py, explicit_tdiff.py, time_stepping.py:16-384)
# ----------------------------------------------------------------------------


def _modeuler(st: OracleState, s: Static, forcflag: bool, diffflag: bool):
    """Modified-Euler spectral updates with the locked quirks (modEuler_tdiff.py:30-268)."""
    Pw, Hw, nn, mJ = s.Pmnw, s.Hmnw, s.narray, s.mJarray
    dt, a, Phibar = s.dt, s.a, s.Phibar
    Am, Bm, Cm, Dm, Em = st.Am_curr, st.Bm_curr, st.Cm_curr, st.Dm_curr, st.Em_curr
    Fm, Gm, PhiFm = st.Fm_curr, st.Gm_curr, st.PhiFm_curr
    etam1, deltam1, Phim1 = st.etam_curr, st.deltam_curr, st.Phim_curr
    # phi_timestep (:30-117): /4 quirk on both coefficients.
    t1 = s.tstepcoeff / 4.0
    t2 = s.tstepcoeff2 / 4.0
    Beff = Bm + Fm if forcflag else Bm
    Aeff = Am - Gm if forcflag else Am
    c1, c2, dleg, d2, d5 = fwd_leg(np.stack((Phim1, t1 * 1j * mJ * Cm, deltam1, t1 * 1j * mJ * Beff, t2 * Em)), Pw)
    c3, d3 = fwd_leg(np.stack((t1 * Dm, t1 * Aeff)), Hw)
    Phimn = c1 - c2 + c3 - dt * Phibar * dleg - Phibar * 0.5 * (d2 + d3 + nn * d5 + (1.0 / a ** 2) * (nn * c1))
    if forcflag:
        Phimn = Phimn + fwd_leg(dt * PhiFm, Pw)
    if diffflag:
        Phimn = Phimn * s.sigmaPhi
    # delta_timestep (:120-208): always the forced A/B combination.
    Bf = Bm + Fm
    Af = Am - Gm
    e1, e2, e4, e5, p2 = fwd_leg(np.stack((deltam1, t1 * 1j * mJ * Bf, t2 * Phim1, t2 * Em, t1 * 1j * mJ * Cm)), Pw)
    e3, p3 = fwd_leg(np.stack((t1 * Af, t1 * Dm)), Hw)
    deltamn = e1 + e2 + e3 + nn * e4 + nn * e5 + (nn * (p2 + p3)) / (2.0 * a ** 2) - Phibar * (nn * e1) / (a ** 2)
    if forcflag:
        deltamn = deltamn + (nn * fwd_leg((dt / 2.0) * PhiFm, Pw)) / (a ** 2)
    if diffflag:
        deltamn = deltamn * s.sigma
    # eta_timestep (:211-268): unscaled coefficient when forced, /2 otherwise.
    te = s.tstepcoeff if forcflag else s.tstepcoeff / 2.0
    h1, h2 = fwd_leg(np.stack((etam1, te * 1j * mJ * Aeff)), Pw)
    etamn = h1 - h2 + fwd_leg(te * Beff, Hw)
    if diffflag:
        etamn = etamn * s.sigma
    return etamn, deltamn, Phimn


def _explicit(st: OracleState, s: Static, forcflag: bool, diffflag: bool):
    """Explicit leapfrog updates with quirks 4-5 (explicit_tdiff.py:20-233)."""
    Pw, Hw, mJ = s.Pmnw, s.Hmnw, s.mJarray
    dt, Phibar, tau = s.dt, s.Phibar, s.taudrag
    t1 = s.tstepcoeff
    Am, Bm, Cm, Dm = st.Am_curr, st.Bm_curr, st.Cm_curr, st.Dm_curr
    Fm, Gm, PhiFm, Um, Vm = st.Fm_curr, st.Gm_curr, st.PhiFm_curr, st.Um_curr, st.Vm_curr
    c1, c2, dleg = fwd_leg(np.stack((st.Phim_prev, t1 * 1j * mJ * Cm, st.deltam_curr)), Pw)
    Phimn = c1 - c2 + fwd_leg(t1 * Dm, Hw) - 2.0 * dt * Phibar * dleg
    if forcflag:
        Phimn = Phimn + fwd_leg(2.0 * dt * PhiFm, Pw)
    if diffflag:
        Phimn = Phimn * s.sigmaPhi
    deltamn = fwd_leg(st.deltam_prev, Pw)
    if forcflag:
        f1, f3 = fwd_leg(np.stack(((t1 * 1j * mJ * Um) / tau, t1 * 1j * mJ * Fm)), Pw)
        f2, f4 = fwd_leg(np.stack(((t1 * Vm) / tau, t1 * Gm)), Hw)
        deltamn = deltamn + (-f1 + f2 + f3 - f4)
    if diffflag:
        deltamn = deltamn * s.sigma
    h1, h2 = fwd_leg(np.stack((st.etam_prev, t1 * 1j * mJ * Am)), Pw)
    etamn = h1 - h2 + fwd_leg(t1 * Bm, Hw)
    if forcflag:
        g1, g3 = fwd_leg(np.stack(((t1 * 1j * mJ * Vm) / tau, t1 * 1j * mJ * Gm)), Pw)
        g2, g4 = fwd_leg(np.stack(((t1 * Um) / tau, t1 * Fm)), Hw)
        etamn = etamn + (-g1 + g2 + g3 + g4)
    if diffflag:
        etamn = etamn * s.sigma
    return etamn, deltamn, Phimn


def si_solve(delta_star, Phi_star, b, c, kd, kr):
    """Closed-form per-degree implicit solve of (1 + kd) delta - b Phi = delta*, c delta + (1 + kr) Phi = Phi*
    (semi_implicit_tdiff.py:232-236)."""
    delta = (delta_star + b * Phi_star / (1.0 + kr)) / ((1.0 + kd) + b * c / (1.0 + kr))
    return delta, (Phi_star - c * delta) / (1.0 + kr)


def _semi_implicit(st: OracleState, s: Static, forcflag: bool, diffflag: bool, relax_implicit: bool,
                   si_alpha: float):
    """Semi-implicit leapfrog with implicit gravity waves and linear relaxation/drag, closed-form
    per-degree solve, exponential hyperdiffusion (semi_implicit_tdiff.py:79-264)."""
    Pw, Hw, mJ, fmn = s.Pmnw, s.Hmnw, s.mJarray, s.fmn
    dt, a, Phibar = s.dt, s.a, s.Phibar
    two_dt = 2.0 * dt
    lam = s.narray / a ** 2
    if relax_implicit:
        inv_td = 0.0 if s.taudrag == -1.0 else 1.0 / s.taudrag
        inv_tr = 1.0 / s.taurad
        Fe, Ge = st.Rum_lag, st.Rvm_lag
        PhiFe = st.PhiFm_curr + st.Phim_curr * inv_tr
    else:
        inv_td = inv_tr = 0.0
        Fe, Ge, PhiFe = st.Fm_curr, st.Gm_curr, st.PhiFm_curr
    Aeff = st.Am_curr - Ge if forcflag else st.Am_curr
    Beff = st.Bm_curr + Fe if forcflag else st.Bm_curr
    tc = s.tstepcoeff
    etamn0, deltamn0, Phimn0, imA, imB, imC, Emn = fwd_leg(np.stack((
        st.etam_prev, st.deltam_prev, st.Phim_prev, tc * 1j * mJ * Aeff, tc * 1j * mJ * Beff,
        tc * 1j * mJ * st.Cm_curr, st.Em_curr)), Pw)
    HA, HB, HD = fwd_leg(np.stack((tc * Aeff, tc * Beff, tc * st.Dm_curr)), Hw)
    xi = two_dt * si_alpha
    expl = two_dt * (1.0 - si_alpha)
    kd, kr = xi * inv_td, xi * inv_tr
    gd, gr = expl * inv_td, expl * inv_tr
    eta_star = etamn0 - imA + HB - gd * (etamn0 - fmn)
    etamn = (eta_star + kd * fmn) / (1.0 + kd)
    delta_star = deltamn0 + imB + HA + two_dt * lam * Emn + expl * lam * Phimn0 - gd * deltamn0
    Phi_star = Phimn0 - imC + HD - expl * Phibar * deltamn0 - gr * Phimn0
    if forcflag:
        Phi_star = Phi_star + fwd_leg(two_dt * PhiFe, Pw)
    deltamn, Phimn = si_solve(delta_star, Phi_star, xi * lam, xi * Phibar, kd, kr)
    if diffflag:
        etamn, deltamn, Phimn = etamn * s.sigma_exp, deltamn * s.sigma_exp, Phimn * s.sigmaPhi_exp
    return etamn, deltamn, Phimn


def tstepping(st: OracleState, s: Static, forcflag: bool, diffflag: bool, expflag: bool, semi_implicit=False,
              relax_implicit=False, si_alpha=0.5):
    """13-tuple scheme dispatch (time_stepping.py:16-384, :387-462)."""
    if semi_implicit:
        etamn, deltamn, Phimn = _semi_implicit(st, s, forcflag, diffflag, relax_implicit, si_alpha)
    else:
        etamn, deltamn, Phimn = (_explicit if expflag else _modeuler)(st, s, forcflag, diffflag)
    I, J, M, N = s.I, s.J, s.M, s.N
    out = {}
    for name, mn in (("eta", etamn), ("delta", deltamn), ("Phi", Phimn)):
        full = invrs_leg(mn, I, J, M, N, s.Pmn)
        out[name + "m"] = full[:, : M + 1]
        out[name] = invrs_fft(full, I)
        out[name + "mn"] = mn
    U, V, Um, Vm = invrs_uv(deltamn, etamn, s.fmn, I, J, M, N, s.Pmn, s.Hmn, s.tstepcoeffmn, s.marray)
    out.update(U=U, V=V, Um=Um, Vm=Vm)
    return out


# ----------------------------------------------------------------------------
# Step + drivers (model.py:795-1194, :1237-1765, :1954-2219)
# ----------------------------------------------------------------------------


def step_once(st: OracleState, t: int, s: Static, flags: dict, test, Uic, Vic):
    """One step with the RA / RAW filter, blow-up gate, lagged-forcing pipeline and next-step
    forcing (model.py:795-1194)."""
    diag = flags["diagnostics"]
    si = flags.get("semi_implicit", False)
    raw = flags.get("raw_filter", False)
    if diag:
        rms = rms_winds(s.a, s.I, s.J, s.lambdas, s.mus, st.U_curr, st.V_curr)
        spin_min = float(np.min(np.sqrt(st.U_curr * st.U_curr + st.V_curr * st.V_curr)))
        dead_next = bool(st.dead or rms > flags["blowup_rms"])
    else:
        rms, spin_min, dead_next = 0.0, 0.0, st.dead
    if diag and dead_next:
        out = dict(t=t, dead=True, eta=st.eta_curr, delta=st.delta_curr, Phi=st.Phi_curr, U=st.U_curr,
                   V=st.V_curr, rms=rms, spin_min=spin_min, phi_min=float(np.min(st.Phi_curr)),
                   phi_max=float(np.max(st.Phi_curr)))
        return replace(st, dead=True), out
    r = tstepping(st, s, flags["forcflag"], flags["diffflag"], flags["expflag"], semi_implicit=si,
                  relax_implicit=bool(flags["forcflag"] and test is None), si_alpha=flags.get("si_alpha", 0.5))
    neweta, newdelta, newPhi = np.real(r["eta"]), np.real(r["delta"]), np.real(r["Phi"])
    newU, newV, newUm, newVm = np.real(r["U"]), np.real(r["V"]), r["Um"], r["Vm"]
    newetam, newdeltam, newPhim = r["etam"], r["deltam"], r["Phim"]
    if test == 1:
        newU, newV, newUm, newVm = Uic, Vic, st.Um_curr, st.Vm_curr
    alpha = flags["alpha"]
    do_ra = flags["modalflag"] and t > 2
    levels = (("eta", neweta, newetam), ("delta", newdelta, newdeltam), ("Phi", newPhi, newPhim))
    mid, eff, mid_m, eff_m = {}, {}, {}, {}
    if not raw:
        # Classic Robert-Asselin on the physical prev level (model.py:991-1006).
        for name, new, newm in levels:
            cur, prv = getattr(st, name + "_curr"), getattr(st, name + "_prev")
            mid[name] = cur + alpha * (prv - 2.0 * cur + new) if do_ra else cur
            eff[name], eff_m[name] = new, newm
            curm, prvm = getattr(st, name + "m_curr"), getattr(st, name + "m_prev")
            if si:
                # The leapfrog reads the Fourier prev carry: filter its lineage too (model.py:1089-1105).
                mid_m[name] = curm + alpha * (prvm - 2.0 * curm + newm) if do_ra else curm
            else:
                mid_m[name] = curm  # locked default: unfiltered current level (model.py:1107-1110)
    else:
        # Robert-Asselin-Williams on the physical and Fourier lineages (model.py:1007-1087).
        w = flags.get("williams_alpha", 0.53)
        for name, new, newm in levels:
            cur, prv = getattr(st, name + "_curr"), getattr(st, name + "_prev")
            curm, prvm = getattr(st, name + "m_curr"), getattr(st, name + "m_prev")
            if do_ra:
                d = prv - 2.0 * cur + new
                dm = prvm - 2.0 * curm + newm
                mid[name], eff[name] = cur + alpha * w * d, new - alpha * (1.0 - w) * d
                mid_m[name], eff_m[name] = curm + alpha * w * dm, newm - alpha * (1.0 - w) * dm
            else:
                mid[name], eff[name], mid_m[name], eff_m[name] = cur, new, curm, newm
    phi_min = float(np.min(mid["Phi"])) if diag else 0.0
    phi_max = float(np.max(mid["Phi"])) if diag else 0.0
    PhiF2, F2, G2 = forcing_phys(s, test, eff["Phi"], newU, newV)
    PhiFm2, Fm2, Gm2 = fwd_fft_trunc(np.stack((PhiF2, F2, G2)), s.I, s.M)
    Am2, Bm2, Cm2, Dm2, Em2 = nonlinear_spectral(s, eff["eta"], eff["Phi"], newU, newV)
    if si:
        # Lagged momentum-forcing remainder of the incoming level (model.py:775-792, :1135-1143).
        inv_td = 0.0 if s.taudrag == -1.0 else 1.0 / s.taudrag
        Rum_next, Rvm_next = st.Fm_curr + st.Um_curr * inv_td, st.Gm_curr + st.Vm_curr * inv_td
    else:
        Rum_next, Rvm_next = st.Rum_lag, st.Rvm_lag
    new = OracleState(
        etam_prev=mid_m["eta"], etam_curr=eff_m["eta"], deltam_prev=mid_m["delta"], deltam_curr=eff_m["delta"],
        Phim_prev=mid_m["Phi"], Phim_curr=eff_m["Phi"],
        eta_prev=mid["eta"], eta_curr=eff["eta"], delta_prev=mid["delta"], delta_curr=eff["delta"],
        Phi_prev=mid["Phi"], Phi_curr=eff["Phi"], U_curr=newU, V_curr=newV, Um_curr=newUm, Vm_curr=newVm,
        Am_curr=Am2, Bm_curr=Bm2, Cm_curr=Cm2, Dm_curr=Dm2, Em_curr=Em2,
        PhiFm_curr=PhiFm2, Fm_curr=Fm2, Gm_curr=Gm2, Rum_lag=Rum_next, Rvm_lag=Rvm_next,
        dead=dead_next, spec=dict(eta=r["etamn"], delta=r["deltamn"], Phi=r["Phimn"]),
    )
    out = dict(t=t, dead=dead_next, eta=eff["eta"], delta=eff["delta"], Phi=eff["Phi"], U=newU, V=newV,
               rms=rms, spin_min=spin_min, phi_min=phi_min, phi_max=phi_max)
    return new, out


def run_model_scan(*, M, dt, tmax, Phibar, omega, a, test=None, g=9.8, forcflag=True, taurad=86400.0,
                   taudrag=86400.0, DPhieq=4e6, a1=0.05, diffflag=True, modalflag=True, alpha=0.01,
                   expflag=False, K6=1.24e33, K6Phi=None, starttime=None, eta0_init=None, delta0_init=None,
                   Phi0_init=None, U0_init=None, V0_init=None, diagnostics=True, return_history=True,
                   blowup_rms=8000.0, semi_implicit=False, si_alpha=0.5, raw_filter=False, williams_alpha=0.53):
    """Scan driver: validation, IC selection, time loop, result dict (model.py:1345-1765)."""
    if tmax < 2:
        raise ValueError("tmax must be >= 2 (SWAMPE uses a 2-level initialization).")
    if float(dt) <= 0.0:
        raise ValueError("dt must be positive.")
    if semi_implicit and expflag:
        raise ValueError("semi_implicit=True is incompatible with expflag=True (pick one scheme).")
    if raw_filter and expflag:
        raise ValueError(
            "raw_filter=True is incompatible with expflag=True (the explicit scheme reads the "
            "unfiltered spectral prev carry, so the RAW lineage change would alter its parity behavior).")
    s = build_static(M=M, dt=dt, a=a, omega=omega, g=g, Phibar=Phibar, taurad=taurad, taudrag=taudrag,
                     DPhieq=DPhieq, K6=K6, K6Phi=K6Phi, test=test)
    start = 2 if starttime is None else int(starttime)
    if start > tmax:
        raise ValueError(f"starttime={start} must be <= tmax={tmax}.")
    if eta0_init is not None or delta0_init is not None or Phi0_init is not None:
        if eta0_init is None or delta0_init is None or Phi0_init is None:
            raise ValueError("If providing explicit ICs, eta0_init, delta0_init, and Phi0_init must all be provided.")
        eta0, delta0, Phi0 = (np.asarray(x, np.float64) for x in (eta0_init, delta0_init, Phi0_init))
        if (U0_init is None) != (V0_init is None):
            raise ValueError("Provide both U0_init and V0_init, or neither.")
        if U0_init is not None:
            U0, V0 = np.asarray(U0_init, np.float64), np.asarray(V0_init, np.float64)
        else:
            U0, V0 = diagnose_winds(eta0, delta0, s)
    else:
        eta0, delta0, Phi0, U0, V0 = analytic_ic(s, test, a1)
    flags = dict(forcflag=forcflag, diffflag=diffflag, expflag=expflag, modalflag=modalflag,
                 diagnostics=diagnostics, alpha=alpha, blowup_rms=blowup_rms, semi_implicit=bool(semi_implicit),
                 si_alpha=float(si_alpha), raw_filter=bool(raw_filter), williams_alpha=float(williams_alpha))
    st = init_state_from_fields(s, test, eta0, delta0, Phi0, U0, V0)
    t_seq = np.arange(start, tmax, dtype=np.int32)
    outs = {k: [] for k in ("t", "dead", "eta", "delta", "Phi", "U", "V", "rms", "spin_min", "phi_min", "phi_max")}
    for t in t_seq:
        st, out = step_once(st, int(t), s, flags, test, U0, V0)
        if return_history:
            for k in outs:
                outs[k].append(out[k])
    if not return_history:
        return dict(static=s, t_seq=t_seq, last_state=st, starttime=start)
    outs = {k: np.asarray(v) for k, v in outs.items()}
    dead = outs["dead"].astype(bool)
    first = int(np.argmax(dead)) if dead.size and dead.any() else -1
    return dict(static=s, t_seq=t_seq, outs=outs, last_state=st, starttime=start, dead_first_idx=first)


def run_model(M, dt, tmax, Phibar, omega, a, test=None, g=9.8, forcflag=True, taurad=86400, taudrag=86400,
              DPhieq=4e6, a1=0.05, diffflag=True, modalflag=True, alpha=0.01, expflag=False, K6=1.24e33,
              K6Phi=None):
    """Legacy (tmax, 2) spinup/geopot diagnostics and final fields (model.py:1954-2219; no I/O)."""
    res = run_model_scan(M=M, dt=dt, tmax=tmax, Phibar=Phibar, omega=omega, a=a, test=test, g=g,
                         forcflag=forcflag, taurad=taurad, taudrag=taudrag, DPhieq=DPhieq, a1=a1,
                         diffflag=diffflag, modalflag=modalflag, alpha=alpha, expflag=expflag, K6=K6, K6Phi=K6Phi)
    s, outs, t_seq = res["static"], res["outs"], res["t_seq"]
    spin = np.zeros((tmax, 2))
    geo = np.zeros((tmax, 2))
    for k, t in enumerate(t_seq):
        idx = int(t) - 1
        if 0 <= idx < tmax:
            spin[idx] = (outs["spin_min"][k], outs["rms"][k])
            geo[idx] = (outs["phi_min"][k], outs["phi_max"][k])
    _, _, Phi0, U0, V0 = analytic_ic(s, test, a1)
    spin[0] = (np.min(np.sqrt(U0 * U0 + V0 * V0)), rms_winds(s.a, s.I, s.J, s.lambdas, s.mus, U0, V0))
    geo[0] = (np.min(Phi0), np.max(Phi0))
    n = outs["eta"].shape[0]
    return dict(eta=outs["eta"][-1] if n else None, delta=outs["delta"][-1] if n else None,
                Phi=outs["Phi"][-1] if n else None, U=outs["U"][-1] if n else None,
                V=outs["V"][-1] if n else None, spinup=spin, geopot=geo, lambdas=s.lambdas, mus=s.mus, t_seq=t_seq)


# ----------------------------------------------------------------------------
# Synthetic "rest + noise" initial condition used by the benchmark workloads
# ----------------------------------------------------------------------------


def rest_noise_ic(M: int, seed: int, omega: float, phi_std: float = 100.0, nmax: int = 10):
    """eta0 = 2 omega mu, delta0 = 0, U0 = V0 = 0, Phi0 = seeded noise on modes n <= nmax.

    Spectral coefficients c_mn ~ N(0,1) + i N(0,1) for 0 <= m <= n <= nmax (imaginary
    part zeroed at m=0), synthesized to the grid and rescaled to std(Phi0) = phi_std.
    """
    N, I, J, lambdas, mus, w, Pmn, Hmn = cached_geometry(int(M))
    rng = np.random.default_rng(seed)
    c = np.zeros((M + 1, N + 1), dtype=np.complex128)
    for m in range(0, nmax + 1):
        for n in range(m, nmax + 1):
            c[m, n] = rng.standard_normal() + 1j * rng.standard_normal()
    c[0, :] = c[0, :].real
    Phi0 = np.real(invrs_fft(invrs_leg(c, I, J, M, N, Pmn), I))
    Phi0 = Phi0 * (phi_std / np.std(Phi0))
    eta0 = np.broadcast_to(2.0 * omega * mus[:, None], (J, I)).copy()
    z = np.zeros((J, I))
    return eta0, z.copy(), Phi0, z.copy(), z.copy()


# ----------------------------------------------------------------------------
# Phase-curve observable (unit_tests/test_parity_reference_regression.py:67-82)
# ----------------------------------------------------------------------------


def phase_curve(phi, phase_angles, normalize: bool = True, raw: bool = False):
    """Visibility-weighted disc integral of a (J, I) geopotential grid per phase angle
    (unit_tests/test_parity_reference_regression.py:67-82): longitudes linspace(-pi, pi, I,
    endpoint=False), nodes numpy leggauss(J), weight sqrt(clip(1 - mu^2, 0)), visibility
    max(cos(lambda - phase), 0), summed over the whole grid per phase; then divided by
    mean|flux| when that is > 0.  raw=True also returns the flux before normalization."""
    phi = np.asarray(phi, dtype=np.float64)
    phase_angles = np.asarray(phase_angles, dtype=np.float64)
    J, I = phi.shape
    lambdas = np.linspace(-np.pi, np.pi, I, endpoint=False)[None, :]
    mus = np.polynomial.legendre.leggauss(J)[0][:, None]
    area_weight = np.sqrt(np.clip(1.0 - mus * mus, 0.0, None))
    flux = np.empty_like(phase_angles, dtype=np.float64)
    for k, phase in enumerate(phase_angles):
        visible = np.maximum(np.cos(lambdas - phase), 0.0)
        flux[k] = float(np.sum(phi * visible * area_weight))
    out = flux
    if normalize:
        norm = np.mean(np.abs(flux))
        if norm > 0:
            out = flux / norm
    return (out, flux) if raw else out


def phase_curve_vjp(g_flux, flux_raw, J: int, I: int, phase_angles, normalize: bool = True):
    """Cotangent of the (J, I) grid from the cotangent g_flux of phase_curve's output.

    With n = mean|f| and output f/n: df_k = g_k/n - sign(f_k) (sum_l g_l f_l)/(K n^2); the grid
    cotangent is sqrt(1 - mu_j^2) sum_k df_k max(cos(lambda_i - phase_k), 0).  (The derivative
    of the reference's projection under reverse-mode differentiation; no reference counterpart.)
    """
    g = np.asarray(g_flux, dtype=np.float64)
    f = np.asarray(flux_raw, dtype=np.float64)
    phases = np.asarray(phase_angles, dtype=np.float64)
    K = f.size
    gf = g.copy()
    if normalize:
        n = np.mean(np.abs(f))
        if n > 0:
            gf = g / n - np.sign(f) * float(np.dot(g, f)) / (K * n * n)
    lambdas = np.linspace(-np.pi, np.pi, I, endpoint=False)
    mus = np.polynomial.legendre.leggauss(J)[0]
    area_weight = np.sqrt(np.clip(1.0 - mus * mus, 0.0, None))
    vis = np.maximum(np.cos(lambdas[None, :] - phases[:, None]), 0.0)  # (K, I)
    return area_weight[:, None] * (gf @ vis)[None, :]
