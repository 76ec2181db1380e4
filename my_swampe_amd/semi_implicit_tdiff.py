"""The reference's semi_implicit_tdiff module (src/my_swampe/semi_implicit_tdiff.py:79-264): one
semi-implicit leapfrog step of one member's Fourier fields, with the reference's 46 arguments and
the 13-array return of ``time_stepping.tstepping``.

Gravity waves and, when ``relax_implicit``, the linear radiative relaxation and Rayleigh drag are
implicit with weight ``si_alpha``, solved in closed form per degree; the momentum forcing's
nonlinear remainders enter at the lagged level (``Rum_lag``, ``Rvm_lag``); diffusion is the
exponential factor pair (``filters.sigma6_exponential``).  Plans with ``semi_implicit=True`` run
this update fused in the Legendre kernel; this function composes the standalone FP64 device
transforms with elementwise GPU arithmetic in the reference's order.  Results are complex128
torch tensors on the GPU.
"""

from __future__ import annotations

from . import _tendency as td
from . import spectral_transform as st
from .plan import require_gpu


def si_timestep(Rum_lag, Rvm_lag, etam0, etam1, deltam0, deltam1, Phim0, Phim1, I, J, M, N, Am, Bm, Cm, Dm, Em, Fm,
                Gm, Um, Vm, fmn, Pmn, Hmn, Pmnw, Hmnw, tstepcoeff, tstepcoeff2, tstepcoeffmn, marray, mJarray, narray,
                PhiFm, dt, a, Phibar, taurad, taudrag, forcflag, diffflag, relax_implicit, sigma_exp, sigmaPhi_exp,
                si_alpha, test, t):
    """(etamn, eta, etam, deltamn, delta, deltam, Phimn, Phi, Phim, U, V, Um, Vm) after one
    semi-implicit step (semi_implicit_tdiff.py:79-264)."""
    torch = require_gpu()
    v = td.bind((etam0, etam1, deltam0, deltam1, Phim0, Phim1, I, J, M, N, Am, Bm, Cm, Dm, Em, Fm, Gm, Um, Vm, Pmn,
                 Pmnw, Hmnw, tstepcoeff, tstepcoeff2, mJarray, narray, PhiFm, dt, a, Phibar, taurad, taudrag,
                 forcflag, diffflag, sigma_exp, sigmaPhi_exp, test, t))
    ru = st._dev(Rum_lag, torch, torch.complex128)
    rv = st._dev(Rvm_lag, torch, torch.complex128)
    f = st._dev(fmn, torch, torch.complex128)
    two_dt = 2.0 * v.dt
    alpha = float(si_alpha)
    lam = v.narray / v.a ** 2
    if relax_implicit:
        inv_td = 0.0 if v.taudrag == -1.0 else 1.0 / v.taudrag
        inv_tr = 1.0 / v.taurad
        F_e, G_e, PhiF_e = ru, rv, v.PhiFm + v.Phim1 * inv_tr
    else:
        inv_td = inv_tr = 0.0
        F_e, G_e, PhiF_e = v.Fm, v.Gm, v.PhiFm
    A_eff = v.Am - G_e if v.forcflag else v.Am
    B_eff = v.Bm + F_e if v.forcflag else v.Bm
    tc = v.tstepcoeff1
    eta0, delta0, Phi0, imA, imB, imC, Emn = td.leg_w(
        [v.etam0, v.deltam0, v.Phim0, tc * 1j * v.mJarray * A_eff, tc * 1j * v.mJarray * B_eff,
         tc * 1j * v.mJarray * v.Cm, v.Em], v.Pmnw)
    HA, HB, HD = td.leg_w([tc * A_eff, tc * B_eff, tc * v.Dm], v.Hmnw)
    xi = two_dt * alpha
    expl = two_dt * (1.0 - alpha)
    kd, kr = xi * inv_td, xi * inv_tr
    gd, gr = expl * inv_td, expl * inv_tr
    eta_star = eta0 - imA + HB - gd * (eta0 - f)
    etamn = (eta_star + kd * f) / (1.0 + kd)
    delta_star = delta0 + imB + HA + two_dt * lam * Emn + expl * lam * Phi0 - gd * delta0
    Phi_star = Phi0 - imC + HD - expl * v.Phibar * delta0 - gr * Phi0
    if v.forcflag:
        Phi_star = Phi_star + td.leg_w([two_dt * PhiF_e], v.Pmnw)[0]
    b = xi * lam
    c = xi * v.Phibar
    deltamn = (delta_star + b * Phi_star / (1.0 + kr)) / ((1.0 + kd) + b * c / (1.0 + kr))
    Phimn = (Phi_star - c * deltamn) / (1.0 + kr)
    if v.diffflag:
        etamn, deltamn, Phimn = etamn * v.sigma, deltamn * v.sigma, Phimn * v.sigmaPhi
    out = []
    for X in (etamn, deltamn, Phimn):
        out += list(td.finish(X, v))
    U, V, Um_new, Vm_new = st.invrsUV_with_coeffs(deltamn, etamn, f, v.I, v.J, v.M, v.N, v.Pmn, Hmn, tstepcoeffmn,
                                                  marray)
    return tuple(out) + (U, V, Um_new, Vm_new)
