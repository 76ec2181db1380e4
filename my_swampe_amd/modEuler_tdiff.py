"""The reference's modEuler_tdiff module (src/my_swampe/modEuler_tdiff.py): the modified-Euler
updates of Phi, delta and eta (Hack and Jakob 1992) with the reference's 38 arguments and its
three-array return (spectrum (M+1, N+1), grid (J, I), positive-m Fourier coefficients (J, M+1)).

The time step runs these updates fused in the Legendre kernel's FP64 spectral update; the
functions here compose the standalone FP64 transforms (``spectral_transform``) with elementwise
complex arithmetic on the GPU for one member's Fourier fields, keeping the reference's
coefficient quirks: Phi and delta take tstepcoeff / 4 and tstepcoeff2 / 4, delta always uses the
forced (B + F, A - G), and eta takes the unscaled tstepcoeff when forced and tstepcoeff / 2 when
not.  Results are complex128 torch tensors on the GPU.
"""

from __future__ import annotations

from . import _tendency as td


def phi_timestep(*args):
    """Modified-Euler update of Phi (modEuler_tdiff.py:30-117)."""
    v = td.bind(args)
    t1, t2 = v.tstepcoeff1 / 4.0, v.tstepcoeff2 / 4.0
    B = v.Bm + v.Fm if v.forcflag else v.Bm
    A = v.Am - v.Gm if v.forcflag else v.Am
    c1, c2, dleg, d2, d5 = td.leg_w([v.Phim1, t1 * 1j * v.mJarray * v.Cm, v.deltam1, t1 * 1j * v.mJarray * B,
                                     t2 * v.Em], v.Pmnw)
    c3, d3 = td.leg_w([t1 * v.Dm, t1 * A], v.Hmnw)
    X = c1 - c2 + c3 - v.dt * v.Phibar * dleg - v.Phibar * 0.5 * (d2 + d3 + v.narray * d5 +
                                                                   (1.0 / v.a ** 2) * (v.narray * c1))
    if v.forcflag:
        X = X + td.leg_w([v.dt * v.PhiFm], v.Pmnw)[0]
    if v.diffflag:
        X = X * v.sigmaPhi
    return td.finish(X, v)


def delta_timestep(*args):
    """Modified-Euler update of delta (modEuler_tdiff.py:120-208); the forced (B + F, A - G) enter
    whatever forcflag says, as in the reference."""
    v = td.bind(args)
    t1, t2 = v.tstepcoeff1 / 4.0, v.tstepcoeff2 / 4.0
    d1, d2, d4, d5, p2 = td.leg_w([v.deltam1, t1 * 1j * v.mJarray * (v.Bm + v.Fm), t2 * v.Phim1, t2 * v.Em,
                                   t1 * 1j * v.mJarray * v.Cm], v.Pmnw)
    d3, p3 = td.leg_w([t1 * (v.Am - v.Gm), t1 * v.Dm], v.Hmnw)
    n = v.narray
    X = d1 + d2 + d3 + n * d4 + n * d5 + (n * (p2 + p3)) / (2.0 * v.a ** 2) - v.Phibar * (n * d1) / v.a ** 2
    if v.forcflag:
        X = X + (n * td.leg_w([(v.dt / 2.0) * v.PhiFm], v.Pmnw)[0]) / v.a ** 2
    if v.diffflag:
        X = X * v.sigma
    return td.finish(X, v)


def eta_timestep(*args):
    """Modified-Euler update of eta (modEuler_tdiff.py:211-268): tstepcoeff when forced,
    tstepcoeff / 2 when not."""
    v = td.bind(args)
    t1 = v.tstepcoeff1 if v.forcflag else v.tstepcoeff1 / 2.0
    A = v.Am - v.Gm if v.forcflag else v.Am
    B = v.Bm + v.Fm if v.forcflag else v.Bm
    e1, e2 = td.leg_w([v.etam1, t1 * 1j * v.mJarray * A], v.Pmnw)
    X = e1 - e2 + td.leg_w([t1 * B], v.Hmnw)[0]
    if v.diffflag:
        X = X * v.sigma
    return td.finish(X, v)
