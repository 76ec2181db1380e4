"""The reference's explicit_tdiff module (src/my_swampe/explicit_tdiff.py): the explicit
(leapfrog-style) updates of Phi, delta and eta with the reference's 38 arguments and its
three-array return (spectrum (M+1, N+1), grid (J, I), positive-m Fourier coefficients (J, M+1)).

The time step (``expflag=True`` plans) runs these updates fused in the Legendre kernel; the
functions here compose the standalone FP64 transforms with elementwise complex arithmetic on the
GPU for one member's Fourier fields, keeping the reference's quirks: delta is its carried level
alone plus forcing, and the forcing adds U / taudrag and V / taudrag terms beside F and G.
Results are complex128 torch tensors on the GPU.
"""

from __future__ import annotations

from . import _tendency as td


def phi_timestep(*args):
    """Explicit update of Phi from the previous level Phim0 (explicit_tdiff.py:20-85)."""
    v = td.bind(args)
    t1 = v.tstepcoeff1
    c1, c2, dleg = td.leg_w([v.Phim0, t1 * 1j * v.mJarray * v.Cm, v.deltam1], v.Pmnw)
    X = c1 - c2 + td.leg_w([t1 * v.Dm], v.Hmnw)[0] - 2.0 * v.dt * v.Phibar * dleg
    if v.forcflag:
        X = X + td.leg_w([2.0 * v.dt * v.PhiFm], v.Pmnw)[0]
    if v.diffflag:
        X = X * v.sigmaPhi
    return td.finish(X, v)


def delta_timestep(*args):
    """Explicit update of delta (explicit_tdiff.py:88-162): the carried level deltam0 only, plus
    the forcing terms."""
    v = td.bind(args)
    X = td.leg_w([v.deltam0], v.Pmnw)[0]
    if v.forcflag:
        t1 = v.tstepcoeff1
        f1, f3 = td.leg_w([(t1 * 1j * v.mJarray * v.Um) / v.taudrag, t1 * 1j * v.mJarray * v.Fm], v.Pmnw)
        f2, f4 = td.leg_w([(t1 * v.Vm) / v.taudrag, t1 * v.Gm], v.Hmnw)
        X = X + (-f1 + f2 + f3 - f4)
    if v.diffflag:
        X = X * v.sigma
    return td.finish(X, v)


def eta_timestep(*args):
    """Explicit update of eta from the previous level etam0 (explicit_tdiff.py:165-233)."""
    v = td.bind(args)
    t1 = v.tstepcoeff1
    e1, e2 = td.leg_w([v.etam0, t1 * 1j * v.mJarray * v.Am], v.Pmnw)
    X = e1 - e2 + td.leg_w([t1 * v.Bm], v.Hmnw)[0]
    if v.forcflag:
        f1, f3 = td.leg_w([(t1 * 1j * v.mJarray * v.Vm) / v.taudrag, t1 * 1j * v.mJarray * v.Gm], v.Pmnw)
        f2, f4 = td.leg_w([(t1 * v.Um) / v.taudrag, t1 * v.Fm], v.Hmnw)
        X = X + (-f1 + f2 + f3 + f4)
    if v.diffflag:
        X = X * v.sigma
    return td.finish(X, v)
