"""Shared plumbing of the standalone scheme tendencies (modEuler_tdiff, explicit_tdiff): the
reference's 38 arguments as device tensors, and the inverse transforms that end every update."""

from __future__ import annotations

from types import SimpleNamespace

from . import spectral_transform as st
from .plan import require_gpu

ARGS = ("etam0", "etam1", "deltam0", "deltam1", "Phim0", "Phim1", "I", "J", "M", "N", "Am", "Bm", "Cm", "Dm", "Em",
        "Fm", "Gm", "Um", "Vm", "Pmn", "Pmnw", "Hmnw", "tstepcoeff1", "tstepcoeff2", "mJarray", "narray", "PhiFm",
        "dt", "a", "Phibar", "taurad", "taudrag", "forcflag", "diffflag", "sigma", "sigmaPhi", "test", "t")
_FOURIER = ("etam0", "etam1", "deltam0", "deltam1", "Phim0", "Phim1", "Am", "Bm", "Cm", "Dm", "Em", "Fm", "Gm", "Um",
            "Vm", "PhiFm")
_REAL = ("Pmn", "Pmnw", "Hmnw", "tstepcoeff1", "tstepcoeff2", "mJarray", "narray", "sigma", "sigmaPhi")
_SCALARS = ("dt", "a", "Phibar", "taurad", "taudrag")


def bind(values):
    """The reference's positional arguments (ARGS order) as a namespace: Fourier fields (J, M+1)
    as complex128 and tables as float64 device tensors, sizes as ints, flags as bools."""
    torch = require_gpu()
    if len(values) != len(ARGS):
        raise TypeError(f"expected the reference's {len(ARGS)} arguments, got {len(values)}")
    v = dict(zip(ARGS, values))
    ns = SimpleNamespace(**v)
    for k in ("I", "J", "M", "N"):
        setattr(ns, k, int(v[k]))
    for k in _FOURIER:
        setattr(ns, k, st._dev(v[k], torch, torch.complex128))
        if tuple(getattr(ns, k).shape) != (ns.J, ns.M + 1):
            raise ValueError(f"{k} must have shape ({ns.J}, {ns.M + 1}); got {tuple(getattr(ns, k).shape)}")
    for k in _REAL:
        setattr(ns, k, st._dev(v[k], torch, torch.float64))
    for k in _SCALARS:  # Python, numpy or 0-d torch scalars (float() reads a device tensor too)
        setattr(ns, k, float(v[k]))
    ns.forcflag, ns.diffflag = bool(v["forcflag"]), bool(v["diffflag"])
    return ns


def leg_w(fields, basis):
    """Forward Legendre transforms of the (J, M+1) fields with a preweighted basis, in one launch:
    the list of (M+1, N+1) spectra (fwd_leg_w_batch)."""
    torch = require_gpu()
    out = st.fwd_leg_w_batch(torch.stack(fields, dim=0), basis)
    return list(out.unbind(0))


def finish(Xmn, ns):
    """(Xmn, its grid, its positive-m Fourier coefficients): invrs_leg, then invrs_fft and the
    first M+1 columns, as every reference update ends (modEuler_tdiff.py:112-117)."""
    Xm = st.invrs_leg(Xmn, ns.I, ns.J, ns.M, ns.N, ns.Pmn)
    return Xmn, st.invrs_fft(Xm, ns.I), Xm[:, : ns.M + 1]
