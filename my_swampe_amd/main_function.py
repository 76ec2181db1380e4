"""The reference's legacy entry point (src/my_swampe/main_function.py:23-211): ``main`` with the
original numpy SWAMPE positional signature, delegating to ``my_swampe_amd.model.run_model``.

0/1 flags become booleans, ``test=0`` means forced (``test=None``), the legacy physics constants
k1, k2, pressure, R, Cp and sigmaSB are accepted and unused, and test codes other than None, 0, 1,
2 raise NotImplementedError before any device work, as in the reference.  ``plotflag`` defaults to
1 there and here; this package draws no figures, so a plotting call raises NotImplementedError
(pass ``plotflag=0``).  The reference's command-line front end is not provided.
"""

from __future__ import annotations

from typing import Optional

from .model import run_model


def main(M, dt, tmax, Phibar, omega, a, test, g=9.8, forcflag=1, taurad=86400, taudrag=86400,
         DPhieq=4 * (10 ** 6), a1=0.05, plotflag=1, plotfreq=5, minlevel=6, maxlevel=7, diffflag=1, modalflag=1,
         alpha=0.01, contflag=0, saveflag=1, expflag=0, savefreq=150, k1=2 * 10 ** (-4), k2=4 * 10 ** (-4),
         pressure=100 * 250 * 9.8 / 10, R=3000, Cp=13000, sigmaSB=5.7 * 10 ** (-8), K6=1.24 * 10 ** 33,
         custompath=None, contTime=None, timeunits="hours", verbose=True, *, K6Phi: Optional[float] = None):
    """run_model with the legacy signature and conventions (main_function.py:23-211); returns the
    run_model result dictionary."""
    del k1, k2, pressure, R, Cp, sigmaSB  # legacy physics constants: accepted, not used
    if test == 0:
        test = None
    if test not in (None, 1, 2):
        raise NotImplementedError(
            f"test={test!r} is not supported. Supported: None (or 0 for forced), 1, 2.")
    return run_model(M=int(M), dt=float(dt), tmax=int(tmax), Phibar=float(Phibar), omega=float(omega), a=float(a),
                     test=test, g=float(g), forcflag=bool(forcflag), taurad=float(taurad), taudrag=float(taudrag),
                     DPhieq=float(DPhieq), a1=float(a1), plotflag=bool(plotflag), plotfreq=int(plotfreq),
                     minlevel=None if minlevel is None else float(minlevel),
                     maxlevel=None if maxlevel is None else float(maxlevel), diffflag=bool(diffflag),
                     modalflag=bool(modalflag), alpha=float(alpha), contflag=bool(contflag), saveflag=bool(saveflag),
                     expflag=bool(expflag), savefreq=int(savefreq), K6=float(K6), custompath=custompath,
                     contTime=contTime, timeunits=str(timeunits), verbose=bool(verbose), K6Phi=K6Phi)
