"""Drop-in host API for the spectral time-step core (mirrors my_swampe.model).

The public functions keep the reference's names, keyword arguments, result
dictionaries and error behaviour (src/my_swampe/model.py:1345-2251).  The
time integration itself runs in the HIP kernels of libswampe_hip.so through
:class:`my_swampe_amd.plan.Plan`; arrays are returned as torch tensors on the
GPU (the counterpart of JAX device arrays).

Continuation (contflag / contTime) and run_model's saveflag use the reference's restart-file
layout (continuation.py).  Out of scope here (NotImplementedError): the legacy plotting side
effects of run_model (plotflag).
"""

from __future__ import annotations

import math
from dataclasses import dataclass
from typing import Any, Dict, NamedTuple, Optional

import numpy as np

from . import _capi, continuation
from .plan import Plan, require_gpu

Scalar = Any


def _is_python_scalar(x: Any) -> bool:
    return isinstance(x, (int, float, np.floating))


@dataclass(frozen=True)
class RunFlags:
    """Run switches (model.py:156-226).  Bools select compile-time kernel variants."""

    forcflag: bool = True
    diffflag: bool = True
    expflag: bool = False
    modalflag: bool = True
    diagnostics: bool = True
    semi_implicit: bool = False
    raw_filter: bool = False
    alpha: Scalar = 0.01
    blowup_rms: Scalar = 8000.0
    williams_alpha: Scalar = 0.53
    si_alpha: Scalar = 0.5


@dataclass(frozen=True)
class Static:
    """Geometry, operators and constants (model.py:228-378), as float64 device tensors."""

    M: int
    N: int
    I: int
    J: int
    dt: Any
    a: Any
    omega: Any
    g: Any
    Phibar: Any
    taurad: Any
    taudrag: Any
    lambdas: Any
    mus: Any
    w: Any
    Pmn: Any
    Hmn: Any
    Pmnw: Any
    Hmnw: Any
    fmn: Any
    tstepcoeff: Any
    tstepcoeff2: Any
    tstepcoeffmn: Any
    marray: Any
    mJarray: Any
    narray: Any
    sigma: Any
    sigmaPhi: Any
    sigma_exp: Any
    sigmaPhi_exp: Any
    Phieq: Any
    # Per-member physical parameters as handed to the native plan.
    params: Any = None


class State(NamedTuple):
    """Scan carry with the reference's 27 field names (model.py:380-425)."""

    etam_prev: Any
    etam_curr: Any
    deltam_prev: Any
    deltam_curr: Any
    Phim_prev: Any
    Phim_curr: Any
    eta_prev: Any
    eta_curr: Any
    delta_prev: Any
    delta_curr: Any
    Phi_prev: Any
    Phi_curr: Any
    U_curr: Any
    V_curr: Any
    Um_curr: Any
    Vm_curr: Any
    Am_curr: Any
    Bm_curr: Any
    Cm_curr: Any
    Dm_curr: Any
    Em_curr: Any
    PhiFm_curr: Any
    Fm_curr: Any
    Gm_curr: Any
    Rum_lag: Any
    Rvm_lag: Any
    dead: Any


_GEOM: Dict[int, tuple] = {}


def _geometry(M: int):
    """(mus, w, lambdas, Pmn, Hmn): the reference's Gauss quadrature (what the plans use) and the
    library's basis."""
    if M not in _GEOM:
        mus, w, lam, P, H = _capi.geometry(M)
        mus, w = _capi.reference_quadrature(mus.size)
        _GEOM[M] = (mus, w, lam, P, H)
    return _GEOM[M]


def param_row(*, dt, a, omega, g, Phibar, taurad, taudrag, DPhieq, K6, K6Phi, alpha, blowup_rms=8000.0,
              si_alpha=0.5, williams_alpha=0.53) -> np.ndarray:
    """One member's SWX_NPARAM parameter vector (order of include/swampe_capi.h)."""
    return np.array([float(dt), float(a), float(omega), float(g), float(Phibar), float(taurad), float(taudrag),
                     float(DPhieq), float(K6), -1.0 if K6Phi is None else float(K6Phi), float(alpha),
                     float(blowup_rms), float(si_alpha), float(williams_alpha)], dtype=np.float64)


def build_static(*, M: int, dt: Scalar, a: Scalar, omega: Scalar, g: Scalar, Phibar: Scalar, taurad: Scalar,
                 taudrag: Scalar, DPhieq: Scalar, K6: Scalar, K6Phi: Optional[Scalar], test: Optional[int],
                 alpha: Scalar = 0.01) -> Static:
    """Static arrays of one run (model.py:457-557); the kernels use the plan's own tables."""
    torch = require_gpu()
    M = int(M)
    _capi.grid_size(M)  # ValueError for unsupported M
    mus_np, w_np, lam_np, P_np, H_np = _geometry(M)
    dev = torch.device("cuda", torch.cuda.current_device())
    f64 = dict(dtype=torch.float64, device=dev)
    N, J, I = M, mus_np.size, lam_np.size
    mus = torch.as_tensor(mus_np, **f64)
    w = torch.as_tensor(w_np, **f64)
    lam = torch.as_tensor(lam_np, **f64)
    Pmn = torch.as_tensor(P_np, **f64)
    Hmn = torch.as_tensor(H_np, **f64)
    dt_, a_, K6_ = float(dt), float(a), float(K6)
    K6Phi_ = K6_ if K6Phi is None else float(K6Phi)
    n = torch.arange(N + 1, **f64)
    fmn = torch.zeros((M + 1, M + 1), **f64)
    fmn[0, 1] = float(omega) / math.sqrt(0.375)
    cmn = n * (n + 1)
    cmn[0] = 1.0
    tcmn = a_ / cmn
    tcmn[0] = 0.0
    ncoeff = ((n * n * n) / a_ ** 3) * (((n + 1) * (n + 1) * (n + 1)) / a_ ** 3)
    sig = 1.0 / (1.0 + 2.0 * dt_ * K6_ * (ncoeff - 8.0 / a_ ** 6))
    sigP = 1.0 / (1.0 + 2.0 * dt_ * K6Phi_ * ncoeff)
    sig_e = torch.exp(-2.0 * dt_ * K6_ * torch.clamp(ncoeff - 8.0 / a_ ** 6, min=0.0))
    sigP_e = torch.exp(-2.0 * dt_ * K6Phi_ * ncoeff)
    bc = lambda v: v[None, :].expand(M + 1, N + 1).contiguous()  # noqa: E731
    if test is None:
        day = (lam > -math.pi / 2) & (lam < math.pi / 2)
        term = float(DPhieq) * torch.cos(lam)[None, :] * torch.sqrt(1 - mus[:, None] ** 2)
        Phieq = torch.where(day[None, :], float(Phibar) + term, torch.full_like(term, float(Phibar)))
    else:
        Phieq = torch.zeros((J, I), **f64)
    params = param_row(dt=dt, a=a, omega=omega, g=g, Phibar=Phibar, taurad=taurad, taudrag=taudrag,
                       DPhieq=DPhieq, K6=K6, K6Phi=K6Phi, alpha=alpha)
    return Static(
        M=M, N=N, I=I, J=J,
        dt=torch.tensor(dt_, **f64), a=torch.tensor(a_, **f64), omega=torch.tensor(float(omega), **f64),
        g=torch.tensor(float(g), **f64), Phibar=torch.tensor(float(Phibar), **f64),
        taurad=torch.tensor(float(taurad), **f64), taudrag=torch.tensor(float(taudrag), **f64),
        lambdas=lam, mus=mus, w=w, Pmn=Pmn, Hmn=Hmn, Pmnw=w[:, None, None] * Pmn, Hmnw=w[:, None, None] * Hmn,
        fmn=fmn,
        tstepcoeff=((2.0 * dt_) / (a_ * (1.0 - mus[:, None] ** 2))).expand(J, M + 1).contiguous(),
        tstepcoeff2=torch.full((J, M + 1), 2.0 * dt_ / a_ ** 2, **f64),
        tstepcoeffmn=bc(tcmn),
        marray=torch.arange(M + 1, **f64)[:, None].expand(M + 1, N + 1).contiguous(),
        mJarray=torch.arange(M + 1, **f64)[None, :].expand(J, M + 1).contiguous(),
        narray=bc(n * (n + 1)), sigma=bc(sig), sigmaPhi=bc(sigP), sigma_exp=bc(sig_e), sigmaPhi_exp=bc(sigP_e),
        Phieq=Phieq, params=params,
    )


def _analytic_ic(static: Static, test: Optional[int], a1: Scalar):
    """Analytic (eta0, delta0, Phi0, U0, V0) (model.py:669-691; initial_conditions.py:18-239)."""
    torch = require_gpu()
    a = float(static.a)
    omega = float(static.omega)
    Phibar = float(static.Phibar)
    SU0 = 2.0 * math.pi * a / (3600.0 * 24.0 * 12.0)
    sina, cosa = math.sin(float(a1)), math.cos(float(a1))
    etaamp = 2.0 * ((SU0 / a) + omega)
    Phiamp = SU0 * a * omega + 0.5 * SU0 ** 2
    mu = static.mus[:, None]
    lam = static.lambdas[None, :]
    J, I = static.J, static.I
    s1m = torch.sqrt(1.0 - mu ** 2)
    zeros = torch.zeros((J, I), dtype=torch.float64, device=mu.device)
    delta0 = zeros.clone()
    if test == 1:
        arg = -torch.cos(lam) * s1m * sina + mu * cosa
        eta0 = etaamp * arg
        bumpr = a / 3.0
        dist_arg = 0.0 * mu + math.cos(math.asin(0.0)) * torch.cos(torch.asin(mu)) * torch.cos(lam - 3.0 * math.pi / 2.0)
        dist = a * torch.arccos(torch.clamp(dist_arg, -1.0, 1.0))
        bump = (Phibar / 2.0) * (1.0 + torch.cos(math.pi * dist / bumpr))
        Phi0 = torch.where(dist < bumpr, bump, zeros)
        U0 = SU0 * (s1m * cosa + mu * torch.cos(lam) * sina) * s1m
        V0 = -SU0 * torch.sin(lam) * sina * s1m
    elif test == 2:
        arg = -torch.cos(lam) * s1m * sina + mu * cosa
        eta0 = etaamp * arg
        Phi0 = (Phibar - Phiamp) * (arg ** 2)
        U0 = SU0 * (s1m * cosa + torch.cos(lam) * mu * sina)
        V0 = (-SU0 * (torch.sin(lam) * sina)).expand(J, I)
    else:
        eta0 = (etaamp * mu).expand(J, I)
        Phi0 = zeros.clone()
        U0 = zeros.clone()
        V0 = zeros.clone()
    c = lambda t: t.contiguous()  # noqa: E731
    return c(eta0), c(delta0), c(Phi0), c(U0), c(V0)


def _rms_winds(static: Static, U, V):
    """Area-weighted RMS wind (time_stepping.py:537-576)."""
    torch = require_gpu()
    phis = torch.arcsin(static.mus)[:, None]
    dl = static.lambdas[2] - static.lambdas[1]
    dp = phis[2, 0] - phis[1, 0]
    a = static.a
    area = (a ** 2) * (torch.sin(phis + math.pi / 2.0) ** 2) * dp * dl
    integrand = (U / torch.cos(phis)) ** 2 + (V / torch.cos(phis)) ** 2
    return torch.sqrt(torch.sum(area * integrand / (4.0 * math.pi * a ** 2)))


def _validate(*, tmax, dt, test, semi_implicit, expflag, raw_filter, contflag, contTime, starttime, timeunits):
    """Host-side validation before any device work (model.py:1520-1613).

    Returns (test, start step, continuation file token, its fallback spelling)."""
    if tmax < 2:
        raise ValueError("tmax must be >= 2 (SWAMPE uses a 2-level initialization).")
    if _is_python_scalar(dt) and float(dt) <= 0.0:
        raise ValueError("dt must be positive.")
    if test is not None:
        test = int(test)
    if semi_implicit and expflag:
        raise ValueError("semi_implicit=True is incompatible with expflag=True (pick one scheme).")
    if raw_filter and expflag:
        raise ValueError(
            "raw_filter=True is incompatible with expflag=True (the explicit scheme reads the "
            "unfiltered spectral prev carry, so the RAW lineage change would alter its parity behavior).")
    token = fallback = None
    stamp = None
    if contflag:
        if contTime is None:
            raise ValueError("contflag=True requires contTime.")
        try:
            stamp = float(contTime)
        except (TypeError, ValueError) as e:
            raise ValueError(f"contTime must be numeric (int/float or numeric string), got {contTime!r}.") from e
        # File token: the integer spelling when the timestamp is integer-like (model.py:1588-1595).
        fallback = str(contTime)
        token = str(int(round(stamp))) if abs(stamp - round(stamp)) < 1e-12 else fallback
    if starttime is not None:
        start = int(starttime)
    elif not contflag:
        start = 2
    else:
        try:
            dt_float = float(dt)
        except TypeError as e:
            raise TypeError("dt must be a Python float when contflag=True (continuation uses Python I/O).") from e
        start = continuation.compute_t_from_timestamp(timeunits, stamp, dt_float)
    if start > tmax:
        raise ValueError(f"starttime={start} must be <= tmax={tmax}.")
    if test not in (None, 1, 2):
        raise NotImplementedError(f"test={test} is not implemented (None, 1 or 2).")
    return test, start, token, fallback


def _continuation_fields(token: str, fallback: str, custompath):
    """eta, delta, Phi of a restart (model.py:1645-1679): the canonical token's files, else the
    token as given; the winds are then diagnosed from eta and delta."""
    def read(prefix):
        try:
            return continuation.read_pickle(f"{prefix}-{token}", custompath=custompath)
        except FileNotFoundError:
            if fallback != token:
                return continuation.read_pickle(f"{prefix}-{fallback}", custompath=custompath)
            raise
    return read("eta"), read("delta"), read("Phi")


def _check_ic_args(shape, eta0_init, delta0_init, Phi0_init, U0_init, V0_init):
    """Presence and shape checks of explicit initial fields, on the host (model.py:1616-1640)."""
    have = (eta0_init is not None) or (delta0_init is not None) or (Phi0_init is not None)
    if not have:
        return
    if eta0_init is None or delta0_init is None or Phi0_init is None:
        raise ValueError("If providing explicit ICs, eta0_init, delta0_init, and Phi0_init must all be provided.")
    for name, arr in (("eta0_init", eta0_init), ("delta0_init", delta0_init), ("Phi0_init", Phi0_init)):
        if tuple(np.shape(arr)) != tuple(shape):
            raise ValueError(f"{name} must have shape {tuple(shape)}, got {tuple(np.shape(arr))}.")
    if (U0_init is None) != (V0_init is None):
        raise ValueError("Provide both U0_init and V0_init, or neither.")
    if U0_init is not None:
        for name, arr in (("U0_init", U0_init), ("V0_init", V0_init)):
            if tuple(np.shape(arr)) != tuple(shape):
                raise ValueError(f"{name} must have shape {tuple(shape)}, got {tuple(np.shape(arr))}.")


def _initial_fields(static: Static, test, a1, eta0_init, delta0_init, Phi0_init, U0_init, V0_init):
    """Explicit or analytic initial fields; U0/V0 None means 'diagnose spectrally' (model.py:1616-1682)."""
    torch = require_gpu()
    have = (eta0_init is not None) or (delta0_init is not None) or (Phi0_init is not None)
    if have:
        if eta0_init is None or delta0_init is None or Phi0_init is None:
            raise ValueError("If providing explicit ICs, eta0_init, delta0_init, and Phi0_init must all be provided.")
        dev = static.mus.device
        conv = lambda x: torch.as_tensor(np.asarray(x) if not torch.is_tensor(x) else x,  # noqa: E731
                                         dtype=torch.float64).to(dev)
        eta0, delta0, Phi0 = conv(eta0_init), conv(delta0_init), conv(Phi0_init)
        shape = (static.J, static.I)
        for name, arr in (("eta0_init", eta0), ("delta0_init", delta0), ("Phi0_init", Phi0)):
            if tuple(arr.shape) != shape:
                raise ValueError(f"{name} must have shape {shape}, got {tuple(arr.shape)}.")
        if (U0_init is None) != (V0_init is None):
            raise ValueError("Provide both U0_init and V0_init, or neither.")
        if U0_init is not None:
            U0, V0 = conv(U0_init), conv(V0_init)
            for name, arr in (("U0_init", U0), ("V0_init", V0)):
                if tuple(arr.shape) != shape:
                    raise ValueError(f"{name} must have shape {shape}, got {tuple(arr.shape)}.")
        else:
            U0 = V0 = None
        return eta0, delta0, Phi0, U0, V0
    return _analytic_ic(static, test, a1)


def _state_from_export(ex: Dict[str, Any], squeeze: bool) -> State:
    vals = {}
    for name in State._fields:
        v = ex[name]
        if name == "dead":
            v = v.bool()
        vals[name] = v[0] if squeeze else v
    return State(**vals)


def run_model_scan(*, M: int, dt: Scalar, tmax: int, Phibar: Scalar, omega: Scalar, a: Scalar,
                   test: Optional[int] = None, g: Scalar = 9.8, forcflag: bool = True, taurad: Scalar = 86400.0,
                   taudrag: Scalar = 86400.0, DPhieq: Scalar = 4 * (10 ** 6), a1: Scalar = 0.05,
                   diffflag: bool = True, modalflag: bool = True, alpha: Scalar = 0.01, expflag: bool = False,
                   K6: Scalar = 1.24 * (10 ** 33), K6Phi: Optional[Scalar] = None, contflag: bool = False,
                   custompath: Optional[str] = None, contTime: Optional[str] = None, timeunits: str = "hours",
                   starttime: Optional[int] = None, eta0_init=None, delta0_init=None, Phi0_init=None,
                   U0_init=None, V0_init=None, semi_implicit: bool = False, si_alpha: Scalar = 0.5,
                   raw_filter: bool = False, williams_alpha: Scalar = 0.53, diagnostics: bool = True,
                   return_history: bool = True, remat_step: bool = False, jit_scan: bool = True,
                   donate_state: bool = False, precision: str = "fp64") -> Dict[str, Any]:
    """Full run on the GPU, same signature and result dict as model.py:1345-1765.

    ``precision`` (extension) selects the transform precision: "fp64" (parity
    path), "fp32" or "bf16x3"; the spectral state update stays FP64.
    ``remat_step``, ``jit_scan`` and ``donate_state`` are accepted for
    signature compatibility and have no effect.
    """
    test, start, token, fallback = _validate(tmax=tmax, dt=dt, test=test, semi_implicit=semi_implicit,
                                             expflag=expflag, raw_filter=raw_filter, contflag=contflag,
                                             contTime=contTime, starttime=starttime, timeunits=timeunits)
    J, I = _capi.grid_size(int(M))  # ValueError for an unsupported truncation (initial_conditions.py:80-81)
    _check_ic_args((J, I), eta0_init, delta0_init, Phi0_init, U0_init, V0_init)
    if precision not in _capi.PRECISIONS:
        raise ValueError(f"precision must be one of {tuple(_capi.PRECISIONS)}, got {precision!r}")
    torch = require_gpu()
    flags = RunFlags(forcflag=bool(forcflag), diffflag=bool(diffflag), expflag=bool(expflag),
                     modalflag=bool(modalflag), diagnostics=bool(diagnostics), semi_implicit=bool(semi_implicit),
                     raw_filter=bool(raw_filter), alpha=alpha, williams_alpha=williams_alpha, si_alpha=si_alpha)
    static = build_static(M=int(M), dt=dt, a=a, omega=omega, g=g, Phibar=Phibar, taurad=taurad, taudrag=taudrag,
                          DPhieq=DPhieq, K6=K6, K6Phi=K6Phi, test=test, alpha=alpha)
    have_ic = (eta0_init is not None) or (delta0_init is not None) or (Phi0_init is not None)
    if contflag and not have_ic:
        eta0_init, delta0_init, Phi0_init = _continuation_fields(token, fallback, custompath)
        U0_init = V0_init = None
        _check_ic_args((J, I), eta0_init, delta0_init, Phi0_init, None, None)
    eta0, delta0, Phi0, U0, V0 = _initial_fields(static, test, a1, eta0_init, delta0_init, Phi0_init,
                                                 U0_init, V0_init)
    params = static.params.copy()
    params[_capi.PARAM_NAMES.index("si_alpha")] = float(si_alpha)
    params[_capi.PARAM_NAMES.index("williams_alpha")] = float(williams_alpha)
    plan = _run_plan(int(static.M), precision, flags, test, params)
    try:
        plan.init(eta0, delta0, Phi0, U0, V0)
        nsteps = int(tmax) - start
        t_seq = torch.arange(start, tmax, dtype=torch.int32, device=plan.device)
        if return_history:
            hist = plan.run(start, nsteps, history=True)
        else:
            _advance(plan, start, nsteps)
        ex = plan.export()
        last_state = _state_from_export(ex, squeeze=True)
        if not return_history:
            return dict(static=static, t_seq=t_seq, last_state=last_state, starttime=start)
        outs = _outs_from_history(hist, t_seq, flags.diagnostics, nsteps, plan, squeeze=True)
        dead = outs["dead"]
        if dead.numel() and bool(dead.any()):
            first = torch.argmax(dead.to(torch.int32)).to(torch.int32)
        else:
            first = torch.tensor(-1, dtype=torch.int32, device=plan.device)
        return dict(static=static, t_seq=t_seq, outs=outs, last_state=last_state, starttime=start,
                    dead_first_idx=first)
    finally:
        torch.cuda.current_stream(plan.device).synchronize()


# Plans of run_model_scan, kept across calls per structure (truncation, precision, scheme flags,
# test) as the reference keeps its compiled scan per static configuration (model.py:88-153):
# a call loads its parameters (swx_plan_set_params) and initial state into the cached plan, so
# the basis tables, the tuned tile sizes and the captured step graphs are built once.
_RUN_PLANS: "Dict[tuple, Plan]" = {}
_RUN_PLANS_MAX = 4
_GRAPH_STEPS = 32


def _run_plan(M: int, precision: str, flags: RunFlags, test: Optional[int], params: np.ndarray) -> Plan:
    key = (M, precision, bool(flags.expflag), bool(flags.forcflag), bool(flags.diffflag), bool(flags.modalflag),
           bool(flags.diagnostics), bool(flags.semi_implicit), bool(flags.raw_filter), test)
    plan = _RUN_PLANS.pop(key, None)
    if plan is None:
        plan = Plan(M, params[None, :], precision=precision, expflag=flags.expflag, forcflag=flags.forcflag,
                    diffflag=flags.diffflag, modalflag=flags.modalflag, diagnostics=flags.diagnostics, test=test,
                    semi_implicit=flags.semi_implicit, raw_filter=flags.raw_filter)
        while len(_RUN_PLANS) >= _RUN_PLANS_MAX:
            _RUN_PLANS.pop(next(iter(_RUN_PLANS))).close()
    else:
        plan.set_params(params[None, :])
    _RUN_PLANS[key] = plan  # most recently used last
    return plan


def _advance(plan: Plan, t0: int, nsteps: int) -> None:
    """nsteps steps without history from step index t0: the graph driver (captured step graphs
    replayed, swx_run_graph; bitwise the eager steps) on a stream of the plan's own, ordered after
    and before the caller's current stream."""
    if nsteps <= 0:
        return
    torch = require_gpu()
    cur = torch.cuda.current_stream(plan.device)
    s = getattr(plan, "_graph_stream", None)
    if s is None:
        s = torch.cuda.Stream(device=plan.device)
        plan._graph_stream = s
    s.wait_stream(cur)
    plan.run_graph(t0, nsteps, graph_steps=_GRAPH_STEPS, stream=s)
    cur.wait_stream(s)


def _outs_from_history(hist, t_seq, diagnostics, nsteps, plan: Plan, squeeze: bool):
    torch = require_gpu()
    B, J, I = plan.B, plan.J, plan.I
    if nsteps == 0:
        z = torch.zeros((B, 0), dtype=torch.float64, device=plan.device)
        f = torch.zeros((B, 0, J, I), dtype=torch.float64, device=plan.device)
        outs = dict(t=t_seq, dead=z.bool(), eta=f, delta=f, Phi=f, U=f, V=f, rms=z, spin_min=z, phi_min=z,
                    phi_max=z)
    else:
        hf = hist["fields"]
        outs = dict(t=t_seq, eta=hf[:, 0], delta=hf[:, 1], Phi=hf[:, 2], U=hf[:, 3], V=hf[:, 4])
        if diagnostics:
            outs["rms"] = hist["scalars"][:, :, 0]
            outs["spin_min"] = hist["scalars"][:, :, 1]
            outs["phi_min"] = hist["phiext"][..., 0].amin(dim=2)
            outs["phi_max"] = hist["phiext"][..., 1].amax(dim=2)
            outs["dead"] = hist["dead"].bool()
        else:
            z = torch.zeros((B, nsteps), dtype=torch.float64, device=plan.device)
            outs.update(rms=z, spin_min=z, phi_min=z, phi_max=z, dead=torch.zeros((B, nsteps), dtype=torch.bool,
                                                                               device=plan.device))
    if squeeze:
        outs = {k: (v if k == "t" else v[0]) for k, v in outs.items()}
    return outs


def run_model_scan_final(*, M: int, dt: Scalar, tmax: int, Phibar: Scalar, omega: Scalar, a: Scalar,
                         test: Optional[int] = None, g: Scalar = 9.8, forcflag: bool = True,
                         taurad: Scalar = 86400.0, taudrag: Scalar = 86400.0, DPhieq: Scalar = 4 * (10 ** 6),
                         a1: Scalar = 0.05, diffflag: bool = True, modalflag: bool = True, alpha: Scalar = 0.01,
                         expflag: bool = False, K6: Scalar = 1.24 * (10 ** 33), K6Phi: Optional[Scalar] = None,
                         contflag: bool = False, custompath: Optional[str] = None, contTime: Optional[str] = None,
                         timeunits: str = "hours", starttime: Optional[int] = None, eta0_init=None,
                         delta0_init=None, Phi0_init=None, U0_init=None, V0_init=None, semi_implicit: bool = False,
                         si_alpha: Scalar = 0.5, raw_filter: bool = False, williams_alpha: Scalar = 0.53,
                         diagnostics: bool = False, remat_step: bool = False, jit_scan: bool = True,
                         donate_state: bool = False, precision: str = "fp64") -> Dict[str, Any]:
    """Terminal state only (model.py:1844-1951)."""
    return run_model_scan(M=M, dt=dt, tmax=tmax, Phibar=Phibar, omega=omega, a=a, test=test, g=g,
                          forcflag=forcflag, taurad=taurad, taudrag=taudrag, DPhieq=DPhieq, a1=a1,
                          diffflag=diffflag, modalflag=modalflag, alpha=alpha, expflag=expflag, K6=K6,
                          K6Phi=K6Phi, contflag=contflag, custompath=custompath, contTime=contTime,
                          timeunits=timeunits, starttime=starttime, eta0_init=eta0_init,
                          delta0_init=delta0_init, Phi0_init=Phi0_init, U0_init=U0_init, V0_init=V0_init,
                          semi_implicit=semi_implicit, si_alpha=si_alpha, raw_filter=raw_filter,
                          williams_alpha=williams_alpha, diagnostics=diagnostics, return_history=False,
                          remat_step=remat_step, jit_scan=jit_scan, donate_state=donate_state, precision=precision)


def assert_finite_state(last_state: State, *, raise_on_nan: bool = True) -> bool:
    """Host-side NaN/Inf check of the terminal physical fields (model.py:1798-1841)."""
    torch = require_gpu()
    bad = [name for name in ("eta_curr", "delta_curr", "Phi_curr", "U_curr", "V_curr")
           if not bool(torch.all(torch.isfinite(getattr(last_state, name))))]
    if bad:
        if raise_on_nan:
            raise RuntimeError(
                "Final state contains non-finite values in fields: " + ", ".join(bad)
                + ". This usually indicates the integration blew up. Consider enabling diagnostics=True to "
                "gate on RMS-wind blowup, or reducing dt / increasing K6.")
        return False
    return True


def run_model(M: int, dt: float, tmax: int, Phibar: float, omega: float, a: float, test: Optional[int] = None,
              g: float = 9.8, forcflag: bool = True, taurad: float = 86400, taudrag: float = 86400,
              DPhieq: float = 4 * (10 ** 6), a1: float = 0.05, plotflag: bool = True, plotfreq: int = 5,
              minlevel: Optional[float] = None, maxlevel: Optional[float] = None, diffflag: bool = True,
              modalflag: bool = True, alpha: float = 0.01, contflag: bool = False, saveflag: bool = True,
              expflag: bool = False, savefreq: int = 150, K6: float = 1.24 * 10 ** 33,
              custompath: Optional[str] = None, contTime: Optional[str] = None, timeunits: str = "hours",
              verbose: bool = True, *, K6Phi: Optional[float] = None, semi_implicit: bool = False,
              si_alpha: float = 0.5, raw_filter: bool = False, williams_alpha: float = 0.53,
              jit_scan: bool = True, as_numpy: bool = True, precision: str = "fp64") -> Dict[str, Any]:
    """Legacy SWAMPE signature with (tmax, 2) spinup/geopot diagnostics (model.py:1954-2219).

    saveflag writes the restart files of every step t with t % savefreq == 0
    (model.py:2165-2182, continuation.save_data)."""
    if plotflag:
        raise NotImplementedError("plotflag side effects (matplotlib figures) are not provided by my_swampe_amd; "
                                  "pass plotflag=False (or use run_model_gpu).")
    torch = require_gpu()
    res = run_model_scan(M=M, dt=dt, tmax=tmax, Phibar=Phibar, omega=omega, a=a, test=test, g=g,
                         forcflag=forcflag, taurad=taurad, taudrag=taudrag, DPhieq=DPhieq, a1=a1,
                         diffflag=diffflag, modalflag=modalflag, alpha=alpha, expflag=expflag, K6=K6,
                         K6Phi=K6Phi, contflag=contflag, custompath=custompath, contTime=contTime,
                         timeunits=timeunits, semi_implicit=semi_implicit, si_alpha=si_alpha,
                         raw_filter=raw_filter, williams_alpha=williams_alpha, jit_scan=jit_scan,
                         precision=precision)
    static: Static = res["static"]
    outs = res["outs"]
    t_seq = res["t_seq"]
    dev = static.mus.device
    spin = torch.zeros((tmax, 2), dtype=torch.float64, device=dev)
    geo = torch.zeros((tmax, 2), dtype=torch.float64, device=dev)
    idx = (t_seq.to(torch.int64) - 1)
    ok = (idx >= 0) & (idx < tmax)
    idx = idx[ok]
    spin[idx, 0] = outs["spin_min"][ok]
    spin[idx, 1] = outs["rms"][ok]
    geo[idx, 0] = outs["phi_min"][ok]
    geo[idx, 1] = outs["phi_max"][ok]
    if not contflag:
        _, _, Phi0, U0, V0 = _analytic_ic(static, test, a1)
        spin[0, 0] = torch.min(torch.sqrt(U0 * U0 + V0 * V0))
        spin[0, 1] = _rms_winds(static, U0, V0)
        geo[0, 0] = torch.min(Phi0)
        geo[0, 1] = torch.max(Phi0)
    if saveflag:
        t_host = t_seq.cpu().numpy()
        for k, t in enumerate(t_host):
            if int(t) % int(savefreq) == 0:
                ts = continuation.compute_timestamp(timeunits, int(t), dt)
                continuation.save_data(ts, outs["eta"][k], outs["delta"][k], outs["Phi"][k], outs["U"][k],
                                       outs["V"][k], spin, geo, custompath=custompath)
    if verbose:
        print("GCM run completed!")
    n = outs["eta"].shape[0]
    conv = (lambda t: t.cpu().numpy()) if as_numpy else (lambda t: t)
    last = lambda k: conv(outs[k][-1]) if n else None  # noqa: E731
    return dict(eta=last("eta"), delta=last("delta"), Phi=last("Phi"), U=last("U"), V=last("V"),
                spinup=conv(spin), geopot=conv(geo), lambdas=conv(static.lambdas), mus=conv(static.mus),
                t_seq=conv(t_seq))


def run_model_gpu(*args, **kwargs) -> Dict[str, Any]:
    """run_model with plotflag/saveflag/as_numpy off by default (model.py:2222-2251)."""
    kwargs = dict(kwargs)
    kwargs.setdefault("plotflag", False)
    kwargs.setdefault("saveflag", False)
    kwargs.setdefault("as_numpy", False)
    kwargs.setdefault("jit_scan", True)
    return run_model(*args, **kwargs)


# ---------------------------------------------------------------------------------------------
# Operator-level seams (model.py:644-772, :795-1342): one step or one scan on a State carry.
#
# A State is a view of a native plan's carry.  Each (Static, RunFlags, test, members) keeps a
# plan; a State this module returned for that plan continues from the plan's carry without any
# copy, and any other State is loaded into the plan from its physical grids (the current level's
# eta, delta, Phi, U, V give the spectral state and the Fourier products exactly as the
# reference's carry holds them, since those are transforms of the same grids; the physical prev
# level and the dead flags are copied).  Loading a mid-run carry needs the modified-Euler scheme,
# whose step reads no spectral prev level; a two-level carry (prev == curr) loads in every
# scheme.  Fields may carry a leading member axis (B, J, I): B members step in one plan.
# ---------------------------------------------------------------------------------------------

class _Session:
    __slots__ = ("plan", "last", "batched")

    def __init__(self, plan: Plan, batched: bool):
        self.plan, self.last, self.batched = plan, None, batched


_SESSIONS: "Dict[tuple, _Session]" = {}
_SESSIONS_MAX = 4


def _plan_params(static: Static, flags: RunFlags) -> np.ndarray:
    """Plan parameters of a Static: its scalar fields (so that dataclasses.replace(static,
    taurad=...) takes effect) over the values it was built with, and the flags' filter scalars."""
    p = np.array(static.params, dtype=np.float64).copy()
    for name in ("dt", "a", "omega", "g", "Phibar", "taurad", "taudrag"):
        p[_capi.PARAM_NAMES.index(name)] = float(getattr(static, name))
    for name in ("alpha", "blowup_rms", "si_alpha", "williams_alpha"):
        p[_capi.PARAM_NAMES.index(name)] = float(getattr(flags, name))
    return p


def _session(static: Static, flags: RunFlags, test: Optional[int], B: int, batched: bool,
             precision: str = "fp64") -> _Session:
    params = _plan_params(static, flags)
    key = (int(static.M), params.tobytes(), bool(flags.forcflag), bool(flags.diffflag), bool(flags.expflag),
           bool(flags.modalflag), bool(flags.diagnostics), bool(flags.semi_implicit), bool(flags.raw_filter),
           test, int(B), bool(batched), precision)
    sess = _SESSIONS.pop(key, None)
    if sess is None:
        plan = Plan(static.M, np.broadcast_to(params, (B, params.size)).copy(), precision=precision,
                    expflag=flags.expflag, forcflag=flags.forcflag, diffflag=flags.diffflag,
                    modalflag=flags.modalflag, diagnostics=flags.diagnostics, test=test,
                    semi_implicit=flags.semi_implicit, raw_filter=flags.raw_filter)
        sess = _Session(plan, batched)
        while len(_SESSIONS) >= _SESSIONS_MAX:
            old = _SESSIONS.pop(next(iter(_SESSIONS)))
            old.plan.close()
    _SESSIONS[key] = sess  # most recently used last
    return sess


def _members(state_or_field, static: Static):
    shape = tuple(state_or_field.shape)
    if shape == (static.J, static.I):
        return 1, False
    if len(shape) == 3 and shape[1:] == (static.J, static.I):
        return shape[0], True
    raise ValueError(f"fields must have shape {(static.J, static.I)} or (B, {static.J}, {static.I}), got {shape}.")


def _as_field(x, static: Static):
    torch = require_gpu()
    return torch.as_tensor(np.asarray(x) if not torch.is_tensor(x) else x, dtype=torch.float64).to(static.mus.device)


def _load_state(sess: _Session, state: State, static: Static, flags: RunFlags, test, Uic, Vic) -> None:
    if state is sess.last:
        return
    torch = require_gpu()
    f = {n: _as_field(getattr(state, n), static)
         for n in ("eta_prev", "eta_curr", "delta_prev", "delta_curr", "Phi_prev", "Phi_curr", "U_curr", "V_curr")}
    two_level = all(torch.equal(f[n + "_prev"], f[n + "_curr"]) for n in ("eta", "delta", "Phi"))
    if not two_level and (flags.expflag or flags.semi_implicit or flags.raw_filter):
        raise NotImplementedError("loading a mid-run State (prev level != current level) is supported for the "
                                  "modified-Euler scheme only; step it from the State this module returned")
    U, V = f["U_curr"], f["V_curr"]
    if test == 1 and Uic is not None:
        U, V = _as_field(Uic, static), _as_field(Vic, static)
    plan = sess.plan
    plan.init(f["eta_curr"], f["delta_curr"], f["Phi_curr"], U, V)
    dead = None if state.dead is None else torch.as_tensor(state.dead, device=plan.device)
    if not two_level:
        plan.set_prev_level(f["eta_prev"], f["delta_prev"], f["Phi_prev"], dead)
    elif dead is not None and bool(dead.any()):
        if flags.expflag or flags.semi_implicit or flags.raw_filter:
            raise NotImplementedError("dead members can be loaded in the modified-Euler scheme only")
        plan.set_prev_level(dead=dead)
    sess.last = None


def _export_state(sess: _Session) -> State:
    st = _state_from_export(sess.plan.export(), squeeze=not sess.batched)
    sess.last = st
    return st


def _init_state_from_fields(*, static: Static, flags: RunFlags, test: Optional[int], eta0, delta0, Phi0,
                            U0=None, V0=None) -> State:
    """Two-level initial carry (prev == curr) from physical fields (model.py:694-772); U0/V0 None
    diagnoses the winds spectrally from eta0/delta0."""
    e = _as_field(eta0, static)
    B, batched = _members(e, static)
    sess = _session(static, flags, test, B, batched)
    sess.plan.init(e, _as_field(delta0, static), _as_field(Phi0, static),
                   None if U0 is None else _as_field(U0, static), None if V0 is None else _as_field(V0, static))
    return _export_state(sess)


def _diagnose_winds(eta, delta, static: Static):
    """(U, V) diagnosed from eta and delta through the spectral transforms (model.py:644-666)."""
    e = _as_field(eta, static)
    B, batched = _members(e, static)
    sess = _session(static, RunFlags(diagnostics=False), None, B, batched)
    torch = require_gpu()
    sess.plan.init(e, _as_field(delta, static), torch.zeros_like(e))
    sess.last = None
    ex = sess.plan.export(fields=("U_curr", "V_curr"))
    U, V = ex["U_curr"], ex["V_curr"]
    return (U, V) if batched else (U[0], V[0])


def _t_list(t_seq):
    torch = require_gpu()
    if torch.is_tensor(t_seq):
        return [int(x) for x in t_seq.detach().cpu().reshape(-1).tolist()]
    return [int(x) for x in np.asarray(t_seq).reshape(-1)]


def _runs(ts):
    """Split step indices into runs of consecutive values: [(t0, n), ...]."""
    out = []
    for t in ts:
        if out and out[-1][0] + out[-1][1] == t:
            out[-1][1] += 1
        else:
            out.append([t, 1])
    return [(a, n) for a, n in out]


def _step_once(state: State, t, static: Static, flags: RunFlags, test: Optional[int], Uic=None, Vic=None):
    """One step: (new State, out) with out = t, dead, eta, delta, Phi, U, V (the new fields) and rms,
    spin_min, phi_min, phi_max (model.py:795-1194)."""
    B, batched = _members(state.Phi_curr, static)
    sess = _session(static, flags, test, B, batched)
    _load_state(sess, state, static, flags, test, Uic, Vic)
    torch = require_gpu()
    tt = int(t)
    hist = sess.plan.run(tt, 1, history=True)
    t_seq = torch.tensor([tt], dtype=torch.int32, device=sess.plan.device)
    outs = _outs_from_history(hist, t_seq, flags.diagnostics, 1, sess.plan, squeeze=False)
    out = {k: (torch.as_tensor(t, dtype=torch.int32, device=sess.plan.device) if k == "t" else
               (v[:, 0] if batched else v[0, 0])) for k, v in outs.items()}
    return _export_state(sess), out


def _step_once_state_only(state: State, t, static: Static, flags: RunFlags, test: Optional[int], Uic=None,
                          Vic=None) -> State:
    """One step returning only the new State (model.py:1197-1235)."""
    B, batched = _members(state.Phi_curr, static)
    sess = _session(static, flags, test, B, batched)
    _load_state(sess, state, static, flags, test, Uic, Vic)
    _advance(sess.plan, int(t), 1)
    return _export_state(sess)


def simulate_scan(*, static: Static, flags: RunFlags, state0: State, t_seq, test: Optional[int], Uic=None,
                  Vic=None):
    """(last State, outs) over the step indices t_seq, outs stacked along the leading time axis as
    the reference's scan returns them (model.py:1237-1278)."""
    torch = require_gpu()
    B, batched = _members(state0.Phi_curr, static)
    sess = _session(static, flags, test, B, batched)
    _load_state(sess, state0, static, flags, test, Uic, Vic)
    ts = _t_list(t_seq)
    parts = []
    for t0, n in _runs(ts):
        parts.append(sess.plan.run(t0, n, history=True))
    t_dev = torch.tensor(ts, dtype=torch.int32, device=sess.plan.device)
    if parts:
        hist = {k: torch.cat([p[k] for p in parts], dim=1 if k != "fields" else 2) for k in parts[0]}
    else:
        hist = {}
    outs = _outs_from_history(hist, t_dev, flags.diagnostics, len(ts), sess.plan, squeeze=False)
    # Time-major like lax.scan's stacked outputs: (T, ...) or (T, B, ...).
    outs = {k: (v if k == "t" else (v.transpose(0, 1) if batched else v[0])) for k, v in outs.items()}
    return _export_state(sess), outs


def simulate_scan_last(*, static: Static, flags: RunFlags, state0: State, t_seq, test: Optional[int], Uic=None,
                       Vic=None, remat_step: bool = False) -> State:
    """Final State after the step indices t_seq, no history (model.py:1281-1342).  remat_step is
    accepted for signature compatibility (the hand-written adjoint has its own checkpointing)."""
    B, batched = _members(state0.Phi_curr, static)
    sess = _session(static, flags, test, B, batched)
    _load_state(sess, state0, static, flags, test, Uic, Vic)
    for t0, n in _runs(_t_list(t_seq)):
        _advance(sess.plan, t0, n)
    return _export_state(sess)
