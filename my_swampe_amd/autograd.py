"""Differentiable rollout: hand-written adjoint behind torch.autograd.Function.

``rollout_phi(c, taurad, taudrag, DPhieq, ...)`` advances spectral initial
coefficients ``c`` (B, 3, M+1, N+1) of eta, delta, Phi by ``nsteps`` forced
modified-Euler steps on the GPU and returns the final geopotential grid
(B, J, I).  Its backward runs the native reverse-mode kernels
(swx_adjoint_*): the forward pass keeps checkpoints of the spectral state every
K steps, and the backward pass recomputes each segment from its checkpoint,
keeping that segment's states, then steps the adjoint back through it.  Memory:
(nsteps/K + K) spectral states per member.  By default K = 1 (every state kept,
no recompute) when the whole trajectory fits RolloutConfig.checkpoint_budget
(32 GiB of device memory: 1400 steps of T42 x 256 members), else
K = ceil(sqrt(nsteps)); checkpoint_every sets K explicitly.

Gradients are returned for c (complex, d/dRe + i d/dIm), for the
per-member taurad, taudrag and DPhieq, and for the hyperdiffusion coefficients
K6 and K6Phi when rollout_phi is given them as tensors (their partials come
from the diffusion factor that ends each step); the other physical constants
are fixed.
This replaces reverse-mode differentiation of the reference's scan
(model.py:1281-1342 under jax.grad, unit_tests/test_autodiff.py:53-185).
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

from . import _capi
from .model import param_row
from .plan import Plan, require_gpu


@dataclass
class RolloutConfig:
    M: int
    nsteps: int
    dt: float
    a: float
    omega: float
    Phibar: float
    K6: float = 1.24e33
    K6Phi: Optional[float] = None
    alpha: float = 0.01
    g: float = 9.8
    precision: str = "fp64"
    checkpoint_every: Optional[int] = None
    diffflag: bool = True
    modalflag: bool = True
    forcflag: bool = True
    t0: int = 2
    stats: dict = field(default_factory=dict)

    members: int = 1
    # Device memory the default checkpoint policy may fill with saved states (bytes).
    checkpoint_budget: int = 32 << 30

    def state_bytes(self) -> int:
        return self.members * 3 * (self.M + 1) * (self.M + 1) * 16

    def every(self) -> int:
        """Steps per checkpoint.  Default: every step when the whole trajectory of spectral
        states fits the budget (no recompute in the reverse sweep), else ceil(sqrt(nsteps))."""
        if self.checkpoint_every:
            return max(1, int(self.checkpoint_every))
        if self.nsteps * self.state_bytes() <= self.checkpoint_budget:
            return 1
        return max(1, int(math.ceil(math.sqrt(max(self.nsteps, 1)))))


def _params(cfg: RolloutConfig, taurad, taudrag, DPhieq, K6=None, K6Phi=None) -> np.ndarray:
    """Parameter rows; K6 / K6Phi (per member or shared) default to the configuration's, and a
    K6Phi of None follows K6."""
    tr = np.atleast_1d(np.asarray(taurad, dtype=np.float64))
    td = np.atleast_1d(np.asarray(taudrag, dtype=np.float64))
    dp = np.atleast_1d(np.asarray(DPhieq, dtype=np.float64))
    k6 = np.atleast_1d(np.asarray(cfg.K6 if K6 is None else K6, dtype=np.float64))
    if K6Phi is None and cfg.K6Phi is None:
        k6p = None
    else:
        k6p = np.atleast_1d(np.asarray(cfg.K6Phi if K6Phi is None else K6Phi, dtype=np.float64))
    B = max(x.size for x in (tr, td, dp, k6) + (() if k6p is None else (k6p,)))
    tr, td, dp, k6 = (np.broadcast_to(x, (B,)) for x in (tr, td, dp, k6))
    k6p = None if k6p is None else np.broadcast_to(k6p, (B,))
    return np.stack([param_row(dt=cfg.dt, a=cfg.a, omega=cfg.omega, g=cfg.g, Phibar=cfg.Phibar, taurad=tr[b],
                               taudrag=td[b], DPhieq=dp[b], K6=k6[b], K6Phi=None if k6p is None else k6p[b],
                               alpha=cfg.alpha)
                     for b in range(B)])


# Idle plans by configuration: a rollout takes one (or creates it), sets the members'
# parameters, and its backward hands it back, so repeated rollouts of one configuration (an
# optimizer loop) skip the plan's table construction and kernel tuning.  A rollout whose
# backward never runs keeps its plan until the autograd graph is freed.
_POOL: dict = {}
_POOL_MAX = 4  # idle plans kept per configuration


def _plan_key(cfg: RolloutConfig, B: int):
    return (cfg.M, B, cfg.precision, cfg.forcflag, cfg.diffflag, cfg.modalflag)


def _make_plan(cfg: RolloutConfig, params) -> Plan:
    idle = _POOL.get(_plan_key(cfg, params.shape[0]))
    if idle:
        plan = idle.pop()
        plan.set_params(params)
    else:
        plan = Plan(cfg.M, params, precision=cfg.precision, forcflag=cfg.forcflag, diffflag=cfg.diffflag,
                    modalflag=cfg.modalflag, diagnostics=False)
    # The rollout reads the spectral state and the final Phi_curr only: the physical grid carry
    # is written in the last step of each run call.
    plan.set_option(_capi.OPT_PHYS_STEPS, 1)
    return plan


def _release_plan(cfg: RolloutConfig, plan: Plan) -> None:
    idle = _POOL.setdefault(_plan_key(cfg, plan.B), [])
    if len(idle) < _POOL_MAX:
        idle.append(plan)
    else:
        plan.close()


def forward_with_checkpoints(plan: Plan, cfg: RolloutConfig, c):
    """Run the rollout from c, returning the checkpoints (spectral states before each segment):
    with one checkpoint per step, the whole trajectory as one tape tensor (swx_run_tape)."""
    K = cfg.every()
    plan.init_spectral(c)
    if K == 1:
        return plan.run_tape(cfg.t0, cfg.nsteps)
    ckpts: List = []
    t = cfg.t0
    done = 0
    while done < cfg.nsteps:
        n = min(K, cfg.nsteps - done)
        ckpts.append(plan.state_get())
        # Only the last segment's final step needs the grid fields (the exported Phi_curr).
        plan.set_option(_capi.OPT_PHYS_STEPS, 1 if done + n == cfg.nsteps else -1)
        plan.run(t, n)
        t += n
        done += n
    plan.set_option(_capi.OPT_PHYS_STEPS, 1)
    return ckpts


def adjoint_through_checkpoints(plan: Plan, cfg: RolloutConfig, ckpts, gPhi, final=None):
    """Reverse sweep: terminal cotangent gPhi (B, J, I) -> (dL/dc, dL/d(taurad, taudrag, DPhieq)).
    With `final` (the state after the last step) the sweep also accumulates the diffusion partials
    (dL/dK6, dL/dK6Phi) of every state a step produced, read by plan.adjoint_diffusion_result()."""
    K = cfg.every()
    plan.adjoint_begin()
    plan.adjoint_terminal(g_Phi=gPhi)
    diffusion = final is not None
    if diffusion:
        plan.adjoint_diffusion(final)
        plan.set_option(_capi.OPT_ADJ_DIFFUSION, 1)
    try:
        return _sweep_checkpoints(plan, cfg, ckpts, K, diffusion)
    finally:
        plan.set_option(_capi.OPT_ADJ_DIFFUSION, 0)


def _sweep_checkpoints(plan: Plan, cfg: RolloutConfig, ckpts, K: int, diffusion: bool):
    if not isinstance(ckpts, list):  # a run_tape trajectory: one native reverse sweep
        plan.adjoint_sweep(ckpts)
        cfg.stats.update(checkpoints=int(ckpts.shape[0]), segment_states=1)
        return plan.adjoint_result()
    plan.set_option(_capi.OPT_PHYS_STEPS, -1)  # the recompute needs the spectral states only
    peak = 0
    seg_tape = None
    for seg in reversed(range(len(ckpts))):
        start = seg * K
        n = min(K, cfg.nsteps - start)
        if n == 1:
            plan.adjoint_step(ckpts[seg])
            peak = max(peak, 1)
            if diffusion and start > 0:  # the segment's first state came out of a step
                plan.adjoint_diffusion(ckpts[seg])
            continue
        # Recompute the segment's states from its checkpoint in one native call (the states
        # before its first n - 1 steps), take the last one, and sweep the segment natively.
        if seg_tape is None:
            seg_tape = ckpts[seg].new_empty((K,) + tuple(ckpts[seg].shape))
        plan.state_set(ckpts[seg])
        plan.run_tape(cfg.t0 + start, n - 1, out=seg_tape)
        seg_tape[n - 1].copy_(plan.state_get())
        peak = max(peak, n)
        plan.adjoint_sweep(seg_tape[:n])
        if diffusion and start > 0:
            plan.adjoint_diffusion(ckpts[seg])
    plan.set_option(_capi.OPT_PHYS_STEPS, 1)
    cfg.stats.update(checkpoints=len(ckpts), segment_states=peak)
    return plan.adjoint_result()


def _forward_observed(plan: Plan, cfg: RolloutConfig, c, obs_steps, observe):
    """Run the rollout from c, calling observe(k) when the state has advanced obs_steps[k] steps;
    returns the checkpoints (spectral states before each segment)."""
    K = cfg.every()
    plan.init_spectral(c)
    ckpts: List = []
    pending = list(enumerate(obs_steps))
    if pending and pending[0][1] == 0:
        observe(pending.pop(0)[0])
    done = 0
    while done < cfg.nsteps:
        seg_end = min(done + K, cfg.nsteps)
        ckpts.append(plan.state_get())
        while done < seg_end:
            stop = min(seg_end, pending[0][1]) if pending else seg_end
            # The grid fields are read only where an observation falls.
            observed = bool(pending) and pending[0][1] == stop
            plan.set_option(_capi.OPT_PHYS_STEPS, 1 if observed else -1)
            plan.run(cfg.t0 + done, stop - done)
            done = stop
            if observed:
                observe(pending.pop(0)[0])
    plan.set_option(_capi.OPT_PHYS_STEPS, 1)
    return ckpts


def _adjoint_observed(plan: Plan, cfg: RolloutConfig, ckpts, obs_steps, cotangent, final=None):
    """Reverse sweep whose observation k adds cotangent(k) (the (B, J, I) cotangent of Phi after
    obs_steps[k] steps) when the sweep reaches that state.  With `final` (the state after the last
    step) it also accumulates the diffusion partials of every state a step produced, each once its
    cotangent is whole (after that state's observation)."""
    K = cfg.every()
    where = {int(n): k for k, n in enumerate(obs_steps)}
    diffusion = final is not None and cfg.nsteps > 0
    plan.adjoint_begin()
    if cfg.nsteps in where:
        plan.adjoint_terminal(g_Phi=cotangent(where[cfg.nsteps]))
    if diffusion:
        plan.adjoint_diffusion(final)
    plan.set_option(_capi.OPT_PHYS_STEPS, -1)
    for seg in reversed(range(len(ckpts))):
        start = seg * K
        n = min(K, cfg.nsteps - start)
        if n > 1:  # recompute the segment's states from its checkpoint
            plan.state_set(ckpts[seg])
        tape = [ckpts[seg]]
        for s in range(n - 1):
            plan.run(cfg.t0 + start + s, 1)
            tape.append(plan.state_get())
        for s in reversed(range(n)):
            plan.adjoint_step(tape[s])
            if start + s in where:
                plan.adjoint_observe(g_Phi=cotangent(where[start + s]))
            if diffusion and start + s > 0:
                plan.adjoint_diffusion(tape[s])
    plan.set_option(_capi.OPT_PHYS_STEPS, 1)
    cfg.stats.update(checkpoints=len(ckpts))
    return plan.adjoint_result()


def forward_series(plan: Plan, cfg: RolloutConfig, c, obs_steps):
    """Run the rollout from c, returning the checkpoints and Phi after each step count in obs_steps."""
    torch = require_gpu()
    obs = []
    ckpts = _forward_observed(plan, cfg, c, obs_steps,
                              lambda k: obs.append(plan.export(fields=("Phi_curr",))["Phi_curr"]))
    return ckpts, torch.stack(obs, dim=1)


def adjoint_series(plan: Plan, cfg: RolloutConfig, ckpts, obs_steps, gPhi, final=None):
    """Reverse sweep with the cotangents gPhi (B, n_obs, J, I) of Phi after each observed step count
    added when the sweep reaches that state."""
    return _adjoint_observed(plan, cfg, ckpts, obs_steps, lambda k: gPhi[:, k].contiguous(), final)


def forward_phase_curve(plan: Plan, cfg: RolloutConfig, c, obs_steps, phases, normalize: bool):
    """Run the rollout from c, keeping only the phase curves of Phi at the observed step counts:
    returns the checkpoints, the curves (B, n_obs, K) and the un-normalized curves (n_obs, B, K).
    phases: float64 device tensor (n_obs, K)."""
    torch = require_gpu()
    n_obs, K = phases.shape
    flux = torch.empty((n_obs, plan.B, K), dtype=torch.float64, device=plan.device)
    raw = torch.empty_like(flux)

    def observe(k):
        f, r = plan.phase_curve(phases[k], normalize)
        flux[k].copy_(f)
        raw[k].copy_(r)

    ckpts = _forward_observed(plan, cfg, c, obs_steps, observe)
    return ckpts, flux.permute(1, 0, 2).contiguous(), raw


def adjoint_phase_curve(plan: Plan, cfg: RolloutConfig, ckpts, obs_steps, phases, normalize: bool, raw, g_flux,
                        final=None):
    """Reverse sweep of a loss of the phase curves: g_flux (B, n_obs, K); each observation's grid
    cotangent is formed on the GPU (swx_phase_curve_vjp) when the sweep reaches it."""
    from .observables import phase_curve_vjp

    mus = plan.mus_device()
    return _adjoint_observed(
        plan, cfg, ckpts, obs_steps,
        lambda k: phase_curve_vjp(g_flux[:, k], raw[k], phases[k], plan.J, plan.I, normalize, mus), final)


def _check_steps(obs_steps):
    steps = [int(n) for n in obs_steps]
    if not steps or any(n < 0 for n in steps) or any(b <= a_ for a_, b in zip(steps, steps[1:])):
        raise ValueError("obs_steps must be a non-empty strictly increasing sequence of step counts >= 0")
    return steps


def _fold(g, shape):
    """Per-member gradient g (B,) folded onto the parameter's shape (a shared scalar sums members)."""
    if len(shape) == 0:
        return g.sum()
    if shape[0] == 1 and g.numel() > 1:
        return g.sum().reshape(shape)
    return g.reshape(shape)


def _host(x):
    return None if x is None else x.detach().cpu().numpy()


def _diffusion_result(plan: Plan, ctx):
    """(dL/dK6, dL/dK6Phi) folded onto the parameters' shapes, or (None, None) when not asked."""
    if ctx.final is None:
        return None, None
    dg = plan.adjoint_diffusion_result()
    sk, skp = ctx.k6_shapes
    # With K6Phi following K6, K6 also scales the Phi factor.
    gk6 = _fold(dg[:, 0] + dg[:, 1] if skp is None else dg[:, 0], sk)
    return gk6, None if skp is None else _fold(dg[:, 1], skp)


def _save_k6(ctx, plan: Plan, K6, K6Phi) -> None:
    # The diffusion partials need the state after the last step.
    ctx.final = plan.state_get() if any(ctx.needs_input_grad[4:6]) else None
    ctx.k6_shapes = (K6.shape, None if K6Phi is None else K6Phi.shape)


class _Rollout:
    """torch.autograd.Function factory (torch imported lazily)."""

    _cls = None

    @classmethod
    def get(cls):
        if cls._cls is not None:
            return cls._cls
        torch = require_gpu()

        class Fn(torch.autograd.Function):
            @staticmethod
            def forward(ctx, c, taurad, taudrag, DPhieq, K6, K6Phi, cfg: RolloutConfig):
                params = _params(cfg, _host(taurad), _host(taudrag), _host(DPhieq), _host(K6), _host(K6Phi))
                plan = _make_plan(cfg, params)
                ckpts = forward_with_checkpoints(plan, cfg, c.detach())
                phi = plan.export(fields=("Phi_curr",))["Phi_curr"]
                _save_k6(ctx, plan, K6, K6Phi)
                torch.cuda.current_stream(plan.device).synchronize()
                ctx.plan, ctx.cfg, ctx.ckpts = plan, cfg, ckpts
                ctx.shapes = (taurad.shape, taudrag.shape, DPhieq.shape)
                return phi

            @staticmethod
            def backward(ctx, gPhi):
                plan, cfg = ctx.plan, ctx.cfg
                lam, pg = adjoint_through_checkpoints(plan, cfg, ctx.ckpts, gPhi.contiguous(), ctx.final)
                gk6, gk6p = _diffusion_result(plan, ctx)
                torch.cuda.current_stream(plan.device).synchronize()
                cfg.stats["plan_device_bytes"] = plan.device_bytes()
                _release_plan(cfg, plan)
                ctx.plan = None
                ctx.ckpts = None
                ctx.final = None
                sr, sd, sp = ctx.shapes
                return lam, _fold(pg[:, 0], sr), _fold(pg[:, 1], sd), _fold(pg[:, 2], sp), gk6, gk6p, None

        cls._cls = Fn
        return Fn


class _RolloutSeries:
    """torch.autograd.Function factory of the time-series rollout (torch imported lazily)."""

    _cls = None

    @classmethod
    def get(cls):
        if cls._cls is not None:
            return cls._cls
        torch = require_gpu()

        class Fn(torch.autograd.Function):
            @staticmethod
            def forward(ctx, c, taurad, taudrag, DPhieq, K6, K6Phi, cfg: RolloutConfig, obs_steps):
                params = _params(cfg, _host(taurad), _host(taudrag), _host(DPhieq), _host(K6), _host(K6Phi))
                plan = _make_plan(cfg, params)
                ckpts, phis = forward_series(plan, cfg, c.detach(), obs_steps)
                _save_k6(ctx, plan, K6, K6Phi)
                torch.cuda.current_stream(plan.device).synchronize()
                ctx.plan, ctx.cfg, ctx.ckpts, ctx.obs = plan, cfg, ckpts, obs_steps
                ctx.shapes = (taurad.shape, taudrag.shape, DPhieq.shape)
                return phis

            @staticmethod
            def backward(ctx, gPhi):
                plan, cfg = ctx.plan, ctx.cfg
                lam, pg = adjoint_series(plan, cfg, ctx.ckpts, ctx.obs, gPhi.contiguous(), ctx.final)
                gk6, gk6p = _diffusion_result(plan, ctx)
                torch.cuda.current_stream(plan.device).synchronize()
                _release_plan(cfg, plan)
                ctx.plan = None
                ctx.ckpts = None
                ctx.final = None
                sr, sd, sp = ctx.shapes
                return lam, _fold(pg[:, 0], sr), _fold(pg[:, 1], sd), _fold(pg[:, 2], sp), gk6, gk6p, None, None

        cls._cls = Fn
        return Fn


class _RolloutPhaseCurve:
    """torch.autograd.Function factory of the phase-curve rollout (torch imported lazily)."""

    _cls = None

    @classmethod
    def get(cls):
        if cls._cls is not None:
            return cls._cls
        torch = require_gpu()

        class Fn(torch.autograd.Function):
            @staticmethod
            def forward(ctx, c, taurad, taudrag, DPhieq, K6, K6Phi, cfg: RolloutConfig, obs_steps, phases, normalize):
                params = _params(cfg, _host(taurad), _host(taudrag), _host(DPhieq), _host(K6), _host(K6Phi))
                plan = _make_plan(cfg, params)
                ckpts, flux, raw = forward_phase_curve(plan, cfg, c.detach(), obs_steps, phases, normalize)
                _save_k6(ctx, plan, K6, K6Phi)
                torch.cuda.current_stream(plan.device).synchronize()
                ctx.plan, ctx.cfg, ctx.ckpts, ctx.obs = plan, cfg, ckpts, obs_steps
                ctx.phases, ctx.normalize, ctx.raw = phases, normalize, raw
                ctx.shapes = (taurad.shape, taudrag.shape, DPhieq.shape)
                return flux

            @staticmethod
            def backward(ctx, g_flux):
                plan, cfg = ctx.plan, ctx.cfg
                lam, pg = adjoint_phase_curve(plan, cfg, ctx.ckpts, ctx.obs, ctx.phases, ctx.normalize, ctx.raw,
                                              g_flux.contiguous(), ctx.final)
                gk6, gk6p = _diffusion_result(plan, ctx)
                torch.cuda.current_stream(plan.device).synchronize()
                _release_plan(cfg, plan)
                ctx.plan = None
                ctx.ckpts = None
                ctx.final = None
                sr, sd, sp = ctx.shapes
                return (lam, _fold(pg[:, 0], sr), _fold(pg[:, 1], sd), _fold(pg[:, 2], sp), gk6, gk6p, None, None,
                        None, None)

        cls._cls = Fn
        return Fn


def rollout_phi(c, taurad, taudrag, DPhieq, *, M: int, nsteps: int, dt: float, a: float, omega: float,
                Phibar: float, K6: float = 1.24e33, K6Phi: Optional[float] = None, alpha: float = 0.01,
                precision: str = "fp64", checkpoint_every: Optional[int] = None, diffflag: bool = True,
                modalflag: bool = True, forcflag: bool = True, stats: Optional[dict] = None):
    """Final Phi grid (B, J, I) of a differentiable rollout (see module docstring).

    c: complex128 tensor (B, 3, M+1, N+1); taurad, taudrag, DPhieq, and K6 / K6Phi: float64
    tensors of shape (B,) or scalars (shared by all members).  K6Phi = None follows K6 (then
    dL/dK6 includes the Phi factor's share).
    """
    torch = require_gpu()
    dev = torch.device("cuda", torch.cuda.current_device())
    cfg = RolloutConfig(members=int(c.shape[0]), M=M, nsteps=nsteps, dt=dt, a=a, omega=omega, Phibar=Phibar, K6=K6, K6Phi=K6Phi,
                        alpha=alpha, precision=precision, checkpoint_every=checkpoint_every, diffflag=diffflag,
                        modalflag=modalflag, forcflag=forcflag, stats=stats if stats is not None else {})
    conv = lambda x: x if torch.is_tensor(x) else torch.tensor(x, dtype=torch.float64, device=dev)  # noqa: E731
    return _Rollout.get().apply(c, conv(taurad), conv(taudrag), conv(DPhieq), conv(K6),
                                None if K6Phi is None else conv(K6Phi), cfg)


def rollout_phi_series(c, taurad, taudrag, DPhieq, *, obs_steps, M: int, dt: float, a: float, omega: float,
                       Phibar: float, K6: float = 1.24e33, K6Phi: Optional[float] = None, alpha: float = 0.01,
                       precision: str = "fp64", checkpoint_every: Optional[int] = None, diffflag: bool = True,
                       modalflag: bool = True, forcflag: bool = True, stats: Optional[dict] = None):
    """Phi grids (B, len(obs_steps), J, I) after each step count in obs_steps (ascending, >= 0) of one
    differentiable rollout of max(obs_steps) steps: a loss over the whole time series (a phase-curve
    likelihood) gets its gradient from one reverse sweep whose observation cotangents are added
    as the sweep passes them (swx_adjoint_observe), with the checkpoints and the parameter tensors
    (K6 / K6Phi included) of rollout_phi.
    """
    torch = require_gpu()
    steps = _check_steps(obs_steps)
    dev = torch.device("cuda", torch.cuda.current_device())
    cfg = RolloutConfig(members=int(c.shape[0]), M=M, nsteps=steps[-1], dt=dt, a=a, omega=omega, Phibar=Phibar, K6=K6, K6Phi=K6Phi,
                        alpha=alpha, precision=precision, checkpoint_every=checkpoint_every, diffflag=diffflag,
                        modalflag=modalflag, forcflag=forcflag, stats=stats if stats is not None else {})
    conv = lambda x: x if torch.is_tensor(x) else torch.tensor(x, dtype=torch.float64, device=dev)  # noqa: E731
    return _RolloutSeries.get().apply(c, conv(taurad), conv(taudrag), conv(DPhieq), conv(K6),
                                      None if K6Phi is None else conv(K6Phi), cfg, tuple(steps))


def rollout_phase_curve(c, taurad, taudrag, DPhieq, *, obs_steps, phase_angles, normalize: bool = True, M: int,
                        dt: float, a: float, omega: float, Phibar: float, K6: float = 1.24e33,
                        K6Phi: Optional[float] = None, alpha: float = 0.01, precision: str = "fp64",
                        checkpoint_every: Optional[int] = None, diffflag: bool = True, modalflag: bool = True,
                        forcflag: bool = True, stats: Optional[dict] = None):
    """Phase curves (B, len(obs_steps), K) of Phi after each step count in obs_steps of one
    differentiable rollout: the observable of the phase-curve likelihood evaluated inside the
    rollout, so that neither pass keeps a grid per observation (the forward keeps the curves, the
    reverse sweep forms each observation's grid cotangent on the GPU when it reaches it).

    phase_angles: (K,) angles used at every observation, or (len(obs_steps), K) per observation.
    The projection is observables.phase_curve's with the plan's Gauss nodes.
    """
    torch = require_gpu()
    steps = _check_steps(obs_steps)
    dev = torch.device("cuda", torch.cuda.current_device())
    ph = torch.as_tensor(phase_angles, dtype=torch.float64, device=dev)
    if ph.dim() == 1:
        ph = ph.expand(len(steps), -1)
    if ph.dim() != 2 or ph.shape[0] != len(steps) or ph.shape[1] == 0:
        raise ValueError("phase_angles must have shape (K,) or (len(obs_steps), K) with K >= 1")
    cfg = RolloutConfig(members=int(c.shape[0]), M=M, nsteps=steps[-1], dt=dt, a=a, omega=omega, Phibar=Phibar, K6=K6, K6Phi=K6Phi,
                        alpha=alpha, precision=precision, checkpoint_every=checkpoint_every, diffflag=diffflag,
                        modalflag=modalflag, forcflag=forcflag, stats=stats if stats is not None else {})
    conv = lambda x: x if torch.is_tensor(x) else torch.tensor(x, dtype=torch.float64, device=dev)  # noqa: E731
    return _RolloutPhaseCurve.get().apply(c, conv(taurad), conv(taudrag), conv(DPhieq), conv(K6),
                                          None if K6Phi is None else conv(K6Phi), cfg, tuple(steps),
                                          ph.contiguous(), bool(normalize))
