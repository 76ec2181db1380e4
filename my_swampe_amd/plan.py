"""Torch-facing wrapper of a native plan (one truncation group, B members).

PyTorch provides device memory and the stream; all arithmetic happens in the
HIP kernels of libswampe_hip.so.
"""

from __future__ import annotations

import ctypes
from typing import Dict, Optional

import numpy as np

from . import _capi


def _torch():
    import torch

    return torch


def require_gpu():
    """Return torch after checking that a ROCm GPU is usable; raise otherwise."""
    torch = _torch()
    if not torch.cuda.is_available():
        raise RuntimeError("my_swampe_amd needs a ROCm GPU (torch.cuda.is_available() is False); "
                           "there is no CPU fallback.")
    return torch


def _ptr(t) -> Optional[int]:
    return None if t is None else ctypes.c_void_p(t.data_ptr())


class Plan:
    """One native plan: geometry, basis tables and the device carry of B members."""

    def __init__(self, M: int, params: np.ndarray, *, precision: str = "fp64", expflag: bool = False,
                 forcflag: bool = True, diffflag: bool = True, modalflag: bool = True,
                 diagnostics: bool = False, test: Optional[int] = None, exact_basis: bool = False,
                 semi_implicit: bool = False, raw_filter: bool = False, quadrature="reference"):
        """quadrature: "reference" (the reference's scipy Gauss nodes and weights, the default),
        None (the library's built-in quadrature) or a (mus, w) pair of length-J arrays."""
        torch = require_gpu()
        lib = _capi.load()
        params = np.ascontiguousarray(np.asarray(params, dtype=np.float64))
        if params.ndim == 1:
            params = params[None, :]
        if params.shape[1] != _capi.NPARAM:
            raise ValueError(f"params must have {_capi.NPARAM} columns {_capi.PARAM_NAMES}")
        if precision not in _capi.PRECISIONS:
            raise ValueError(f"precision must be one of {tuple(_capi.PRECISIONS)}")
        self.M = int(M)
        self.J, self.I = _capi.grid_size(self.M)
        self.B = params.shape[0]
        self.precision = precision
        self.test = test
        self.device = torch.device("cuda", torch.cuda.current_device())
        quad = None
        if isinstance(quadrature, str):
            if quadrature != "reference":
                raise ValueError('quadrature must be "reference", None or a (mus, w) pair')
            quad = _capi.reference_quadrature(self.J)
        elif quadrature is not None:
            quad = tuple(np.ascontiguousarray(np.asarray(q, dtype=np.float64)) for q in quadrature)
            if len(quad) != 2 or any(q.shape != (self.J,) for q in quad):
                raise ValueError(f"quadrature must be two arrays of length J={self.J}")
        desc = _capi.PlanDesc(M=self.M, n_members=self.B, precision=_capi.PRECISIONS[precision],
                              expflag=int(bool(expflag)), forcflag=int(bool(forcflag)),
                              diffflag=int(bool(diffflag)), modalflag=int(bool(modalflag)),
                              diagnostics=int(bool(diagnostics)), test=-1 if test is None else int(test),
                              exact_basis=int(bool(exact_basis)), semi_implicit=int(bool(semi_implicit)),
                              raw_filter=int(bool(raw_filter)),
                              quad_mus=None if quad is None else quad[0].ctypes.data,
                              quad_w=None if quad is None else quad[1].ctypes.data)
        handle = ctypes.c_void_p()
        _capi.check(lib.swx_plan_create(ctypes.byref(desc), params.ctypes.data_as(ctypes.c_void_p),
                                        ctypes.byref(handle)), "swx_plan_create")
        self._h = handle
        self._lib = lib
        self._quad_mus = None if quad is None else quad[0]
        self._mus_d = None

    # -- lifetime ---------------------------------------------------------
    def close(self) -> None:
        """Destroy the native plan, after closing the ensembles and pipelines that step it (their
        native objects hold the plan)."""
        for user in list(getattr(self, "_users", ())):
            user.close()
        if getattr(self, "_h", None) is not None and self._h.value:
            self._lib.swx_plan_destroy(self._h)
            self._h = None

    def _add_user(self, user) -> None:
        import weakref

        if getattr(self, "_users", None) is None:
            self._users = weakref.WeakSet()
        self._users.add(user)

    def __del__(self):  # pragma: no cover - best effort
        try:
            self.close()
        except Exception:
            pass

    def _stream(self, stream=None):
        torch = _torch()
        s = stream if stream is not None else torch.cuda.current_stream(self.device)
        return ctypes.c_void_p(s.cuda_stream)

    def _field(self, x, name):
        torch = _torch()
        if x is None:
            return None
        t = torch.as_tensor(x, dtype=torch.float64, device=self.device)
        if t.dim() == 2:
            t = t.expand(self.B, -1, -1)
        if tuple(t.shape) != (self.B, self.J, self.I):
            raise ValueError(f"{name} must have shape {(self.J, self.I)} or {(self.B, self.J, self.I)}, "
                             f"got {tuple(t.shape)}.")
        return t.contiguous()

    # -- operations -------------------------------------------------------
    def set_params(self, params: np.ndarray, stream=None) -> None:
        params = np.ascontiguousarray(np.asarray(params, dtype=np.float64).reshape(self.B, _capi.NPARAM))
        _capi.check(self._lib.swx_plan_set_params(self._h, params.ctypes.data_as(ctypes.c_void_p),
                                                  self._stream(stream)), "swx_plan_set_params")

    def set_option(self, option: int, value: int) -> None:
        """Plan option (see SWX_OPT_* in include/swampe_capi.h)."""
        _capi.check(self._lib.swx_plan_set_option(self._h, int(option), int(value)), "swx_plan_set_option")

    def init(self, eta0, delta0, Phi0, U0=None, V0=None, stream=None) -> None:
        if (U0 is None) != (V0 is None):
            raise ValueError("Provide both U0_init and V0_init, or neither.")
        e, d, p = self._field(eta0, "eta0"), self._field(delta0, "delta0"), self._field(Phi0, "Phi0")
        u, v = self._field(U0, "U0"), self._field(V0, "V0")
        self._keep = (e, d, p, u, v)  # inputs must outlive the asynchronous launches
        _capi.check(self._lib.swx_init(self._h, _ptr(e), _ptr(d), _ptr(p), _ptr(u), _ptr(v),
                                       self._stream(stream)), "swx_init")

    def set_prev_level(self, eta_prev=None, delta_prev=None, Phi_prev=None, dead=None, stream=None) -> None:
        """After init() from a carry's current-level grids: set the physical prev level (all three
        fields or none) and the per-member dead flags (modified-Euler plans)."""
        torch = _torch()
        if (eta_prev is None) != (delta_prev is None) or (eta_prev is None) != (Phi_prev is None):
            raise ValueError("give all of eta_prev, delta_prev, Phi_prev or none")
        e, d, p = (self._field(x, n) for x, n in ((eta_prev, "eta_prev"), (delta_prev, "delta_prev"),
                                                  (Phi_prev, "Phi_prev")))
        dd = None
        if dead is not None:
            dd = torch.as_tensor(dead, device=self.device).to(torch.uint8).reshape(-1).expand(self.B).contiguous()
        self._keep_prev = (e, d, p, dd)
        _capi.check(self._lib.swx_set_prev_level(self._h, _ptr(e), _ptr(d), _ptr(p), _ptr(dd), self._stream(stream)),
                    "swx_set_prev_level")

    def run(self, t0: int, nsteps: int, *, history: bool = False, stream=None) -> Dict[str, object]:
        """Advance nsteps steps; with history, return per-step fields and diagnostics (device tensors)."""
        torch = _torch()
        out: Dict[str, object] = {}
        hf = hs = hp = hd = None
        if history and nsteps > 0:
            hf = torch.empty((self.B, 5, nsteps, self.J, self.I), dtype=torch.float64, device=self.device)
            hs = torch.zeros((self.B, nsteps, 4), dtype=torch.float64, device=self.device)
            hp = torch.empty((self.B, nsteps, self.J, 2), dtype=torch.float64, device=self.device)
            hd = torch.zeros((self.B, nsteps), dtype=torch.uint8, device=self.device)
            out = dict(fields=hf, scalars=hs, phiext=hp, dead=hd)
        _capi.check(self._lib.swx_run(self._h, int(t0), int(nsteps), _ptr(hf), _ptr(hs), _ptr(hp), _ptr(hd),
                                      int(nsteps) if history else 0, 0, self._stream(stream)), "swx_run")
        return out

    def run_graph(self, t0: int, nsteps: int, graph_steps: int = 32, stream=None) -> None:
        """Forward-only stepping that replays a captured HIP graph (needs a non-default stream)."""
        _capi.check(self._lib.swx_run_graph(self._h, int(t0), int(nsteps), int(graph_steps),
                                            self._stream(stream)), "swx_run_graph")

    # -- spectral states, checkpoints and the adjoint ------------------------
    def _spec(self, c, name="state"):
        torch = _torch()
        t = torch.as_tensor(c, dtype=torch.complex128, device=self.device)
        Mp = self.M + 1
        if t.dim() == 3:
            t = t.expand(self.B, -1, -1, -1)
        if tuple(t.shape) != (self.B, 3, Mp, Mp):
            raise ValueError(f"{name} must have shape {(self.B, 3, Mp, Mp)}, got {tuple(t.shape)}")
        return t.contiguous()

    def init_spectral(self, c, stream=None) -> None:
        """Start from spectral coefficients c (B, 3, M+1, N+1) of eta, delta, Phi."""
        t = self._spec(c, "c")
        self._keep = (t,)
        _capi.check(self._lib.swx_init_spectral(self._h, _ptr(t), self._stream(stream)), "swx_init_spectral")

    def state_get(self, stream=None):
        torch = _torch()
        out = torch.empty((self.B, 3, self.M + 1, self.M + 1), dtype=torch.complex128, device=self.device)
        _capi.check(self._lib.swx_state_get(self._h, _ptr(out), self._stream(stream)), "swx_state_get")
        return out

    def state_set(self, S, stream=None) -> None:
        t = self._spec(S)
        self._keep = (t,)
        _capi.check(self._lib.swx_state_set(self._h, _ptr(t), self._stream(stream)), "swx_state_set")

    def adjoint_begin(self, stream=None) -> None:
        _capi.check(self._lib.swx_adjoint_begin(self._h, self._stream(stream)), "swx_adjoint_begin")

    def adjoint_terminal(self, g_eta=None, g_delta=None, g_Phi=None, g_U=None, g_V=None, stream=None) -> None:
        fields = [self._field(g, n) for g, n in ((g_eta, "g_eta"), (g_delta, "g_delta"), (g_Phi, "g_Phi"),
                                                  (g_U, "g_U"), (g_V, "g_V"))]
        self._keep_adj = fields
        _capi.check(self._lib.swx_adjoint_terminal(self._h, *[_ptr(f) for f in fields], self._stream(stream)),
                    "swx_adjoint_terminal")

    def adjoint_observe(self, g_eta=None, g_delta=None, g_Phi=None, g_U=None, g_V=None, stream=None) -> None:
        """Add the cotangents of the grid fields observed at the sweep's current state."""
        fields = [self._field(g, n) for g, n in ((g_eta, "g_eta"), (g_delta, "g_delta"), (g_Phi, "g_Phi"),
                                                  (g_U, "g_U"), (g_V, "g_V"))]
        self._keep_obs = fields
        _capi.check(self._lib.swx_adjoint_observe(self._h, *[_ptr(f) for f in fields], self._stream(stream)),
                    "swx_adjoint_observe")

    def adjoint_step(self, state_t, stream=None) -> None:
        t = self._spec(state_t)
        self._keep_adj_state = t
        _capi.check(self._lib.swx_adjoint_step(self._h, _ptr(t), self._stream(stream)), "swx_adjoint_step")

    def run_tape(self, t0: int, nsteps: int, stream=None, out=None):
        """Advance nsteps steps from step index t0, returning the spectral state before each step as
        a complex128 tape (nsteps, B, 3, M+1, N+1); only the last step writes the grid carry.  With
        `out` (a contiguous complex128 tensor whose first nsteps entries take the states) the tape is
        written there and `out` is returned."""
        torch = _torch()
        Mp = self.M + 1
        if out is None:
            tape = torch.empty((int(nsteps), self.B, 3, Mp, Mp), dtype=torch.complex128, device=self.device)
        else:
            if (out.dtype != torch.complex128 or not out.is_contiguous() or out.shape[0] < int(nsteps)
                    or tuple(out.shape[1:]) != (self.B, 3, Mp, Mp)):
                raise ValueError("out must be a contiguous complex128 (>= nsteps, B, 3, M+1, N+1) tensor")
            tape = out
        _capi.check(self._lib.swx_run_tape(self._h, int(t0), int(nsteps), _ptr(tape), self._stream(stream)),
                    "swx_run_tape")
        return tape

    def adjoint_sweep(self, tape, stream=None) -> None:
        """Reverse steps over a run_tape trajectory (after adjoint_begin and the terminal cotangents)."""
        torch = _torch()
        if tape.dtype != torch.complex128 or tuple(tape.shape[1:]) != (self.B, 3, self.M + 1, self.M + 1):
            raise ValueError("tape must be complex128 (nsteps, B, 3, M+1, N+1)")
        t = tape.contiguous()
        self._keep_tape = t
        _capi.check(self._lib.swx_adjoint_sweep(self._h, int(t.shape[0]), _ptr(t), self._stream(stream)),
                    "swx_adjoint_sweep")

    def adjoint_result(self, stream=None):
        """(dL/dS_0 as complex128 (B, 3, M+1, N+1), dL/d(taurad, taudrag, DPhieq) as float64 (B, 3))."""
        torch = _torch()
        lam = torch.empty((self.B, 3, self.M + 1, self.M + 1), dtype=torch.complex128, device=self.device)
        pg = torch.empty((self.B, 3), dtype=torch.float64, device=self.device)
        _capi.check(self._lib.swx_adjoint_result(self._h, _ptr(lam), _ptr(pg), self._stream(stream)),
                    "swx_adjoint_result")
        return lam, pg

    def adjoint_diffusion(self, state, stream=None) -> None:
        """Add the diffusion partials of the pair (the held dL/dS, S = state): state is the spectral
        state whose cotangent the plan holds, produced by a step."""
        t = self._spec(state)
        self._keep_diff_state = t
        _capi.check(self._lib.swx_adjoint_diffusion(self._h, _ptr(t), self._stream(stream)), "swx_adjoint_diffusion")

    def adjoint_diffusion_result(self, stream=None):
        """Accumulated (dL/dK6 over eta and delta, dL/dK6Phi over Phi) as float64 (B, 2)."""
        torch = _torch()
        out = torch.empty((self.B, 2), dtype=torch.float64, device=self.device)
        _capi.check(self._lib.swx_adjoint_diffusion_result(self._h, _ptr(out), self._stream(stream)),
                    "swx_adjoint_diffusion_result")
        return out

    def mus_device(self):
        """The plan's Gauss nodes (J,) as a float64 device tensor."""
        if self._mus_d is None:
            torch = _torch()
            mus = self._quad_mus if self._quad_mus is not None else _capi.geometry(self.M)[0]
            self._mus_d = torch.as_tensor(np.ascontiguousarray(mus, dtype=np.float64), device=self.device)
        return self._mus_d

    def phase_curve(self, phases, normalize: bool = True, stream=None):
        """(flux, flux_raw), each float64 (B, K): the phase curve of every member's current Phi
        grid (the exported Phi_curr) at the K phase angles, normalized by mean|flux| per member
        when normalize is set (flux_raw: before normalization)."""
        torch = _torch()
        ph = torch.as_tensor(phases, dtype=torch.float64, device=self.device).reshape(-1).contiguous()
        K = ph.numel()
        flux = torch.empty((self.B, K), dtype=torch.float64, device=self.device)
        raw = torch.empty((self.B, K), dtype=torch.float64, device=self.device)
        _capi.check(self._lib.swx_plan_phase_curve(self._h, _ptr(ph), K, int(bool(normalize)), _ptr(flux), _ptr(raw),
                                                   self._stream(stream)), "swx_plan_phase_curve")
        self._keep_pc = ph
        return flux, raw

    def get_option(self, option: int) -> int:
        v = ctypes.c_int()
        _capi.check(self._lib.swx_plan_get_option(self._h, int(option), ctypes.byref(v)), "swx_plan_get_option")
        return int(v.value)

    def device_bytes(self) -> int:
        """Device memory held by the plan (tables, carry, work buffers, adjoint workspace)."""
        n = ctypes.c_size_t()
        _capi.check(self._lib.swx_plan_device_bytes(self._h, ctypes.byref(n)), "swx_plan_device_bytes")
        return int(n.value)

    def step_counts(self) -> Dict[str, int]:
        """Steps run since creation: replayed from a graph, launched eagerly, advancing the spectral
        prev lineage, writing the physical grid carry."""
        c = (ctypes.c_longlong * 4)()
        _capi.check(self._lib.swx_plan_step_counts(self._h, c), "swx_plan_step_counts")
        return dict(graph=c[0], eager=c[1], lineage=c[2], grid_carry=c[3])

    def time_kernels(self, t0: int, nsteps: int, stream=None):
        """Average (ms) of the Legendre and grid kernels: nsteps eager steps timed as one block
        (HIP events), then nsteps back-to-back grid launches on the final state (state unchanged);
        the Legendre figure is the step average minus the grid average."""
        a, b = ctypes.c_float(), ctypes.c_float()
        _capi.check(self._lib.swx_time_kernels(self._h, int(t0), int(nsteps), ctypes.byref(a), ctypes.byref(b),
                                               self._stream(stream)), "swx_time_kernels")
        return a.value, b.value

    def export(self, fields=None, spec: bool = False, stream=None, noalloc: bool = False) -> Dict[str, object]:
        """Copy the carry out as complex128/float64 device tensors with leading member axis.
        noalloc: through swx_export_state_noalloc (the library allocates nothing; the output
        tensors are still allocated here)."""
        torch = _torch()
        names = _capi.STATE_FIELDS if fields is None else tuple(fields)
        out = {}
        st = _capi.StateOut()
        Mp = self.M + 1
        for name in names:
            if name.endswith("m_prev") or name.endswith("m_curr") or name.endswith("_lag"):
                t = torch.empty((self.B, self.J, Mp), dtype=torch.complex128, device=self.device)
            else:
                t = torch.empty((self.B, self.J, self.I), dtype=torch.float64, device=self.device)
            out[name] = t
            setattr(st, name, t.data_ptr())
        dead = torch.zeros((self.B,), dtype=torch.uint8, device=self.device)
        st.dead = dead.data_ptr()
        out["dead"] = dead
        if spec:
            sp = torch.empty((self.B, 3, Mp, Mp), dtype=torch.complex128, device=self.device)
            st.spec = sp.data_ptr()
            out["spec"] = sp
        if noalloc:
            _capi.check(self._lib.swx_export_state_noalloc(self._h, ctypes.byref(st), self._stream(stream)),
                        "swx_export_state_noalloc")
        else:
            _capi.check(self._lib.swx_export_state(self._h, ctypes.byref(st), self._stream(stream)),
                        "swx_export_state")
        return out


class Ensemble:
    """Two FP32 plans of different truncations (42/63, 42/106, 63/106) stepped together: one
    Legendre launch and one grid launch per step cover both plans' members.  Each plan keeps
    its own state, parameters and tables; export them with Plan.export as usual."""

    def __init__(self, a: Plan, b: Plan):
        self._lib = _capi.load()
        self.plans = (a, b)  # keep both alive while the ensemble exists
        if a.device != b.device:
            raise ValueError("ensemble plans must be on the same device")
        h = ctypes.c_void_p()
        _capi.check(self._lib.swx_ensemble_create(a._h, b._h, ctypes.byref(h)), "swx_ensemble_create")
        self._h = h
        self.device = a.device
        for p in self.plans:
            p._add_user(self)

    def run_graph(self, t0: int, nsteps: int, graph_steps: int = 32, stream=None) -> None:
        """Advance both plans by nsteps steps from t0 (as Plan.run_graph does for each)."""
        _capi.check(self._lib.swx_ensemble_run_graph(self._h, int(t0), int(nsteps), int(graph_steps),
                                                     self.plans[0]._stream(stream)), "swx_ensemble_run_graph")

    def close(self) -> None:
        if getattr(self, "_h", None) is not None and self._h.value:
            self._lib.swx_ensemble_destroy(self._h)
            self._h = None

    def __del__(self):  # pragma: no cover - best effort
        try:
            self.close()
        except Exception:
            pass


class Pipeline:
    """2 to 4 initialized plans stepped concurrently (swx_pipeline): each plan on an internal
    stream of its own (distinct hardware queues), forked from and joined back to the caller's
    stream, so that one plan's Legendre launch overlaps another's grid launch.  Each plan's
    results are those it gives alone."""

    def __init__(self, plans):
        self._lib = _capi.load()
        self.plans = tuple(plans)  # kept alive while the pipeline exists
        if len({p.device for p in self.plans}) != 1:
            raise ValueError("pipeline plans must be on the same device")
        arr = (ctypes.c_void_p * len(self.plans))(*[p._h.value for p in self.plans])
        h = ctypes.c_void_p()
        _capi.check(self._lib.swx_pipeline_create(arr, len(self.plans), ctypes.byref(h)), "swx_pipeline_create")
        self._h = h
        self.device = self.plans[0].device
        for p in self.plans:
            p._add_user(self)

    def run_graph(self, t0: int, nsteps: int, graph_steps: int = 32, stream=None) -> None:
        """Advance every plan by nsteps steps from t0 (as Plan.run_graph does for each)."""
        _capi.check(self._lib.swx_pipeline_run_graph(self._h, int(t0), int(nsteps), int(graph_steps),
                                                     self.plans[0]._stream(stream)), "swx_pipeline_run_graph")

    def close(self) -> None:
        if getattr(self, "_h", None) is not None and self._h.value:
            self._lib.swx_pipeline_destroy(self._h)
            self._h = None

    def __del__(self):  # pragma: no cover - best effort
        try:
            self.close()
        except Exception:
            pass


def split_plans(M: int, params: np.ndarray, parts: int = 2, **kw):
    """The members of params ([B, NPARAM]) split into `parts` plans of consecutive members (sizes
    differ by at most one), each created with Plan(M, ..., **kw): the plans of a Pipeline."""
    params = np.asarray(params, dtype=np.float64)
    if params.ndim == 1:
        params = params[None, :]
    bounds = split_bounds(params.shape[0], parts)
    return [Plan(M, params[bounds[k]:bounds[k + 1]], **kw) for k in range(parts)]


def split_bounds(B: int, parts: int):
    """Member boundaries [0, ..., B] of `parts` consecutive groups whose sizes differ by at most one."""
    if not 1 <= parts <= min(B, 4):
        raise ValueError("parts must be between 1 and min(members, 4)")
    return [int(x) for x in np.linspace(0, B, parts + 1).round().astype(int)]


def measure_ceilings(nbytes: int = 1 << 30, stream=None) -> Dict[str, float]:
    """Ceilings of the current GPU measured with the step kernels' instructions and access patterns
    (swx_measure_ceilings): matrix-core TFLOP/s of the FP32 16x16x4, FP64 16x16x4 and BF16 16x16x16
    MFMA instructions, and HBM TB/s of a device copy (bytes read + written), a read-only and a
    write-only pass over `nbytes` bytes."""
    torch = _torch()
    require_gpu()
    lib = _capi.load()
    s = stream if stream is not None else torch.cuda.current_stream()
    out = (ctypes.c_double * 6)()
    _capi.check(lib.swx_measure_ceilings(out, ctypes.c_size_t(int(nbytes)), ctypes.c_void_p(s.cuda_stream)),
                "swx_measure_ceilings")
    keys = ("mfma_fp32_tflops", "mfma_fp64_tflops", "mfma_bf16_tflops", "hbm_copy_tbps", "hbm_read_tbps",
            "hbm_write_tbps")
    return {k: float(v) for k, v in zip(keys, out)}
