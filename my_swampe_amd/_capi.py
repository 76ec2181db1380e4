"""ctypes bindings of libswampe_hip.so (declarations in include/swampe_capi.h).

Loading fails loudly: there is no CPU fallback for any entry point.
"""

from __future__ import annotations

import ctypes
from ctypes import POINTER, c_char_p, c_double, c_int, c_ubyte, c_void_p
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "_lib" / "libswampe_hip.so"

SWX_OK = 0
SWX_ERR_ARG = -1
SWX_ERR_HIP = -2
SWX_ERR_STATE = -3

PREC_FP64 = 0
PREC_FP32 = 1
PREC_BF16X3 = 2

OPT_GRID_KERNEL = 1
OPT_FULL_PHYS_CARRY = 2
OPT_LEG_KERNEL = 3
OPT_PHYS_STEPS = 4
OPT_LEG_MB = 5
OPT_GRID_FOLD = 6
OPT_CONCURRENT = 7
OPT_PHYS_LINEAGE = 8
OPT_LEG_FOLD = 9
OPT_ADJ_DIFFUSION = 10
PRECISIONS = {"fp64": PREC_FP64, "fp32": PREC_FP32, "bf16x3": PREC_BF16X3}

PARAM_NAMES = ("dt", "a", "omega", "g", "Phibar", "taurad", "taudrag", "DPhieq", "K6", "K6Phi", "alpha", "blowup_rms",
               "si_alpha", "williams_alpha")
NPARAM = len(PARAM_NAMES)

# Every extern "C" symbol declared in include/swampe_capi.h.
EXPORTED_SYMBOLS = (
    "swx_geometry", "swx_grid_size", "swx_plan_create", "swx_plan_set_params", "swx_plan_destroy",
    "swx_init", "swx_run", "swx_export_state", "swx_run_graph", "swx_time_kernels", "swx_plan_set_option",
    "swx_init_spectral", "swx_state_get", "swx_state_set", "swx_adjoint_begin", "swx_adjoint_terminal", "swx_adjoint_observe",
    "swx_phase_curve", "swx_phase_curve_vjp", "swx_plan_phase_curve", "swx_set_prev_level",
    "swx_adjoint_step", "swx_adjoint_result", "swx_adjoint_diffusion", "swx_adjoint_diffusion_result",
    "swx_last_error", "swx_version", "swx_plan_device_bytes",
    "swx_plan_step_counts", "swx_export_state_noalloc", "swx_ensemble_create", "swx_ensemble_run_graph", "swx_ensemble_destroy", "swx_plan_get_option",
    "swx_pipeline_create", "swx_pipeline_run_graph", "swx_pipeline_destroy", "swx_run_tape", "swx_adjoint_sweep",
    "swx_diagnostic_eta_delta", "swx_measure_ceilings", "swx_gauss_legendre", "swx_lambdas", "swx_legendre_basis",
    "swx_fwd_fft_trunc", "swx_invrs_fft", "swx_fwd_leg", "swx_invrs_leg", "swx_invrs_uv",
)

STATE_FIELDS = (
    "etam_prev", "etam_curr", "deltam_prev", "deltam_curr", "Phim_prev", "Phim_curr",
    "eta_prev", "eta_curr", "delta_prev", "delta_curr", "Phi_prev", "Phi_curr", "U_curr", "V_curr",
    "Um_curr", "Vm_curr", "Am_curr", "Bm_curr", "Cm_curr", "Dm_curr", "Em_curr",
    "PhiFm_curr", "Fm_curr", "Gm_curr", "Rum_lag", "Rvm_lag",
)


class PlanDesc(ctypes.Structure):
    _fields_ = [
        ("M", c_int), ("n_members", c_int), ("precision", c_int), ("expflag", c_int), ("forcflag", c_int),
        ("diffflag", c_int), ("modalflag", c_int), ("diagnostics", c_int), ("test", c_int),
        ("exact_basis", c_int), ("semi_implicit", c_int), ("raw_filter", c_int),
        ("quad_mus", c_void_p), ("quad_w", c_void_p),
    ]


class StateOut(ctypes.Structure):
    _fields_ = [(name, c_void_p) for name in STATE_FIELDS] + [("dead", c_void_p), ("spec", c_void_p)]


class SwxError(RuntimeError):
    """Error reported by the native library (message from swx_last_error)."""


_lib = None


def load() -> ctypes.CDLL:
    """Load (once) and return the native library; raise if it was not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `python -m my_swampe_amd.build` "
            "(or __graft_entry__.build()). There is no CPU fallback."
        )
    # The library needs libamdhip64.so.7 by soname.  torch ships its own copy under that soname:
    # loading torch first makes the one runtime torch uses serve the library too (loading the
    # library first would put a second HIP runtime beside torch's, and device calls then fail).
    import torch  # noqa: F401

    lib = ctypes.CDLL(str(LIB_PATH))
    lib.swx_last_error.restype = c_char_p
    lib.swx_version.restype = c_char_p
    lib.swx_geometry.argtypes = [c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]
    lib.swx_grid_size.argtypes = [c_int, POINTER(c_int), POINTER(c_int)]
    lib.swx_plan_create.argtypes = [POINTER(PlanDesc), c_void_p, POINTER(c_void_p)]
    lib.swx_plan_set_params.argtypes = [c_void_p, c_void_p, c_void_p]
    lib.swx_plan_destroy.argtypes = [c_void_p]
    lib.swx_plan_destroy.restype = None
    lib.swx_init.argtypes = [c_void_p] + [c_void_p] * 5 + [c_void_p]
    lib.swx_set_prev_level.argtypes = [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]
    lib.swx_run.argtypes = [c_void_p, c_int, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_int, c_void_p]
    lib.swx_export_state.argtypes = [c_void_p, POINTER(StateOut), c_void_p]
    lib.swx_run_graph.argtypes = [c_void_p, c_int, c_int, c_int, c_void_p]
    lib.swx_plan_set_option.argtypes = [c_void_p, c_int, c_int]
    lib.swx_init_spectral.argtypes = [c_void_p, c_void_p, c_void_p]
    lib.swx_state_get.argtypes = [c_void_p, c_void_p, c_void_p]
    lib.swx_state_set.argtypes = [c_void_p, c_void_p, c_void_p]
    lib.swx_adjoint_begin.argtypes = [c_void_p, c_void_p]
    lib.swx_adjoint_terminal.argtypes = [c_void_p] + [c_void_p] * 5 + [c_void_p]
    lib.swx_adjoint_observe.argtypes = [c_void_p] + [c_void_p] * 5 + [c_void_p]
    lib.swx_adjoint_step.argtypes = [c_void_p, c_void_p, c_void_p]
    lib.swx_run_tape.argtypes = [c_void_p, c_int, c_int, c_void_p, c_void_p]
    lib.swx_adjoint_sweep.argtypes = [c_void_p, c_int, c_void_p, c_void_p]
    lib.swx_adjoint_result.argtypes = [c_void_p, c_void_p, c_void_p, c_void_p]
    lib.swx_adjoint_diffusion.argtypes = [c_void_p, c_void_p, c_void_p]
    lib.swx_adjoint_diffusion_result.argtypes = [c_void_p, c_void_p, c_void_p]
    lib.swx_phase_curve.argtypes = [c_void_p, ctypes.c_longlong, c_int, c_int, c_int, c_void_p, c_void_p, c_int, c_int,
                                    c_void_p, c_void_p, c_void_p]
    lib.swx_phase_curve_vjp.argtypes = [c_void_p, c_void_p, c_int, c_int, c_int, c_void_p, c_void_p, c_int, c_int,
                                        c_void_p, c_void_p]
    lib.swx_plan_phase_curve.argtypes = [c_void_p, c_void_p, c_int, c_int, c_void_p, c_void_p, c_void_p]
    lib.swx_diagnostic_eta_delta.argtypes = ([c_void_p] * 3 + [c_int] * 5 + [c_void_p] * 5 + [ctypes.c_double] +
                                             [c_void_p] * 5)
    lib.swx_time_kernels.argtypes = [c_void_p, c_int, c_int, POINTER(ctypes.c_float), POINTER(ctypes.c_float), c_void_p]
    lib.swx_measure_ceilings.argtypes = [c_void_p, ctypes.c_size_t, c_void_p]
    lib.swx_gauss_legendre.argtypes = [c_int, c_void_p, c_void_p]
    lib.swx_lambdas.argtypes = [c_int, c_void_p]
    lib.swx_legendre_basis.argtypes = [c_int, c_int, c_int, c_void_p, c_int, c_void_p, c_void_p]
    lib.swx_fwd_fft_trunc.argtypes = [c_void_p] + [c_int] * 6 + [c_void_p, c_void_p]
    lib.swx_invrs_fft.argtypes = [c_void_p] + [c_int] * 4 + [c_void_p, c_void_p]
    lib.swx_fwd_leg.argtypes = [c_void_p] * 3 + [c_int] * 4 + [c_void_p, c_void_p]
    lib.swx_invrs_leg.argtypes = [c_void_p] * 2 + [c_int] * 5 + [c_void_p, c_void_p]
    lib.swx_invrs_uv.argtypes = [c_void_p] * 3 + [c_int] * 5 + [c_void_p] * 9
    lib.swx_plan_device_bytes.argtypes = [c_void_p, POINTER(ctypes.c_size_t)]
    lib.swx_plan_step_counts.argtypes = [c_void_p, POINTER(ctypes.c_longlong)]
    lib.swx_export_state_noalloc.argtypes = lib.swx_export_state.argtypes
    lib.swx_plan_get_option.argtypes = [c_void_p, c_int, POINTER(c_int)]
    lib.swx_ensemble_create.argtypes = [c_void_p, c_void_p, POINTER(c_void_p)]
    lib.swx_ensemble_run_graph.argtypes = [c_void_p, c_int, c_int, c_int, c_void_p]
    lib.swx_ensemble_destroy.argtypes = [c_void_p]
    lib.swx_ensemble_destroy.restype = None
    lib.swx_pipeline_create.argtypes = [POINTER(c_void_p), c_int, POINTER(c_void_p)]
    lib.swx_pipeline_run_graph.argtypes = [c_void_p, c_int, c_int, c_int, c_void_p]
    lib.swx_pipeline_destroy.argtypes = [c_void_p]
    lib.swx_pipeline_destroy.restype = None
    for name in EXPORTED_SYMBOLS:
        getattr(lib, name)  # AttributeError when a declared symbol is missing
    for name in ("swx_geometry", "swx_grid_size", "swx_plan_create", "swx_plan_set_params", "swx_init",
                 "swx_run", "swx_export_state", "swx_run_graph", "swx_time_kernels", "swx_plan_set_option",
                 "swx_init_spectral", "swx_state_get", "swx_state_set", "swx_adjoint_begin",
                 "swx_adjoint_terminal", "swx_adjoint_step", "swx_adjoint_result", "swx_adjoint_diffusion",
                 "swx_adjoint_diffusion_result", "swx_plan_device_bytes",
                 "swx_plan_step_counts", "swx_export_state_noalloc",
                 "swx_ensemble_create", "swx_ensemble_run_graph", "swx_plan_get_option", "swx_adjoint_observe",
                 "swx_phase_curve", "swx_phase_curve_vjp", "swx_plan_phase_curve", "swx_set_prev_level",
                 "swx_pipeline_create", "swx_pipeline_run_graph", "swx_run_tape", "swx_adjoint_sweep",
                 "swx_diagnostic_eta_delta", "swx_measure_ceilings", "swx_gauss_legendre", "swx_lambdas",
                 "swx_legendre_basis", "swx_fwd_fft_trunc", "swx_invrs_fft", "swx_fwd_leg", "swx_invrs_leg",
                 "swx_invrs_uv"):
        getattr(lib, name).restype = c_int
    _lib = lib
    return lib


def check(rc: int, what: str) -> None:
    """Raise SwxError (ValueError for argument errors) when rc is not SWX_OK."""
    if rc == SWX_OK:
        return
    msg = load().swx_last_error().decode(errors="replace")
    if rc == SWX_ERR_ARG:
        raise ValueError(f"{what}: {msg}")
    raise SwxError(f"{what} failed ({rc}): {msg}")


def grid_size(M: int):
    """(J, I) of truncation M; ValueError for an unsupported M."""
    J, I = c_int(), c_int()
    check(load().swx_grid_size(int(M), ctypes.byref(J), ctypes.byref(I)), "swx_grid_size")
    return J.value, I.value


def reference_quadrature(J: int):
    """Gauss nodes (ascending) and weights as the reference takes them: scipy.special.roots_legendre,
    or numpy's leggauss where SciPy is missing (spectral_transform.py:80-114)."""
    import numpy as np

    try:
        from scipy import special as sp

        mus, w = sp.roots_legendre(int(J))
    except ImportError:  # pragma: no cover - SciPy is part of the image
        from numpy.polynomial.legendre import leggauss

        mus, w = leggauss(int(J))
    return np.ascontiguousarray(mus, dtype=np.float64), np.ascontiguousarray(w, dtype=np.float64)


def geometry(M: int, exact_basis: bool = False):
    """Host FP64 geometry (mus, w, lambdas, Pmn, Hmn) as numpy arrays (no GPU needed)."""
    import numpy as np

    J, I = grid_size(M)
    mus = np.zeros(J)
    w = np.zeros(J)
    lam = np.zeros(I)
    P = np.zeros((J, M + 1, M + 1))
    H = np.zeros((J, M + 1, M + 1))
    ptr = lambda a: a.ctypes.data_as(c_void_p)  # noqa: E731
    check(load().swx_geometry(int(M), int(bool(exact_basis)), ptr(mus), ptr(w), ptr(lam), ptr(P), ptr(H)), "swx_geometry")
    return mus, w, lam, P, H
