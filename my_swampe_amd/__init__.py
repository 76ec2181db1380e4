"""MI355X-native spectral time-step core of SWAMPE-JAX (drop-in for my_swampe.model).

Public API mirrors my_swampe/__init__.py:64-71 (run_model, run_model_gpu,
run_model_scan, run_model_scan_final, assert_finite_state) and its submodules
continuation, spectral_transform, initial_conditions, filters, forcing,
explicit_tdiff, modEuler_tdiff, semi_implicit_tdiff, time_stepping, main_function (the legacy main) and
model, plus the batched ensemble driver
(plan) and the differentiable rollouts (autograd).  All time stepping runs in the HIP kernels of
my_swampe_amd/_lib/libswampe_hip.so.
"""

from . import (  # noqa: F401
    autodiff_utils,
    continuation,
    explicit_tdiff,
    filters,
    forcing,
    initial_conditions,
    main_function,
    modEuler_tdiff,
    model,
    semi_implicit_tdiff,
    spectral_transform,
    time_stepping,
)
from .model import (  # noqa: F401
    RunFlags,
    State,
    Static,
    assert_finite_state,
    build_static,
    run_model,
    run_model_gpu,
    run_model_scan,
    run_model_scan_final,
)
from .autodiff_utils import fwd_grad  # noqa: F401

__version__ = "0.1.0"

__all__ = [
    "RunFlags", "State", "Static", "assert_finite_state", "build_static", "run_model", "run_model_gpu",
    "run_model_scan", "run_model_scan_final", "__version__", "autodiff_utils", "continuation", "fwd_grad", "explicit_tdiff", "filters",
    "forcing", "initial_conditions", "main_function", "modEuler_tdiff", "model", "semi_implicit_tdiff", "spectral_transform",
    "time_stepping",
]
