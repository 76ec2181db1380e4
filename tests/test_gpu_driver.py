"""The drop-in driver (GPU): cached plans across run_model_scan calls, graph-replayed stepping of
calls without history, deterministic kernel choices, plan lifetimes."""

from __future__ import annotations

import numpy as np
import pytest

from oracle import swampe_oracle as orc

pytestmark = pytest.mark.gpu

BASE = dict(M=42, dt=120.0, Phibar=3e5, omega=3.2e-5, a=8.2e7, DPhieq=3e5, diffflag=True, modalflag=True)


def _ics():
    eta0, delta0, Phi0, U0, V0 = orc.rest_noise_ic(42, 0, 3.2e-5)
    return dict(eta0_init=eta0, delta0_init=delta0, Phi0_init=Phi0, U0_init=U0, V0_init=V0)


def _fields(st):
    return {f: getattr(st, f).detach().cpu().numpy() for f in ("eta_prev", "eta_curr", "Phi_prev", "Phi_curr",
                                                               "U_curr", "V_curr", "etam_curr")}


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_cached_plans_equal_fresh_plans(precision, torch_gpu):
    """run_model_scan_final keeps one plan per structure across calls and loads each call's
    parameters into it: calls with other parameters in between give bitwise what a fresh plan
    gives; the graph-stepped final state equals the eagerly stepped one of run_model_scan with
    history (bitwise, FP64: the same kernels)."""
    from my_swampe_amd import model

    print(f"[gpu-case] driver cache {precision}", flush=True)
    kw = dict(BASE, tmax=60, precision=precision, **_ics())
    model._RUN_PLANS.clear()
    a1 = _fields(model.run_model_scan_final(taurad=86400.0, taudrag=86400.0, **kw)["last_state"])
    b = _fields(model.run_model_scan_final(taurad=3e4, taudrag=2e5, **kw)["last_state"])
    a2 = _fields(model.run_model_scan_final(taurad=86400.0, taudrag=86400.0, **kw)["last_state"])
    assert len(model._RUN_PLANS) == 1
    for k in a1:
        np.testing.assert_array_equal(a2[k], a1[k], err_msg=k)
    for p in list(model._RUN_PLANS.values()):
        p.close()
    model._RUN_PLANS.clear()
    b_fresh = _fields(model.run_model_scan_final(taurad=3e4, taudrag=2e5, **kw)["last_state"])
    for k in b:
        np.testing.assert_array_equal(b[k], b_fresh[k], err_msg=k)
    assert not np.array_equal(a1["Phi_curr"], b["Phi_curr"])
    if precision == "fp64":
        h = _fields(model.run_model_scan(taurad=86400.0, taudrag=86400.0, diagnostics=False, **kw)["last_state"])
        for k in ("eta_curr", "Phi_curr", "U_curr", "V_curr", "etam_curr"):
            np.testing.assert_array_equal(h[k], a1[k], err_msg=k)


def test_fresh_fp32_plans_round_alike(torch_gpu):
    """Two fresh FP32 plans with every option left at its default give bitwise-equal states: the
    grid kernel variant is a fixed rule and the Legendre member-group size does not change results."""
    from my_swampe_amd.model import param_row
    from my_swampe_amd.plan import Plan

    torch = torch_gpu
    print("[gpu-case] fresh fp32 plans", flush=True)
    rng = np.random.default_rng(3)
    params = np.stack([param_row(dt=120.0, a=8.2e7, omega=3.2e-5, g=9.8, Phibar=3e5, taurad=10 ** rng.uniform(4, 6),
                                 taudrag=10 ** rng.uniform(4, 6), DPhieq=rng.uniform(1e5, 5e5), K6=1.24e33,
                                 K6Phi=None, alpha=0.01) for _ in range(6)])
    ics = orc.rest_noise_ic(42, 0, 3.2e-5)
    out = []
    for _ in range(2):
        plan = Plan(42, params, precision="fp32")
        plan.init(*ics)
        plan.run(2, 30)
        out.append(plan.state_get().cpu().numpy())
        plan.close()
    np.testing.assert_array_equal(out[0], out[1])


def test_closing_a_plan_closes_its_pipelines_and_ensembles(torch_gpu):
    from my_swampe_amd.model import param_row
    from my_swampe_amd.plan import Ensemble, Pipeline, Plan

    torch = torch_gpu
    print("[gpu-case] plan lifetime", flush=True)
    p = param_row(dt=120.0, a=8.2e7, omega=3.2e-5, g=9.8, Phibar=3e5, taurad=86400.0, taudrag=86400.0, DPhieq=3e5,
                  K6=1.24e33, K6Phi=None, alpha=0.01)
    ics = orc.rest_noise_ic(42, 0, 3.2e-5)
    a, b = Plan(42, np.stack([p, p]), precision="fp32"), Plan(42, np.stack([p]), precision="fp32")
    c = Plan(63, np.stack([p]), precision="fp32")
    for pl in (a, b):
        pl.init(*ics)
    c.init(*orc.rest_noise_ic(63, 0, 3.2e-5))
    torch.cuda.synchronize()
    pipe = Pipeline([a, b])
    ens = Ensemble(b, c)
    a.close()
    assert pipe._h is None  # closed with its plan
    b.close()
    assert ens._h is None
    with pytest.raises(Exception):
        pipe.run_graph(2, 3, stream=torch.cuda.Stream())
    c.close()


def test_measured_ceilings_are_plausible(torch_gpu):
    """swx_measure_ceilings: positive figures below the datasheet ceilings plus a margin (FP32 matrix
    157.3, FP64 matrix 78.6, BF16 matrix 2500 TFLOP/s; HBM 8 TB/s), above a tenth of them; the copy
    counts its bytes twice; sizes below 256 MiB are refused before any device work."""
    from my_swampe_amd.plan import measure_ceilings

    c = measure_ceilings(1 << 28)
    print("[gpu-case] ceilings", c, flush=True)
    for key, sheet in (("mfma_fp32_tflops", 157.3), ("mfma_fp64_tflops", 78.6), ("mfma_bf16_tflops", 2500.0),
                       ("hbm_copy_tbps", 8.0), ("hbm_read_tbps", 8.0), ("hbm_write_tbps", 8.0)):
        assert 0.1 * sheet < c[key] < 1.1 * sheet, (key, c[key])
    with pytest.raises(ValueError, match="256 MiB"):
        measure_ceilings(1 << 20)


def test_legacy_main_equals_run_model(torch_gpu):
    """main_function.main with the legacy conventions (test=0 for forced, 0/1 flags, the unused
    physics constants) returns what run_model returns for the same run, array for array, on the
    forced fixture case (unit_tests/fixtures/reference_case_forced_default.npz parameters)."""
    from my_swampe_amd.main_function import main
    from my_swampe_amd.model import run_model

    print("[gpu-case] legacy main", flush=True)
    kw = dict(taurad=36000.0, taudrag=21600.0, DPhieq=1.0e6, verbose=False)
    got = main(42, 1200.0, 6, 3.0e5, 3.2e-5, 8.2e7, 0, plotflag=0, saveflag=0, diffflag=1, modalflag=1, k1=1.0,
               sigmaSB=0.0, **kw)
    want = run_model(42, 1200.0, 6, 3.0e5, 3.2e-5, 8.2e7, None, plotflag=False, saveflag=False, **kw)
    assert set(got) == set(want)
    for k in want:
        assert np.array_equal(np.asarray(got[k]), np.asarray(want[k])), k
