"""Per-step history and diagnostics on the matrix-core kernels (GPU, FP32 transforms): the
reference's default run_model_scan (diagnostics=True, return_history=True; model.py:1345-1388,
outs at :1175-1189, blow-up gate at :838-864) with precision="fp32" steps through the folded MFMA
grid kernel's history variant and the MFMA Legendre kernel behind the blow-up gate kernel."""

from __future__ import annotations

import numpy as np
import pytest

from oracle import swampe_oracle as orc

pytestmark = pytest.mark.gpu

KW = dict(M=42, dt=120.0, tmax=14, Phibar=3e5, omega=3.2e-5, a=8.2e7, taurad=86400.0, taudrag=86400.0, DPhieq=3e5)


def _rel(g, r):
    return float(np.linalg.norm(g - r) / np.linalg.norm(r))


def _balanced_flow():
    """Forced run from the balanced zonal flow of Williamson test 2 (U ~ 40 m/s): winds well above
    the FP32 transforms' rounding (in a state near rest the winds are the small residual of large
    balanced terms, which FP32 rounding dominates in every FP32 path, plain kernels included)."""
    kw = dict(M=42, dt=120.0, tmax=14, Phibar=3e4, omega=7.2921159e-5, a=6.37122e6, taurad=86400.0,
              taudrag=86400.0, DPhieq=3e4)
    s = orc.build_static(M=42, dt=120.0, a=kw["a"], omega=kw["omega"], g=9.8, Phibar=kw["Phibar"], taurad=86400.0,
                         taudrag=86400.0, DPhieq=3e4, K6=1.24e33, K6Phi=None, test=2)
    eta0, delta0, Phi0, U0, V0 = orc.analytic_ic(s, 2, 0.3)
    return dict(kw, eta0_init=eta0, delta0_init=delta0, Phi0_init=Phi0, U0_init=U0, V0_init=V0)


def test_fp32_history_and_diagnostics_vs_oracle(torch_gpu):
    """12 steps: every step's eta, Phi, U, V and the rms, spin_min, phi_min, phi_max series against
    the FP64 oracle (relative L2 <= 5e-6 on the fields, 1e-6 relative on rms, the Phi extrema within
    1e-6 of the largest |Phi|, spin_min within 1e-5 of the largest wind); the final state's fields
    are the history's last entries bit for bit."""
    from my_swampe_amd.model import run_model_scan

    print("[gpu-case] fp32 history+diagnostics", flush=True)
    kw = _balanced_flow()
    ref = orc.run_model_scan(**kw)
    got = run_model_scan(**kw, precision="fp32")
    outs = {k: v.detach().cpu().numpy() for k, v in got["outs"].items()}
    for f in ("eta", "Phi", "U", "V"):
        worst = max(_rel(outs[f][k], ref["outs"][f][k]) for k in range(outs[f].shape[0]))
        print(f, "worst step relL2 vs oracle", worst)
        assert worst <= 5e-6, (f, worst)
    np.testing.assert_allclose(outs["rms"], ref["outs"]["rms"], rtol=1e-6, err_msg="rms")
    # The extrema are single grid values (spin_min where the winds nearly cancel): their FP32 error
    # is bounded against the field's scale, not against themselves.
    phi_scale = float(np.abs(ref["outs"]["Phi"]).max())
    for f in ("phi_min", "phi_max"):
        np.testing.assert_allclose(outs[f], ref["outs"][f], rtol=0, atol=1e-6 * phi_scale, err_msg=f)
    scale = float(np.abs(ref["outs"]["U"]).max())
    np.testing.assert_allclose(outs["spin_min"], ref["outs"]["spin_min"], rtol=0, atol=1e-5 * scale)
    assert not outs["dead"].any() and int(got["dead_first_idx"]) == -1
    last = got["last_state"]
    for f in ("eta", "delta", "Phi", "U", "V"):
        np.testing.assert_array_equal(getattr(last, f + "_curr").cpu().numpy(), outs[f][-1], err_msg=f)


def test_fp32_history_kernels_differ_from_plain_and_agree(torch_gpu):
    """The same 12 history steps through the matrix-core kernels and through the plain FP32 kernels
    (SWX_OPT_GRID_KERNEL = SWX_OPT_LEG_KERNEL = 1): not the same arithmetic (the results differ),
    the same steps (relative L2 <= 5e-6 apart)."""
    from my_swampe_amd import _capi
    from my_swampe_amd.model import param_row
    from my_swampe_amd.plan import Plan

    torch = torch_gpu
    print("[gpu-case] fp32 history mfma vs plain", flush=True)
    kw = _balanced_flow()
    p = param_row(dt=120.0, a=kw["a"], omega=kw["omega"], g=9.8, Phibar=kw["Phibar"], taurad=86400.0,
                  taudrag=86400.0, DPhieq=3e4, K6=1.24e33, K6Phi=None, alpha=0.01)
    ics = tuple(kw[k] for k in ("eta0_init", "delta0_init", "Phi0_init", "U0_init", "V0_init"))
    res = {}
    for plain in (0, 1):
        plan = Plan(42, np.stack([p, p, p]), precision="fp32", diagnostics=True)
        plan.set_option(_capi.OPT_GRID_KERNEL, plain)
        plan.set_option(_capi.OPT_LEG_KERNEL, plain)
        plan.init(*ics)
        hist = plan.run(2, 12, history=True)
        torch.cuda.synchronize()
        res[plain] = {k: v.cpu().numpy() for k, v in hist.items()}
        plan.close()
    a, b = res[0]["fields"], res[1]["fields"]
    assert not np.array_equal(a, b)
    for f in (0, 2, 3, 4):  # eta, Phi, U, V
        rel = _rel(a[:, f], b[:, f])
        print("field", f, "mfma vs plain relL2", rel)
        assert rel <= 5e-6, (f, rel)
    np.testing.assert_allclose(res[0]["scalars"][..., 0], res[1]["scalars"][..., 0], rtol=1e-6)


def test_fp32_blowup_gate_freezes_member(torch_gpu):
    """Huge initial winds trip rms > blowup_rms at the first step (model.py:838-864) on the matrix-core
    path: every step dead, the history frozen at the initial Phi, rms as the oracle's."""
    from my_swampe_amd.model import run_model_scan

    print("[gpu-case] fp32 blowup", flush=True)
    s = orc.build_static(M=42, dt=120.0, a=8.2e7, omega=3.2e-5, g=9.8, Phibar=3e5, taurad=86400.0, taudrag=86400.0,
                         DPhieq=3e5, K6=1.24e33, K6Phi=None, test=None)
    eta0, delta0, Phi0, U0, V0 = orc.analytic_ic(s, None, 0.05)
    U0 = U0 + 2.0e4
    kw = dict(KW, tmax=6, eta0_init=eta0, delta0_init=delta0, Phi0_init=Phi0, U0_init=U0, V0_init=V0)
    ref = orc.run_model_scan(**kw)
    got = run_model_scan(**kw, precision="fp32")
    assert int(got["dead_first_idx"]) == ref["dead_first_idx"] == 0
    assert bool(got["outs"]["dead"].all())
    np.testing.assert_array_equal(got["outs"]["Phi"][-1].cpu().numpy(), Phi0)
    np.testing.assert_allclose(got["outs"]["rms"].cpu().numpy(), ref["outs"]["rms"], rtol=1e-6)
