"""Operator-level seams on the GPU: _init_state_from_fields, _step_once, _step_once_state_only,
simulate_scan, simulate_scan_last, _diagnose_winds, build_static + dataclasses.replace (the imports
of the reference's external consumers, model.py:644-772, :795-1342).

Bars: the seams stepping their own carry are bit-identical to run_model_scan (same kernels, same
order); a State made elsewhere (the oracle's mid-run carry) is loaded and stepped to the reference
parity contract against the oracle's own step (eta/delta 1e-10, Phi 5e-8, U/V 1e-9).
"""

from __future__ import annotations

import dataclasses

import numpy as np
import pytest

from conftest import ATOL, FIXTURE_CASES
from oracle import swampe_oracle as orc

pytestmark = pytest.mark.gpu

FORCED = FIXTURE_CASES["reference_case_forced_default"]
FIELDS = ("eta", "delta", "Phi", "U", "V")


def _static_flags(case=FORCED, diagnostics=True):
    from my_swampe_amd.model import RunFlags, build_static

    static = build_static(M=case["M"], dt=case["dt"], a=case["a"], omega=case["omega"], g=case["g"],
                          Phibar=case["Phibar"], taurad=case["taurad"], taudrag=case["taudrag"],
                          DPhieq=case["DPhieq"], K6=case["K6"], K6Phi=None, test=case["test"], alpha=case["alpha"])
    flags = RunFlags(forcflag=case["forcflag"], diffflag=case["diffflag"], expflag=case["expflag"],
                     modalflag=case["modalflag"], diagnostics=diagnostics, alpha=case["alpha"])
    return static, flags


def _ic(static, case=FORCED):
    from my_swampe_amd.model import _analytic_ic

    return _analytic_ic(static, case["test"], 0.05)


def test_simulate_scan_last_equals_run_model_scan_final(torch_gpu):
    from my_swampe_amd.model import _init_state_from_fields, run_model_scan_final, simulate_scan_last

    torch = torch_gpu
    print("[gpu-case] seams simulate_scan_last", flush=True)
    static, flags = _static_flags(diagnostics=False)
    eta0, delta0, Phi0, U0, V0 = _ic(static)
    st0 = _init_state_from_fields(static=static, flags=flags, test=None, eta0=eta0, delta0=delta0, Phi0=Phi0,
                                  U0=U0, V0=V0)
    last = simulate_scan_last(static=static, flags=flags, state0=st0, t_seq=torch.arange(2, 12), test=None,
                              Uic=U0, Vic=V0)
    ref = run_model_scan_final(**{**FORCED, "tmax": 12})["last_state"]
    for name in ("eta_curr", "delta_curr", "Phi_curr", "U_curr", "V_curr", "Phi_prev", "Phim_curr", "Am_curr"):
        assert torch.equal(getattr(last, name), getattr(ref, name)), name


def test_step_once_chain_equals_simulate_scan_and_run_model_scan(torch_gpu):
    from my_swampe_amd.model import _init_state_from_fields, _step_once, run_model_scan, simulate_scan

    torch = torch_gpu
    print("[gpu-case] seams step_once chain", flush=True)
    static, flags = _static_flags()
    eta0, delta0, Phi0, U0, V0 = _ic(static)
    st = _init_state_from_fields(static=static, flags=flags, test=None, eta0=eta0, delta0=delta0, Phi0=Phi0,
                                 U0=U0, V0=V0)
    st0 = st
    steps = []
    for t in range(2, 8):
        st, out = _step_once(st, t, static, flags, None, U0, V0)
        steps.append(out)
    ref = run_model_scan(**{**FORCED, "tmax": 8})
    for k, out in enumerate(steps):
        assert int(out["t"]) == k + 2
        for f in FIELDS:
            assert torch.equal(out[f], ref["outs"][f][k]), (k, f)
        for f in ("rms", "spin_min", "phi_min", "phi_max"):
            assert torch.equal(out[f], ref["outs"][f][k]), (k, f)
    assert torch.equal(st.Phi_curr, ref["last_state"].Phi_curr)
    # The same scan from the saved initial carry (a reload of st0: no longer the plan's carry).
    last, outs = simulate_scan(static=static, flags=flags, state0=st0, t_seq=np.arange(2, 8), test=None,
                               Uic=U0, Vic=V0)
    assert outs["Phi"].shape == (6, static.J, static.I)
    assert torch.equal(outs["Phi"], ref["outs"]["Phi"])
    assert torch.equal(last.eta_curr, ref["last_state"].eta_curr)


def test_foreign_mid_run_state_vs_oracle_step(torch_gpu):
    """A mid-run carry made by the oracle (prev != curr after the Robert-Asselin filter) is loaded
    from its grids and stepped; the result matches the oracle's step at the parity contract."""
    from my_swampe_amd.model import State, _step_once, _step_once_state_only

    torch = torch_gpu
    print("[gpu-case] seams foreign state", flush=True)
    static, flags = _static_flags()
    o = orc.run_model_scan(**{**FORCED, "tmax": 6}, return_history=False)
    ost, os_ = o["last_state"], o["static"]
    assert not np.array_equal(ost.Phi_prev, ost.Phi_curr)
    dev = torch.device("cuda")
    conv = lambda x: torch.as_tensor(np.asarray(x), device=dev)  # noqa: E731
    state = State(**{n: (torch.tensor(bool(ost.dead), device=dev) if n == "dead" else conv(getattr(ost, n)))
                     for n in State._fields})
    eta0, delta0, Phi0, U0, V0 = orc.analytic_ic(os_, None, 0.05)
    oflags = dict(forcflag=True, diffflag=True, expflag=False, modalflag=True, diagnostics=True, alpha=0.01,
                  blowup_rms=8000.0)
    ref, ref_out = orc.step_once(ost, 6, os_, oflags, None, U0, V0)
    got, out = _step_once(state, 6, static, flags, None, U0, V0)
    for f in FIELDS:
        np.testing.assert_allclose(getattr(got, f + "_curr").cpu().numpy(), getattr(ref, f + "_curr"), rtol=0,
                                   atol=ATOL[f], err_msg=f)
        np.testing.assert_allclose(out[f].cpu().numpy(), ref_out[f], rtol=0, atol=ATOL[f], err_msg=f)
    for f in ("eta", "delta", "Phi"):
        np.testing.assert_allclose(getattr(got, f + "_prev").cpu().numpy(), getattr(ref, f + "_prev"), rtol=0,
                                   atol=ATOL[f], err_msg=f + "_prev")
    np.testing.assert_allclose(float(out["rms"]), ref_out["rms"], rtol=1e-12)
    # Continue two more steps from the loaded carry, against the oracle.
    for t in (7, 8):
        ref, _ = orc.step_once(ref, t, os_, oflags, None, U0, V0)
        got = _step_once_state_only(got, t, static, flags, None, U0, V0)
    for f in FIELDS:
        np.testing.assert_allclose(getattr(got, f + "_curr").cpu().numpy(), getattr(ref, f + "_curr"), rtol=0,
                                   atol=ATOL[f], err_msg=f)


def test_batched_state_members_equal_single_runs(torch_gpu):
    from my_swampe_amd.model import _init_state_from_fields, simulate_scan_last

    torch = torch_gpu
    print("[gpu-case] seams batched", flush=True)
    static, flags = _static_flags(diagnostics=False)
    eta0, delta0, Phi0, U0, V0 = _ic(static)
    rng = np.random.default_rng(2)
    bump = torch.tensor(rng.standard_normal((3, static.J, static.I)) * 50.0, device="cuda")
    Phis = Phi0[None] + bump
    st = _init_state_from_fields(static=static, flags=flags, test=None, eta0=eta0.expand(3, -1, -1),
                                 delta0=delta0.expand(3, -1, -1), Phi0=Phis, U0=U0.expand(3, -1, -1),
                                 V0=V0.expand(3, -1, -1))
    last = simulate_scan_last(static=static, flags=flags, state0=st, t_seq=range(2, 7), test=None)
    assert last.Phi_curr.shape == (3, static.J, static.I)
    for b in range(3):
        one = _init_state_from_fields(static=static, flags=flags, test=None, eta0=eta0, delta0=delta0,
                                      Phi0=Phis[b], U0=U0, V0=V0)
        ref = simulate_scan_last(static=static, flags=flags, state0=one, t_seq=range(2, 7), test=None)
        np.testing.assert_allclose(last.Phi_curr[b].cpu().numpy(), ref.Phi_curr.cpu().numpy(), rtol=0, atol=1e-10)


def test_replace_static_parameter_takes_effect(torch_gpu):
    from my_swampe_amd.model import _init_state_from_fields, run_model_scan_final, simulate_scan_last

    torch = torch_gpu
    print("[gpu-case] seams replace(static)", flush=True)
    static, flags = _static_flags(diagnostics=False)
    eta0, delta0, Phi0, U0, V0 = _ic(static)
    s2 = dataclasses.replace(static, taurad=torch.tensor(5.0e4, dtype=torch.float64, device="cuda"))
    st = _init_state_from_fields(static=s2, flags=flags, test=None, eta0=eta0, delta0=delta0, Phi0=Phi0, U0=U0, V0=V0)
    last = simulate_scan_last(static=s2, flags=flags, state0=st, t_seq=range(2, 9), test=None)
    ref = run_model_scan_final(**{**FORCED, "tmax": 9, "taurad": 5.0e4})["last_state"]
    assert torch.equal(last.Phi_curr, ref.Phi_curr)


def test_diagnose_winds_vs_oracle(torch_gpu):
    from my_swampe_amd.model import _diagnose_winds

    print("[gpu-case] seams diagnose winds", flush=True)
    static, _ = _static_flags()
    os_ = orc.build_static(M=42, dt=1200.0, a=FORCED["a"], omega=FORCED["omega"], g=9.8, Phibar=FORCED["Phibar"],
                           taurad=FORCED["taurad"], taudrag=FORCED["taudrag"], DPhieq=FORCED["DPhieq"],
                           K6=FORCED["K6"], K6Phi=None, test=None)
    o = orc.run_model_scan(**{**FORCED, "tmax": 5}, return_history=False)["last_state"]
    U, V = _diagnose_winds(o.eta_curr, o.delta_curr, static)
    Ur, Vr = orc.diagnose_winds(o.eta_curr, o.delta_curr, os_)
    np.testing.assert_allclose(U.cpu().numpy(), Ur, rtol=0, atol=ATOL["U"])
    np.testing.assert_allclose(V.cpu().numpy(), Vr, rtol=0, atol=ATOL["V"])


def test_dead_member_is_frozen_and_mid_run_load_scheme_limits(torch_gpu):
    from my_swampe_amd.model import RunFlags, _init_state_from_fields, _step_once

    torch = torch_gpu
    print("[gpu-case] seams dead / scheme limits", flush=True)
    static, flags = _static_flags()
    eta0, delta0, Phi0, U0, V0 = _ic(static)
    st = _init_state_from_fields(static=static, flags=flags, test=None, eta0=eta0, delta0=delta0, Phi0=Phi0,
                                 U0=U0, V0=V0)
    st, _ = _step_once(st, 2, static, flags, None, U0, V0)
    st, _ = _step_once(st, 3, static, flags, None, U0, V0)
    dead = st._replace(dead=torch.tensor(True, device="cuda"))
    nxt, out = _step_once(dead, 4, static, flags, None, U0, V0)
    assert bool(out["dead"]) and bool(nxt.dead)
    assert torch.equal(nxt.Phi_curr, st.Phi_curr) and torch.equal(out["Phi"], st.Phi_curr)
    # A mid-run carry of the explicit scheme reads a spectral prev level the grids do not hold.
    ex_flags = RunFlags(expflag=True, diagnostics=True)
    mid = st._replace(Phi_prev=st.Phi_prev + 1.0)
    with pytest.raises(NotImplementedError):
        _step_once(mid, 4, static, ex_flags, None, U0, V0)
