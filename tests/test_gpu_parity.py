"""HIP path vs the upstream-SWAMPE fixtures and the FP64 oracle (GPU).

Tolerances: the reference's parity contract (eta/delta 1e-10, Phi 5e-8,
U/V 1e-9 atol; unit_tests/test_parity_reference_regression.py:58-64) for the
FP64 path on short runs; relative L2 bounds, written in each test, for the
FP32 transform path.
"""

from __future__ import annotations

import numpy as np
import pytest

from conftest import ATOL, FIXTURE_CASES, GOLDEN, phase_curve_from_phi
from oracle import swampe_oracle as orc

pytestmark = pytest.mark.gpu

FIELDS = ("eta", "delta", "Phi", "U", "V")


def _np(t):
    return t.detach().cpu().numpy()


def _print_case(tag, **kw):
    print(f"[gpu-case] {tag} {kw}", flush=True)


@pytest.mark.parametrize("case", list(FIXTURE_CASES))
def test_fixture_terminal_and_snapshots(case, torch_gpu):
    """unit_tests/test_parity_reference_regression.py:96-130 on the HIP FP64 path."""
    from my_swampe_amd.model import run_model_scan

    _print_case("fixture", case=case)
    ref = np.load(GOLDEN / f"{case}.npz")
    out = run_model_scan(**FIXTURE_CASES[case], diagnostics=True, jit_scan=False)
    last = out["last_state"]
    for f in FIELDS:
        np.testing.assert_allclose(_np(getattr(last, f + "_curr")), ref[f"final_{f}"], rtol=0.0, atol=ATOL[f])
    t_seq = _np(out["t_seq"])
    for i, step in enumerate(ref["snapshot_steps"]):
        idx = int(np.where(t_seq == step)[0][0])
        for f in FIELDS:
            np.testing.assert_allclose(_np(out["outs"][f][idx]), ref[f"snapshot_{f}"][i], rtol=0.0, atol=ATOL[f])


@pytest.mark.parametrize("case", list(FIXTURE_CASES))
def test_fixture_phase_curve(case, torch_gpu):
    """unit_tests/test_parity_reference_regression.py:133-148 on the HIP FP64 path."""
    from my_swampe_amd.model import run_model_scan_final

    _print_case("phase", case=case)
    ref = np.load(GOLDEN / f"{case}.npz")
    out = run_model_scan_final(**FIXTURE_CASES[case], diagnostics=False, jit_scan=False)
    curve = phase_curve_from_phi(_np(out["last_state"].Phi_curr), ref["phase_angles"])
    np.testing.assert_allclose(curve, ref["phase_curve"], rtol=0.0, atol=1.0e-10)


@pytest.mark.parametrize("case", list(FIXTURE_CASES))
def test_fixture_legacy_diagnostics(case, torch_gpu):
    """Legacy (tmax, 2) spinup/geopot arrays of run_model (model.py:2105-2158) vs the fixtures."""
    from my_swampe_amd.model import run_model_gpu

    _print_case("legacy", case=case)
    ref = np.load(GOLDEN / f"{case}.npz")
    kw = dict(FIXTURE_CASES[case])
    res = run_model_gpu(kw.pop("M"), kw.pop("dt"), kw.pop("tmax"), kw.pop("Phibar"), kw.pop("omega"), kw.pop("a"),
                        verbose=False, **kw)
    np.testing.assert_allclose(_np(res["spinup"]), ref["spinup"], rtol=1e-11, atol=1e-9)
    np.testing.assert_allclose(_np(res["geopot"]), ref["geopot"], rtol=1e-11, atol=5e-8)


BASE = dict(M=42, dt=1200.0, tmax=9, Phibar=3.0e5, omega=3.2e-5, a=8.2e7, test=None, g=9.8, forcflag=True,
            taurad=36000.0, taudrag=21600.0, DPhieq=1.0e6, diffflag=True, modalflag=True, alpha=0.01,
            expflag=False, K6=1.24e33)
VARIANTS = {
    "modeuler_forced": dict(),
    "modeuler_unforced": dict(forcflag=False),
    "explicit_forced": dict(expflag=True, dt=120.0),
    "explicit_unforced": dict(expflag=True, forcflag=False, dt=120.0),
    "no_diffusion_no_modal": dict(diffflag=False, modalflag=False),
    "taudrag_sentinel": dict(taudrag=-1.0),
    "test2": dict(test=2, dt=30.0, Phibar=3e3 * 9.8, omega=7.2921159e-5, a=6.37122e6),
    "test1_explicit": dict(test=1, expflag=True, forcflag=False, diffflag=False, dt=30.0, Phibar=3e3,
                           omega=7.2921159e-5, a=6.37122e6),
}


def _scale_atol(ref_arr, atol_rel):
    return atol_rel * max(1.0, float(np.max(np.abs(ref_arr))))


@pytest.mark.parametrize("name", list(VARIANTS))
def test_scheme_variants_vs_oracle(name, torch_gpu):
    """Flag/scheme variants vs the oracle over the fixtures' run length (4 steps).

    Physical fields: the reference's parity contract atols.  Fourier carries and
    diagnostics: 1e-11 relative to the field's largest magnitude.
    """
    from my_swampe_amd.model import run_model_scan

    kw = dict(BASE, tmax=6)
    kw.update(VARIANTS[name])
    _print_case("variant", name=name)
    ref = orc.run_model_scan(**kw)
    got = run_model_scan(**kw, diagnostics=True)
    last, rl = got["last_state"], ref["last_state"]
    for f in FIELDS:
        for lvl in ("_curr", "_prev") if f in ("eta", "delta", "Phi") else ("_curr",):
            r = getattr(rl, f + lvl)
            np.testing.assert_allclose(_np(getattr(last, f + lvl)), r, rtol=0, atol=ATOL[f], err_msg=f + lvl)
        r = ref["outs"][f]
        np.testing.assert_allclose(_np(got["outs"][f]), r, rtol=0, atol=ATOL[f], err_msg=f)
    for fld in ("etam_curr", "deltam_curr", "Phim_curr", "etam_prev", "Phim_prev", "Um_curr", "Vm_curr",
                "Am_curr", "Bm_curr", "Cm_curr", "Dm_curr", "Em_curr", "PhiFm_curr", "Fm_curr", "Gm_curr",
                "Rum_lag", "Rvm_lag"):
        r = getattr(rl, fld)
        np.testing.assert_allclose(_np(getattr(last, fld)), r, rtol=0, atol=_scale_atol(r, 1e-11), err_msg=fld)
    for f in ("rms", "spin_min", "phi_min", "phi_max"):
        np.testing.assert_allclose(_np(got["outs"][f]), ref["outs"][f], rtol=1e-10, atol=1e-10, err_msg=f)


@pytest.mark.parametrize("M", [63, 106])
def test_higher_truncations_vs_oracle(M, torch_gpu):
    """T63 / T106 forced modified-Euler, 4 steps, vs the oracle at the parity-contract atols.

    At T106 this includes the reference's underflowed basis normalization
    (modes with n + m >= ~170 vanish or are attenuated each step).
    """
    from my_swampe_amd.model import run_model_scan

    kw = dict(BASE, M=M, tmax=6, dt=600.0)
    _print_case("truncation", M=M)
    ref = orc.run_model_scan(**kw, diagnostics=False, return_history=False)
    got = run_model_scan(**kw, diagnostics=False, return_history=False)
    for f in FIELDS:
        r = getattr(ref["last_state"], f + "_curr")
        np.testing.assert_allclose(_np(getattr(got["last_state"], f + "_curr")), r, rtol=0, atol=ATOL[f], err_msg=f)
    r = ref["last_state"].Phim_curr
    np.testing.assert_allclose(_np(got["last_state"].Phim_curr), r, rtol=0, atol=_scale_atol(r, 1e-11))


def test_diagnosed_winds_vs_oracle(torch_gpu):
    """Explicit eta/delta/Phi ICs without winds: spectral wind diagnosis (model.py:644-666)."""
    from my_swampe_amd.model import run_model_scan

    s = orc.build_static(M=42, dt=600.0, a=6.37122e6, omega=7.2921159e-5, g=9.8, Phibar=3e3, taurad=86400.0,
                         taudrag=86400.0, DPhieq=4e6, K6=1.24e33, K6Phi=None, test=2)
    eta0, delta0, Phi0, _, _ = orc.analytic_ic(s, 2, np.pi / 4)
    rng = np.random.default_rng(3)
    delta0 = delta0 + 1e-7 * rng.standard_normal(delta0.shape)
    kw = dict(M=42, dt=600.0, tmax=4, Phibar=3e3, omega=7.2921159e-5, a=6.37122e6, test=2,
              eta0_init=eta0, delta0_init=delta0, Phi0_init=Phi0)
    _print_case("diag-winds")
    ref = orc.run_model_scan(**kw)
    got = run_model_scan(**kw)
    for f in FIELDS:
        r = ref["outs"][f]
        np.testing.assert_allclose(_np(got["outs"][f]), r, rtol=0, atol=_scale_atol(r, 1e-12), err_msg=f)


def test_blowup_gate_freezes_member(torch_gpu):
    """Huge initial winds trip rms > blowup_rms at the first step (model.py:838-864)."""
    from my_swampe_amd.model import run_model_scan

    s = orc.build_static(**{k: BASE[k] for k in ("M", "dt", "a", "omega", "g", "Phibar", "taurad", "taudrag",
                                                  "DPhieq", "K6")}, K6Phi=None, test=None)
    eta0, delta0, Phi0, U0, V0 = orc.analytic_ic(s, None, 0.05)
    U0 = U0 + 2.0e4
    kw = dict(BASE, tmax=6, eta0_init=eta0, delta0_init=delta0, Phi0_init=Phi0, U0_init=U0, V0_init=V0)
    _print_case("blowup")
    ref = orc.run_model_scan(**kw)
    got = run_model_scan(**kw)
    assert int(got["dead_first_idx"]) == ref["dead_first_idx"] == 0
    assert bool(got["outs"]["dead"].all())
    np.testing.assert_allclose(_np(got["outs"]["Phi"][-1]), Phi0, rtol=0, atol=0)
    np.testing.assert_allclose(_np(got["outs"]["rms"]), ref["outs"]["rms"], rtol=1e-12)


def test_ensemble_members_equal_direct_runs(torch_gpu):
    """Batched members (per-member DPhieq/taurad/taudrag) == direct single runs (test_vmap_smoke.py:84-93)."""
    from my_swampe_amd.model import param_row
    from my_swampe_amd.plan import Plan

    torch = torch_gpu
    base = dict(dt=600.0, a=6.37122e6, omega=7.292e-5, g=9.8, Phibar=3.0e5, K6=1.24e33, K6Phi=None, alpha=0.01)
    members = [dict(DPhieq=2e6, taurad=86400.0, taudrag=86400.0), dict(DPhieq=4e6, taurad=5e4, taudrag=2e5),
               dict(DPhieq=6e6, taurad=1e5, taudrag=-1.0)]
    params = np.stack([param_row(**base, **m) for m in members])
    _print_case("ensemble", B=len(members))
    s = orc.build_static(M=42, test=None, **{k: v for k, v in base.items() if k not in ("alpha",)},
                         taurad=86400.0, taudrag=86400.0, DPhieq=4e6)
    eta0, delta0, Phi0, U0, V0 = orc.analytic_ic(s, None, 0.05)
    plan = Plan(42, params)
    plan.init(eta0, delta0, Phi0, U0, V0)
    plan.run(2, 5)
    ens = _np(plan.export(fields=("Phi_curr",))["Phi_curr"])
    plan.close()
    assert np.min([np.max(np.abs(ens[i] - ens[j])) for i in range(3) for j in range(i + 1, 3)]) > 0.0
    for i, m in enumerate(members):
        ref = orc.run_model_scan(M=42, tmax=7, test=None, diagnostics=False, return_history=False,
                                 eta0_init=eta0, delta0_init=delta0, Phi0_init=Phi0, U0_init=U0, V0_init=V0,
                                 **{k: v for k, v in base.items()}, **m)
        np.testing.assert_allclose(ens[i], ref["last_state"].Phi_curr, rtol=0, atol=1e-7)
        single = Plan(42, params[i:i + 1])
        single.init(eta0, delta0, Phi0, U0, V0)
        single.run(2, 5)
        np.testing.assert_allclose(ens[i], _np(single.export(fields=("Phi_curr",))["Phi_curr"][0]), rtol=0, atol=1e-10)
        single.close()
    torch.cuda.synchronize()


def test_graph_replay_matches_eager(torch_gpu):
    """The HIP-graph forward driver gives the same state as eager stepping (bitwise)."""
    from my_swampe_amd.plan import Plan
    from my_swampe_amd.model import param_row

    torch = torch_gpu
    p = param_row(dt=120.0, a=8.2e7, omega=3.2e-5, g=9.8, Phibar=3e5, taurad=86400.0, taudrag=86400.0,
                  DPhieq=3e5, K6=1.24e33, K6Phi=None, alpha=0.01)
    eta0, delta0, Phi0, U0, V0 = orc.rest_noise_ic(42, 0, 3.2e-5)
    _print_case("graph")
    outs = []
    for use_graph in (False, True):
        plan = Plan(42, np.stack([p, p]))
        s = torch.cuda.Stream()
        with torch.cuda.stream(s):
            plan.init(eta0, delta0, Phi0, U0, V0, stream=s)  # initialization ordered before the steps
            if use_graph:
                plan.run_graph(2, 40, graph_steps=8, stream=s)
            else:
                plan.run(2, 40, stream=s)
            ex = plan.export(fields=("Phi_curr", "Phi_prev", "U_curr"), spec=True, stream=s)
        s.synchronize()
        outs.append({k: _np(v) for k, v in ex.items()})
        plan.close()
    for k in outs[0]:
        np.testing.assert_array_equal(outs[0][k], outs[1][k], err_msg=k)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_graph_calls_of_many_lengths_match_eager(precision, torch_gpu):
    """Calls of the graph driver shorter than graph_steps (one captured sequence each, window and
    last step included), repeated lengths (cached graphs replayed) and more distinct lengths than
    the plan caches (least recently replayed graphs replaced): every call's state equals that of
    eager calls of the same lengths (bitwise), with exports and synchronizations between the calls;
    the calls after the first step replay graphs only."""
    from my_swampe_amd.plan import Plan
    from my_swampe_amd.model import param_row

    torch = torch_gpu
    p = param_row(dt=120.0, a=8.2e7, omega=3.2e-5, g=9.8, Phibar=3e5, taurad=86400.0, taudrag=86400.0,
                  DPhieq=3e5, K6=1.24e33, K6Phi=None, alpha=0.01)
    ics = orc.rest_noise_ic(42, 0, 3.2e-5)
    lengths = [20, 20, 20, 5, 7, 9, 11, 13, 3, 4, 6, 17, 20, 40, 20]
    _print_case("graph-lengths", precision=precision)
    outs = []
    for use_graph in (False, True):
        plan = Plan(42, np.stack([p, p]), precision=precision)
        s = torch.cuda.Stream()
        got = []
        with torch.cuda.stream(s):
            plan.init(*ics, stream=s)
            plan.run(2, 1, stream=s)
            t = 3
            c0 = plan.step_counts()
            for n in lengths:
                if use_graph:
                    plan.run_graph(t, n, graph_steps=16, stream=s)
                else:
                    plan.run(t, n, stream=s)
                t += n
                ex = plan.export(fields=("Phi_curr", "eta_prev", "U_curr"), spec=True, stream=s)
                s.synchronize()
                got.append({k: _np(v) for k, v in ex.items()})
            c1 = plan.step_counts()
        if use_graph:
            assert c1["graph"] - c0["graph"] == sum(lengths), (c0, c1)
            assert c1["eager"] == c0["eager"], (c0, c1)
        outs.append(got)
        plan.close()
    for i, (a, b) in enumerate(zip(*outs)):
        for k in a:
            np.testing.assert_array_equal(a[k], b[k], err_msg=f"call {i} ({lengths[i]} steps): {k}")


def test_graph_eviction_without_synchronization_matches_eager(torch_gpu):
    """Calls of 12 distinct lengths issued back to back with no synchronization (more step-sequence
    shapes than the plan caches, so graphs are replaced while earlier replays may still run): the
    final state equals that of eager calls of the same lengths (bitwise)."""
    from my_swampe_amd.plan import Plan
    from my_swampe_amd.model import param_row

    torch = torch_gpu
    p = param_row(dt=120.0, a=8.2e7, omega=3.2e-5, g=9.8, Phibar=3e5, taurad=86400.0, taudrag=86400.0,
                  DPhieq=3e5, K6=1.24e33, K6Phi=None, alpha=0.01)
    ics = orc.rest_noise_ic(42, 0, 3.2e-5)
    lengths = [3, 5, 7, 9, 11, 13, 6, 4, 8, 10, 12, 14, 3, 5, 7]
    _print_case("graph-eviction")
    outs = []
    for use_graph in (False, True):
        plan = Plan(42, np.stack([p] * 8), precision="fp32")
        s = torch.cuda.Stream()
        with torch.cuda.stream(s):
            plan.init(*ics, stream=s)
            plan.run(2, 1, stream=s)
            t = 3
            for n in lengths:
                if use_graph:
                    plan.run_graph(t, n, graph_steps=16, stream=s)
                else:
                    plan.run(t, n, stream=s)
                t += n
            ex = plan.export(fields=("Phi_curr", "eta_prev", "U_curr"), spec=True, stream=s)
            s.synchronize()
        outs.append({k: _np(v) for k, v in ex.items()})
        plan.close()
    for k in outs[0]:
        np.testing.assert_array_equal(outs[0][k], outs[1][k], err_msg=k)


def test_fp32_transforms_close_to_fp64(torch_gpu):
    """FP32 transforms with the FP64 state update vs the FP64 oracle, 60 forced steps at T42
    (dt = 120 s, a = 8.2e7, omega = 3.2e-5, Phibar = 3e5, DPhieq = 3e5, rest + noise).

    Bounds (relative L2): Phi and eta <= 1e-5 (the stated target for the
    reduced-precision path); delta, a small residual of nearly balanced
    terms, <= 1e-3.
    """
    from my_swampe_amd.model import run_model_scan_final

    eta0, delta0, Phi0, U0, V0 = orc.rest_noise_ic(42, 0, 3.2e-5)
    kw = dict(M=42, dt=120.0, tmax=62, Phibar=3e5, omega=3.2e-5, a=8.2e7, taurad=86400.0, taudrag=86400.0,
              DPhieq=3e5, eta0_init=eta0, delta0_init=delta0, Phi0_init=Phi0, U0_init=U0, V0_init=V0)
    _print_case("fp32")
    ref = orc.run_model_scan(**kw, diagnostics=False, return_history=False)["last_state"]
    got = run_model_scan_final(**kw, precision="fp32")["last_state"]
    rel = {}
    for f in ("eta_curr", "delta_curr", "Phi_curr"):
        r = getattr(ref, f)
        g = _np(getattr(got, f))
        rel[f] = float(np.linalg.norm(g - r) / np.linalg.norm(r))
    print("relL2", rel)
    assert rel["Phi_curr"] <= 1e-5 and rel["eta_curr"] <= 1e-5 and rel["delta_curr"] <= 1e-3, rel


@pytest.mark.parametrize("M", [42, 63, 106])
def test_mfma_grid_kernel_matches_plain_fp32(M, torch_gpu):
    """The MFMA grid (full and half-period folded DFTs) and Legendre kernels (FP32; equatorially folded
    contractions, and over all latitudes in "mfma_legfull") vs the plain FP32 kernels and the FP64
    oracle, 12 forced steps from a balanced zonal flow (Williamson test 2 fields, U ~ 40 m/s).
    Relative L2 <= 5e-6 for Phi, eta and U."""
    from my_swampe_amd import _capi
    from my_swampe_amd.model import param_row
    from my_swampe_amd.plan import Plan

    dt = 120.0 if M == 42 else 60.0
    a, omega, Phibar = 6.37122e6, 7.2921159e-5, 3e4
    p = param_row(dt=dt, a=a, omega=omega, g=9.8, Phibar=Phibar, taurad=86400.0, taudrag=86400.0,
                  DPhieq=3e4, K6=1.24e33, K6Phi=None, alpha=0.01)
    s = orc.build_static(M=M, dt=dt, a=a, omega=omega, g=9.8, Phibar=Phibar, taurad=86400.0, taudrag=86400.0,
                         DPhieq=3e4, K6=1.24e33, K6Phi=None, test=2)
    ics = orc.analytic_ic(s, 2, 0.3)
    _print_case("mfma-vs-plain", M=M)
    res = {}
    kinds = ("plain", "mfma", "mfma_fold", "mfma_legfull")
    for kind in kinds:
        plan = Plan(M, np.stack([p] * 3), precision="fp32")  # 3 members: a partial member group
        plan.set_option(_capi.OPT_GRID_KERNEL, 1 if kind == "plain" else 0)
        plan.set_option(_capi.OPT_LEG_KERNEL, 1 if kind == "plain" else 0)
        plan.set_option(_capi.OPT_GRID_FOLD, 1 if kind == "mfma" else 2)
        assert plan.get_option(_capi.OPT_LEG_FOLD) == 2  # the folded Legendre tables are built
        plan.set_option(_capi.OPT_LEG_FOLD, 1 if kind == "mfma_legfull" else 2)
        plan.init(*ics)
        plan.run(2, 12)
        res[kind] = {k: _np(v[0]) for k, v in plan.export(fields=("Phi_curr", "U_curr", "eta_curr")).items()
                     if k != "dead"}
        plan.close()
    ref = orc.run_model_scan(M=M, dt=dt, tmax=14, Phibar=Phibar, omega=omega, a=a, taurad=86400.0,
                             taudrag=86400.0, DPhieq=3e4, eta0_init=ics[0], delta0_init=ics[1], Phi0_init=ics[2],
                             U0_init=ics[3], V0_init=ics[4], diagnostics=False, return_history=False)["last_state"]
    for f in ("Phi_curr", "U_curr", "eta_curr"):
        r = getattr(ref, f)
        for kind in kinds:
            rel = np.linalg.norm(res[kind][f] - r) / np.linalg.norm(r)
            print(M, kind, f, "relL2 vs oracle", rel)
            assert rel < 5e-6, (M, kind, f, rel)


_CARRY_FIELDS = ("eta_prev", "delta_prev", "Phi_prev", "eta_curr", "Phi_curr", "U_curr", "V_curr")


def _carry_run(precision, options, nsteps=60, scheme="modeuler", members=3):
    """Export of a plan initialized from the rest + noise grids and stepped `nsteps` times by the graph
    driver, with plan options `options` ({option: value})."""
    from my_swampe_amd.model import param_row
    from my_swampe_amd.plan import Plan

    import torch

    si_raw = scheme == "si_raw"
    p = param_row(dt=600.0 if si_raw else 120.0, a=8.2e7, omega=3.2e-5, g=9.8, Phibar=3e5, taurad=86400.0,
                  taudrag=86400.0, DPhieq=3e5, K6=1.24e33, K6Phi=None, alpha=0.05 if si_raw else 0.01,
                  williams_alpha=0.53)
    ics = orc.rest_noise_ic(42, 0, 3.2e-5)
    plan = Plan(42, np.stack([p] * members), precision=precision, semi_implicit=si_raw, raw_filter=si_raw)
    for opt, val in options.items():
        plan.set_option(opt, val)
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        plan.init(*ics, stream=s)
        plan.run_graph(2, nsteps, graph_steps=8, stream=s)
        ex = plan.export(fields=_CARRY_FIELDS, stream=s)
    s.synchronize()
    out = {k: _np(ex[k]) for k in _CARRY_FIELDS}
    plan.close()
    return out


@pytest.mark.parametrize("precision,scheme", [("fp64", "modeuler"), ("fp32", "modeuler"), ("fp64", "si_raw"),
                                              ("fp32", "si_raw")])
def test_physical_carry_window_matches_full_carry(precision, scheme, torch_gpu):
    """The grid carry written only in the final Robert-Asselin window exports the same state as the
    grid carry written every step (graph driver, 60 steps): classic RA on modified Euler, and the
    semi-implicit leapfrog with the RAW filter (alpha = 0.05, williams_alpha = 0.53), whose window
    uses the contraction alpha w / (1 - alpha (1 - w))."""
    from my_swampe_amd import _capi

    _print_case("phys-window", precision=precision, scheme=scheme)
    full = _carry_run(precision, {_capi.OPT_FULL_PHYS_CARRY: 1}, scheme=scheme)
    win = _carry_run(precision, {_capi.OPT_PHYS_LINEAGE: -1}, scheme=scheme)
    for k in full:
        np.testing.assert_allclose(win[k], full[k], rtol=1e-15, atol=0, err_msg=k)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_lineage_window_matches_longer_lineage(precision, torch_gpu):
    """The spectral prev lineage advanced in the default final window (16 steps at alpha = 0.01)
    exports the same state as one advanced over the last 40 of the 60 steps."""
    from my_swampe_amd import _capi

    _print_case("lineage-window", precision=precision)
    dflt = _carry_run(precision, {})
    longer = _carry_run(precision, {_capi.OPT_PHYS_LINEAGE: 40})
    for k in dflt:
        np.testing.assert_allclose(dflt[k], longer[k], rtol=1e-15, atol=0, err_msg=k)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_lineage_carry_matches_grid_carry(precision, torch_gpu):
    """The fields rebuilt from the spectral lineage and the current Fourier rows equal the grid carry
    written every step to the transforms' rounding: FP64 within 1e-13 of each field's largest
    magnitude, FP32 within 1e-5 (the grid carry of an FP32 plan sums FP32-rounded grids)."""
    from my_swampe_amd import _capi

    _print_case("lineage-vs-grid", precision=precision)
    full = _carry_run(precision, {_capi.OPT_FULL_PHYS_CARRY: 1})
    lin = _carry_run(precision, {})
    tol = 1e-13 if precision == "fp64" else 1e-5
    for k in full:
        scale = float(np.max(np.abs(full[k])))
        err = float(np.max(np.abs(lin[k] - full[k])))
        print(k, "max|diff| / max|field|", err / scale)
        assert err <= tol * scale, (k, err / scale)


def test_lineage_export_vs_oracle(torch_gpu):
    """FP64, forced modified Euler, 20 steps without diagnostics or history (16 of them advance the
    spectral lineage; the grid fields are rebuilt at export): every exported grid field, prev levels
    included, against the oracle within the parity contract atol or 1e-12 of the field's largest
    magnitude, whichever is larger."""
    from my_swampe_amd.model import run_model_scan

    _print_case("lineage-oracle")
    eta0, delta0, Phi0, U0, V0 = orc.rest_noise_ic(42, 0, 3.2e-5)
    kw = dict(M=42, dt=120.0, tmax=22, Phibar=3e5, omega=3.2e-5, a=8.2e7, taurad=86400.0, taudrag=86400.0,
              DPhieq=3e5, eta0_init=eta0, delta0_init=delta0, Phi0_init=Phi0, U0_init=U0, V0_init=V0,
              diagnostics=False, return_history=False)
    ref = orc.run_model_scan(**kw)["last_state"]
    got = run_model_scan(**kw)["last_state"]
    for name in ("eta_prev", "eta_curr", "delta_prev", "delta_curr", "Phi_prev", "Phi_curr", "U_curr", "V_curr"):
        r = getattr(ref, name)
        g = _np(getattr(got, name))
        tol = max(ATOL[name.split("_")[0]], _scale_atol(r, 1e-12))
        print(name, "max|diff|", float(np.max(np.abs(g - r))), "atol", tol)
        np.testing.assert_allclose(g, r, rtol=0, atol=tol, err_msg=name)


def _noise_spectral(M, B, seed, omega):
    rng = np.random.default_rng(seed)
    c = np.zeros((B, 3, M + 1, M + 1), complex)
    c[:, 0, 0, 1] = omega / np.sqrt(0.375)
    for m in range(9):
        for n in range(m, 9):
            c[:, 2, m, n] = 40.0 * (rng.standard_normal(B) + (1j * rng.standard_normal(B) if m else 0.0))
    return c


def test_ensemble_recaptures_after_plan_option_change(torch_gpu):
    """An option set on an ensemble's plan after the ensemble captured its step graphs (here the
    Legendre contractions over all latitudes instead of the folded ones, which round differently)
    takes effect in the ensemble's next call of the same length: bitwise the states the two plans
    reach stepped separately with the same option change between the calls."""
    from my_swampe_amd import _capi
    from my_swampe_amd.model import param_row
    from my_swampe_amd.plan import Ensemble, Plan

    torch = torch_gpu
    _print_case("ensemble-option-change")
    omega = 3.2e-5
    specs = []
    for M, B in ((42, 3), (63, 2)):
        rng = np.random.default_rng(M + 1)
        params = np.stack([param_row(dt=120.0, a=8.2e7, omega=omega, g=9.8, Phibar=3e5,
                                     taurad=10 ** rng.uniform(4, 6), taudrag=10 ** rng.uniform(4, 6),
                                     DPhieq=rng.uniform(1e5, 5e5), K6=1.24e33, K6Phi=None, alpha=0.01)
                           for _ in range(B)])
        specs.append((M, params, _noise_spectral(M, B, M + 1, omega)))
    stream = torch.cuda.Stream()
    out = {}
    for mode in ("separate", "ensemble"):
        with torch.cuda.stream(stream):
            plans = []
            for M, params, c in specs:
                pl = Plan(M, params, precision="fp32")
                pl.init_spectral(torch.tensor(c, device="cuda"), stream=stream)
                plans.append(pl)
            ens = Ensemble(plans[0], plans[1]) if mode == "ensemble" else None
            for t0, fold in ((2, 2), (22, 1)):
                for pl in plans:
                    pl.set_option(_capi.OPT_LEG_FOLD, fold)
                if ens is None:
                    for pl in plans:
                        pl.run_graph(t0, 20, graph_steps=8, stream=stream)
                else:
                    ens.run_graph(t0, 20, graph_steps=8, stream=stream)
            res = [pl.export(fields=("Phi_curr", "etam_curr"), spec=True, stream=stream) for pl in plans]
            stream.synchronize()
            if ens is not None:
                ens.close()
            for pl in plans:
                pl.close()
        out[mode] = res
    for q in range(2):
        for k in ("Phi_curr", "etam_curr", "spec"):
            a, b = out["separate"][q][k], out["ensemble"][q][k]
            assert torch.equal(a, b), (q, k)


@pytest.mark.parametrize("pair", [(42, 63), (42, 106)])
def test_mixed_truncation_ensemble_matches_separate_plans(pair, torch_gpu):
    """Two FP32 plans of different truncations stepped together (one Legendre and one grid launch
    per step for both) reach exactly the state each plan reaches alone: bitwise, since every
    workgroup runs the same device code.  40 steps: eager start, graph replays, the physical
    carry window and the last step; partial member groups (3 and 2 members)."""
    from my_swampe_amd.model import param_row
    from my_swampe_amd.plan import Ensemble, Plan

    torch = torch_gpu
    _print_case("mixed-ensemble", pair=pair)
    omega = 3.2e-5
    specs = []
    for M, B in zip(pair, (3, 2)):
        dt = 60.0 if M == 106 else 120.0
        rng = np.random.default_rng(M)
        params = np.stack([param_row(dt=dt, a=8.2e7, omega=omega, g=9.8, Phibar=3e5, taurad=10 ** rng.uniform(4, 6),
                                     taudrag=10 ** rng.uniform(4, 6), DPhieq=rng.uniform(1e5, 5e5), K6=1.24e33,
                                     K6Phi=None, alpha=0.01) for _ in range(B)])
        specs.append((M, params, _noise_spectral(M, B, M + 1, omega)))
    stream = torch.cuda.Stream()
    out = {}
    for mode in ("separate", "ensemble"):
        with torch.cuda.stream(stream):
            plans = []
            for M, params, c in specs:
                pl = Plan(M, params, precision="fp32")
                pl.init_spectral(torch.tensor(c, device="cuda"), stream=stream)
                plans.append(pl)
            if mode == "separate":
                for pl in plans:
                    pl.run_graph(2, 40, graph_steps=8, stream=stream)
            else:
                ens = Ensemble(plans[0], plans[1])
                ens.run_graph(2, 40, graph_steps=8, stream=stream)
            res = [pl.export(fields=("Phi_curr", "U_curr", "etam_curr", "Phi_prev"), spec=True, stream=stream)
                   for pl in plans]
            stream.synchronize()
            if mode == "ensemble":
                ens.close()
            for pl in plans:
                pl.close()
        out[mode] = res
    for q, M in enumerate(pair):
        for k in ("Phi_curr", "U_curr", "etam_curr", "Phi_prev", "spec"):
            a, b = out["separate"][q][k], out["ensemble"][q][k]
            assert bool(torch.isfinite(torch.view_as_real(a) if a.is_complex() else a).all()), (M, k)
            assert torch.equal(a, b), (M, k, float((a - b).abs().max()))


@pytest.mark.parametrize("M", [42, 63, 106])
def test_bf16x3_grid_kernel_vs_oracle(M, torch_gpu):
    """Split-BF16x3 transforms (the BF16x3 grid kernel and the MFMA Legendre kernel's BF16x3
    contractions, FP64 update) vs the FP64 oracle and vs the FP32 MFMA path, 12 forced steps from
    the balanced zonal flow of test_mfma_grid_kernel_matches_plain_fp32 (3 members).  Relative
    L2 <= 2e-5 for Phi, U, eta (the dropped lo*lo term and the 16-bit split leave ~2^-17 per
    product; FP32 reaches 2e-7)."""
    from my_swampe_amd.model import param_row
    from my_swampe_amd.plan import Plan

    dt = 120.0 if M == 42 else 60.0
    a, omega, Phibar = 6.37122e6, 7.2921159e-5, 3e4
    p = param_row(dt=dt, a=a, omega=omega, g=9.8, Phibar=Phibar, taurad=86400.0, taudrag=86400.0,
                  DPhieq=3e4, K6=1.24e33, K6Phi=None, alpha=0.01)
    s = orc.build_static(M=M, dt=dt, a=a, omega=omega, g=9.8, Phibar=Phibar, taurad=86400.0, taudrag=86400.0,
                         DPhieq=3e4, K6=1.24e33, K6Phi=None, test=2)
    ics = orc.analytic_ic(s, 2, 0.3)
    _print_case("bf16x3-vs-oracle", M=M)
    res = {}
    for prec in ("fp32", "bf16x3"):
        plan = Plan(M, np.stack([p] * 3), precision=prec)
        plan.init(*ics)
        plan.run(2, 12)
        res[prec] = {k: _np(v[0]) for k, v in plan.export(fields=("Phi_curr", "U_curr", "eta_curr")).items()
                     if k != "dead"}
        plan.close()
    ref = orc.run_model_scan(M=M, dt=dt, tmax=14, Phibar=Phibar, omega=omega, a=a, taurad=86400.0,
                             taudrag=86400.0, DPhieq=3e4, eta0_init=ics[0], delta0_init=ics[1], Phi0_init=ics[2],
                             U0_init=ics[3], V0_init=ics[4], diagnostics=False, return_history=False)["last_state"]
    for f in ("Phi_curr", "U_curr", "eta_curr"):
        r = getattr(ref, f)
        rel = {k: float(np.linalg.norm(res[k][f] - r) / np.linalg.norm(r)) for k in res}
        print(M, f, "relL2 vs oracle", rel)
        assert rel["bf16x3"] < 2e-5, (M, f, rel)


@pytest.mark.parametrize("M", [42, 63, 106])
def test_bf16x3_legendre_contractions(M, torch_gpu):
    """The Legendre kernel's split-BF16x3 contractions alone: a BF16x3 plan and an FP32 plan both
    on the plain FP32 grid kernel (SWX_OPT_GRID_KERNEL = 1), so they differ only in the Legendre
    products.  40 forced steps (the call's last steps advance the prev lineage, the others not:
    both specialised-update variants run), 3 members.  The two differ (the BF16x3 path runs) and
    stay within relative L2 2e-5 of each other and of the FP64 oracle on Phi, U, eta."""
    from my_swampe_amd import _capi
    from my_swampe_amd.model import param_row
    from my_swampe_amd.plan import Plan

    dt = 120.0 if M == 42 else 60.0
    a, omega, Phibar = 6.37122e6, 7.2921159e-5, 3e4
    p = param_row(dt=dt, a=a, omega=omega, g=9.8, Phibar=Phibar, taurad=86400.0, taudrag=86400.0,
                  DPhieq=3e4, K6=1.24e33, K6Phi=None, alpha=0.01)
    s = orc.build_static(M=M, dt=dt, a=a, omega=omega, g=9.8, Phibar=Phibar, taurad=86400.0, taudrag=86400.0,
                         DPhieq=3e4, K6=1.24e33, K6Phi=None, test=2)
    ics = orc.analytic_ic(s, 2, 0.3)
    _print_case("bf16x3-legendre", M=M)
    res = {}
    for prec in ("fp32", "bf16x3"):
        plan = Plan(M, np.stack([p] * 3), precision=prec)
        plan.set_option(_capi.OPT_GRID_KERNEL, 1)
        plan.init(*ics)
        plan.run(2, 40)
        res[prec] = {k: _np(v[0]) for k, v in plan.export(fields=("Phi_curr", "U_curr", "eta_curr")).items()
                     if k != "dead"}
        plan.close()
    ref = orc.run_model_scan(M=M, dt=dt, tmax=42, Phibar=Phibar, omega=omega, a=a, taurad=86400.0,
                             taudrag=86400.0, DPhieq=3e4, eta0_init=ics[0], delta0_init=ics[1], Phi0_init=ics[2],
                             U0_init=ics[3], V0_init=ics[4], diagnostics=False, return_history=False)["last_state"]
    for f in ("Phi_curr", "U_curr", "eta_curr"):
        r = getattr(ref, f)
        d = float(np.linalg.norm(res["bf16x3"][f] - res["fp32"][f]) / np.linalg.norm(res["fp32"][f]))
        rel = float(np.linalg.norm(res["bf16x3"][f] - r) / np.linalg.norm(r))
        print(M, f, "bf16x3 vs fp32", d, "vs oracle", rel)
        assert 0.0 < d < 2e-5 and rel < 2e-5, (M, f, d, rel)


@pytest.mark.parametrize("M", [42, 63, 106])
@pytest.mark.parametrize("fold", [1, 2])
def test_legendre_member_group_sizes_agree_bitwise(M, fold, torch_gpu):
    """Every member-group size of the MFMA Legendre kernel (the autotuner's candidates) gives the
    same state bitwise: a member's arithmetic does not depend on which members share its tile.
    5 members (partial groups for every size), 6 forced steps, FP32 transforms."""
    from my_swampe_amd import _capi
    from my_swampe_amd.model import param_row
    from my_swampe_amd.plan import Plan

    torch = torch_gpu
    _print_case("leg-mb", M=M, fold=fold)
    dt = 60.0 if M == 106 else 120.0
    rng = np.random.default_rng(5)
    params = np.stack([param_row(dt=dt, a=8.2e7, omega=3.2e-5, g=9.8, Phibar=3e5, taurad=10 ** rng.uniform(4, 6),
                                 taudrag=10 ** rng.uniform(4, 6), DPhieq=rng.uniform(1e5, 5e5), K6=1.24e33,
                                 K6Phi=None, alpha=0.01) for _ in range(5)])
    c = torch.tensor(_noise_spectral(M, 5, 11, 3.2e-5), device="cuda")
    states = {}
    tuned = None
    for mb in ((4, 8) if M == 42 else (2, 4, 8)):
        plan = Plan(M, params, precision="fp32")
        tuned = plan.get_option(_capi.OPT_LEG_MB)
        plan.set_option(_capi.OPT_LEG_MB, mb)
        plan.set_option(_capi.OPT_LEG_FOLD, fold)
        plan.init_spectral(c)
        plan.run(2, 6)
        states[mb] = plan.state_get()
        torch.cuda.synchronize()
        plan.close()
    print(M, "autotuned member-group size", tuned)
    ref = next(iter(states.values()))
    for mb, st in states.items():
        assert torch.equal(st, ref), (M, mb)
