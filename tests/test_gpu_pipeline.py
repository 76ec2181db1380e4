"""Concurrent stepping of several plans (swx_pipeline): each plan's results equal those it gives
when stepped alone, and an ensemble split into plans equals the one-plan ensemble bit for bit
(members are independent; the kernel variants are pinned to the same choice on every plan)."""

from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _params(B, seed=7):
    from my_swampe_amd.model import param_row

    rng = np.random.default_rng(seed)
    return np.stack([param_row(dt=120.0, a=8.2e7, omega=3.2e-5, g=9.8, Phibar=3e5, taurad=10 ** rng.uniform(4, 6),
                               taudrag=10 ** rng.uniform(4, 6), DPhieq=rng.uniform(1e5, 5e5), K6=1.24e33, K6Phi=None,
                               alpha=0.01) for _ in range(B)])


def _ics(M, B):
    rng = np.random.default_rng(1000)
    c = np.zeros((B, 3, M + 1, M + 1), complex)
    c[:, 0, 0, 1] = 3.2e-5 / np.sqrt(0.375)
    for m in range(11):
        for n in range(m, 11):
            c[:, 2, m, n] = 30.0 * (rng.standard_normal(B) + (1j * rng.standard_normal(B) if m else 0))
    return c


def _pin(plan, precision):
    from my_swampe_amd import _capi

    if precision == "fp32":
        plan.set_option(_capi.OPT_GRID_FOLD, 2)
        plan.set_option(_capi.OPT_LEG_MB, 4)


@pytest.mark.parametrize("precision,parts", [("fp32", 2), ("fp32", 3), ("fp64", 2)])
def test_split_ensemble_equals_one_plan(torch_gpu, precision, parts):
    from my_swampe_amd.plan import Pipeline, Plan, split_plans

    torch = torch_gpu
    print(f"[gpu-case] pipeline {precision} parts={parts}", flush=True)
    M, B, steps = 42, 40, 100
    params, c0 = _params(B), torch.tensor(_ics(M, B), device="cuda")
    s = torch.cuda.Stream()
    one = Plan(M, params, precision=precision)
    _pin(one, precision)
    plans = split_plans(M, params, parts, precision=precision)
    sizes = [p.B for p in plans]
    assert sum(sizes) == B and max(sizes) - min(sizes) <= 1, sizes
    with torch.cuda.stream(s):
        one.init_spectral(c0, stream=s)
        one.run_graph(2, steps, graph_steps=16, stream=s)
        want = one.export(fields=("Phi_curr", "eta_curr", "U_curr"), spec=True, stream=s)
        b0 = 0
        for p in plans:
            _pin(p, precision)
            p.init_spectral(c0[b0:b0 + p.B].contiguous(), stream=s)
            b0 += p.B
        pipe = Pipeline(plans)
        pipe.run_graph(2, 40, graph_steps=16, stream=s)
        pipe.run_graph(42, steps - 40, graph_steps=16, stream=s)
        got = [p.export(fields=("Phi_curr", "eta_curr", "U_curr"), spec=True, stream=s) for p in plans]
    s.synchronize()
    for name in ("Phi_curr", "eta_curr", "U_curr", "spec"):
        cat = torch.cat([g[name] for g in got])
        assert torch.equal(cat, want[name]), name
    pipe.close()
    for p in plans + [one]:
        p.close()


def test_pipeline_arguments(torch_gpu):
    from my_swampe_amd import _capi
    from my_swampe_amd.plan import Pipeline, Plan, split_plans

    torch = torch_gpu
    print("[gpu-case] pipeline arguments", flush=True)
    a = Plan(42, _params(2), precision="fp32")
    with pytest.raises(ValueError):
        Pipeline([a])
    with pytest.raises(ValueError):
        Pipeline([a, a])
    with pytest.raises(ValueError):
        split_plans(42, _params(3), 5)
    b = Plan(42, _params(2, seed=3), precision="fp32")
    pipe = Pipeline([a, b])
    # The pipeline marks its plans as stepped concurrently, and releases them when closed.
    assert a.get_option(_capi.OPT_CONCURRENT) == 1 and b.get_option(_capi.OPT_CONCURRENT) == 1
    with pytest.raises(Exception):  # run before the plans are initialized
        pipe.run_graph(2, 10, stream=torch.cuda.Stream())
    pipe.close()
    assert a.get_option(_capi.OPT_CONCURRENT) == 0
    a.set_option(_capi.OPT_CONCURRENT, 1)
    assert a.get_option(_capi.OPT_CONCURRENT) == 1
    a.close()
    b.close()


def test_pipeline_on_the_default_stream(torch_gpu):
    """The caller's stream only forks and joins (the plans replay their graphs on the pipeline's
    own streams), so the default stream serves as well; results equal each plan run alone."""
    from my_swampe_amd.plan import Pipeline, Plan

    torch = torch_gpu
    print("[gpu-case] pipeline on the default stream", flush=True)
    M, B = 42, 3
    c0 = torch.tensor(_ics(M, 2 * B), device="cuda")
    alone = []
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        for k in range(2):
            p = Plan(M, _params(B, seed=11 + k), precision="fp32")
            _pin(p, "fp32")
            p.init_spectral(c0[k * B:(k + 1) * B].contiguous(), stream=s)
            p.run_graph(2, 30, graph_steps=8, stream=s)
            alone.append(p.export(fields=("Phi_curr",), spec=True, stream=s))
            p.close()
    s.synchronize()
    plans = []
    for k in range(2):
        p = Plan(M, _params(B, seed=11 + k), precision="fp32")
        _pin(p, "fp32")
        p.init_spectral(c0[k * B:(k + 1) * B].contiguous())
        plans.append(p)
    pipe = Pipeline(plans)
    pipe.run_graph(2, 30, graph_steps=8)
    got = [p.export(fields=("Phi_curr",), spec=True) for p in plans]
    torch.cuda.synchronize()
    for want, have in zip(alone, got):
        for name in ("Phi_curr", "spec"):
            assert torch.equal(want[name], have[name]), name
    pipe.close()
    for p in plans:
        p.close()


def test_pipeline_of_two_truncations(torch_gpu):
    """A T42 plan and a T63 plan stepped concurrently give what each gives alone (bitwise, with
    the kernel variants pinned to the same choice)."""
    from my_swampe_amd.plan import Pipeline, Plan

    torch = torch_gpu
    print("[gpu-case] pipeline T42 + T63", flush=True)
    s = torch.cuda.Stream()
    cases = ((42, 4, 21), (63, 2, 22))
    ics = {M: torch.tensor(_ics(M, B), device="cuda") for M, B, _ in cases}
    alone = {}
    with torch.cuda.stream(s):
        for M, B, seed in cases:
            p = Plan(M, _params(B, seed=seed), precision="fp32")
            _pin(p, "fp32")
            p.init_spectral(ics[M], stream=s)
            p.run_graph(2, 24, graph_steps=8, stream=s)
            alone[M] = p.export(fields=("Phi_curr", "U_curr"), spec=True, stream=s)
            p.close()
        plans = []
        for M, B, seed in cases:
            p = Plan(M, _params(B, seed=seed), precision="fp32")
            _pin(p, "fp32")
            p.init_spectral(ics[M], stream=s)
            plans.append(p)
        pipe = Pipeline(plans)
        pipe.run_graph(2, 24, graph_steps=8, stream=s)
        got = {M: p.export(fields=("Phi_curr", "U_curr"), spec=True, stream=s) for (M, _, _), p in zip(cases, plans)}
    s.synchronize()
    for M in alone:
        for name in ("Phi_curr", "U_curr", "spec"):
            assert torch.equal(alone[M][name], got[M][name]), (M, name)
    pipe.close()
    for p in plans:
        p.close()


def test_headline_arrangement_equals_one_plan_at_full_size(torch_gpu):
    """The benchmark's arrangement at its size: 256 T42 members in FP32 as two concurrent plans of
    128 (each plan's kernel variants and Legendre member group left to its own choice, as the
    benchmark leaves them) export bit for bit what one 256-member plan exports, over graph-replayed
    calls of 50, 20 and 20 steps (each call ending in its lineage window and row save)."""
    from my_swampe_amd.plan import Pipeline, Plan, split_plans

    torch = torch_gpu
    print("[gpu-case] pipeline 2 x 128 vs 256, FP32", flush=True)
    M, B = 42, 256
    params, c0 = _params(B), torch.tensor(_ics(M, B), device="cuda")
    fields = ("Phi_curr", "Phi_prev", "eta_curr", "delta_curr", "U_curr", "V_curr")
    s = torch.cuda.Stream()
    one = Plan(M, params, precision="fp32", diagnostics=False)
    plans = split_plans(M, params, 2, precision="fp32", diagnostics=False)
    with torch.cuda.stream(s):
        one.init_spectral(c0, stream=s)
        b0 = 0
        for p in plans:
            p.init_spectral(c0[b0:b0 + p.B].contiguous(), stream=s)
            b0 += p.B
        pipe = Pipeline(plans)
        t = 2
        for n in (50, 20, 20):
            one.run_graph(t, n, stream=s)
            pipe.run_graph(t, n, stream=s)
            t += n
        want = one.export(fields=fields, spec=True, stream=s)
        got = [p.export(fields=fields, spec=True, stream=s) for p in plans]
    s.synchronize()
    for name in fields + ("spec",):
        cat = torch.cat([g[name] for g in got])
        assert torch.isfinite(cat.real if cat.is_complex() else cat).all(), name
        assert torch.equal(cat, want[name]), name
    pipe.close()
    for p in plans + [one]:
        p.close()
