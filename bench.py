#!/usr/bin/env python3
"""Benchmark: member-steps/s of the forced modified-Euler spectral step, T42 x 256 members.

Workload (synthetic, seeded): T42 (J=64, I=128), 256 ensemble members, forced
modified Euler, dt=120 s, a=8.2e7, omega=3.2e-5, Phibar=3e5, K6=1.24e33,
diffusion and modal filter on; per-member taurad, taudrag log-uniform in
[1e4, 1e6] s and DPhieq uniform in [1e5, 5e5] (seed 7); rest + seeded
spectral Phi noise initial condition (n <= 10), set through swx_init_spectral.  Transforms in FP32, spectral
state update and diffusion in FP64.  Inputs are resident in HBM before the
timed region; each step replays the HIP graph driver.  `value` is the members split into two
plans of 128 stepped concurrently on two streams (plan.Pipeline), the fastest arrangement of
this batch on the GPU; one plan of all 256 members is timed beside it, and its isolated kernel
launches give the roofline.

Protocol (scripts/benchmark_scan.py:116-141): W untimed warmup steps, then K
timed steps bracketed by device synchronization; one JSON line on stdout.
Beside it: the same batch in FP64, the plain FP64 path at one member, the
per-kernel roofline from HIP events, and the FP64 CPU oracle timed on the host.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "member-steps/s at T42×256 ens; fwd+adjoint time vs plain FP64 GPU, same GPU"
PEAK = {"fp32": 157.3, "fp64": 78.6, "bf16x3": 2500.0 / 3}  # TFLOP/s, MI355X dense matrix peaks (bf16x3: 3 bf16 products)
HBM_PEAK = 8000.0  # GB/s


def member_params(B: int, seed: int = 7, dt: float = 120.0, alpha: float = 0.01):
    from my_swampe_amd.model import param_row

    rng = np.random.default_rng(seed)
    taurad = 10.0 ** rng.uniform(4.0, 6.0, B)
    taudrag = 10.0 ** rng.uniform(4.0, 6.0, B)
    dphieq = rng.uniform(1e5, 5e5, B)
    return np.stack([param_row(dt=dt, a=8.2e7, omega=3.2e-5, g=9.8, Phibar=3e5, taurad=taurad[b],
                               taudrag=taudrag[b], DPhieq=dphieq[b], K6=1.24e33, K6Phi=None, alpha=alpha)
                     for b in range(B)])


def flops_per_member_step(M: int, J: int, I: int):
    """Algorithmic work per member-step (fixed formula; FP64-carry forward step).

    Legendre: 18 triangular contractions (11 forward, 7 inverse), DFT: 13 dense
    real contractions (8 forward, 5 inverse), complex x real = 4 flops per
    multiply-add site.
    """
    L = (M + 1) * (M + 2) // 2
    leg = 4 * J * 18 * L
    dft = 4 * J * 13 * I * (M + 1)
    return leg, dft


def fold_grid_mfma_flops(M: int, J: int, I: int, B: int) -> int:
    """Matrix-core flops one launch of the steady half-period folded grid kernel executes (padded
    tiles included): per 16-row block, the inverse GEMM's K1 / 4 k-steps x 4 fields (eta, Phi, U,
    V: the steady products do not read delta) x H / 16 longitude tiles and the forward GEMM's 4
    waves x H / 16 k-groups x 4 k-steps x NT2 output tiles x 2 fields, each a 16x16x4 MFMA of 2048
    flops (K1 = even-m and odd-m rows each padded to 4, NT2 = even-m and odd-m output tiles of 16
    rows, H = I / 2).  T42: 1120 MFMAs per block, 280 per wave, as SQ_INSTS_MFMA counts them."""
    Mp = M + 1
    ne, no = (Mp + 1) // 2, Mp // 2
    k1 = (2 * ne + 3) // 4 * 4 + (2 * no + 3) // 4 * 4
    nt2 = (2 * ne + 15) // 16 + (2 * no + 15) // 16
    h16 = I // 2 // 16
    per_block = (k1 // 4) * 4 * h16 + 4 * h16 * 4 * nt2 * 2
    return B * J // 16 * per_block * 2048


def run_batch(M, B, precision, steps, warmup, graph_steps=32, full_phys=False, si_raw=False, parts=1, reps=1,
              clock_warmup_s=0.0):
    """si_raw: the semi-implicit leapfrog with the RAW filter in its production setting
    (dt = 600 s, alpha = 0.05, williams_alpha = 0.53), else forced modified Euler at dt = 120 s.
    parts > 1: the members split into that many plans stepped concurrently (plan.Pipeline); the
    kernel timings are then not taken.  The timed call of `steps` steps runs `reps` times in a row
    (the state advancing); `wall` is the median, `walls` all of them.  clock_warmup_s: untimed
    stepping before the warmup steps, so the timed calls find the GPU clocks ramped."""
    import torch
    from my_swampe_amd import _capi
    from my_swampe_amd.plan import Pipeline, split_plans

    params = member_params(B, dt=600.0, alpha=0.05) if si_raw else member_params(B)
    c0 = torch.tensor(spectral_ics(M, B), device="cuda")
    plans = split_plans(M, params, parts, precision=precision, diagnostics=False, semi_implicit=si_raw,
                        raw_filter=si_raw)
    for plan in plans:
        if full_phys:
            plan.set_option(_capi.OPT_FULL_PHYS_CARRY, 1)
    runner = Pipeline(plans) if parts > 1 else plans[0]
    stream = torch.cuda.Stream()
    torch.cuda.synchronize()
    ms_leg = ms_grid = None
    with torch.cuda.stream(stream):
        b0 = 0
        for plan in plans:
            plan.init_spectral(c0[b0:b0 + plan.B].contiguous(), stream=stream)
            b0 += plan.B
        t = 2
        if clock_warmup_s > 0:
            t1 = time.perf_counter()
            while time.perf_counter() - t1 < clock_warmup_s:
                runner.run_graph(t, 50, graph_steps=graph_steps, stream=stream)
                stream.synchronize()
                t += 50
        runner.run_graph(t, warmup, graph_steps=graph_steps, stream=stream)
        t += warmup
        # One untimed call of the timed length: its step sequences are captured and instantiated
        # here, so the timed calls replay them (the compile run of scripts/benchmark_scan.py).
        runner.run_graph(t, steps, graph_steps=graph_steps, stream=stream)
        t += steps
        stream.synchronize()
        clock_before = gpu_sclk_mhz()
        c0 = [plan.step_counts() for plan in plans]
        walls = []
        for _ in range(reps):
            t0 = time.perf_counter()
            runner.run_graph(t, steps, graph_steps=graph_steps, stream=stream)
            stream.synchronize()
            walls.append(time.perf_counter() - t0)
            t += steps
        clock_after = gpu_sclk_mhz()
        c1 = [plan.step_counts() for plan in plans]
        counts = {k: sum(b[k] - a[k] for a, b in zip(c0, c1)) for k in c0[0]}
        wall = float(np.median(walls))
        if parts == 1:
            ms_leg, ms_grid = plans[0].time_kernels(t, 20, stream=stream)
        exs = [plan.export(fields=("Phi_curr",), stream=stream) for plan in plans]
        stream.synchronize()
    finite = all(bool(torch.isfinite(ex["Phi_curr"]).all()) for ex in exs)
    fold = plans[0].get_option(_capi.OPT_GRID_FOLD) if precision == "fp32" else 1
    leg_mb = plans[0].get_option(_capi.OPT_LEG_MB) if precision != "fp64" else 0
    if parts > 1:
        runner.close()
    for plan in plans:
        plan.close()
    return dict(wall=wall, walls=walls, ms_leg=ms_leg, ms_grid=ms_grid, finite=finite, fold=fold, leg_mb=leg_mb,
                counts=counts, clock_mhz=[clock_before, clock_after])


def gpu_sclk_mhz():
    """Current shader clock (MHz) of the visible GPU from the kernel driver's DPM table (the
    entry marked '*'), or None where it cannot be read."""
    import glob

    vals = []
    for path in sorted(glob.glob("/sys/class/drm/card*/device/pp_dpm_sclk")):
        try:
            with open(path) as f:
                for line in f:
                    if line.rstrip().endswith("*"):
                        vals.append(int(line.split(":")[1].strip().lower().split("mhz")[0]))
        except (OSError, ValueError, IndexError):
            continue
    return vals[0] if len(vals) == 1 else (vals or None)


def spectral_ics(M: int, B: int, omega: float = 3.2e-5):
    """Spectral (eta, delta, Phi) of the rest + noise state: the planetary vorticity (m=0, n=1 coefficient of
    2 omega mu) plus Phi noise of std 30 per real component on n <= 10 (one generator, seed 1000)."""
    rng = np.random.default_rng(1000)
    c = np.zeros((B, 3, M + 1, M + 1), complex)
    c[:, 0, 0, 1] = omega / np.sqrt(0.375)
    for m in range(11):
        for n in range(m, 11):
            c[:, 2, m, n] = 30.0 * (rng.standard_normal(B) + (1j * rng.standard_normal(B) if m else 0))
    return c


def run_fwd_adjoint(M, B, precision, steps, reps=2):
    """Wall time of forward + hand-written adjoint (sqrt(N) checkpointing) through the autograd rollout."""
    import torch
    from my_swampe_amd.autograd import rollout_phi

    dev = torch.device("cuda")
    params = member_params(B)
    c0 = torch.tensor(spectral_ics(M, B), device=dev)
    tr0 = torch.tensor(params[:, 5], device=dev)
    td0 = torch.tensor(params[:, 6], device=dev)
    dp0 = torch.tensor(params[:, 7], device=dev)
    J, I = (64, 128) if M == 42 else ((96, 192) if M == 63 else (160, 320))
    tgt = torch.zeros((J, I), dtype=torch.float64, device=dev)
    times = []
    stats = {}
    for r in range(reps + 1):
        c = c0.clone().requires_grad_(True)
        tr, td, dp = (x.clone().requires_grad_(True) for x in (tr0, td0, dp0))
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        phi = rollout_phi(c, tr, td, dp, M=M, nsteps=steps, dt=120.0, a=8.2e7, omega=3.2e-5, Phibar=3e5,
                          precision=precision, stats=stats)
        loss = ((phi - tgt) ** 2).mean(dim=(1, 2)).sum()
        loss.backward()
        torch.cuda.synchronize()
        if r > 0:  # first repetition warms up plan creation and caches
            times.append(time.perf_counter() - t0)
    finite = bool(torch.isfinite(c.grad).all()) and bool(torch.isfinite(tr.grad).all())
    return dict(wall=float(np.median(times)), finite=finite, checkpoints=stats.get("checkpoints"),
                segment_states=stats.get("segment_states"))


def run_t63_fp64_adjoint(steps=2000):
    """T63, 1 member, FP64, `steps` steps of forced modified Euler (dt = 120 s): the scalar loss
    mean((Phi_final - Phi_target)^2) with the target from a seeded run at taurad = 5e4 s,
    taudrag = 2e5 s; forward + adjoint to the gradients w.r.t. taurad, taudrag, DPhieq and the
    initial spectral state, with a checkpoint every ceil(sqrt(steps)) steps (recompute in the
    reverse sweep).  Wall time of the forward + adjoint (median of 2 after a warm-up)."""
    import math

    import torch
    from my_swampe_amd.autograd import rollout_phi

    dev = torch.device("cuda")
    every = math.isqrt(steps - 1) + 1
    kw = dict(M=63, nsteps=steps, dt=120.0, a=8.2e7, omega=3.2e-5, Phibar=3e5, precision="fp64",
              checkpoint_every=every)
    c0 = torch.tensor(spectral_ics(63, 1), device=dev)
    with torch.no_grad():
        tgt = rollout_phi(c0, 5e4, 2e5, 3e5, **kw)
    times, stats = [], {}
    for r in range(3):
        c = c0.clone().requires_grad_(True)
        tr, td, dp = (torch.tensor([v], dtype=torch.float64, device=dev, requires_grad=True)
                      for v in (86400.0, 86400.0, 3e5))
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        loss = ((rollout_phi(c, tr, td, dp, stats=stats, **kw) - tgt) ** 2).mean()
        loss.backward()
        torch.cuda.synchronize()
        if r > 0:
            times.append(time.perf_counter() - t0)
    wall = float(np.median(times))
    finite = all(bool(torch.isfinite(x.grad).all()) for x in (c, tr, td, dp))
    return dict(members=1, steps=steps, precision="fp64", checkpoint_every=every,
                checkpoints=stats.get("checkpoints"), wall_s=wall, member_steps_per_s=steps / wall, finite=finite)


def run_mixed(steps, warmup, B42=256, B63=64, graph_steps=32):
    """256 x T42 + 64 x T63 members, FP32 transforms, dt = 120 s, per-member forcing parameters
    (seeds 7 and 8): both truncations in one Legendre and one grid launch per step (Ensemble),
    the same two plans stepped one after the other, and the two plans stepped concurrently on two
    streams (Pipeline)."""
    import torch
    from my_swampe_amd.plan import Ensemble, Pipeline, Plan

    stream = torch.cuda.Stream()
    torch.cuda.synchronize()
    walls = {}
    with torch.cuda.stream(stream):
        plans = [Plan(42, member_params(B42, seed=7), precision="fp32"),
                 Plan(63, member_params(B63, seed=8), precision="fp32")]
        for pl, (M, B) in zip(plans, ((42, B42), (63, B63))):
            pl.init_spectral(torch.tensor(spectral_ics(M, B), device="cuda"), stream=stream)
        ens = Ensemble(plans[0], plans[1])
        t = 2

        def timed(name, run):
            # warmup steps, one untimed call of the timed length (graphs captured), the timed call
            nonlocal t
            run(t, warmup)
            run(t + warmup, steps)
            stream.synchronize()
            t += warmup + steps
            t0 = time.perf_counter()
            run(t, steps)
            stream.synchronize()
            walls[name] = time.perf_counter() - t0
            t += steps

        timed("one_launch", lambda t0, n: ens.run_graph(t0, n, graph_steps=graph_steps, stream=stream))

        def separate(t0, n):
            for pl in plans:
                pl.run_graph(t0, n, graph_steps=graph_steps, stream=stream)

        timed("separate", separate)
        pipe = Pipeline(plans)
        timed("concurrent", lambda t0, n: pipe.run_graph(t0, n, graph_steps=graph_steps, stream=stream))
        pipe.close()
        ex = [pl.export(fields=("Phi_curr",), stream=stream)["Phi_curr"] for pl in plans]
        stream.synchronize()
        finite = all(bool(torch.isfinite(x).all()) for x in ex)
        ens.close()
        for pl in plans:
            pl.close()
    n = B42 + B63
    return dict(members={"T42": B42, "T63": B63}, steps=steps, ms_per_step=walls["one_launch"] / steps * 1e3,
                member_steps_per_s=steps * n / walls["one_launch"],
                separate_ms_per_step=walls["separate"] / steps * 1e3,
                concurrent_ms_per_step=walls["concurrent"] / steps * 1e3,
                concurrent_member_steps_per_s=steps * n / walls["concurrent"],
                separate_member_steps_per_s=steps * n / walls["separate"], finite=finite)


def _t106_rollout(c0, params, steps, every, precision):
    import torch
    from my_swampe_amd.autograd import rollout_phi

    dev = c0.device
    c = c0.clone().requires_grad_(True)
    tr, td, dp = (torch.tensor(params[:, k], device=dev).requires_grad_(True) for k in (5, 6, 7))
    tgt = torch.zeros((160, 320), dtype=torch.float64, device=dev)
    torch.cuda.synchronize()
    torch.cuda.reset_peak_memory_stats()
    base = torch.cuda.memory_allocated()
    stats = {}
    t0 = time.perf_counter()
    phi = rollout_phi(c, tr, td, dp, M=106, nsteps=steps, dt=60.0, a=8.2e7, omega=3.2e-5, Phibar=3e5,
                      precision=precision, checkpoint_every=every, stats=stats)
    loss = ((phi - tgt) ** 2).mean(dim=(1, 2)).sum()
    loss.backward()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    torch_peak = torch.cuda.max_memory_allocated() - base
    return dict(wall=wall, stats=stats, torch_peak=int(torch_peak), phi=phi.detach(),
                grads=[c.grad.detach(), tr.grad.detach(), td.grad.detach(), dp.grad.detach()])


def run_t106_fwd_adjoint(B=64, steps=5000, every=70, precision="fp32", compare=("bf16x3", "fp64")):
    """T106 x 64 members, dt = 60 s, 5000 steps forward + adjoint with a spectral checkpoint every
    70 steps: wall time and peak device memory (torch-held checkpoints and segment states at their
    peak plus the plan's own buffers); the same run with the other transform precisions timed
    beside it, and the final Phi and gradients of the reduced-precision runs against FP64
    (relative L2: initial spectral state, taurad, taudrag, DPhieq)."""
    import torch

    dev = torch.device("cuda")
    params = member_params(B, dt=60.0)
    c0 = torch.tensor(spectral_ics(106, B), device=dev)
    runs = {precision: _t106_rollout(c0, params, steps, every, precision)}
    for prec in compare:
        if prec != precision:
            runs[prec] = _t106_rollout(c0, params, steps, every, prec)
    main = runs[precision]
    stats = main["stats"]
    finite = all(bool(torch.isfinite(g).all()) for g in main["grads"])
    out = dict(members=B, steps=steps, dt=60.0, checkpoint_every=every, precision=precision, wall_s=main["wall"],
               member_steps_per_s=steps * B / main["wall"], checkpoints=stats.get("checkpoints"),
               segment_states=stats.get("segment_states"), torch_peak_bytes=main["torch_peak"],
               plan_device_bytes=int(stats.get("plan_device_bytes", 0)),
               peak_hbm_bytes=int(main["torch_peak"] + stats.get("plan_device_bytes", 0)), finite=finite)
    rel = lambda a, b: float(torch.linalg.vector_norm(a - b) / torch.linalg.vector_norm(b))  # noqa: E731
    ref = runs.get("fp64")
    for prec, r in runs.items():
        if prec == precision:
            continue
        out[f"{prec}_wall_s"] = r["wall"]
        out[f"{prec}_member_steps_per_s"] = steps * B / r["wall"]
    if ref is not None:
        for prec, r in runs.items():
            if prec == "fp64":
                continue
            names = ("init_state", "taurad", "taudrag", "DPhieq")
            out[f"{prec}_vs_fp64_rel_l2"] = dict(
                final_Phi=rel(r["phi"], ref["phi"]),
                **{f"grad_{n}": rel(g, gr) for n, g, gr in zip(names, r["grads"], ref["grads"])})
    return out


CPU_STEPS = 240  # oracle steps per timed run (one member)


def _cpu_worker(steps: int, start_at: float) -> None:
    """One CPU worker: the FP64 numpy oracle, T42, 1 member, forced modified Euler, dt = 120 s.
    Warms its geometry caches, waits for the common start time, runs `steps` steps and prints
    {"t0", "t1"} (wall-clock seconds) as JSON."""
    from oracle import swampe_oracle as orc

    eta0, delta0, Phi0, U0, V0 = orc.rest_noise_ic(42, 0, 3.2e-5)
    kw = dict(M=42, dt=120.0, Phibar=3e5, omega=3.2e-5, a=8.2e7, taurad=86400.0, taudrag=86400.0, DPhieq=3e5,
              eta0_init=eta0, delta0_init=delta0, Phi0_init=Phi0, U0_init=U0, V0_init=V0, diagnostics=False,
              return_history=False)
    orc.run_model_scan(tmax=4, **kw)  # geometry and basis caches
    while time.time() < start_at:
        time.sleep(0.005)
    t0 = time.time()
    orc.run_model_scan(tmax=steps + 2, **kw)
    print(json.dumps({"t0": t0, "t1": time.time()}), flush=True)


def _cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _cpu_round(workers: int, steps: int) -> float:
    """`workers` single-threaded oracle processes started together: member-steps/s of the round
    (workers x steps over the span from the common start to the last finish)."""
    import subprocess

    env = dict(os.environ, OMP_NUM_THREADS="1", OPENBLAS_NUM_THREADS="1", MKL_NUM_THREADS="1")
    start_at = time.time() + 4.0  # after every worker has imported and warmed up
    procs = [subprocess.Popen([sys.executable, os.path.abspath(__file__), "--cpu-worker", str(steps), repr(start_at)],
                              env=env, stdout=subprocess.PIPE, text=True) for _ in range(workers)]
    ends = []
    for pr in procs:
        outp, _ = pr.communicate()
        if pr.returncode != 0:
            raise RuntimeError(f"CPU worker failed ({pr.returncode})")
        ends.append(json.loads(outp.strip().splitlines()[-1]))
    t0 = min(r["t0"] for r in ends)
    return workers * steps / (max(r["t1"] for r in ends) - t0)


def cpu_baseline(runs: int = 3, steps: int = CPU_STEPS):
    """FP64 numpy oracle on the host (scripts/benchmark_scan.py:116-141 protocol: median of
    repeated runs): T42, forced modified Euler, dt = 120 s, one member per single-threaded process;
    one process (single core) and one process per available core, at most 16 (all cores, members
    in parallel), `runs` rounds each.  Run before the GPU is initialised (the workers are separate
    interpreters)."""
    nproc = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    workers = max(1, min(16, nproc))
    single = [_cpu_round(1, steps) for _ in range(runs)]
    allc = [_cpu_round(workers, steps) for _ in range(runs)]
    med = lambda v: float(np.median(v))  # noqa: E731
    return dict(value=med(allc), unit="member-steps/s", cores=workers, kind="port",
                sample=(f"FP64 numpy oracle, T42 forced modEuler, dt=120 s, {steps} steps per member; {workers} "
                        f"single-threaded processes (one member each) started together, median of {runs} rounds"),
                single_core=dict(value=med(single), runs=single, cores=1),
                all_core_runs=allc, nproc=nproc, cpu_model=_cpu_model())


def accuracy_vs_oracle(steps: int = 2000):
    """T42, 1 member, forced modified Euler, dt = 120 s, rest + seeded Phi noise (std 100, n <= 10):
    the final eta, delta, Phi of the HIP path in FP64, FP32 and split-BF16x3 transforms against the
    FP64 oracle over `steps` steps (max-abs and relative L2 per field), and the oracle's time."""
    import torch
    from my_swampe_amd.model import run_model_scan
    from oracle import swampe_oracle as orc

    eta0, delta0, Phi0, U0, V0 = orc.rest_noise_ic(42, 0, 3.2e-5)
    kw = dict(M=42, dt=120.0, tmax=steps + 2, Phibar=3e5, omega=3.2e-5, a=8.2e7, taurad=86400.0, taudrag=86400.0,
              DPhieq=3e5, eta0_init=eta0, delta0_init=delta0, Phi0_init=Phi0, U0_init=U0, V0_init=V0,
              diagnostics=False, return_history=False)
    t0 = time.perf_counter()
    ref = orc.run_model_scan(**kw)["last_state"]
    oracle_s = time.perf_counter() - t0
    out = dict(steps=steps, members=1, oracle_seconds=oracle_s)
    for prec in ("fp64", "fp32", "bf16x3"):
        st = run_model_scan(**kw, precision=prec)["last_state"]
        torch.cuda.synchronize()
        res = {}
        for f in ("eta", "delta", "Phi"):
            g = getattr(st, f + "_curr").cpu().numpy()
            r = getattr(ref, f + "_curr")
            res[f] = dict(max_abs=float(np.max(np.abs(g - r))),
                          rel_l2=float(np.linalg.norm(g - r) / max(np.linalg.norm(r), 1e-300)))
        out[prec] = res
    return out


def measured_traffic(kernel: str, members: int, precision: str):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 counter passes (FETCH_SIZE,
    WRITE_SIZE, each scaled by the copy-kernel calibration), or None for another workload."""
    path = os.path.join(ROOT, "profiles", f"hbm_traffic_T42x{members}_{precision}.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        rec = json.load(f)
    k = rec["kernels"].get(kernel)
    return None if k is None else int(k["hbm_bytes"])


def copy_calibration(mib: int):
    """A plain copy of known size: FETCH_SIZE / WRITE_SIZE counters read on it give the scale
    factors applied to the step kernels' counter values."""
    import torch

    n = mib * (1 << 20) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes": n * 4, "copies": 3}), flush=True)


def main():
    if len(sys.argv) == 4 and sys.argv[1] == "--cpu-worker":
        _cpu_worker(int(sys.argv[2]), float(sys.argv[3]))
        return
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--members", type=int, default=256)
    ap.add_argument("--precision", default="fp32", choices=("fp32", "fp64", "bf16x3"))
    ap.add_argument("--reps", type=int, default=3, help="timed calls of --steps steps each (median reported)")
    ap.add_argument("--clock-warmup", type=float, default=0.3, metavar="S",
                    help="seconds of untimed stepping before the warmup steps of the first run")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-extra", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--adjoint-steps", type=int, default=100)
    ap.add_argument("--no-t106", action="store_true", help="skip the T106 x 64, 5000-step fwd+adjoint run")
    ap.add_argument("--t106-steps", type=int, default=5000, help="steps of the T106 x 64 fwd+adjoint run")
    ap.add_argument("--t106-compare", default="bf16x3,fp64",
                    help="transform precisions timed beside the T106 run (comma list, or 'none')")
    ap.add_argument("--no-mixed", action="store_true", help="skip the 256 x T42 + 64 x T63 one-launch run")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MiB",
                    help="only run a plain device-to-device copy of this size (HBM counter calibration)")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; members shard across GPUs with no per-step "
                                   "exchange, multi-GPU runs are not provided"}))
        sys.exit(2)
    # The CPU baseline's worker processes start before anything touches the GPU.
    cpu = None if args.no_cpu_baseline else cpu_baseline()
    import torch

    if not torch.cuda.is_available():
        print(json.dumps({"error": "no ROCm GPU visible"}))
        sys.exit(1)
    M, J, I, B = 42, 64, 128, args.members
    from my_swampe_amd.plan import measure_ceilings

    ceil = measure_ceilings()
    # The B members split into two plans stepped concurrently on two streams (plan.Pipeline):
    # `value`.  One plan of all B members (one Legendre and one grid launch per step) is timed
    # beside it, and its kernels, timed alone, give the roofline (while two streams overlap, a
    # launch's duration is not its own).
    main_run = run_batch(M, B, args.precision, args.steps, args.warmup, reps=args.reps,
                         clock_warmup_s=args.clock_warmup)
    parts = 2 if B >= 2 else 1
    pipe_run = (run_batch(M, B, args.precision, args.steps, args.warmup, parts=parts, reps=args.reps,
                          clock_warmup_s=args.clock_warmup)
                if parts > 1 else main_run)
    one_value = args.steps * B / main_run["wall"]
    value = args.steps * B / pipe_run["wall"]
    leg_f, dft_f = flops_per_member_step(M, J, I)
    grid_tflops = dft_f * B / (main_run["ms_grid"] * 1e-3) / 1e12
    leg_tflops = leg_f * B / (main_run["ms_leg"] * 1e-3) / 1e12
    peak = PEAK[args.precision]
    peak_measured = {"fp32": ceil["mfma_fp32_tflops"], "fp64": ceil["mfma_fp64_tflops"],
                     "bf16x3": ceil["mfma_bf16_tflops"] / 3}[args.precision]
    grid_kernel = {"fp32": f"k_grid_mfma<{M}>", "bf16x3": f"k_grid_bf16x3<{M}>", "fp64": "k_grid<double>"}[args.precision]
    if args.precision == "fp32" and main_run["fold"] == 2:
        grid_kernel = f"k_grid_mfma_fold_w2<{M}, true>"
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "member-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": pipe_run["wall"] / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": f"{args.precision} transforms, fp64 state",
        "data": "synthetic (seeded rest+noise ICs, per-member forcing params)",
        "config": {"workload": f"T42x{B} forced modified-Euler forward, dt=120 s, FP32 transforms",
                   "members": B, "M": M, "J": J, "I": I, "precision": args.precision,
                   "plans": parts, "arrangement": f"{parts} plans of {B // parts} members stepped concurrently "
                   "on two streams (plan.Pipeline)" if parts > 1 else "one plan",
                   "timed_calls": args.reps, "value_from": "median of the timed calls",
                   "call_ms": [w * 1e3 for w in pipe_run["walls"]],
                   "timed_step_counts": pipe_run["counts"],
                   "gpu_sclk_mhz_before_after": pipe_run["clock_mhz"],
                   "one_plan": {"plans": 1, "member_steps_per_s": one_value,
                                "ms_per_step": main_run["wall"] / args.steps * 1e3,
                                "call_ms": [w * 1e3 for w in main_run["walls"]],
                                "timed_step_counts": main_run["counts"],
                                "gpu_sclk_mhz_before_after": main_run["clock_mhz"]}},
        "roofline": {"bound": "mfma", "achieved": grid_tflops, "peak": peak, "unit": "TFLOP/s",
                     "frac": grid_tflops / peak,
                     "traffic": measured_traffic(grid_kernel, B, args.precision),
                     "kernel": f"{grid_kernel} (inverse DFT + grid physics + forward DFT)",
                     "note": ("achieved = the dense-DFT algorithmic flops (4 J 13 I (M+1) per member, no "
                              "symmetry folding) / kernel time; the folded kernel executes half of them"
                              if grid_kernel.startswith("k_grid_mfma_fold") else "achieved = algorithmic DFT flops / kernel time"),
                     "leg_member_group": main_run["leg_mb"],
                     "algorithmic_flops_per_launch": dft_f * B,
                     "kernel_ms": main_run["ms_grid"], "leg_kernel_ms": main_run["ms_leg"],
                     "leg_tflops": leg_tflops,
                     "peak_measured": peak_measured, "frac_of_measured": grid_tflops / peak_measured},
        "ceilings_measured": dict(ceil, how="swx_measure_ceilings: 8 independent accumulators per wave, "
                                  "8 waves per SIMD; HBM passes over 1 GiB; shortest of 5 launches"),
        "finite": main_run["finite"] and pipe_run["finite"],
    }
    leg_kernel = f"k_leg_mfma_w5<{M}, {main_run['leg_mb']}, false, false, true>"
    hbm = {}
    for name, kern, ms in (("grid", grid_kernel, main_run["ms_grid"]), ("legendre", leg_kernel, main_run["ms_leg"])):
        tb = measured_traffic(kern, B, args.precision)
        if tb is not None and ms:
            rate = tb / (ms * 1e-3) / 1e12
            hbm[name] = {"kernel": kern, "traffic": tb, "tbps": rate, "frac_of_peak": rate / (HBM_PEAK / 1000),
                         "frac_of_measured_copy": rate / ceil["hbm_copy_tbps"]}
    out["roofline"]["hbm"] = hbm
    if grid_kernel.startswith("k_grid_mfma_fold"):
        ex = fold_grid_mfma_flops(M, J, I, B)
        ex_tflops = ex / (main_run["ms_grid"] * 1e-3) / 1e12
        out["roofline"]["executed"] = {"flops_per_launch": ex, "tflops": ex_tflops, "frac_of_peak": ex_tflops / peak,
                                       "frac_of_measured": ex_tflops / peak_measured}
    if not args.no_extra:
        full = run_batch(M, B, args.precision, max(args.steps // 2, 100), 10, full_phys=True)
        fp64 = run_batch(M, B, "fp64", max(args.steps // 4, 50), 10)
        plain = run_batch(M, 1, "fp64", max(args.steps // 2, 200), 20)
        bf = run_batch(M, B, "bf16x3", max(args.steps // 2, 100), 10)
        sr = run_batch(M, B, args.precision, max(args.steps // 2, 100), 10, si_raw=True)
        fp64_v = max(args.steps // 4, 50) * B / fp64["wall"]
        plain_v = max(args.steps // 2, 200) / plain["wall"]
        out["extra"] = {
            "full_phys_carry_member_steps_per_s": max(args.steps // 2, 100) * B / full["wall"],
            "full_phys_carry_kernel_ms": {"leg": full["ms_leg"], "grid": full["ms_grid"]},
            "fp64_batched_member_steps_per_s": fp64_v,
            "plain_fp64_1member_member_steps_per_s": plain_v,
            "ratio_vs_plain_fp64_1member": value / plain_v,
            "ratio_vs_fp64_same_batch": value / fp64_v,
            "fp64_batched_kernel_ms": {"leg": fp64["ms_leg"], "grid": fp64["ms_grid"]},
            "plain_kernel_ms": {"leg": plain["ms_leg"], "grid": plain["ms_grid"]},
            "bf16x3_member_steps_per_s": max(args.steps // 2, 100) * B / bf["wall"],
            "bf16x3_kernel_ms": {"leg": bf["ms_leg"], "grid": bf["ms_grid"]},
            "bf16x3_finite": bf["finite"],
            "si_raw_production_member_steps_per_s": max(args.steps // 2, 100) * B / sr["wall"],
            "si_raw_production_kernel_ms": {"leg": sr["ms_leg"], "grid": sr["ms_grid"]},
            "si_raw_production_finite": sr["finite"],
        }
    if not args.no_adjoint:
        na = args.adjoint_steps
        fa = run_fwd_adjoint(M, B, args.precision, na)
        pa = run_fwd_adjoint(M, 1, "fp64", na)
        fa_v = na * B / fa["wall"]
        pa_v = na / pa["wall"]
        out["fwd_adjoint"] = {
            "steps": na, "members": B, "checkpoints": fa["checkpoints"], "segment_states": fa["segment_states"],
            "checkpoint_policy": "every state kept when the trajectory fits 32 GiB, else every ceil(sqrt(steps))",
            "wall_s": fa["wall"], "member_steps_per_s": fa_v, "finite": fa["finite"],
            "plain_fp64_1member_wall_s": pa["wall"], "plain_fp64_1member_member_steps_per_s": pa_v,
            "ratio_vs_plain_fp64_1member": fa_v / pa_v,
        }
        out["t63_fp64_adjoint"] = run_t63_fp64_adjoint()
    if not args.no_mixed and args.precision == "fp32":
        out["mixed_T42x256_T63x64"] = run_mixed(args.steps, args.warmup)
    if not args.no_t106:
        cmp = () if args.t106_compare == "none" else tuple(x for x in args.t106_compare.split(",") if x)
        out["t106_fwd_adjoint"] = run_t106_fwd_adjoint(steps=args.t106_steps, precision=args.precision, compare=cmp)
    if cpu is not None:
        out["cpu_baseline"] = cpu
        out["accuracy_T42_1member"] = accuracy_vs_oracle()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
