"""The mixed-truncation ensemble at full size (GPU): 256 members at T42 and 64 at T63 stepped in one
Legendre and one grid launch per step (Ensemble), FP32 transforms, per-member forcing (taurad,
taudrag log-uniform in [1e4, 1e6] s, DPhieq uniform in [1e5, 5e5]), dt = 120 s, rest + seeded noise
initial states.  Sampled members of each truncation against the FP64 oracle over 12 steps (relative
L2 <= 5e-6 on eta and Phi, the FP32-transform bound of the other oracle comparisons; 2e-5 on delta,
which starts from zero in the rest state and after 12 steps is still small beside the tendencies
whose FP32 rounding it collects), then
1000 steps in all with every member finite (after unit_tests/test_vmap_smoke.py:84-93: a batched
member equals its direct run)."""

from __future__ import annotations

import numpy as np
import pytest

from oracle import swampe_oracle as orc

pytestmark = pytest.mark.gpu

SAMPLES = {42: (0, 97, 255), 63: (0, 31, 63)}


def _params(B, seed):
    from my_swampe_amd.model import param_row

    rng = np.random.default_rng(seed)
    tr = 10.0 ** rng.uniform(4.0, 6.0, B)
    td = 10.0 ** rng.uniform(4.0, 6.0, B)
    dp = rng.uniform(1e5, 5e5, B)
    return np.stack([param_row(dt=120.0, a=8.2e7, omega=3.2e-5, g=9.8, Phibar=3e5, taurad=tr[b], taudrag=td[b],
                               DPhieq=dp[b], K6=1.24e33, K6Phi=None, alpha=0.01) for b in range(B)])


def _ics(M, B, seed):
    rng = np.random.default_rng(seed)
    c = np.zeros((B, 3, M + 1, M + 1), complex)
    c[:, 0, 0, 1] = 3.2e-5 / np.sqrt(0.375)
    for m in range(11):
        for n in range(m, 11):
            c[:, 2, m, n] = 30.0 * (rng.standard_normal(B) + (1j * rng.standard_normal(B) if m else 0))
    return c


def _oracle_final(M, c, p, nsteps):
    N, I, J, lam, mus, w, P, H = orc.cached_geometry(M)
    fields = [np.real(orc.invrs_fft(orc.invrs_leg(c[f], I, J, M, N, P), I)) for f in range(3)]
    out = orc.run_model_scan(M=M, dt=120.0, tmax=2 + nsteps, Phibar=3e5, omega=3.2e-5, a=8.2e7, taurad=p[5],
                             taudrag=p[6], DPhieq=p[7], eta0_init=fields[0], delta0_init=fields[1],
                             Phi0_init=fields[2], diagnostics=False, return_history=False)
    return out["last_state"]


def test_full_size_one_launch_ensemble(torch_gpu):
    from my_swampe_amd.plan import Ensemble, Plan

    torch = torch_gpu
    print("[gpu-case] ensemble 256xT42 + 64xT63", flush=True)
    sizes = {42: 256, 63: 64}
    params = {42: _params(256, 7), 63: _params(64, 8)}
    ics = {42: _ics(42, 256, 1000), 63: _ics(63, 64, 2000)}
    plans = {M: Plan(M, params[M], precision="fp32") for M in sizes}
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        for M, pl in plans.items():
            pl.init_spectral(torch.tensor(ics[M], device="cuda"), stream=s)
        ens = Ensemble(plans[42], plans[63])
        ens.run_graph(2, 12, graph_steps=32, stream=s)
        ex = {M: pl.export(fields=("eta_curr", "delta_curr", "Phi_curr"), stream=s) for M, pl in plans.items()}
    s.synchronize()
    for M, idx in SAMPLES.items():
        for b in idx:
            ref = _oracle_final(M, ics[M][b], params[M][b], 12)
            for f in ("eta", "delta", "Phi"):
                g = ex[M][f + "_curr"][b].cpu().numpy()
                r = getattr(ref, f + "_curr")
                rel = float(np.linalg.norm(g - r) / np.linalg.norm(r))
                print(M, b, f, "relL2 vs oracle", rel)
                assert rel <= (2e-5 if f == "delta" else 5e-6), (M, b, f, rel)
    counts = {M: pl.step_counts() for M, pl in plans.items()}
    with torch.cuda.stream(s):
        ens.run_graph(14, 988, graph_steps=32, stream=s)
        ex = {M: pl.export(fields=("Phi_curr", "U_curr"), stream=s) for M, pl in plans.items()}
    s.synchronize()
    for M, e in ex.items():
        for k, v in e.items():
            if k != "dead":
                assert bool(torch.isfinite(v).all()), (M, k)
        done = plans[M].step_counts()
        assert done["graph"] - counts[M]["graph"] >= 900, done  # the long call replays captured steps
    ens.close()
    for pl in plans.values():
        pl.close()
