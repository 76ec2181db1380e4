"""swx_export_state_noalloc (GPU): the export that allocates nothing, its rebuild rows in the
library's static device scratch, against swx_export_state on identical plans; two calls in a row,
the second with a smaller plan (fewer rows than the scratch holds from the first), and the size
limit refused on the host."""

from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FIELDS = ("eta_prev", "eta_curr", "delta_prev", "delta_curr", "Phi_prev", "Phi_curr", "U_curr", "V_curr",
          "etam_curr", "Phim_prev", "Um_curr")


def _plan(M, B, precision, seed):
    import torch

    from my_swampe_amd import _capi
    from my_swampe_amd.model import param_row
    from my_swampe_amd.plan import Plan

    rng = np.random.default_rng(seed)
    params = np.stack([param_row(dt=120.0, a=8.2e7, omega=3.2e-5, g=9.8, Phibar=3e5, taurad=10 ** rng.uniform(4, 6),
                                 taudrag=10 ** rng.uniform(4, 6), DPhieq=rng.uniform(1e5, 5e5), K6=1.24e33,
                                 K6Phi=None, alpha=0.01) for _ in range(B)])
    c = np.zeros((B, 3, M + 1, M + 1), complex)
    c[:, 0, 0, 1] = 3.2e-5 / np.sqrt(0.375)
    for m in range(9):
        for n in range(m, 9):
            c[:, 2, m, n] = 30.0 * (rng.standard_normal(B) + (1j * rng.standard_normal(B) if m else 0))
    plan = Plan(M, params, precision=precision)
    if precision == "fp32":
        plan.set_option(_capi.OPT_GRID_FOLD, 2)  # the same grid kernel in both plans of a pair
    plan.init_spectral(torch.tensor(c, device="cuda"))
    torch.cuda.synchronize()  # the caller steps on a stream of its own
    return plan


@pytest.mark.parametrize("first,second", [((42, 96, "fp32"), (42, 5, "fp32")), ((63, 40, "fp64"), (42, 3, "fp64"))])
def test_noalloc_export_matches_export(first, second, torch_gpu):
    torch = torch_gpu
    for M, B, prec in (first, second):
        print(f"[gpu-case] noalloc export M={M} B={B} {prec}", flush=True)
        out = {}
        for noalloc in (False, True):
            plan = _plan(M, B, prec, seed=M + B)
            s = torch.cuda.Stream()
            with torch.cuda.stream(s):
                plan.run_graph(2, 40, graph_steps=8, stream=s)  # the lineage carry: a rebuild on export
                assert plan.step_counts()["lineage"] > 0
                out[noalloc] = {k: v.cpu().numpy() for k, v in plan.export(fields=FIELDS, stream=s,
                                                                            noalloc=noalloc).items()}
            s.synchronize()
            plan.close()
        for k in out[False]:
            np.testing.assert_array_equal(out[True][k], out[False][k], err_msg=f"{M} {B} {prec} {k}")


def test_noalloc_export_refuses_plans_beyond_the_scratch(torch_gpu):
    plan = _plan(106, 49, "fp64", seed=1)  # 49 * 1,369,600 bytes of rows > 64 MiB
    with pytest.raises(ValueError, match="scratch"):
        plan.export(fields=("Phi_curr",), noalloc=True)
    ok = plan.export(fields=("Phi_curr",))
    assert bool(torch_gpu.isfinite(ok["Phi_curr"]).all())
    plan.close()
