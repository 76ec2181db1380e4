"""Phase-curve observable on the GPU (swx_phase_curve, swx_phase_curve_vjp, swx_plan_phase_curve,
rollout_phase_curve) against the oracle's projection (unit_tests/test_parity_reference_regression.py:67-82)
and its pinned vector-Jacobian product.

Bars (FP64): the projection and its VJP within 1e-12 of the oracle relative to the largest
value; the fixtures' phase curves at the reference's 1e-10; the rollout's curves identical to
the projection of the rollout's fields; its gradients equal to those through the grid series
(1e-10 relative) and to central differences (1e-6 relative).
"""

from __future__ import annotations

import numpy as np
import pytest

from conftest import FIXTURE_CASES, GOLDEN
from oracle import swampe_oracle as orc

pytestmark = pytest.mark.gpu

GRIDS = {42: (64, 128), 63: (96, 192), 106: (160, 320)}


def _grids(n, J, I, seed):
    rng = np.random.default_rng(seed)
    lam = np.linspace(-np.pi, np.pi, I, endpoint=False)
    # A positive day-side bulge plus noise: realistic curves with a well-defined normalization.
    base = 3e5 + 1e5 * np.maximum(np.cos(lam), 0)[None, :]
    return base[None] + 2e4 * rng.standard_normal((n, J, I))


@pytest.mark.parametrize("normalize", [True, False])
@pytest.mark.parametrize("M,K", [(42, 1), (42, 37), (42, 2048), (63, 64), (106, 50)])
def test_projection_vs_oracle(torch_gpu, M, K, normalize):
    from my_swampe_amd.observables import phase_curve

    torch = torch_gpu
    J, I = GRIDS[M]
    print(f"[gpu-case] phase curve M={M} K={K} normalize={normalize}", flush=True)
    phi = _grids(3, J, I, seed=M + K)
    phases = np.random.default_rng(K).uniform(-np.pi, np.pi, K)
    got = phase_curve(torch.tensor(phi, device="cuda"), phases, normalize=normalize).cpu().numpy()
    for b in range(3):
        ref = orc.phase_curve(phi[b], phases, normalize=normalize)
        np.testing.assert_allclose(got[b], ref, rtol=0, atol=1e-12 * np.abs(ref).max(), err_msg=str(b))


@pytest.mark.parametrize("normalize", [True, False])
@pytest.mark.parametrize("M,K", [(42, 1), (42, 40), (106, 9)])
def test_vjp_vs_oracle(torch_gpu, M, K, normalize):
    from my_swampe_amd.observables import phase_curve

    torch = torch_gpu
    J, I = GRIDS[M]
    print(f"[gpu-case] phase curve vjp M={M} K={K} normalize={normalize}", flush=True)
    phi = _grids(2, J, I, seed=7 + K)
    rng = np.random.default_rng(K + 1)
    phases = rng.uniform(-np.pi, np.pi, K)
    g = rng.standard_normal((2, K))
    x = torch.tensor(phi, device="cuda", requires_grad=True)
    f = phase_curve(x, phases, normalize=normalize)
    (f * torch.tensor(g, device="cuda")).sum().backward()
    got = x.grad.cpu().numpy()
    for b in range(2):
        _, raw = orc.phase_curve(phi[b], phases, normalize=normalize, raw=True)
        ref = orc.phase_curve_vjp(g[b], raw, J, I, phases, normalize=normalize)
        np.testing.assert_allclose(got[b], ref, rtol=0, atol=1e-12 * np.abs(ref).max(), err_msg=str(b))


@pytest.mark.parametrize("case", sorted(FIXTURE_CASES))
def test_fixture_phase_curve_through_the_kernel(torch_gpu, case):
    from my_swampe_amd.model import run_model_scan
    from my_swampe_amd.observables import phase_curve

    print(f"[gpu-case] phase curve fixture {case}", flush=True)
    ref = np.load(GOLDEN / f"{case}.npz")
    out = run_model_scan(**FIXTURE_CASES[case], return_history=False)
    curve = phase_curve(out["last_state"].Phi_curr, ref["phase_angles"]).cpu().numpy()
    np.testing.assert_allclose(curve, ref["phase_curve"], rtol=0.0, atol=1.0e-10)


def test_plan_projection_reads_the_exported_grid(torch_gpu):
    from my_swampe_amd.model import param_row
    from my_swampe_amd.observables import phase_curve
    from my_swampe_amd.plan import Plan

    torch = torch_gpu
    print("[gpu-case] plan phase curve", flush=True)
    params = np.stack([param_row(dt=600.0, a=6.37122e6, omega=7.292e-5, g=9.8, Phibar=3e5, taurad=tr, taudrag=8e4,
                                 DPhieq=4e6, K6=1.24e33, K6Phi=None, alpha=0.01) for tr in (5e4, 9e4, 2e5)])
    plan = Plan(42, params)
    eta0, delta0, Phi0, U0, V0 = orc.rest_noise_ic(42, seed=3, omega=7.292e-5, phi_std=2e3)
    plan.init(eta0, delta0, Phi0, U0, V0)
    plan.run(2, 5)
    phases = np.linspace(-np.pi, np.pi, 24, endpoint=False)
    for normalize in (True, False):
        flux, raw = plan.phase_curve(phases, normalize)
        phi = plan.export(fields=("Phi_curr",))["Phi_curr"]
        want = phase_curve(phi, phases, normalize=normalize, mus=plan.mus_device())
        assert torch.equal(flux, want)
        assert torch.equal(raw, phase_curve(phi, phases, normalize=False, mus=plan.mus_device()))
    plan.close()


def test_argument_errors(torch_gpu):
    from my_swampe_amd import _capi
    from my_swampe_amd.observables import phase_curve, phase_curve_raw

    torch = torch_gpu
    print("[gpu-case] phase curve errors", flush=True)
    phi = torch.zeros((2, 64, 128), dtype=torch.float64, device="cuda")
    with pytest.raises(ValueError):
        phase_curve(phi, [])
    with pytest.raises(ValueError, match="phases"):
        phase_curve(phi, np.zeros(2049))
    with pytest.raises(TypeError):
        phase_curve(phi.float(), [0.0])
    empty = torch.zeros((0, 64, 128), dtype=torch.float64, device="cuda")
    flux, raw = phase_curve_raw(empty, torch.zeros(3, dtype=torch.float64, device="cuda"))
    assert flux.shape == (0, 3)
    # An all-zero grid has zero mean |flux|: the reference returns the raw (zero) curve.
    assert torch.count_nonzero(phase_curve(phi, [0.0, 1.0])) == 0
    assert _capi.SWX_ERR_ARG == -1


# ---- phase curve inside the differentiable rollout ------------------------------------------

CFG = dict(M=42, dt=600.0, a=6.37122e6, omega=7.292e-5, Phibar=3.0e5)
TAURAD = np.array([86400.0, 5.0e4])
TAUDRAG = np.array([86400.0, 2.0e5])
DPHIEQ = np.array([4.0e6, 3.0e6])
OBS = (0, 2, 5, 8)


def _ic(M=42, seed=5):
    rng = np.random.default_rng(seed)
    c = np.zeros((3, M + 1, M + 1), complex)
    c[0, 0, 1] = 2 * CFG["omega"] / np.sqrt(0.375) * 0.5
    for f, amp in ((0, 2e-6), (1, 5e-7), (2, 5e2)):
        for m in range(0, 9):
            for n in range(m, 9):
                c[f, m, n] += amp * (rng.standard_normal() + (1j * rng.standard_normal() if m else 0))
    return c


def _phases():
    # One orbital phase per observation plus two neighbours: (n_obs, 3).
    base = np.linspace(-2.5, 2.0, len(OBS))
    return np.stack([base - 0.3, base, base + 0.4], axis=1)


def _data():
    rng = np.random.default_rng(31)
    return 1.0 + 0.1 * rng.standard_normal((2, len(OBS), 3))


def _chi2(flux, data):
    return (((flux - data) / 0.05) ** 2).sum()


def test_rollout_phase_curve_matches_the_grid_series(torch_gpu):
    """The rollout's curves equal the projection of rollout_phi_series's grids bit for bit, and the
    gradient of a chi-square of the curves equals the gradient through the grid series."""
    from my_swampe_amd import _capi
    from my_swampe_amd.autograd import rollout_phase_curve, rollout_phi_series
    from my_swampe_amd.observables import phase_curve

    torch = torch_gpu
    print("[gpu-case] rollout phase curve vs grid series", flush=True)
    dev = torch.device("cuda")
    c0 = torch.tensor(np.stack([_ic(seed=5), _ic(seed=6)]), device=dev)
    mus = _capi.reference_quadrature(64)[0]
    ph = torch.tensor(_phases(), device=dev)
    data = torch.tensor(_data(), device=dev)

    def leaves():
        return (c0.clone().requires_grad_(True),) + tuple(
            torch.tensor(x, device=dev, requires_grad=True) for x in (TAURAD, TAUDRAG, DPHIEQ))

    x = leaves()
    flux = rollout_phase_curve(*x, obs_steps=OBS, phase_angles=ph, checkpoint_every=3, **CFG)
    _chi2(flux, data).backward()
    fused = [t.grad.detach().clone() for t in x]

    y = leaves()
    phis = rollout_phi_series(*y, obs_steps=OBS, checkpoint_every=3, **CFG)
    curves = torch.stack([phase_curve(phis[:, k], ph[k], mus=mus) for k in range(len(OBS))], dim=1)
    assert torch.equal(flux.detach(), curves.detach())
    _chi2(curves, data).backward()
    for i, (a, t) in enumerate(zip(fused, y)):
        a, b = a.cpu().numpy(), t.grad.detach().cpu().numpy()
        np.testing.assert_allclose(a, b, rtol=0, atol=1e-10 * np.abs(b).max(), err_msg=str(i))


def test_rollout_phase_curve_parameter_gradients_vs_central_differences(torch_gpu):
    from my_swampe_amd.autograd import rollout_phase_curve

    torch = torch_gpu
    print("[gpu-case] rollout phase curve finite differences", flush=True)
    dev = torch.device("cuda")
    c = torch.tensor(np.stack([_ic(seed=5), _ic(seed=6)]), device=dev)
    ph = torch.tensor(_phases(), device=dev)
    data = torch.tensor(_data(), device=dev)
    tr = torch.tensor(TAURAD, device=dev, requires_grad=True)
    td = torch.tensor(TAUDRAG, device=dev, requires_grad=True)
    dp = torch.tensor(DPHIEQ, device=dev, requires_grad=True)
    _chi2(rollout_phase_curve(c, tr, td, dp, obs_steps=OBS, phase_angles=ph, **CFG), data).backward()
    grads = {"taurad": tr.grad.cpu().numpy(), "taudrag": td.grad.cpu().numpy(), "DPhieq": dp.grad.cpu().numpy()}
    base = {"taurad": TAURAD, "taudrag": TAUDRAG, "DPhieq": DPHIEQ}

    def L(name, b, h):
        vals = {k: v.copy() for k, v in base.items()}
        vals[name][b] += h
        with torch.no_grad():
            f = rollout_phase_curve(c, *(torch.tensor(vals[k], device=dev) for k in ("taurad", "taudrag", "DPhieq")),
                                    obs_steps=OBS, phase_angles=ph, **CFG)
        return float(_chi2(f, data))

    for name in ("taurad", "taudrag", "DPhieq"):
        for b in range(2):
            h = 1e-5 * base[name][b]
            fd = (8.0 * (L(name, b, h) - L(name, b, -h)) - (L(name, b, 2 * h) - L(name, b, -2 * h))) / (12.0 * h)
            an = grads[name][b]
            assert abs(fd - an) <= 1e-6 * max(abs(an), 1e-30) + 1e-9 * abs(fd), (name, b, fd, an)


def test_rollout_phase_curve_hyperdiffusion_gradient_vs_central_differences(torch_gpu):
    """dL/dK6 of the phase-curve chi^2 for a K6 shared by both members (a 0-d tensor: the members'
    partials summed; K6Phi follows K6) against fourth-order central differences, within 1e-6."""
    from my_swampe_amd.autograd import rollout_phase_curve

    torch = torch_gpu
    print("[gpu-case] rollout phase curve K6 finite differences", flush=True)
    dev = torch.device("cuda")
    c = torch.tensor(np.stack([_ic(seed=5), _ic(seed=6)]), device=dev)
    ph = torch.tensor(_phases(), device=dev)
    data = torch.tensor(_data(), device=dev)
    params = [torch.tensor(x, device=dev) for x in (TAURAD, TAUDRAG, DPHIEQ)]
    K6 = 1.24e33
    k6 = torch.tensor(K6, dtype=torch.float64, device=dev, requires_grad=True)
    _chi2(rollout_phase_curve(c, *params, K6=k6, obs_steps=OBS, phase_angles=ph, checkpoint_every=2, **CFG),
          data).backward()
    an = float(k6.grad)

    def L(h):
        with torch.no_grad():
            f = rollout_phase_curve(c, *params, K6=K6 + h, obs_steps=OBS, phase_angles=ph, **CFG)
        return float(_chi2(f, data))

    h = 1e-3 * K6
    fd = (8.0 * (L(h) - L(-h)) - (L(2 * h) - L(-2 * h))) / (12.0 * h)
    print("phase-curve dchi2/dK6 adjoint", an, "fd", fd)
    assert an != 0.0 and abs(fd - an) <= 1e-6 * abs(an), (fd, an)
