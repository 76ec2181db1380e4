"""Hand-written adjoint (GPU, FP64) vs central finite differences of the HIP forward and of the oracle.

Bar: gradients within 1e-6 relative in FP64.  Configuration: T42, two
members with different forcing parameters, 6 forced modified-Euler steps,
loss = mean((Phi_T - Phi_target)^2) per member, summed.
"""

from __future__ import annotations

import numpy as np
import pytest

from oracle import swampe_oracle as orc

pytestmark = pytest.mark.gpu

CFG = dict(M=42, nsteps=6, dt=600.0, a=6.37122e6, omega=7.292e-5, Phibar=3.0e5)
TAURAD = np.array([86400.0, 5.0e4])
TAUDRAG = np.array([86400.0, 2.0e5])
DPHIEQ = np.array([4.0e6, 3.0e6])


def _ic(M=42, seed=5):
    """Spectral eta/delta/Phi: solid-body Coriolis vorticity plus low-degree noise (m=0 real)."""
    rng = np.random.default_rng(seed)
    c = np.zeros((3, M + 1, M + 1), complex)
    c[0, 0, 1] = 2 * CFG["omega"] / np.sqrt(0.375) * 0.5  # eta: partially balanced vorticity
    for f, amp in ((0, 2e-6), (1, 5e-7), (2, 5e2)):
        for m in range(0, 9):
            for n in range(m, 9):
                c[f, m, n] += amp * (rng.standard_normal() + (1j * rng.standard_normal() if m else 0))
    return c


def _target(J=64, I=128):
    rng = np.random.default_rng(11)
    return 2e3 * rng.standard_normal((J, I))


def _loss(phi, tgt):
    return ((phi - tgt) ** 2).mean(dim=(1, 2)).sum()


def _forward(torch, c, tr, td, dp, **kw):
    from my_swampe_amd.autograd import rollout_phi

    with torch.no_grad():
        return rollout_phi(c, tr, td, dp, **CFG, **kw)


@pytest.fixture(scope="module")
def grads():
    import torch

    from my_swampe_amd.autograd import rollout_phi

    print("[gpu-case] adjoint grads", flush=True)
    dev = torch.device("cuda")
    c = torch.tensor(np.stack([_ic(seed=5), _ic(seed=6)]), device=dev, requires_grad=True)
    tr = torch.tensor(TAURAD, device=dev, requires_grad=True)
    td = torch.tensor(TAUDRAG, device=dev, requires_grad=True)
    dp = torch.tensor(DPHIEQ, device=dev, requires_grad=True)
    tgt = torch.tensor(_target(), device=dev)
    stats = {}
    phi = rollout_phi(c, tr, td, dp, **CFG, stats=stats)
    L = _loss(phi, tgt)
    L.backward()
    return dict(c=c.detach(), tgt=tgt, L=float(L.detach()), gc=c.grad.detach(), gtr=tr.grad.detach().cpu().numpy(),
                gtd=td.grad.detach().cpu().numpy(), gdp=dp.grad.detach().cpu().numpy(), stats=stats)


def _richardson(f, h):
    """Fourth-order central difference: [8(f(h) - f(-h)) - (f(2h) - f(-2h))] / (12 h)."""
    return (8.0 * (f(h) - f(-h)) - (f(2 * h) - f(-2 * h))) / (12.0 * h)


def test_parameter_gradients_vs_hip_central_differences(grads):
    import torch

    dev = torch.device("cuda")
    base = {"tr": TAURAD, "td": TAUDRAG, "dp": DPHIEQ}
    got = {"tr": grads["gtr"], "td": grads["gtd"], "dp": grads["gdp"]}
    for name in ("tr", "td", "dp"):
        for b in range(2):
            def f(dh):
                p = {k: v.copy() for k, v in base.items()}
                p[name][b] += dh
                phi = _forward(torch, grads["c"], *(torch.tensor(p[k], device=dev) for k in ("tr", "td", "dp")))
                return float(_loss(phi, grads["tgt"]))
            fd = _richardson(f, 1e-4 * abs(base[name][b]))
            rel = abs(got[name][b] - fd) / max(abs(fd), 1e-300)
            print(name, b, "adjoint", got[name][b], "fd", fd, "rel", rel)
            assert rel < 1e-6, (name, b, got[name][b], fd, rel)


K6S = np.array([1.24e33, 1.24e33])
K6PHIS = np.array([1.5e33, 0.9e33])


def _diffusion_grads(torch, grads, shared_phi, every=None):
    from my_swampe_amd.autograd import rollout_phi

    dev = torch.device("cuda")
    c = grads["c"].clone().requires_grad_(True)
    tr, td, dp = (torch.tensor(x, device=dev, requires_grad=True) for x in (TAURAD, TAUDRAG, DPHIEQ))
    k6 = torch.tensor(K6S, device=dev, requires_grad=True)
    k6p = None if shared_phi else torch.tensor(K6PHIS, device=dev, requires_grad=True)
    L = _loss(rollout_phi(c, tr, td, dp, **CFG, K6=k6, K6Phi=k6p, checkpoint_every=every), grads["tgt"])
    L.backward()
    return (c.grad, tr.grad, td.grad, dp.grad, k6.grad.cpu().numpy(),
            None if shared_phi else k6p.grad.cpu().numpy())


@pytest.mark.parametrize("shared_phi", [True, False])
def test_hyperdiffusion_gradients_vs_hip_central_differences(grads, shared_phi):
    """dL/dK6 and dL/dK6Phi per member (the reverse mode of the sigma6 / sigma6Phi factors) against
    fourth-order central differences of the HIP forward, within 1e-6 relative; asking for them
    leaves the other gradients bit for bit as they were, and every checkpoint layout (one tape,
    segments of 4 + 2, 5 + 1) gives the same K6 gradients bit for bit."""
    import torch

    print(f"[gpu-case] K6 gradients shared_phi={shared_phi}", flush=True)
    dev = torch.device("cuda")
    gc, gtr, gtd, gdp, gk6, gk6p = _diffusion_grads(torch, grads, shared_phi)
    if shared_phi:  # K6 = 1.24e33 is the fixture's default: the same rollout
        np.testing.assert_array_equal(gc.cpu().numpy(), grads["gc"].cpu().numpy())
        for g, name in ((gtr, "gtr"), (gtd, "gtd"), (gdp, "gdp")):
            np.testing.assert_array_equal(g.cpu().numpy(), grads[name])
    for every in (4, 5):
        _, _, _, _, k6e, k6pe = _diffusion_grads(torch, grads, shared_phi, every)
        np.testing.assert_array_equal(k6e, gk6)
        if not shared_phi:
            np.testing.assert_array_equal(k6pe, gk6p)
    cases = [("K6", K6S, gk6)] + ([] if shared_phi else [("K6Phi", K6PHIS, gk6p)])
    for name, base, got in cases:
        for b in range(2):
            def f(dh):
                kw = {"K6": torch.tensor(K6S, device=dev),
                      "K6Phi": None if shared_phi else torch.tensor(K6PHIS, device=dev)}
                v = base.copy()
                v[b] += dh
                kw[name] = torch.tensor(v, device=dev)
                phi = _forward(torch, grads["c"], *(torch.tensor(x, device=dev) for x in (TAURAD, TAUDRAG, DPHIEQ)),
                               **kw)
                return float(_loss(phi, grads["tgt"]))
            fd = _richardson(f, 1e-3 * base[b])
            rel = abs(got[b] - fd) / max(abs(fd), 1e-300)
            print(name, b, "adjoint", got[b], "fd", fd, "rel", rel)
            assert got[b] != 0.0 and rel < 1e-6, (name, b, got[b], fd, rel)


def test_hyperdiffusion_gradient_vs_oracle_central_differences(grads):
    """dL/dK6 (K6Phi following K6) against central differences of the FP64 NumPy oracle."""
    import torch

    N, I, J, lam, mus, w, P, H = orc.cached_geometry(42)
    c = grads["c"].cpu().numpy()
    tgt = grads["tgt"].cpu().numpy()
    got = _diffusion_grads(torch, grads, True)[4]
    for b in range(2):
        fields = [np.real(orc.invrs_fft(orc.invrs_leg(c[b, f], I, J, 42, 42, P), I)) for f in range(3)]

        def oracle_loss(dh):
            out = orc.run_model_scan(M=42, dt=CFG["dt"], tmax=2 + CFG["nsteps"], Phibar=CFG["Phibar"],
                                     omega=CFG["omega"], a=CFG["a"], taurad=TAURAD[b], taudrag=TAUDRAG[b],
                                     DPhieq=DPHIEQ[b], K6=K6S[b] + dh, eta0_init=fields[0], delta0_init=fields[1],
                                     Phi0_init=fields[2], diagnostics=False, return_history=False)
            return float(np.mean((out["last_state"].Phi_curr - tgt) ** 2))

        fd = _richardson(oracle_loss, 1e-3 * K6S[b])
        rel = abs(got[b] - fd) / max(abs(fd), 1e-300)
        print("oracle K6", b, got[b], fd, rel)
        assert rel < 1e-6, (b, got[b], fd, rel)


def test_fwd_grad_over_the_rollout_parameters(grads):
    """fwd_grad of the loss as a function of theta = (taurad, taudrag, DPhieq) of both members
    equals the adjoint's parameter gradients, which the central differences above pin."""
    import torch

    from my_swampe_amd import fwd_grad
    from my_swampe_amd.autograd import rollout_phi

    print("[gpu-case] fwd_grad over 6 rollout parameters", flush=True)
    theta = torch.tensor(np.concatenate([TAURAD, TAUDRAG, DPHIEQ]), device="cuda")
    loss = lambda t: _loss(rollout_phi(grads["c"], t[0:2], t[2:4], t[4:6], **CFG), grads["tgt"])  # noqa: E731
    g = fwd_grad(loss, theta, chunk=2).cpu().numpy()
    want = np.concatenate([grads["gtr"], grads["gtd"], grads["gdp"]])
    np.testing.assert_allclose(g, want, rtol=1e-12, atol=0)


def test_initial_state_gradient_vs_directional_difference(grads):
    import torch

    dev = torch.device("cuda")
    rng = np.random.default_rng(21)
    for trial in range(3):
        v = np.zeros((2, 3, 43, 43), complex)
        for f, amp in ((0, 1e-6), (1, 1e-7), (2, 1e2)):
            v[:, f] = amp * (rng.standard_normal((2, 43, 43)) + 1j * rng.standard_normal((2, 43, 43)))
        v[:, :, 0, :] = v[:, :, 0, :].real
        v *= np.triu(np.ones((43, 43)))[None, None]
        vt = torch.tensor(v, device=dev)
        h = 1e-5
        tr, td, dp = (torch.tensor(x, device=dev) for x in (TAURAD, TAUDRAG, DPHIEQ))
        Lp = float(_loss(_forward(torch, grads["c"] + h * vt, tr, td, dp), grads["tgt"]))
        Lm = float(_loss(_forward(torch, grads["c"] - h * vt, tr, td, dp), grads["tgt"]))
        fd = (Lp - Lm) / (2 * h)
        g = grads["gc"]
        ad = float((g.real * vt.real + g.imag * vt.imag).sum())
        rel = abs(ad - fd) / abs(fd)
        print("direction", trial, "adjoint", ad, "fd", fd, "rel", rel)
        assert rel < 1e-6, (trial, ad, fd, rel)


def test_parameter_gradients_vs_oracle_central_differences(grads):
    """Same gradients against finite differences of the FP64 NumPy oracle run from the grid fields of c."""
    N, I, J, lam, mus, w, P, H = orc.cached_geometry(42)
    c = grads["c"].cpu().numpy()
    tgt = grads["tgt"].cpu().numpy()

    def oracle_loss(b, tr, td, dp):
        fields = [np.real(orc.invrs_fft(orc.invrs_leg(c[b, f], I, J, 42, 42, P), I)) for f in range(3)]
        out = orc.run_model_scan(M=42, dt=CFG["dt"], tmax=2 + CFG["nsteps"], Phibar=CFG["Phibar"],
                                 omega=CFG["omega"], a=CFG["a"], taurad=tr, taudrag=td, DPhieq=dp,
                                 eta0_init=fields[0], delta0_init=fields[1], Phi0_init=fields[2],
                                 diagnostics=False, return_history=False)
        return float(np.mean((out["last_state"].Phi_curr - tgt) ** 2))

    got = {"tr": grads["gtr"], "td": grads["gtd"], "dp": grads["gdp"]}
    for b in range(2):
        base = {"tr": TAURAD[b], "td": TAUDRAG[b], "dp": DPHIEQ[b]}
        for name in ("tr", "td", "dp"):
            fd = _richardson(lambda dh: oracle_loss(b, **{k: v + (dh if k == name else 0.0) for k, v in base.items()}),
                             1e-4 * abs(base[name]))
            rel = abs(got[name][b] - fd) / max(abs(fd), 1e-300)
            print("oracle", name, b, got[name][b], fd, rel)
            assert rel < 1e-6, (name, b, got[name][b], fd, rel)


def test_checkpoint_interval_does_not_change_gradients(grads):
    import torch

    from my_swampe_amd.autograd import rollout_phi

    dev = torch.device("cuda")
    for every, nck, seg in ((1, 6, 1), (3, 2, 3), (4, 2, 4), (6, 1, 6)):
        c = grads["c"].clone().requires_grad_(True)
        tr = torch.tensor(TAURAD, device=dev, requires_grad=True)
        td = torch.tensor(TAUDRAG, device=dev)
        dp = torch.tensor(DPHIEQ, device=dev)
        st = {}
        L = _loss(rollout_phi(c, tr, td, dp, **CFG, checkpoint_every=every, stats=st), grads["tgt"])
        L.backward()
        np.testing.assert_array_equal(c.grad.cpu().numpy(), grads["gc"].cpu().numpy())
        np.testing.assert_array_equal(tr.grad.cpu().numpy(), grads["gtr"])
        assert st["checkpoints"] == nck and st["segment_states"] == seg, (every, st)
    # Default policy: the 6-step trajectory fits the checkpoint budget, so every state is kept.
    assert grads["stats"]["checkpoints"] == 6 and grads["stats"]["segment_states"] == 1


@pytest.mark.parametrize("kernel,M", [("grid", 42), ("grid", 63), ("grid", 106), ("grid_full", 42), ("grid_full", 63),
                                      ("leg", 42), ("leg", 63), ("leg", 106)])
def test_mfma_adjoint_matches_plain_fp32(kernel, M, torch_gpu):
    """The matrix-core grid (or Legendre) adjoint kernels vs the plain FP32 ones, selected with
    SWX_OPT_GRID_KERNEL (or SWX_OPT_LEG_KERNEL, which switches the forward and reverse Legendre
    kernels together) on the same reverse sweep; "grid" runs the half-period folded grid kernels,
    "grid_full" the full-DFT ones (SWX_OPT_GRID_FOLD; at T106 only the folded two-pass grid adjoint
    fits the LDS): 3 members with different forcing, 4 steps,
    random terminal Phi cotangent.  Relative L2 of dL/dS_0 and of the parameter gradients
    <= 1e-5 (FP32 transforms in both, different summation orders)."""
    from my_swampe_amd import _capi
    from my_swampe_amd.model import param_row
    from my_swampe_amd.plan import Plan

    torch = torch_gpu
    print(f"[gpu-case] mfma-adjoint kernel={kernel} M={M}", flush=True)
    option = _capi.OPT_LEG_KERNEL if kernel == "leg" else _capi.OPT_GRID_KERNEL
    B, nsteps = 3, 4
    rng = np.random.default_rng(M)
    params = np.stack([param_row(dt=120.0, a=8.2e7, omega=3.2e-5, g=9.8, Phibar=3e5, taurad=10 ** rng.uniform(4.5, 5.5),
                                 taudrag=10 ** rng.uniform(4.5, 5.5), DPhieq=rng.uniform(1e5, 5e5), K6=1.24e33,
                                 K6Phi=None, alpha=0.01) for _ in range(B)])
    c = np.stack([_ic(M, seed=30 + b) for b in range(B)])
    J, I = _capi.grid_size(M)
    g = torch.tensor(1e-3 * rng.standard_normal((B, J, I)), device="cuda")
    out = {}
    for kind, opt in (("plain", 1), ("mfma", 0)):
        plan = Plan(M, params, precision="fp32")
        plan.set_option(option, opt)
        plan.set_option(_capi.OPT_GRID_FOLD, 1 if kernel == "grid_full" else 2)
        plan.init_spectral(torch.tensor(c, device="cuda"))
        states = []
        for k in range(nsteps):
            states.append(plan.state_get())
            plan.run(2 + k, 1)
        plan.adjoint_begin()
        plan.adjoint_terminal(g_Phi=g)
        for k in reversed(range(nsteps)):
            plan.adjoint_step(states[k])
        lam, pg = plan.adjoint_result()
        torch.cuda.synchronize()
        out[kind] = (lam.cpu().numpy(), pg.cpu().numpy())
        plan.close()
    rl = float(np.linalg.norm(out["mfma"][0] - out["plain"][0]) / np.linalg.norm(out["plain"][0]))
    rp = float(np.linalg.norm(out["mfma"][1] - out["plain"][1]) / np.linalg.norm(out["plain"][1]))
    print(kernel, M, "dL/dS0 relL2", rl, "param grad relL2", rp)
    assert np.isfinite(out["mfma"][0]).all() and np.isfinite(out["mfma"][1]).all()
    assert rl <= 1e-5 and rp <= 1e-5, (rl, rp)


def test_time_series_rollout_gradient_is_the_sum_of_terminal_gradients(torch_gpu):
    """rollout_phi_series (observation cotangents added mid-sweep by swx_adjoint_observe) vs one
    rollout per observed step: the observed fields are identical, and the gradient of a loss
    summed over the series equals the sum of the single-observation gradients, K6 and K6Phi
    included (1e-10 relative, FP64; the terminal gradients are the ones pinned by finite
    differences above)."""
    from my_swampe_amd.autograd import rollout_phi, rollout_phi_series

    torch = torch_gpu
    print("[gpu-case] adjoint time series", flush=True)
    dev = torch.device("cuda")
    obs = (0, 2, 5, 8)
    rng = np.random.default_rng(21)
    tgts = torch.tensor(2e3 * rng.standard_normal((len(obs), 64, 128)), device=dev)
    c0 = torch.tensor(np.stack([_ic(seed=5), _ic(seed=6)]), device=dev)
    cfg = {k: v for k, v in CFG.items() if k != "nsteps"}

    def leaves():
        return (c0.clone().requires_grad_(True),) + tuple(
            torch.tensor(x, device=dev, requires_grad=True) for x in (TAURAD, TAUDRAG, DPHIEQ, K6S, K6PHIS))

    x = leaves()
    phis = rollout_phi_series(*x[:4], K6=x[4], K6Phi=x[5], obs_steps=obs, checkpoint_every=3, **cfg)
    sum(_loss(phis[:, k], tgts[k]) for k in range(len(obs))).backward()
    series = [t.grad.detach().clone() for t in x]

    total = [torch.zeros_like(t) for t in series]
    for k, n in enumerate(obs):
        y = leaves()
        if n > 0:
            phi = rollout_phi(*y[:4], K6=y[4], K6Phi=y[5], nsteps=n, checkpoint_every=3, **cfg)
        else:  # the initial field: a zero-step series
            phi = rollout_phi_series(*y[:4], K6=y[4], K6Phi=y[5], obs_steps=(0,), **cfg)[:, 0]
        assert torch.equal(phi.detach(), phis[:, k].detach()), n
        _loss(phi, tgts[k]).backward()
        for acc, t in zip(total, y):
            if t.grad is not None:
                acc += t.grad
    assert np.abs(series[4].cpu().numpy()).min() > 0 and np.abs(series[5].cpu().numpy()).min() > 0
    for i, (a, b) in enumerate(zip(series, total)):
        a, b = a.cpu().numpy(), b.cpu().numpy()
        np.testing.assert_allclose(a, b, rtol=0, atol=1e-10 * np.abs(b).max(), err_msg=str(i))

