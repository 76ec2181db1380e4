"""Hand-written adjoint at T106 (GPU): the workload of the T106 x 64 forward + adjoint benchmark
(dt = 60 s, checkpointed sweep) at test size and at full batch.

* FP64, 2 members with different forcing, 8 steps, a checkpoint every 3 steps (the reverse sweep
  recomputes its segments): parameter gradients and an initial-state direction against
  fourth-order central differences of the HIP forward, and the parameter gradients of one member
  against fourth-order central differences of the FP64 oracle run from the same grids.  Bar 1e-6
  relative (unit_tests/test_autodiff.py:102-129 asks 1e-4 of its finite differences).
* FP32 transforms at the full batch of 64 members, 700 steps with a checkpoint every 70: the
  gradients within 1e-4 (relative L2) of the FP64 run's.
"""

from __future__ import annotations

import numpy as np
import pytest

from oracle import swampe_oracle as orc

pytestmark = pytest.mark.gpu

M = 106
J, I = 160, 320
CFG = dict(M=M, dt=60.0, a=8.2e7, omega=3.2e-5, Phibar=3.0e5)
TAURAD = np.array([86400.0, 3.0e4])
TAUDRAG = np.array([86400.0, 2.0e5])
DPHIEQ = np.array([3.0e5, 4.5e5])
NSTEPS, EVERY = 8, 3


def _ic(seed, nmax=10):
    """Spectral eta/delta/Phi: planetary vorticity plus low-degree noise on all three fields."""
    rng = np.random.default_rng(seed)
    c = np.zeros((3, M + 1, M + 1), complex)
    c[0, 0, 1] = CFG["omega"] / np.sqrt(0.375)
    for f, amp in ((0, 2e-6), (1, 5e-7), (2, 40.0)):
        for m in range(nmax + 1):
            for n in range(m, nmax + 1):
                c[f, m, n] += amp * (rng.standard_normal() + (1j * rng.standard_normal() if m else 0))
    return c


def _target():
    return 1e3 * np.random.default_rng(17).standard_normal((J, I))


def _loss(phi, tgt):
    return ((phi - tgt) ** 2).mean(dim=(1, 2)).sum()


def _richardson(f, h):
    """Fourth-order central difference: [8(f(h) - f(-h)) - (f(2h) - f(-2h))] / (12 h)."""
    return (8.0 * (f(h) - f(-h)) - (f(2 * h) - f(-2 * h))) / (12.0 * h)


@pytest.fixture(scope="module")
def t106():
    import torch

    from my_swampe_amd.autograd import rollout_phi

    print("[gpu-case] t106 adjoint grads fp64", flush=True)
    dev = torch.device("cuda")
    c = torch.tensor(np.stack([_ic(51), _ic(52)]), device=dev, requires_grad=True)
    tr, td, dp = (torch.tensor(x, device=dev, requires_grad=True) for x in (TAURAD, TAUDRAG, DPHIEQ))
    tgt = torch.tensor(_target(), device=dev)
    stats = {}
    L = _loss(rollout_phi(c, tr, td, dp, nsteps=NSTEPS, checkpoint_every=EVERY, stats=stats, **CFG), tgt)
    L.backward()
    assert stats["checkpoints"] == 3 and stats["segment_states"] == 3, stats

    def loss_at(cc, p):
        with torch.no_grad():
            phi = rollout_phi(cc, *(torch.tensor(p[k], device=dev) for k in ("tr", "td", "dp")), nsteps=NSTEPS,
                              **CFG)
        return float(_loss(phi, tgt))

    return dict(c=c.detach(), tgt=tgt, loss_at=loss_at, gc=c.grad.detach(),
                g={"tr": tr.grad.cpu().numpy(), "td": td.grad.cpu().numpy(), "dp": dp.grad.cpu().numpy()})


def test_t106_parameter_gradients_vs_hip_central_differences(t106):
    base = {"tr": TAURAD, "td": TAUDRAG, "dp": DPHIEQ}
    for name in ("tr", "td", "dp"):
        for b in range(2):
            def f(dh):
                p = {k: v.copy() for k, v in base.items()}
                p[name][b] += dh
                return t106["loss_at"](t106["c"], p)
            fd = _richardson(f, 1e-4 * abs(base[name][b]))
            got = t106["g"][name][b]
            rel = abs(got - fd) / max(abs(fd), 1e-300)
            print(name, b, "adjoint", got, "fd", fd, "rel", rel)
            assert rel < 1e-6, (name, b, got, fd, rel)


def test_t106_initial_state_direction_vs_hip_central_differences(t106):
    import torch

    rng = np.random.default_rng(23)
    v = np.zeros((2, 3, M + 1, M + 1), complex)
    for f, amp in ((0, 1e-6), (1, 1e-7), (2, 1e1)):
        v[:, f] = amp * (rng.standard_normal((2, M + 1, M + 1)) + 1j * rng.standard_normal((2, M + 1, M + 1)))
    v[:, :, 0, :] = v[:, :, 0, :].real
    v *= np.triu(np.ones((M + 1, M + 1)))[None, None]
    vt = torch.tensor(v, device="cuda")
    p = {"tr": TAURAD, "td": TAUDRAG, "dp": DPHIEQ}
    fd = _richardson(lambda h: t106["loss_at"](t106["c"] + h * vt, p), 1e-3)
    g = t106["gc"]
    ad = float((g.real * vt.real + g.imag * vt.imag).sum())
    rel = abs(ad - fd) / abs(fd)
    print("direction adjoint", ad, "fd", fd, "rel", rel)
    assert rel < 1e-6, (ad, fd, rel)


def test_t106_parameter_gradients_vs_oracle_central_differences(t106):
    """Member 0 against the FP64 oracle run from the grid fields of its spectral state (the HIP
    run starts from the spectral state itself: the two start from the same grids)."""
    N, I_, J_, lam, mus, w, P, H = orc.cached_geometry(M)
    c = t106["c"][0].cpu().numpy()
    tgt = t106["tgt"].cpu().numpy()
    fields = [np.real(orc.invrs_fft(orc.invrs_leg(c[f], I_, J_, M, N, P), I_)) for f in range(3)]

    def oracle_loss(tr, td, dp):
        out = orc.run_model_scan(M=M, dt=CFG["dt"], tmax=2 + NSTEPS, Phibar=CFG["Phibar"], omega=CFG["omega"],
                                 a=CFG["a"], taurad=tr, taudrag=td, DPhieq=dp, eta0_init=fields[0],
                                 delta0_init=fields[1], Phi0_init=fields[2], diagnostics=False,
                                 return_history=False)
        return float(np.mean((out["last_state"].Phi_curr - tgt) ** 2))

    base = {"tr": TAURAD[0], "td": TAUDRAG[0], "dp": DPHIEQ[0]}
    for name in ("tr", "td", "dp"):
        fd = _richardson(lambda dh: oracle_loss(**{k: v + (dh if k == name else 0.0) for k, v in base.items()}),
                         1e-4 * abs(base[name]))
        got = t106["g"][name][0]
        rel = abs(got - fd) / max(abs(fd), 1e-300)
        print("oracle", name, "adjoint", got, "fd", fd, "rel", rel)
        assert rel < 1e-6, (name, got, fd, rel)


def test_t106x64_fp32_gradients_close_to_fp64(torch_gpu):
    """64 members with per-member forcing (taurad, taudrag log-uniform in [1e4, 1e6] s, DPhieq
    uniform in [1e5, 5e5]), 700 steps, a checkpoint every 70: the FP32-transform gradients (initial
    spectral state, taurad, taudrag, DPhieq) and final Phi within 1e-4 (relative L2) of FP64."""
    from my_swampe_amd.autograd import rollout_phi

    torch = torch_gpu
    print("[gpu-case] t106x64 fp32 vs fp64 gradients", flush=True)
    B = 64
    rng = np.random.default_rng(7)
    tr0 = 10.0 ** rng.uniform(4.0, 6.0, B)
    td0 = 10.0 ** rng.uniform(4.0, 6.0, B)
    dp0 = rng.uniform(1e5, 5e5, B)
    c0 = torch.tensor(np.stack([_ic(100 + b) for b in range(B)]), device="cuda")
    tgt = torch.zeros((J, I), dtype=torch.float64, device="cuda")
    runs = {}
    for prec in ("fp64", "fp32"):
        c = c0.clone().requires_grad_(True)
        tr, td, dp = (torch.tensor(x, device="cuda", requires_grad=True) for x in (tr0, td0, dp0))
        stats = {}
        phi = rollout_phi(c, tr, td, dp, nsteps=700, checkpoint_every=70, precision=prec, stats=stats, **CFG)
        _loss(phi, tgt).backward()
        runs[prec] = [phi.detach(), c.grad, tr.grad, td.grad, dp.grad]
        assert stats["checkpoints"] == 10, stats
    for name, a, b in zip(("Phi", "init_state", "taurad", "taudrag", "DPhieq"), runs["fp32"], runs["fp64"]):
        assert bool(torch.isfinite(a).all()), name
        rel = float(torch.linalg.vector_norm(a - b) / torch.linalg.vector_norm(b))
        print(name, "fp32 vs fp64 relL2", rel)
        assert rel <= 1e-4, (name, rel)
