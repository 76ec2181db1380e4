"""The reference's diagnostic vorticity/divergence transform on the GPU (spectral_transform.py:
599-650, diagnostic_eta_delta) against the FP64 oracle, and the reference's transform round trips
(unit_tests/test_transform_stack.py:100-180) with it running on the HIP kernels."""

from __future__ import annotations

import numpy as np
import pytest

from oracle import swampe_oracle as orc

pytestmark = pytest.mark.gpu


def _np(t):
    return t.detach().cpu().numpy()


def _setup(M, a1, Phibar):
    """The reference tests' case: test-1 solid-body rotation (omega = 7.2921159e-5, a = 6.37122e6,
    dt = 30 s), as in tests/test_oracle.py."""
    omega, a, dt = 7.2921159e-5, 6.37122e6, 30.0
    s = orc.build_static(M=M, dt=dt, a=a, omega=omega, g=9.8, Phibar=Phibar, taurad=86400.0, taudrag=86400.0,
                         DPhieq=4e6, K6=1.24e33, K6Phi=None, test=1)
    eta0, delta0, Phi0, U0, V0 = orc.analytic_ic(s, 1, a1)
    return s, eta0, delta0, U0, V0


def _args(s):
    return (s.fmn, s.I, s.J, s.M, s.N, s.Pmn, s.Hmn, s.w, s.tstepcoeff, s.mJarray, s.dt)


@pytest.mark.parametrize("M", [42, 106])
def test_diagnostic_eta_delta_vs_oracle(M, torch_gpu):
    """Random band-limited winds (the Fourier winds of seeded random vorticity and divergence
    spectra, n <= 20): all four outputs against the oracle within 1e-12 of each array's largest
    magnitude (FP64, different summation orders and FFT against direct synthesis)."""
    from my_swampe_amd.spectral_transform import diagnostic_eta_delta

    print(f"[gpu-case] diagnostic_eta_delta M={M}", flush=True)
    s, *_ = _setup(M, np.pi / 4, 1.0e3)
    rng = np.random.default_rng(M)
    etamn = np.zeros((M + 1, M + 1), complex)
    deltamn = np.zeros_like(etamn)
    for m in range(21):
        for n in range(m, 21):
            etamn[m, n] = 1e-6 * (rng.standard_normal() + (1j * rng.standard_normal() if m else 0.0))
            deltamn[m, n] = 1e-7 * (rng.standard_normal() + (1j * rng.standard_normal() if m else 0.0))
    _, _, Um, Vm = orc.invrs_uv(deltamn, etamn, s.fmn, s.I, s.J, s.M, s.N, s.Pmn, s.Hmn, s.tstepcoeffmn, s.marray)
    ref = orc.diagnostic_eta_delta(Um, Vm, *_args(s))
    got = diagnostic_eta_delta(Um, Vm, *_args(s))
    for name, g, r in zip(("eta", "delta", "etamn", "deltamn"), got, ref):
        g = _np(g)
        err = float(np.max(np.abs(g - r)))
        scale = float(np.max(np.abs(r)))
        print(name, "max|diff| / max|ref|", err / scale)
        assert g.shape == r.shape, name
        assert err <= 1e-12 * scale, (name, err / scale)


def test_wind_transform_round_trip_M106(torch_gpu):
    """diagnostic_eta_delta (GPU) then invrs_uv recovers U, V (test_transform_stack.py:100-136, its
    FP64 tolerances atol = rtol = 1e-10)."""
    from my_swampe_amd.spectral_transform import diagnostic_eta_delta

    print("[gpu-case] wind round trip M=106", flush=True)
    s, eta0, delta0, U0, V0 = _setup(106, np.pi / 4, 1.0e3)
    Um = orc.fwd_fft_trunc(U0, s.I, s.M)
    Vm = orc.fwd_fft_trunc(V0, s.I, s.M)
    _, _, etamn, deltamn = diagnostic_eta_delta(Um, Vm, *_args(s))
    U, V, _, _ = orc.invrs_uv(_np(deltamn), _np(etamn), s.fmn, s.I, s.J, s.M, s.N, s.Pmn, s.Hmn, s.tstepcoeffmn,
                              s.marray)
    np.testing.assert_allclose(np.real(U), U0, atol=1e-10, rtol=1e-10)
    np.testing.assert_allclose(np.real(V), V0, atol=1e-10, rtol=1e-10)


def test_vorticity_divergence_round_trip_M63(torch_gpu):
    """invrs_uv then diagnostic_eta_delta (GPU) recovers eta, delta on the grid
    (test_transform_stack.py:139-180, atol = rtol = 1e-10)."""
    from my_swampe_amd.spectral_transform import diagnostic_eta_delta

    print("[gpu-case] vorticity/divergence round trip M=63", flush=True)
    s, eta0, delta0, U0, V0 = _setup(63, np.pi / 4, 4.0e4)
    etamn = orc.fwd_leg(orc.fwd_fft_trunc(eta0, s.I, s.M), s.Pmnw)
    deltamn = orc.fwd_leg(orc.fwd_fft_trunc(delta0, s.I, s.M), s.Pmnw)
    U, V, _, _ = orc.invrs_uv(deltamn, etamn, s.fmn, s.I, s.J, s.M, s.N, s.Pmn, s.Hmn, s.tstepcoeffmn, s.marray)
    Um = orc.fwd_fft_trunc(np.real(U), s.I, s.M)
    Vm = orc.fwd_fft_trunc(np.real(V), s.I, s.M)
    eta, delta, _, _ = diagnostic_eta_delta(Um, Vm, *_args(s))
    np.testing.assert_allclose(np.real(_np(eta)), eta0, atol=1e-10, rtol=1e-10)
    np.testing.assert_allclose(np.real(_np(delta)), delta0, atol=1e-10, rtol=1e-10)


def test_diagnostic_eta_delta_batched_and_errors(torch_gpu):
    """A batch of 3 members equals three single calls bit for bit; wrong shapes raise ValueError
    before any device work, as the reference's shape errors do."""
    import torch

    from my_swampe_amd.spectral_transform import diagnostic_eta_delta

    print("[gpu-case] diagnostic_eta_delta batch", flush=True)
    s, eta0, delta0, U0, V0 = _setup(42, np.pi / 4, 1.0e3)
    rng = np.random.default_rng(5)
    Ums = [orc.fwd_fft_trunc(U0 * (1 + 0.1 * k) + rng.standard_normal(U0.shape), s.I, s.M) for k in range(3)]
    Vms = [orc.fwd_fft_trunc(V0 * (1 - 0.1 * k) + rng.standard_normal(V0.shape), s.I, s.M) for k in range(3)]
    batch = diagnostic_eta_delta(np.stack(Ums), np.stack(Vms), *_args(s))
    for k in range(3):
        one = diagnostic_eta_delta(Ums[k], Vms[k], *_args(s))
        for a, b in zip(batch, one):
            assert torch.equal(a[k], b)
    with pytest.raises(ValueError):
        diagnostic_eta_delta(Ums[0][:, :-1], Vms[0][:, :-1], *_args(s))
    with pytest.raises(ValueError):
        diagnostic_eta_delta(Ums[0], Vms[0], s.fmn, s.I, s.J, s.M, s.N, s.Pmn[:, :, :-1], s.Hmn, s.w, s.tstepcoeff,
                             s.mJarray, s.dt)
