"""The reference's standalone spectral transforms on the GPU (spectral_transform.py:290-650):
fwd_fft_trunc, invrs_fft, fwd_leg(_w)(_batch), invrs_leg, invrsUV(_with_coeffs) and
diagnostic_eta_delta, against numpy.fft and the FP64 oracle, and the reference's transform tests
(unit_tests/test_transform_stack.py:37-180) with every transform in the chain running on the HIP
kernels."""

from __future__ import annotations

import numpy as np
import pytest

from oracle import swampe_oracle as orc

pytestmark = pytest.mark.gpu

# The reference's FP64 tolerances for chained transforms (test_transform_stack.py:21-26).
RT_ATOL = RT_RTOL = 1e-10


def _np(t):
    return t.detach().cpu().numpy()


def _rel(got, ref):
    got = _np(got) if hasattr(got, "detach") else got
    scale = float(np.max(np.abs(ref)))
    return float(np.max(np.abs(got - ref))) / (scale if scale else 1.0)


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


def _uv_args(s):
    return (s.fmn, s.I, s.J, s.M, s.N, s.Pmn, s.Hmn, s.tstepcoeffmn, s.marray)


def _random_spectrum(rng, M, nmax, scale):
    c = np.zeros((M + 1, M + 1), complex)
    for m in range(nmax + 1):
        for n in range(m, nmax + 1):
            c[m, n] = scale * (rng.standard_normal() + (1j * rng.standard_normal() if m else 0.0))
    return c


@pytest.mark.parametrize("M", [42, 106])
def test_fwd_fft_trunc_matches_numpy_fft(M, torch_gpu):
    """fwd_fft_trunc of real and complex (J, I) fields, of a short row (zero-padded to I, as
    fft(n=I) pads) and fwd_fft_trunc_batch of a (2, 3, J, I) stack against numpy.fft.fft(x / I)
    truncated to m <= M, within 1e-13 of the largest coefficient (direct DFT against FFT order)."""
    from my_swampe_amd.spectral_transform import fwd_fft_trunc, fwd_fft_trunc_batch

    print(f"[gpu-case] fwd_fft_trunc M={M}", flush=True)
    J, I = {42: (64, 128), 106: (160, 320)}[M]
    rng = np.random.default_rng(M)
    x = rng.standard_normal((J, I))
    z = x + 1j * rng.standard_normal((J, I))

    def ref(a):
        return np.fft.fft(a / I, n=I, axis=-1)[..., : M + 1]

    for name, a in (("real", x), ("complex", z), ("padded", x[:, : I - 7])):
        got = fwd_fft_trunc(a, I, M)
        assert tuple(got.shape) == (J, M + 1), name
        assert _rel(got, ref(a)) <= 1e-13, (name, _rel(got, ref(a)))
    stack = rng.standard_normal((2, 3, J, I))
    got = fwd_fft_trunc_batch(stack, I, M)
    assert tuple(got.shape) == (2, 3, J, M + 1)
    assert _rel(got, ref(stack)) <= 1e-13
    import torch

    assert torch.equal(got[1, 2], fwd_fft_trunc(stack[1, 2], I, M))
    with pytest.raises(ValueError):
        fwd_fft_trunc(stack, I, M)


@pytest.mark.parametrize("M", [42, 106])
def test_invrs_fft_matches_numpy_ifft(M, torch_gpu):
    """invrs_fft of full (J, I) and truncated (J, M+1) coefficient rows against
    numpy.fft.ifft(I X, n=I), within 1e-13 of the largest value."""
    from my_swampe_amd.spectral_transform import invrs_fft

    print(f"[gpu-case] invrs_fft M={M}", flush=True)
    J, I = {42: (64, 128), 106: (160, 320)}[M]
    rng = np.random.default_rng(100 + M)
    for C in (I, M + 1):
        X = rng.standard_normal((J, C)) + 1j * rng.standard_normal((J, C))
        got = invrs_fft(X, I)
        ref = np.fft.ifft(I * X, n=I, axis=1)
        assert tuple(got.shape) == (J, I)
        assert _rel(got, ref) <= 1e-13, (C, _rel(got, ref))


@pytest.mark.parametrize("M", [42, 106])
def test_legendre_transforms_vs_oracle(M, torch_gpu):
    """fwd_leg, fwd_leg_w, fwd_leg_w_batch and invrs_leg against the oracle (1e-13 of the largest
    value); invrs_leg's layout: +m at column m, the exact conjugate at column I - m, zeros
    between; a batch of 3 spectra equals three single calls bit for bit."""
    import torch

    from my_swampe_amd.spectral_transform import (PmnHmn, fwd_leg, fwd_leg_w, fwd_leg_w_batch, gauss_legendre,
                                                  invrs_leg, weighted_legendre_basis)

    print(f"[gpu-case] Legendre transforms M={M}", flush=True)
    J, I = {42: (64, 128), 106: (160, 320)}[M]
    N = M
    mus, w = gauss_legendre(J)
    Pmn, Hmn = PmnHmn(J, M, N, mus)
    rng = np.random.default_rng(200 + M)
    data = rng.standard_normal((3, J, M + 1)) + 1j * rng.standard_normal((3, J, M + 1))
    ref = np.stack([orc.fwd_leg(d, w[:, None, None] * Pmn) for d in data])
    assert _rel(fwd_leg(data[0], J, M, N, Pmn, w), ref[0]) <= 1e-13
    Hw = weighted_legendre_basis(Hmn, w)
    refH = orc.fwd_leg(data[1], Hw)
    assert _rel(fwd_leg_w(data[1], Hw), refH) <= 1e-13
    batch = fwd_leg_w_batch(data, weighted_legendre_basis(Pmn, w))
    assert _rel(batch, ref) <= 1e-13
    S = np.stack([_random_spectrum(rng, M, M, 1.0) for _ in range(3)])
    out = invrs_leg(S[0], I, J, M, N, Pmn)
    refS = orc.invrs_leg(S[0], I, J, M, N, Pmn)
    assert tuple(out.shape) == (J, I)
    assert _rel(out, refS) <= 1e-13
    o = _np(out)
    np.testing.assert_array_equal(o[:, I - M:], np.conj(o[:, 1:M + 1][:, ::-1]))
    assert not np.any(o[:, M + 1:I - M])
    outs = invrs_leg(S, I, J, M, N, Pmn)
    for k in range(3):
        assert torch.equal(outs[k], invrs_leg(S[k], I, J, M, N, Pmn))


@pytest.mark.parametrize("M", [42, 106])
def test_invrs_uv_vs_oracle(M, torch_gpu):
    """invrsUV and invrsUV_with_coeffs against the oracle's invrs_uv on random spectra
    (n <= 20), within 1e-12 of the largest value; both return the same winds bit for bit, and a
    batch of 2 equals two single calls bit for bit."""
    import torch

    from my_swampe_amd.spectral_transform import invrsUV, invrsUV_with_coeffs

    print(f"[gpu-case] invrsUV M={M}", flush=True)
    s, *_ = _setup(M, np.pi / 4, 1.0e3)
    rng = np.random.default_rng(300 + M)
    d = np.stack([_random_spectrum(rng, M, 20, 1e-7) for _ in range(2)])
    e = np.stack([_random_spectrum(rng, M, 20, 1e-6) for _ in range(2)])
    ref = orc.invrs_uv(d[0], e[0], *_uv_args(s))
    U, V = invrsUV(d[0], e[0], *_uv_args(s))
    got4 = invrsUV_with_coeffs(d[0], e[0], *_uv_args(s))
    for name, g, r in zip(("U", "V"), (U, V), ref[:2]):
        assert _rel(g, r) <= 1e-12, (name, _rel(g, r))
    for name, g, r in zip(("U", "V", "Um", "Vm"), got4, ref):
        assert tuple(g.shape) == r.shape, name
        assert _rel(g, r) <= 1e-12, (name, _rel(g, r))
    assert torch.equal(got4[0], U) and torch.equal(got4[1], V)
    both = invrsUV_with_coeffs(d, e, *_uv_args(s))
    for k in range(2):
        one = invrsUV_with_coeffs(d[k], e[k], *_uv_args(s))
        for a, b in zip(both, one):
            assert torch.equal(a[k], b)


def _reference_case(M, a1, Phibar, dt=30.0):
    """The setup of the reference's transform tests (test_transform_stack.py:69-180) through this
    package's own functions: test-1 fields, basis, Coriolis term and coefficient arrays."""
    from my_swampe_amd import initial_conditions as ic
    from my_swampe_amd import time_stepping as tstep
    from my_swampe_amd.spectral_transform import PmnHmn

    omega, a = 7.2921159e-5, 6.37122e6
    N, I, J, _dt, lambdas, mus, w = ic.spectral_params(M)
    Pmn, Hmn = PmnHmn(J, M, N, mus)
    SU0, sina, cosa, etaamp, Phiamp = ic.test1_init(a, omega, a1)
    eta0, _, delta0, _, Phi0, _ = ic.state_var_init(I, J, mus, lambdas, 1, etaamp, a, sina, cosa, Phibar, Phiamp)
    U, V = ic.velocity_init(I, J, SU0, cosa, sina, mus, lambdas, 1)
    return dict(M=M, N=N, I=I, J=J, w=w, Pmn=Pmn, Hmn=Hmn, dt=dt, fmn=ic.coriolismn(M, omega), eta0=eta0,
                delta0=delta0, U=U, V=V, tstepcoeffmn=tstep.tstepcoeffmn(M, N, a),
                tstepcoeff=tstep.tstepcoeff(J, M, dt, mus, a), mJarray=tstep.mJarray(J, M), marray=tstep.marray(M, N))


def _diag(c, Um, Vm):
    from my_swampe_amd.spectral_transform import diagnostic_eta_delta

    return diagnostic_eta_delta(Um, Vm, c["fmn"], c["I"], c["J"], c["M"], c["N"], c["Pmn"], c["Hmn"], c["w"],
                                c["tstepcoeff"], c["mJarray"], c["dt"])


def _uv(c, deltamn, etamn):
    from my_swampe_amd.spectral_transform import invrsUV

    return invrsUV(deltamn, etamn, c["fmn"], c["I"], c["J"], c["M"], c["N"], c["Pmn"], c["Hmn"], c["tstepcoeffmn"],
                   c["marray"])


def test_spectral_transform_coriolis_M106(torch_gpu):
    """f = 2 omega mu through fwd_fft_trunc and fwd_leg (GPU) is omega / sqrt(0.375) at (0, 1)
    and zero elsewhere (test_transform_stack.py:51-66)."""
    from my_swampe_amd import initial_conditions as ic
    from my_swampe_amd.spectral_transform import PmnHmn, fwd_fft_trunc, fwd_leg

    print("[gpu-case] Coriolis M=106", flush=True)
    M, omega = 106, 3.2e-5
    N, I, J, _dt, _lambdas, mus, w = ic.spectral_params(M)
    Pmn, _ = PmnHmn(J, M, N, mus)
    f = (2.0 * omega) * mus[:, None] * np.ones((1, I))
    fmn = _np(fwd_leg(fwd_fft_trunc(f, I, M), J, M, N, Pmn, w))
    check = np.zeros((M + 1, N + 1), complex)
    check[0, 1] = omega / np.sqrt(0.375)
    np.testing.assert_allclose(fmn, check, atol=RT_ATOL, rtol=RT_RTOL)


def test_spectral_transform_forward_inverse_M63(torch_gpu):
    """fwd_fft_trunc, fwd_leg, invrs_leg, invrs_fft (all GPU) recover the test-1 zonal wind
    (test_transform_stack.py:69-97, a1 = pi/2, Phibar = 3e3)."""
    from my_swampe_amd.spectral_transform import fwd_fft_trunc, fwd_leg, invrs_fft, invrs_leg

    print("[gpu-case] forward/inverse round trip M=63", flush=True)
    c = _reference_case(63, np.pi / 2, 3.0e3)
    I, J, M, N = c["I"], c["J"], c["M"], c["N"]
    Um = fwd_fft_trunc(c["U"], I, M)
    Umn = fwd_leg(Um, J, M, N, c["Pmn"], c["w"])
    Unew = invrs_fft(invrs_leg(Umn, I, J, M, N, c["Pmn"]), I)
    np.testing.assert_allclose(np.real(_np(Unew)), c["U"], atol=RT_ATOL, rtol=RT_RTOL)


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
        print(name, "max|diff| / max|ref|", _rel(g, r))
        assert tuple(g.shape) == r.shape, name
        assert _rel(g, r) <= 1e-12, (name, _rel(g, r))


def test_wind_transform_round_trip_M106(torch_gpu):
    """fwd_fft_trunc, diagnostic_eta_delta, invrsUV (all GPU) recover U, V
    (test_transform_stack.py:100-136, a1 = pi/4, Phibar = 1e3)."""
    from my_swampe_amd.spectral_transform import fwd_fft_trunc

    print("[gpu-case] wind round trip M=106", flush=True)
    c = _reference_case(106, np.pi / 4, 1.0e3)
    Um = fwd_fft_trunc(c["U"], c["I"], c["M"])
    Vm = fwd_fft_trunc(c["V"], c["I"], c["M"])
    _, _, etamn, deltamn = _diag(c, Um, Vm)
    U, V = _uv(c, deltamn, etamn)
    np.testing.assert_allclose(np.real(_np(U)), c["U"], atol=RT_ATOL, rtol=RT_RTOL)
    np.testing.assert_allclose(np.real(_np(V)), c["V"], atol=RT_ATOL, rtol=RT_RTOL)


def test_vorticity_divergence_round_trip_M63(torch_gpu):
    """fwd_fft_trunc, fwd_leg, invrsUV, fwd_fft_trunc of the complex winds, diagnostic_eta_delta
    (all GPU) recover eta, delta on the grid (test_transform_stack.py:139-180, Phibar = 4e4)."""
    from my_swampe_amd.spectral_transform import fwd_fft_trunc, fwd_leg

    print("[gpu-case] vorticity/divergence round trip M=63", flush=True)
    c = _reference_case(63, np.pi / 4, 4.0e4)
    I, J, M, N = c["I"], c["J"], c["M"], c["N"]
    deltamn = fwd_leg(fwd_fft_trunc(c["delta0"], I, M), J, M, N, c["Pmn"], c["w"])
    etamn = fwd_leg(fwd_fft_trunc(c["eta0"], I, M), J, M, N, c["Pmn"], c["w"])
    U, V = _uv(c, deltamn, etamn)
    eta, delta, _, _ = _diag(c, fwd_fft_trunc(U, I, M), fwd_fft_trunc(V, I, M))
    np.testing.assert_allclose(np.real(_np(eta)), c["eta0"], atol=RT_ATOL, rtol=RT_RTOL)
    np.testing.assert_allclose(np.real(_np(delta)), c["delta0"], atol=RT_ATOL, rtol=RT_RTOL)


def test_setup_functions_on_device_tensors(torch_gpu):
    """state_var_init, velocity_init, ABCDE_init, tstepcoeff and RMS_winds given torch tensors on
    the GPU compute there and agree with their numpy results within 1e-14 of each field's
    largest value (the device's transcendental functions against numpy's)."""
    import torch

    from my_swampe_amd import initial_conditions as ic
    from my_swampe_amd import time_stepping as ts

    print("[gpu-case] setup functions on device tensors", flush=True)
    N, I, J, dt, lam, mus, w = ic.spectral_params(63)
    mus_t, lam_t = torch.tensor(mus, device="cuda"), torch.tensor(lam, device="cuda")
    a, Phibar = 6.37122e6, 4.0e4
    SU0, sina, cosa, etaamp, Phiamp = ic.test1_init(a, 7.2921159e-5, np.pi / 4)
    for test in (None, 1, 2):
        args = () if test is None else (a, sina, cosa, Phibar, Phiamp)
        host = ic.state_var_init(I, J, mus, lam, test, etaamp, *args) + ic.velocity_init(I, J, SU0, cosa, sina, mus,
                                                                                          lam, test)
        dev = ic.state_var_init(I, J, mus_t, lam_t, test, etaamp, *args) + ic.velocity_init(I, J, SU0, cosa, sina,
                                                                                             mus_t, lam_t, test)
        for h, d in zip(host, dev):
            assert d.is_cuda and tuple(d.shape) == (J, I)
            assert _rel(d, h) <= 1e-14
        A = ic.ABCDE_init(dev[6], dev[7], dev[0], dev[4], mus_t, I, J)
        Ah = ic.ABCDE_init(host[6], host[7], host[0], host[4], mus, I, J)
        for d, h in zip(A, Ah):
            assert d.is_cuda and _rel(d, h) <= 1e-14
        r_dev = float(ts.RMS_winds(a, I, J, lam_t, mus_t, dev[6], dev[7]))
        r_host = float(ts.RMS_winds(a, I, J, lam, mus, host[6], host[7]))
        assert abs(r_dev - r_host) <= 1e-13 * max(r_host, 1.0)
    assert _rel(ts.tstepcoeff(J, N, dt, mus_t, a), ts.tstepcoeff(J, N, dt, mus, a)) <= 1e-15


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


def test_transform_argument_errors(torch_gpu):
    """Shape and size errors raise ValueError before any launch: a wrong basis shape, mismatched
    delta/eta spectra, I below 2 M + 1 for the wind inversion, I above the DFT kernels' 2048."""
    from my_swampe_amd.spectral_transform import fwd_fft_trunc, fwd_leg, invrs_leg, invrsUV

    s, *_ = _setup(42, np.pi / 4, 1.0e3)
    rng = np.random.default_rng(9)
    S = _random_spectrum(rng, 42, 42, 1.0)
    with pytest.raises(ValueError):
        fwd_leg(np.zeros((s.J, 43), complex), s.J, 42, 42, s.Pmn[:, :, :-1], s.w)
    with pytest.raises(ValueError):
        invrs_leg(S[:, :-1], s.I, s.J, 42, 42, s.Pmn)
    with pytest.raises(ValueError):
        invrs_leg(S, 40, s.J, 42, 42, s.Pmn)
    with pytest.raises(ValueError):
        invrsUV(S, S[None], *_uv_args(s))
    with pytest.raises(ValueError):
        invrsUV(S, S, s.fmn, 84, s.J, s.M, s.N, s.Pmn, s.Hmn, s.tstepcoeffmn, s.marray)
    with pytest.raises(ValueError):
        fwd_fft_trunc(np.zeros((2, 4096)), 4096, 42)


def test_batched_wind_round_trip_at_full_size(torch_gpu):
    """The headline batch size through the standalone transforms in one launch each: 256 members'
    random vorticity and divergence spectra (n <= 30) -> invrsUV_with_coeffs -> the Fourier winds ->
    diagnostic_eta_delta recovers every member's spectra for n >= 1 (the degrees the wind
    inversion keeps; it zeroes n = 0) within 1e-10 of the largest coefficient, and the grids it
    returns equal the synthesis of those spectra."""
    import torch

    from my_swampe_amd.spectral_transform import diagnostic_eta_delta, invrs_fft, invrs_leg, invrsUV_with_coeffs

    print("[gpu-case] batched wind round trip T42 x 256", flush=True)
    s, *_ = _setup(42, np.pi / 4, 1.0e3)
    B, M = 256, 42
    rng = np.random.default_rng(256)
    eta = np.stack([_random_spectrum(rng, M, 30, 1e-6) for _ in range(B)])
    delta = np.stack([_random_spectrum(rng, M, 30, 1e-7) for _ in range(B)])
    eta = eta + s.fmn[None]
    U, V, Um, Vm = invrsUV_with_coeffs(delta, eta, *_uv_args(s))
    assert tuple(U.shape) == (B, s.J, s.I) and tuple(Um.shape) == (B, s.J, M + 1)
    e_grid, d_grid, etamn, deltamn = diagnostic_eta_delta(Um, Vm, *_args(s))
    keep = np.triu(np.ones((M + 1, M + 1), bool))
    keep[:, 0] = False
    for name, got, want in (("eta", etamn, eta), ("delta", deltamn, delta)):
        g = _np(got)
        scale = float(np.max(np.abs(want[:, keep])))
        assert float(np.max(np.abs(g[:, keep] - want[:, keep]))) <= 1e-10 * scale, name
    k = 137
    synth = invrs_fft(invrs_leg(etamn[k], s.I, s.J, M, s.N, s.Pmn), s.I)
    assert float(torch.max(torch.abs(synth - e_grid[k]))) <= 1e-12 * float(torch.max(torch.abs(e_grid[k])))
