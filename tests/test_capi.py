"""C-ABI library on the CPU: load, exported symbols, host geometry, API validation (no GPU compute)."""

from __future__ import annotations

import re

import numpy as np
import pytest

from conftest import ROOT
from oracle import swampe_oracle as orc


def _declared_symbols():
    text = (ROOT / "include" / "swampe_capi.h").read_text()
    return sorted(set(re.findall(r"\b(swx_[a-z_0-9]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    from my_swampe_amd import _capi

    lib = _capi.load()
    declared = _declared_symbols()
    assert len(declared) >= 10
    for name in declared:
        assert hasattr(lib, name), name
    assert set(declared) == set(_capi.EXPORTED_SYMBOLS)
    assert lib.swx_version().startswith(b"swampe-hip")


def test_grid_size_and_unsupported_truncation():
    from my_swampe_amd import _capi

    assert _capi.grid_size(42) == (64, 128)
    assert _capi.grid_size(63) == (96, 192)
    assert _capi.grid_size(106) == (160, 320)
    with pytest.raises(ValueError, match="M=50"):
        _capi.grid_size(50)


@pytest.mark.parametrize("M", [42, 63, 106])
def test_host_geometry_matches_oracle(M):
    """C++ Gauss-Legendre + stable-recurrence basis with the reference normalization vs the scipy-based oracle."""
    from my_swampe_amd import _capi

    mus, w, lam, P, H = _capi.geometry(M)
    N, I, J, lo, mo, wo, Po, Ho = orc.cached_geometry(M)
    np.testing.assert_allclose(mus, mo, rtol=0, atol=1e-15)
    np.testing.assert_allclose(w, wo, rtol=0, atol=1e-14)
    np.testing.assert_allclose(lam, lo, rtol=0, atol=0)
    np.testing.assert_allclose(P, Po, rtol=0, atol=1e-12)
    np.testing.assert_allclose(H, Ho, rtol=0, atol=1e-11)


@pytest.mark.parametrize("M", [42, 63, 106])
def test_host_geometry_equatorial_parity(M):
    """What the folded Legendre kernels rely on: mirrored nodes mu_{J-1-j} = -mu_j, equal weights,
    P(-mu) = (-1)^{n+m} P(mu) and H(-mu) = -(-1)^{n+m} H(mu) for every (m, n) of the library's basis
    (reference normalization included), to 1e-12 of each block's largest value."""
    from my_swampe_amd import _capi

    mus, w, lam, P, H = _capi.geometry(M)
    J = mus.size
    np.testing.assert_allclose(mus[::-1], -mus, rtol=0, atol=1e-15)
    np.testing.assert_allclose(w[::-1], w, rtol=0, atol=1e-15)
    m, n = np.meshgrid(np.arange(M + 1), np.arange(M + 1), indexing="ij")
    sign = np.where((n + m) % 2 == 0, 1.0, -1.0)
    scale = np.maximum(np.abs(P).max(axis=0), np.abs(H).max(axis=0))
    assert np.all(np.abs(P[::-1] - sign * P) <= 1e-12 * scale)
    assert np.all(np.abs(H[::-1] + sign * H) <= 1e-12 * scale)


def test_exact_basis_is_orthonormal_at_T106():
    from my_swampe_amd import _capi

    mus, w, lam, P, H = _capi.geometry(106, exact_basis=True)
    for m in (0, 53, 90, 106):
        G = np.einsum("j,jn,jk->nk", w, P[:, m, m:], P[:, m, m:])
        np.testing.assert_allclose(G, np.eye(G.shape[0]), atol=1e-13)
    Pr = _capi.geometry(106)[3]
    assert np.all(Pr[:, 106, 106] == 0.0) and not np.any(P[:, 106, 106] == 0.0)


BASE = dict(M=42, dt=600.0, Phibar=3.0e5, omega=7.292e-5, a=6.37122e6, test=None, g=9.8, forcflag=True,
            taurad=86400.0, taudrag=86400.0, DPhieq=4.0e6, diffflag=True, modalflag=True, alpha=0.01,
            expflag=False, K6=1.24e33, diagnostics=False, jit_scan=False)


def _run(**kw):
    from my_swampe_amd.model import run_model_scan

    return run_model_scan(**{**BASE, **kw})


def test_validation_errors_raise_before_device_work():
    """unit_tests/test_invalid_input.py: same exception types and messages, raised on the host."""
    with pytest.raises(ValueError, match="tmax"):
        _run(tmax=1)
    with pytest.raises(ValueError, match="dt"):
        _run(tmax=4, dt=0.0)
    with pytest.raises(ValueError, match="starttime"):
        _run(tmax=4, starttime=10)
    with pytest.raises(ValueError, match="contTime"):
        _run(tmax=4, contflag=True)
    with pytest.raises(ValueError, match="numeric"):
        _run(tmax=4, contflag=True, contTime="not-a-number")
    with pytest.raises(ValueError, match="semi_implicit"):
        _run(tmax=4, semi_implicit=True, expflag=True)
    with pytest.raises(ValueError, match="raw_filter"):
        _run(tmax=4, raw_filter=True, expflag=True)
    with pytest.raises(ValueError, match="M=50"):
        _run(tmax=4, M=50)
    with pytest.raises(ValueError, match="precision"):
        _run(tmax=4, precision="fp16")
    z = np.zeros((64, 128))
    with pytest.raises(ValueError, match=r"eta0_init.*delta0_init.*Phi0_init"):
        _run(tmax=4, eta0_init=z)
    with pytest.raises(ValueError, match=r"U0_init.*V0_init"):
        _run(tmax=4, eta0_init=z, delta0_init=z, Phi0_init=z, U0_init=z)
    with pytest.raises(ValueError, match="Phi0_init"):
        _run(tmax=4, eta0_init=z, delta0_init=z, Phi0_init=np.zeros((65, 128)))


def test_out_of_scope_options_raise_not_implemented():
    with pytest.raises(NotImplementedError):
        _run(tmax=4, test=99)


def test_product_path_fails_loudly_without_gpu():
    """No CPU fallback: on a host without a ROCm GPU the run raises instead of computing."""
    import torch

    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    with pytest.raises(RuntimeError, match="no CPU fallback"):
        _run(tmax=4)


def test_measure_ceilings_refuses_small_or_ragged_sizes():
    """swx_measure_ceilings checks its arguments before any device work: fewer than 256 MiB, a size
    that is not a multiple of 16 and a null output are SWX_ERR_ARG with a message."""
    import ctypes

    from my_swampe_amd import _capi

    lib = _capi.load()
    out = (ctypes.c_double * 6)()
    assert lib.swx_measure_ceilings(out, ctypes.c_size_t(1 << 20), None) == _capi.SWX_ERR_ARG
    assert b"256 MiB" in lib.swx_last_error()
    assert lib.swx_measure_ceilings(out, ctypes.c_size_t((256 << 20) + 8), None) == _capi.SWX_ERR_ARG
    assert lib.swx_measure_ceilings(None, ctypes.c_size_t(1 << 30), None) == _capi.SWX_ERR_ARG


def test_spectral_transform_geometry_functions():
    """The reference's geometry functions from the native library (spectral_transform.py:54-114,
    :219-287): gauss_legendre against scipy.special.roots_legendre at several orders (nodes to
    1e-15; weights to 1e-9 relative, the accuracy of scipy's own weights at J = 333, whose error
    against a long-double Newton solve is 4e-10 there), with weights summing to 2 and integrating
    x^2 exactly, and by default scipy's values themselves, as the reference returns them;
    build_lambdas against numpy.linspace, PmnHmn at the T42 Gauss nodes equal to
    the plan geometry bit for bit and to the closed forms of test_transform_stack.py:37-48, and at
    random nodes against the scipy-based oracle."""
    import scipy.special

    from my_swampe_amd import _capi
    from my_swampe_amd import spectral_transform as st

    for J in (1, 2, 7, 64, 96, 160, 333):
        mus, w = st.gauss_legendre(J, exact=True)
        x, wx = scipy.special.roots_legendre(J)
        mr, wr = st.gauss_legendre(J)
        assert np.array_equal(mr, x) and np.array_equal(wr, wx)
        np.testing.assert_allclose(mus, x, rtol=0, atol=1e-15)
        np.testing.assert_allclose(w, wx, rtol=1e-9, atol=0)
        assert abs(np.sum(w) - 2.0) <= 1e-13
        if J > 1:
            assert abs(np.sum(w * mus**2) - 2.0 / 3.0) <= 1e-13
        assert np.all(np.diff(mus) > 0)
    for I in (1, 128, 192, 320):
        np.testing.assert_allclose(st.build_lambdas(I), np.linspace(-np.pi, np.pi, I, endpoint=False), rtol=0,
                                   atol=1e-15)
    mus, w = st.gauss_legendre(64, exact=True)
    P, H = st.PmnHmn(64, 42, 42, mus)
    m0, w0, l0, P0, H0 = _capi.geometry(42)
    assert np.array_equal(mus, m0) and np.array_equal(w, w0)
    assert np.array_equal(P, P0) and np.array_equal(H, H0)
    np.testing.assert_allclose(P[:, 2, 2], 0.25 * np.sqrt(15.0) * (1.0 - mus**2), atol=1e-10, rtol=1e-10)
    np.testing.assert_allclose(H[:, 0, 1], 0.5 * np.sqrt(6.0) * (1.0 - mus**2), atol=1e-10, rtol=1e-10)
    xr = np.sort(np.random.default_rng(3).uniform(-0.99, 0.99, 11))
    P, H = st.PmnHmn(11, 20, 24, xr)
    Po, Ho = orc.pmn_hmn(11, 20, 24, xr)
    np.testing.assert_allclose(P, Po, rtol=0, atol=1e-12)
    np.testing.assert_allclose(H, Ho, rtol=0, atol=1e-11)
    Pw = st.weighted_legendre_basis(P0, w0)
    assert np.array_equal(Pw, w0[:, None, None] * P0)
    with pytest.raises(ValueError):
        st.gauss_legendre(0)
    with pytest.raises(ValueError):
        st.PmnHmn(2, 3, 3, np.array([0.5, 1.5]))
    with pytest.raises(ValueError):
        st.PmnHmn(3, 3, 3, np.array([0.5, 0.2]))


def test_standalone_transforms_refuse_bad_arguments():
    """The device transforms check their arguments before any launch: negative or oversized batch
    counts, a longitude count outside [1, 2048], I below M + 1 (invrs_leg) or 2 M + 1 (invrsUV), null
    pointers and no requested output are SWX_ERR_ARG; an empty batch is a no-op."""
    import ctypes

    from my_swampe_amd import _capi

    lib = _capi.load()
    E = _capi.SWX_ERR_ARG
    p = ctypes.c_void_p(8)  # never dereferenced: every call below fails or returns before a launch
    assert lib.swx_fwd_fft_trunc(p, 0, -1, 4, 8, 8, 2, p, None) == E
    assert lib.swx_fwd_fft_trunc(p, 0, 70000, 4, 8, 8, 2, p, None) == E
    assert lib.swx_fwd_fft_trunc(p, 0, 1, 4, 8, 4096, 2, p, None) == E
    assert b"2048" in lib.swx_last_error()
    assert lib.swx_fwd_fft_trunc(None, 0, 1, 4, 8, 8, 2, p, None) == E
    assert lib.swx_fwd_fft_trunc(None, 0, 0, 4, 8, 8, 2, None, None) == _capi.SWX_OK
    assert lib.swx_invrs_fft(p, 1, 4, 0, 8, p, None) == E
    assert lib.swx_fwd_leg(p, None, None, 1, 4, 2, 2, p, None) == E
    assert lib.swx_fwd_leg(p, p, None, 1, 0, 2, 2, p, None) == E
    assert lib.swx_invrs_leg(p, p, 1, 2, 4, 2, 2, p, None) == E
    assert b"M + 1" in lib.swx_last_error()
    args = [p, p, p, 1, 4, 4, 2, 2, p, p, p, p]
    assert lib.swx_invrs_uv(*args, p, p, None, None, None) == E
    assert b"2 M + 1" in lib.swx_last_error()
    args[4] = 5
    assert lib.swx_invrs_uv(*args, None, None, None, None, None) == E
    assert b"no output" in lib.swx_last_error()
    args[7] = 1
    assert lib.swx_invrs_uv(*args, p, p, None, None, None) == E


def test_legacy_main_conventions_on_the_host():
    """main_function.main refuses test codes beyond None/0/1/2 and the default plotflag=1 with
    NotImplementedError before any device work (main_function.py:166-176); without a GPU a valid
    call raises rather than computing."""
    import torch

    from my_swampe_amd.main_function import main

    args = (42, 1200.0, 6, 3.0e5, 3.2e-5, 8.2e7)
    with pytest.raises(NotImplementedError, match="test=3"):
        main(*args, 3, plotflag=0)
    with pytest.raises(NotImplementedError, match="plotflag"):
        main(*args, 0)
    if not torch.cuda.is_available():
        with pytest.raises(RuntimeError, match="no CPU fallback"):
            main(*args, 0, plotflag=0, saveflag=0, verbose=False)
