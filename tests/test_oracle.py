"""Pinning of the FP64 oracle (CPU): upstream fixtures, closed forms, round trips, quirk properties.

The oracle is the checker of the HIP path, so every part of it is pinned here by
something other than a retyping of its own formulas: the upstream-SWAMPE fixture
data, textbook closed forms and library routines (numpy FFT / leggauss), brute
force on small inputs, and invariants.
"""

from __future__ import annotations

import numpy as np
import pytest

from conftest import ATOL, FIXTURE_CASES, GOLDEN, phase_curve_from_phi
from oracle import swampe_oracle as orc

FIELDS = ("eta", "delta", "Phi", "U", "V")


# --- end to end: upstream-SWAMPE fixtures (unit_tests/test_parity_reference_regression.py) -------

@pytest.mark.parametrize("case", list(FIXTURE_CASES))
def test_oracle_fixture_terminal_and_snapshots(case):
    ref = np.load(GOLDEN / f"{case}.npz")
    out = orc.run_model_scan(**FIXTURE_CASES[case], diagnostics=True)
    for f in FIELDS:
        np.testing.assert_allclose(getattr(out["last_state"], f + "_curr"), ref[f"final_{f}"], rtol=0, atol=ATOL[f])
    for i, step in enumerate(ref["snapshot_steps"]):
        idx = int(np.where(out["t_seq"] == step)[0][0])
        for f in FIELDS:
            np.testing.assert_allclose(out["outs"][f][idx], ref[f"snapshot_{f}"][i], rtol=0, atol=ATOL[f])


@pytest.mark.parametrize("case", list(FIXTURE_CASES))
def test_oracle_fixture_phase_curve(case):
    ref = np.load(GOLDEN / f"{case}.npz")
    out = orc.run_model_scan(**FIXTURE_CASES[case], diagnostics=False, return_history=False)
    curve = phase_curve_from_phi(out["last_state"].Phi_curr, ref["phase_angles"])
    np.testing.assert_allclose(curve, ref["phase_curve"], rtol=0, atol=1e-10)


@pytest.mark.parametrize("case", list(FIXTURE_CASES))
def test_oracle_fixture_legacy_diagnostics(case):
    ref = np.load(GOLDEN / f"{case}.npz")
    kw = dict(FIXTURE_CASES[case])
    res = orc.run_model(kw.pop("M"), kw.pop("dt"), kw.pop("tmax"), kw.pop("Phibar"), kw.pop("omega"), kw.pop("a"), **kw)
    np.testing.assert_allclose(res["spinup"], ref["spinup"], rtol=1e-11, atol=1e-9)
    np.testing.assert_allclose(res["geopot"], ref["geopot"], rtol=1e-11, atol=5e-8)


# --- geometry ---------------------------------------------------------------------------------

def test_spectral_params_table():
    """initial_conditions.py:66-81 resolution table (test_transform_stack.py:29-34)."""
    N, I, J, dt, lam, mus, w = orc.spectral_params(42)
    assert (N, I, J, dt) == (42, 128, 64, 1200)
    assert orc.spectral_params(63)[1:4] == (192, 96, 900)
    assert orc.spectral_params(106)[1:4] == (320, 160, 600)
    with pytest.raises(ValueError, match="M"):
        orc.spectral_params(50)


@pytest.mark.parametrize("J", [4, 64, 96, 160])
def test_gauss_legendre_matches_numpy_leggauss(J):
    x, w = orc.gauss_legendre(J)
    xr, wr = np.polynomial.legendre.leggauss(J)
    np.testing.assert_allclose(x, xr, rtol=0, atol=1e-14)
    np.testing.assert_allclose(w, wr, rtol=0, atol=1e-14)
    assert abs(w.sum() - 2.0) < 1e-13


def test_lambdas_grid():
    lam = orc.build_lambdas(128)
    assert lam[0] == -np.pi and np.allclose(np.diff(lam), 2 * np.pi / 128)


def test_basis_closed_forms():
    """Pmn[:,2,2] = (sqrt15/4)(1-mu^2), Hmn[:,0,1] = (sqrt6/2)(1-mu^2) (test_transform_stack.py:37-48)."""
    N, I, J, lam, mus, w, P, H = orc.cached_geometry(42)
    np.testing.assert_allclose(P[:, 2, 2], 0.25 * np.sqrt(15.0) * (1 - mus ** 2), atol=1e-13)
    np.testing.assert_allclose(H[:, 0, 1], 0.5 * np.sqrt(6.0) * (1 - mus ** 2), atol=1e-13)
    # P_1^1 without Condon-Shortley phase is +sqrt(3)/2 sqrt(1-mu^2) (the odd-m sign flip).
    np.testing.assert_allclose(P[:, 1, 1], 0.5 * np.sqrt(3.0) * np.sqrt(1 - mus ** 2), atol=1e-13)


@pytest.mark.parametrize("M", [42, 63])
def test_basis_orthonormal_and_parity(M):
    N, I, J, lam, mus, w, P, H = orc.cached_geometry(M)
    for m in range(0, M + 1, 7):
        G = np.einsum("j,jn,jk->nk", w, P[:, m, m:], P[:, m, m:])
        np.testing.assert_allclose(G, np.eye(G.shape[0]), atol=2e-13)
    n = np.arange(N + 1)[None, :]
    m = np.arange(M + 1)[:, None]
    sign = np.where(n >= m, (-1.0) ** (n + m), 0.0)
    np.testing.assert_allclose(P[::-1], P * sign[None], atol=1e-13)


def test_t106_basis_reproduces_underflowed_normalization():
    """At T106 the factorial scale (2n+1)(n-m)!/(2(n+m)!) rounds to 0 for n+m >= 179 (spectral_transform.py:143-145)."""
    s = orc.scaling_table(106, 106)
    assert s[106, 106] == 0.0 and s[90, 90] == 0.0 and s[40, 106] > 0.0
    N, I, J, lam, mus, w, P, H = orc.cached_geometry(106)
    assert np.all(P[:, 106, 106] == 0.0)


# --- transforms --------------------------------------------------------------------------------

def test_fft_trunc_matches_brute_force_dft():
    rng = np.random.default_rng(1)
    I, M = 12, 4
    x = rng.standard_normal((3, I))
    got = orc.fwd_fft_trunc(x, I, M)
    i = np.arange(I)
    ref = np.stack([[np.sum(x[r] / I * np.exp(-2j * np.pi * m * i / I)) for m in range(M + 1)] for r in range(3)])
    np.testing.assert_allclose(got, ref, atol=1e-14)


def test_hermitian_inverse_recovers_real_field():
    """invrs_leg layout + invrs_fft reproduce a band-limited real field (spectral_transform.py:438-481)."""
    N, I, J, lam, mus, w, P, H = orc.cached_geometry(42)
    rng = np.random.default_rng(2)
    c = np.zeros((43, 43), complex)
    for m in range(0, 12):
        for n in range(m, 12):
            c[m, n] = rng.standard_normal() + (1j * rng.standard_normal() if m else 0)
    X = orc.invrs_leg(c, I, J, 42, 42, P)
    x = orc.invrs_fft(X, I)
    assert np.max(np.abs(x.imag)) < 1e-12
    Pw = w[:, None, None] * P
    back = orc.fwd_leg(orc.fwd_fft_trunc(x.real, I, 42), Pw)
    np.testing.assert_allclose(back, c, atol=1e-12)


def test_coriolis_closed_form_T106():
    """f = 2 omega mu transforms to fmn[0,1] = omega/sqrt(0.375) only (test_transform_stack.py:51-66)."""
    M, omega = 106, 3.2e-5
    N, I, J, lam, mus, w, P, H = orc.cached_geometry(M)
    f = 2.0 * omega * mus[:, None] * np.ones((1, I))
    fmn = orc.fwd_leg(orc.fwd_fft_trunc(f, I, M), w[:, None, None] * P)
    ref = np.zeros((M + 1, N + 1), complex)
    ref[0, 1] = omega / np.sqrt(0.375)
    np.testing.assert_allclose(fmn, ref, atol=1e-10)


def _test_setup(M, a1, Phibar):
    omega, a, dt = 7.2921159e-5, 6.37122e6, 30.0
    s = orc.build_static(M=M, dt=dt, a=a, omega=omega, g=9.8, Phibar=Phibar, taurad=86400.0, taudrag=86400.0,
                         DPhieq=4e6, K6=1.24e33, K6Phi=None, test=1)
    eta0, delta0, Phi0, U0, V0 = orc.analytic_ic(s, 1, a1)
    return s, eta0, delta0, U0, V0


def test_forward_inverse_round_trip_M63():
    """test_transform_stack.py:69-97."""
    s, eta0, delta0, U0, V0 = _test_setup(63, np.pi / 2, 3.0e3)
    Um = orc.fwd_fft_trunc(U0, s.I, s.M)
    back = orc.invrs_fft(orc.invrs_leg(orc.fwd_leg(Um, s.Pmnw), s.I, s.J, s.M, s.N, s.Pmn), s.I)
    np.testing.assert_allclose(np.real(back), U0, atol=1e-10, rtol=1e-10)


def test_wind_transform_round_trip_M106():
    """diagnostic_eta_delta then invrs_uv recovers U, V (test_transform_stack.py:100-136)."""
    s, eta0, delta0, U0, V0 = _test_setup(106, np.pi / 4, 1.0e3)
    Um = orc.fwd_fft_trunc(U0, s.I, s.M)
    Vm = orc.fwd_fft_trunc(V0, s.I, s.M)
    _, _, etamn, deltamn = orc.diagnostic_eta_delta(Um, Vm, s.fmn, s.I, s.J, s.M, s.N, s.Pmn, s.Hmn, s.w,
                                                    s.tstepcoeff, s.mJarray, s.dt)
    U, V, _, _ = orc.invrs_uv(deltamn, etamn, s.fmn, s.I, s.J, s.M, s.N, s.Pmn, s.Hmn, s.tstepcoeffmn, s.marray)
    np.testing.assert_allclose(np.real(U), U0, atol=1e-10, rtol=1e-10)
    np.testing.assert_allclose(np.real(V), V0, atol=1e-10, rtol=1e-10)


def test_vorticity_divergence_round_trip_M63():
    """invrs_uv then diagnostic_eta_delta recovers eta, delta (test_transform_stack.py:139-180)."""
    s, eta0, delta0, U0, V0 = _test_setup(63, np.pi / 4, 4.0e4)
    etamn = orc.fwd_leg(orc.fwd_fft_trunc(eta0, s.I, s.M), s.Pmnw)
    deltamn = orc.fwd_leg(orc.fwd_fft_trunc(delta0, s.I, s.M), s.Pmnw)
    U, V, _, _ = orc.invrs_uv(deltamn, etamn, s.fmn, s.I, s.J, s.M, s.N, s.Pmn, s.Hmn, s.tstepcoeffmn, s.marray)
    Um = orc.fwd_fft_trunc(np.real(U), s.I, s.M)
    Vm = orc.fwd_fft_trunc(np.real(V), s.I, s.M)
    eta, delta, _, _ = orc.diagnostic_eta_delta(Um, Vm, s.fmn, s.I, s.J, s.M, s.N, s.Pmn, s.Hmn, s.w,
                                                s.tstepcoeff, s.mJarray, s.dt)
    np.testing.assert_allclose(np.real(eta), eta0, atol=1e-10, rtol=1e-10)
    np.testing.assert_allclose(np.real(delta), delta0, atol=1e-10, rtol=1e-10)


# --- forcing / quirk properties (unit_tests/test_parity_quirks.py) ------------------------------

def test_dayside_mask_is_strict():
    out = orc.phieqfun(10.0, 4.0, np.array([-np.pi / 2, 0.0, np.pi / 2]), np.array([0.0]))
    assert out[0, 0] == 10.0 and out[0, 2] == 10.0 and out[0, 1] > 10.0


def test_rfun_q_clamp_and_no_drag():
    U = np.array([[2.0, -3.0]]); V = np.array([[1.0, 5.0]]); Q = np.array([[-2.0, 4.0]]); Phi = np.array([[10.0, 10.0]])
    F, G = orc.rfun(U, V, Q, Phi, 2.0, -1.0)
    np.testing.assert_allclose(F, [[0.0, 3.0 * 4.0 / 12.0]], atol=1e-12)
    np.testing.assert_allclose(G, [[0.0, -5.0 * 4.0 / 12.0]], atol=1e-12)
    F2, _ = orc.rfun(U, V, Q, Phi, 2.0, 10.0)
    np.testing.assert_allclose(F2, F - U / 10.0, atol=1e-12)


def _tiny():
    """M=N=2, J=4, I=8 setup with seed-7 random Fourier fields (test_parity_quirks.py:7-49)."""
    M = N = 2
    J, I = 4, 8
    mus, w = orc.gauss_legendre(J)
    P, H = orc.pmn_hmn(J, M, N, mus)
    dt, a = 30.0, 6.37122e6
    s = orc.Static(M=M, N=N, I=I, J=J, dt=dt, a=a, omega=7.29e-5, g=9.8, Phibar=3e3, taurad=86400.0,
                   taudrag=86400.0, lambdas=orc.build_lambdas(I), mus=mus, w=w, Pmn=P, Hmn=H,
                   Pmnw=w[:, None, None] * P, Hmnw=w[:, None, None] * H, fmn=orc.coriolismn(M, 7.29e-5),
                   tstepcoeff=orc.tstepcoeff(J, M, dt, mus, a), tstepcoeff2=orc.tstepcoeff2(J, M, dt, a),
                   tstepcoeffmn=orc.tstepcoeffmn(M, N, a), marray=orc.marray(M, N), mJarray=orc.mJarray(J, M),
                   narray=orc.narray(M, N), sigma=np.ones((M + 1, N + 1)), sigmaPhi=np.ones((M + 1, N + 1)),
                   Phieq=np.zeros((J, I)))
    rng = np.random.default_rng(7)
    names = ("etam0", "etam1", "deltam0", "deltam1", "Phim0", "Phim1", "Am", "Bm", "Cm", "Dm", "Em", "Fm", "Gm",
             "Um", "Vm", "PhiFm")
    arr = {k: rng.standard_normal((J, M + 1)) for k in names}
    z = np.zeros((J, I))
    st = orc.OracleState(etam_prev=arr["etam0"], etam_curr=arr["etam1"], deltam_prev=arr["deltam0"],
                         deltam_curr=arr["deltam1"], Phim_prev=arr["Phim0"], Phim_curr=arr["Phim1"],
                         eta_prev=z, eta_curr=z, delta_prev=z, delta_curr=z, Phi_prev=z, Phi_curr=z, U_curr=z, V_curr=z,
                         Um_curr=arr["Um"], Vm_curr=arr["Vm"], Am_curr=arr["Am"], Bm_curr=arr["Bm"], Cm_curr=arr["Cm"],
                         Dm_curr=arr["Dm"], Em_curr=arr["Em"], PhiFm_curr=arr["PhiFm"], Fm_curr=arr["Fm"],
                         Gm_curr=arr["Gm"], Rum_lag=z, Rvm_lag=z)
    return s, st


def test_explicit_delta_is_carry_only():
    """explicit_tdiff.py:134-138: without forcing or diffusion delta = Pw(deltam0) (test_parity_quirks.py:85-135)."""
    s, st = _tiny()
    _, delta, _ = orc._explicit(st, s, forcflag=False, diffflag=False)
    np.testing.assert_allclose(delta, orc.fwd_leg(st.deltam_prev, s.Pmnw), atol=1e-12)
    # and it ignores every tendency input
    st2 = orc.replace(st, Am_curr=st.Am_curr * 3, Cm_curr=st.Cm_curr - 1, Em_curr=st.Em_curr + 2)
    np.testing.assert_allclose(orc._explicit(st2, s, False, False)[1], delta, atol=1e-12)


def test_modeuler_delta_uses_forced_terms_even_when_unforced():
    """modEuler_tdiff.py:166-168: Bm+Fm / Am-Gm in delta regardless of forcflag; forcflag adds only PhiF."""
    s, st = _tiny()
    _, d_forced, _ = orc._modeuler(st, s, True, False)
    _, d_unforced, _ = orc._modeuler(st, s, False, False)
    phif = (s.narray * orc.fwd_leg((s.dt / 2.0) * st.PhiFm_curr, s.Pmnw)) / s.a ** 2
    np.testing.assert_allclose(d_unforced, d_forced - phif, atol=1e-10)
    st0 = orc.replace(st, Fm_curr=0 * st.Fm_curr, Gm_curr=0 * st.Gm_curr)
    assert not np.allclose(orc._modeuler(st0, s, False, False)[1], d_unforced)


def test_modeuler_eta_forced_unforced_asymmetry():
    """modEuler_tdiff.py:253-256: eta uses tstepcoeff when forced and tstepcoeff/2 otherwise."""
    s, st = _tiny()
    st0 = orc.replace(st, Fm_curr=0 * st.Fm_curr, Gm_curr=0 * st.Gm_curr)
    e_f, _, _ = orc._modeuler(st0, s, True, False)
    e_u, _, _ = orc._modeuler(st0, s, False, False)
    base = orc.fwd_leg(st.etam_curr, s.Pmnw)
    np.testing.assert_allclose(e_u - base, 0.5 * (e_f - base), atol=1e-12)
    assert not np.allclose(e_f, e_u)


def test_modeuler_phi_quarter_coefficients():
    """modEuler_tdiff.py:72-74: Phi uses tstepcoeff/4 (a /2 scaling of dt would double the C, D terms)."""
    s, st = _tiny()
    zero = {k: 0 * getattr(st, k) for k in ("Am_curr", "Bm_curr", "Dm_curr", "Em_curr", "Fm_curr", "Gm_curr",
                                            "PhiFm_curr", "deltam_curr", "Phim_curr")}
    stC = orc.replace(st, **zero)
    _, _, Phi = orc._modeuler(stC, s, True, False)
    expect = -orc.fwd_leg((s.tstepcoeff / 4.0) * 1j * s.mJarray * st.Cm_curr, s.Pmnw)
    np.testing.assert_allclose(Phi, expect, atol=1e-12)


def test_members_are_independent():
    """An ensemble is independent members: running a member alone gives the same state (test_vmap_smoke.py)."""
    kw = dict(M=42, dt=600.0, tmax=5, Phibar=3e5, omega=7.292e-5, a=6.37122e6, diagnostics=False,
              return_history=False)
    a = orc.run_model_scan(DPhieq=2e6, **kw)["last_state"].Phi_curr
    b = orc.run_model_scan(DPhieq=6e6, **kw)["last_state"].Phi_curr
    a2 = orc.run_model_scan(DPhieq=2e6, **kw)["last_state"].Phi_curr
    assert np.max(np.abs(a - b)) > 0 and np.array_equal(a, a2)


# ---------------------------------------------------------------------------------------------
# Semi-implicit scheme, exponential hyperdiffusion and the RAW filter (semi_implicit_tdiff.py,
# filters.py:154-213, model.py:1007-1143); properties of unit_tests/test_semi_implicit.py and
# unit_tests/test_raw_filter.py.

def test_si_solve_satisfies_the_coupled_system():
    """The closed form solves (1 + kd) delta - b Phi = delta*, c delta + (1 + kr) Phi = Phi* for every mode."""
    rng = np.random.default_rng(4)
    ds = rng.standard_normal(50) + 1j * rng.standard_normal(50)
    ps = 1e4 * (rng.standard_normal(50) + 1j * rng.standard_normal(50))
    b, c = 10 ** rng.uniform(-8, -2, 50), 10 ** rng.uniform(3, 8, 50)
    kd, kr = 10 ** rng.uniform(-4, 1, 50), 10 ** rng.uniform(-4, 1, 50)
    d, p = orc.si_solve(ds, ps, b, c, kd, kr)
    np.testing.assert_allclose((1 + kd) * d - b * p, ds, rtol=1e-12, atol=1e-12 * np.abs(ds).max())
    np.testing.assert_allclose(c * d + (1 + kr) * p, ps, rtol=1e-12, atol=1e-12 * np.abs(ps).max())


def test_exponential_hyperdiffusion_is_the_exact_decay():
    """sigma6_exponential is exp(-2 dt K6 k_n): a semigroup in dt, neutral at n = 0 and 1, and equal to
    the implicit factor 1/(1 + 2 dt K6 k_n) to first order in the damping."""
    M, K6, a = 42, 1.24e33, 8.2e7
    s1 = orc.sigma6_exponential(M, M, K6, a, 100.0)[0]
    s2 = orc.sigma6_exponential(M, M, K6, a, 200.0)[0]
    np.testing.assert_allclose(s1 * s1, s2, rtol=1e-14)
    assert s1[0] == 1.0 and abs(s1[1] - 1.0) < 1e-15
    tiny = orc.sigma6_exponential(M, M, K6, a, 1e-6)[0]
    impl = orc.sigma6(M, M, K6, a, 1e-6)[0]
    eps = 1.0 - impl
    np.testing.assert_allclose(tiny, impl, atol=2.0 * float(np.max(eps * eps)) + 1e-16)
    sp = orc.sigma6Phi_exponential(M, M, K6, a, 100.0)[0]
    assert sp[0] == 1.0 and sp[1] < 1.0  # no 8/a^6 offset for Phi


_RAW_COMMON = dict(M=42, dt=240.0, tmax=40, Phibar=3.0e5, omega=3.2e-5, a=8.2e7, taurad=36000.0, taudrag=21600.0,
                   DPhieq=1.0e6, test=None, diagnostics=False)


def test_raw_filter_alpha_one_is_bit_identical_to_classic_ra():
    """unit_tests/test_raw_filter.py:37-46: williams_alpha = 1 reproduces the classic filter bit for bit."""
    base = orc.run_model_scan(**_RAW_COMMON)
    raw = orc.run_model_scan(**_RAW_COMMON, raw_filter=True, williams_alpha=1.0)
    for k in ("eta", "delta", "Phi", "U", "V"):
        np.testing.assert_array_equal(base["outs"][k], raw["outs"][k], err_msg=k)


def test_raw_filter_williams_optimum_changes_trajectory():
    """unit_tests/test_raw_filter.py:49-56."""
    base = orc.run_model_scan(**_RAW_COMMON)
    raw = orc.run_model_scan(**_RAW_COMMON, raw_filter=True, williams_alpha=0.53)
    d = np.max(np.abs(base["outs"]["Phi"][-1] - raw["outs"]["Phi"][-1]))
    assert np.isfinite(d) and d > 0.0


def test_semi_implicit_runs_finite_default_regime():
    """unit_tests/test_semi_implicit.py:52-57 (dt = 240 s, 118 steps)."""
    out = orc.run_model_scan(**dict(_RAW_COMMON, tmax=120), semi_implicit=True)
    assert out["dead_first_idx"] == -1
    for k in ("eta", "delta", "Phi", "U", "V"):
        assert np.all(np.isfinite(out["outs"][k][-1])), k


def test_semi_implicit_preserves_steady_advection():
    """unit_tests/test_semi_implicit.py:108-119: Williamson test 1 (solid-body advection) keeps its
    steady vorticity to roundoff over 6 hours at dt = 600 s."""
    out = orc.run_model_scan(M=42, dt=600.0, tmax=2 + 36, Phibar=3.0e3, omega=7.2921159e-5, a=6.37122e6, test=1,
                             forcflag=False, semi_implicit=True)
    eta = out["outs"]["eta"]
    drift = float(np.sqrt(np.mean((eta[-1] - eta[0]) ** 2))) / float(np.std(eta[0]))
    assert drift < 1e-8


def test_semi_implicit_and_raw_reject_the_explicit_scheme():
    with pytest.raises(ValueError):
        orc.run_model_scan(**_RAW_COMMON, semi_implicit=True, expflag=True)
    with pytest.raises(ValueError):
        orc.run_model_scan(**_RAW_COMMON, raw_filter=True, expflag=True)


# ----------------------------------------------------------------------------
# Explicit scheme (explicit_tdiff.py:20-233) pinned by identities the mathematics fixes.
# delta_mn(U, V) and zeta_mn(U, V) come from diagnostic_eta_delta, itself pinned by the round trips
# of unit_tests/test_transform_stack.py:100-180 above; each identity fails on a dropped term, a wrong
# sign or a wrong coefficient of the flux, continuity or drag terms.
# ----------------------------------------------------------------------------


def _explicit_case(seed=3, M=42, nmax=12):
    """T42 static (dt = 120 s, taudrag = 5e4 s) and band-limited Fourier winds U_m, V_m from random
    spectral divergence and vorticity (n <= nmax, real at m = 0)."""
    s = orc.build_static(M=M, dt=120.0, a=8.2e7, omega=3.2e-5, g=9.8, Phibar=3e5, taurad=86400.0,
                         taudrag=5.0e4, DPhieq=3e5, K6=1.24e33, K6Phi=None, test=None)
    rng = np.random.default_rng(seed)

    def spec():
        c = np.zeros((M + 1, M + 1), complex)
        for m in range(nmax + 1):
            for n in range(max(m, 1), nmax + 1):
                c[m, n] = rng.standard_normal() + (1j * rng.standard_normal() if m else 0.0)
        return 1e-6 * c

    d_mn, z_mn = spec(), spec()
    _, _, Um, Vm = orc.invrs_uv(d_mn, z_mn + s.fmn, s.fmn, s.I, s.J, M, M, s.Pmn, s.Hmn, s.tstepcoeffmn, s.marray)
    return s, Um, Vm


def _div_vort(s, Um, Vm):
    _, _, eta_mn, delta_mn = orc.diagnostic_eta_delta(Um, Vm, s.fmn, s.I, s.J, s.M, s.N, s.Pmn, s.Hmn, s.w,
                                                      s.tstepcoeff, s.mJarray, s.dt)
    return delta_mn, eta_mn - s.fmn


def _explicit_state(s, **fields):
    z = np.zeros((s.J, s.M + 1), complex)
    g = np.zeros((s.J, s.I))
    base = {k: z for k in ("etam_prev", "etam_curr", "deltam_prev", "deltam_curr", "Phim_prev", "Phim_curr",
                           "Um_curr", "Vm_curr", "Am_curr", "Bm_curr", "Cm_curr", "Dm_curr", "Em_curr",
                           "PhiFm_curr", "Fm_curr", "Gm_curr", "Rum_lag", "Rvm_lag")}
    base.update({k: g for k in ("eta_prev", "eta_curr", "delta_prev", "delta_curr", "Phi_prev", "Phi_curr",
                                "U_curr", "V_curr")})
    base.update(fields)
    return orc.OracleState(**base)


def test_explicit_flux_of_uniform_scalar_is_divergence():
    """Uniform q0 carried by band-limited winds: the flux terms of Phi (C = U q0, D = V q0) and of eta
    (A = U q0, B = V q0) each equal -2 dt q0 delta_mn(U, V) (explicit_tdiff.py:63-70, :206-212)."""
    s, Um, Vm = _explicit_case()
    delta_mn, _ = _div_vort(s, Um, Vm)
    q0 = 3.7e4
    st = _explicit_state(s, Cm_curr=q0 * Um, Dm_curr=q0 * Vm)
    _, _, phi = orc._explicit(st, s, forcflag=False, diffflag=False)
    scale = np.max(np.abs(2 * s.dt * q0 * delta_mn))
    np.testing.assert_allclose(phi, -2.0 * s.dt * q0 * delta_mn, rtol=0, atol=1e-12 * scale)
    st = _explicit_state(s, Am_curr=q0 * Um, Bm_curr=q0 * Vm)
    eta, _, _ = orc._explicit(st, s, forcflag=False, diffflag=False)
    np.testing.assert_allclose(eta, -2.0 * s.dt * q0 * delta_mn, rtol=0, atol=1e-12 * scale)


def test_explicit_continuity_term():
    """Only the divergence carried: Phi_new - Pw(Phi_prev) = -2 dt Phibar delta_mn (explicit_tdiff.py:71)."""
    s, Um, Vm = _explicit_case(seed=5)
    delta_mn, _ = _div_vort(s, Um, Vm)
    deltam = orc.invrs_leg(delta_mn, s.I, s.J, s.M, s.N, s.Pmn)[:, : s.M + 1]
    rng = np.random.default_rng(9)
    phim_prev = orc.invrs_leg(1e3 * (rng.standard_normal((s.M + 1, s.N + 1)) * np.triu(np.ones((s.M + 1, s.N + 1)))),
                              s.I, s.J, s.M, s.N, s.Pmn)[:, : s.M + 1].real.astype(complex)
    st = _explicit_state(s, deltam_curr=deltam, Phim_prev=phim_prev)
    _, _, phi = orc._explicit(st, s, forcflag=False, diffflag=False)
    lhs = phi - orc.fwd_leg(phim_prev, s.Pmnw)
    rhs = -2.0 * s.dt * s.Phibar * delta_mn
    np.testing.assert_allclose(lhs, rhs, rtol=0, atol=1e-12 * np.max(np.abs(rhs)))


def test_explicit_drag_only_forcing():
    """Quirk 5 (DEVELOPMENT.md:124-128; explicit_tdiff.py:145-152, :216-224): with F = -U/taudrag,
    G = -V/taudrag and no other forcing, the forced delta increment is -(4 dt / taudrag) delta_mn(U, V);
    the forced eta increment is 0 for zonal winds only and -(4 dt / taudrag) zeta_mn(0, V) for
    meridional winds only (the extra U/taudrag, V/taudrag terms double the drag in delta and cancel
    the U part of it in eta)."""
    s, Um, Vm = _explicit_case(seed=11)
    tau = s.taudrag
    z = np.zeros_like(Um)
    for U, V in ((Um, Vm), (Um, z), (z, Vm)):
        st = _explicit_state(s, Um_curr=U, Vm_curr=V, Fm_curr=-U / tau, Gm_curr=-V / tau)
        eta_f, delta_f, _ = orc._explicit(st, s, forcflag=True, diffflag=False)
        eta_u, delta_u, _ = orc._explicit(st, s, forcflag=False, diffflag=False)
        delta_mn, zeta_mn = _div_vort(s, U, V)
        d_inc = delta_f - delta_u
        e_inc = eta_f - eta_u
        k = 4.0 * s.dt / tau
        scale = k * max(np.max(np.abs(delta_mn)), np.max(np.abs(zeta_mn)))
        np.testing.assert_allclose(d_inc, -k * delta_mn, rtol=0, atol=1e-12 * scale)
        if V is z:
            np.testing.assert_allclose(e_inc, 0.0, rtol=0, atol=1e-12 * scale)
        if U is z:
            np.testing.assert_allclose(e_inc, -k * zeta_mn, rtol=0, atol=1e-12 * scale)


def test_robert_asselin_weight_reaches_only_the_prev_level():
    """The modified-Euler scan reads the current levels only, so the Robert-Asselin weight alpha
    (the RA filter of _step_once, model.py:795-1194, on the physical prev level) leaves every current level
    bit for bit unchanged: the gradient of a loss of Phi_curr in alpha is zero.  The prev level
    does move."""
    import numpy as np

    from oracle import swampe_oracle as orc

    runs = [orc.run_model_scan(M=42, dt=600.0, tmax=8, Phibar=3.0e5, omega=7.292e-5, a=6.37122e6, alpha=al,
                               diagnostics=False, return_history=False)["last_state"] for al in (0.01, 0.3)]
    for f in ("Phi_curr", "eta_curr", "delta_curr"):
        assert np.array_equal(getattr(runs[0], f), getattr(runs[1], f)), f
    assert not np.array_equal(runs[0].Phi_prev, runs[1].Phi_prev)
