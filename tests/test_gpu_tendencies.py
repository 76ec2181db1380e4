"""The reference's scheme tendencies on the GPU (modEuler_tdiff.py, explicit_tdiff.py): the quirk
tests of unit_tests/test_parity_quirks.py:85-317 on its small configuration (M = N = 2, J = 4,
I = 8, seed 7), and all six updates against the oracle's modified-Euler and explicit updates (the
restatement the fixtures pin end to end) at that size and at T42, with and without forcing and
diffusion; the one-member step time_stepping.tstepping of both schemes; and the semi-implicit
step (semi_implicit_tdiff.si_timestep, time_stepping.tstepping_semi_implicit)."""

from __future__ import annotations

from types import SimpleNamespace

import numpy as np
import pytest

from oracle import swampe_oracle as orc

pytestmark = pytest.mark.gpu

FIELDS = ("etam0", "etam1", "deltam0", "deltam1", "Phim0", "Phim1", "Am", "Bm", "Cm", "Dm", "Em", "Fm", "Gm", "Um", "Vm")
DT, A, PHIBAR, TAURAD, TAUDRAG = 30.0, 6.37122e6, 3.0e3, 86400.0, 86400.0


def _np(t):
    return t.detach().cpu().numpy()


def _setup(M, J, I, seed=7, sigma_K6=None):
    """test_parity_quirks.py:7-49 through this package's own setup functions; random Fourier fields
    (complex at T42, real as in the reference's small case)."""
    from my_swampe_amd import filters
    from my_swampe_amd import spectral_transform as st
    from my_swampe_amd import time_stepping as ts

    N = M
    mus, w = st.gauss_legendre(J)
    Pmn, Hmn = st.PmnHmn(J, M, N, mus)
    rng = np.random.default_rng(seed)
    shape = (J, M + 1)
    if M == 2:
        arrays = {name: rng.standard_normal(shape) for name in FIELDS}
        arrays["PhiFm"] = rng.standard_normal(shape)
    else:
        arrays = {name: rng.standard_normal(shape) + 1j * rng.standard_normal(shape) for name in FIELDS + ("PhiFm",)}
    if sigma_K6 is None:
        sigma, sigmaPhi = np.ones((M + 1, N + 1)), np.ones((M + 1, N + 1))
    else:
        sigma, sigmaPhi = filters.sigma6(M, N, sigma_K6, A, DT), filters.sigma6Phi(M, N, sigma_K6, A, DT)
    return dict(M=M, N=N, J=J, I=I, Pmn=Pmn, Hmn=Hmn, Pmnw=st.weighted_legendre_basis(Pmn, w),
                Hmnw=st.weighted_legendre_basis(Hmn, w), tstepcoeff1=ts.tstepcoeff(J, M, DT, mus, A),
                tstepcoeff2=ts.tstepcoeff2(J, M, DT, A), mJarray=ts.mJarray(J, M), narray=ts.narray(M, N),
                sigma=sigma, sigmaPhi=sigmaPhi, arrays=arrays)


def _args(s, forcflag, diffflag):
    a = s["arrays"]
    return (a["etam0"], a["etam1"], a["deltam0"], a["deltam1"], a["Phim0"], a["Phim1"], s["I"], s["J"], s["M"],
            s["N"], a["Am"], a["Bm"], a["Cm"], a["Dm"], a["Em"], a["Fm"], a["Gm"], a["Um"], a["Vm"], s["Pmn"],
            s["Pmnw"], s["Hmnw"], s["tstepcoeff1"], s["tstepcoeff2"], s["mJarray"], s["narray"], a["PhiFm"], DT, A,
            PHIBAR, TAURAD, TAUDRAG, forcflag, diffflag, s["sigma"], s["sigmaPhi"], None, 2)


def _oracle(scheme, s, forcflag, diffflag):
    """The oracle's (eta, delta, Phi) spectra for the same fields, through duck-typed state and
    static namespaces."""
    a = s["arrays"]
    state = SimpleNamespace(Am_curr=a["Am"], Bm_curr=a["Bm"], Cm_curr=a["Cm"], Dm_curr=a["Dm"], Em_curr=a["Em"],
                            Fm_curr=a["Fm"], Gm_curr=a["Gm"], PhiFm_curr=a["PhiFm"], Um_curr=a["Um"], Vm_curr=a["Vm"],
                            etam_curr=a["etam1"], deltam_curr=a["deltam1"], Phim_curr=a["Phim1"],
                            etam_prev=a["etam0"], deltam_prev=a["deltam0"], Phim_prev=a["Phim0"])
    static = SimpleNamespace(Pmnw=s["Pmnw"], Hmnw=s["Hmnw"], narray=s["narray"], mJarray=s["mJarray"], dt=DT, a=A,
                             Phibar=PHIBAR, taudrag=TAUDRAG, tstepcoeff=s["tstepcoeff1"], tstepcoeff2=s["tstepcoeff2"],
                             sigma=s["sigma"], sigmaPhi=s["sigmaPhi"])
    fn = orc._modeuler if scheme == "modeuler" else orc._explicit
    return fn(state, static, forcflag, diffflag)


@pytest.mark.parametrize("scheme", ["modeuler", "explicit"])
@pytest.mark.parametrize("size", ["small", "T42"])
@pytest.mark.parametrize("forcflag,diffflag", [(True, True), (False, False), (True, False)])
def test_tendencies_vs_oracle(scheme, size, forcflag, diffflag, torch_gpu):
    """phi/delta/eta_timestep against the oracle's spectra within 1e-12 of each spectrum's largest
    value; the grid and the truncated Fourier coefficients they return are invrs_fft(invrs_leg(.))
    of the spectrum and its first M+1 columns (against the oracle's inverse transforms)."""
    from my_swampe_amd import explicit_tdiff, modEuler_tdiff

    print(f"[gpu-case] tendencies {scheme} {size} forc={forcflag} diff={diffflag}", flush=True)
    s = _setup(2, 4, 8) if size == "small" else _setup(42, 64, 128, seed=11, sigma_K6=1.0e27)
    mod = modEuler_tdiff if scheme == "modeuler" else explicit_tdiff
    args = _args(s, forcflag, diffflag)
    ref = dict(zip(("eta", "delta", "Phi"), _oracle(scheme, s, forcflag, diffflag)))
    for name, fn in (("Phi", mod.phi_timestep), ("delta", mod.delta_timestep), ("eta", mod.eta_timestep)):
        spec, grid, trunc = (_np(x) for x in fn(*args))
        r = ref[name]
        assert spec.shape == (s["M"] + 1, s["N"] + 1) and grid.shape == (s["J"], s["I"])
        assert trunc.shape == (s["J"], s["M"] + 1)
        scale = float(np.max(np.abs(r)))
        assert float(np.max(np.abs(spec - r))) <= 1e-12 * scale, (name, float(np.max(np.abs(spec - r))) / scale)
        Xm = orc.invrs_leg(r, s["I"], s["J"], s["M"], s["N"], s["Pmn"])
        gr = orc.invrs_fft(Xm, s["I"])
        assert float(np.max(np.abs(grid - gr))) <= 1e-11 * float(np.max(np.abs(gr))), name
        assert float(np.max(np.abs(trunc - Xm[:, : s["M"] + 1]))) <= 1e-12 * float(np.max(np.abs(Xm))), name


def test_explicit_delta_is_carry_only_without_forcing_or_diffusion(torch_gpu):
    """test_parity_quirks.py:85-135: the explicit delta update is Pw(deltam0) alone."""
    from my_swampe_amd import explicit_tdiff as exp
    from my_swampe_amd import spectral_transform as st

    s = _setup(2, 4, 8)
    got, _, _ = exp.delta_timestep(*_args(s, False, False))
    expected = st.fwd_leg_w(s["arrays"]["deltam0"], s["Pmnw"])
    np.testing.assert_allclose(_np(got), _np(expected), rtol=0.0, atol=1e-12)


def test_modeuler_delta_uses_forced_terms_even_when_forcflag_false(torch_gpu):
    """test_parity_quirks.py:138-216: the modified-Euler delta takes (B + F, A - G) with
    forcflag = False."""
    import torch

    from my_swampe_amd import modEuler_tdiff as mod
    from my_swampe_amd import spectral_transform as st

    s = _setup(2, 4, 8)
    a = s["arrays"]
    got, _, _ = mod.delta_timestep(*_args(s, False, False))
    t1, t2 = s["tstepcoeff1"] / 4.0, s["tstepcoeff2"] / 4.0
    nn = torch.tensor(s["narray"], device="cuda")
    stack = np.stack((a["deltam1"], t1 * 1j * s["mJarray"] * (a["Bm"] + a["Fm"]), t2 * a["Phim1"], t2 * a["Em"],
                      t1 * 1j * s["mJarray"] * a["Cm"]))
    d1, d2, d4, d5, p2 = st.fwd_leg_w_batch(stack, s["Pmnw"]).unbind(0)
    d3, p3 = st.fwd_leg_w_batch(np.stack((t1 * (a["Am"] - a["Gm"]), t1 * a["Dm"])), s["Hmnw"]).unbind(0)
    expected = (d1 + d2 + d3 + nn * d4 + nn * d5 + (nn * (p2 + p3)) / (2.0 * (A ** 2))
                - PHIBAR * (nn * d1) / (A ** 2))
    np.testing.assert_allclose(_np(got), _np(expected), rtol=0.0, atol=1e-10)


def test_modeuler_eta_forced_unforced_asymmetry(torch_gpu):
    """test_parity_quirks.py:219-317: eta takes tstepcoeff and (A - G, B + F) when forced,
    tstepcoeff / 2 and (A, B) when not, and the two differ."""
    from my_swampe_amd import modEuler_tdiff as mod
    from my_swampe_amd import spectral_transform as st

    s = _setup(2, 4, 8)
    a = s["arrays"]
    forced, _, _ = mod.eta_timestep(*_args(s, True, False))
    unforced, _, _ = mod.eta_timestep(*_args(s, False, False))
    t1, Pw, Hw, mJ = s["tstepcoeff1"], s["Pmnw"], s["Hmnw"], s["mJarray"]
    ef = (st.fwd_leg_w(a["etam1"], Pw) - st.fwd_leg_w(t1 * 1j * mJ * (a["Am"] - a["Gm"]), Pw)
          + st.fwd_leg_w(t1 * (a["Bm"] + a["Fm"]), Hw))
    eu = (st.fwd_leg_w(a["etam1"], Pw) - st.fwd_leg_w((t1 / 2.0) * 1j * mJ * a["Am"], Pw)
          + st.fwd_leg_w((t1 / 2.0) * a["Bm"], Hw))
    np.testing.assert_allclose(_np(forced), _np(ef), rtol=0.0, atol=1e-12)
    np.testing.assert_allclose(_np(unforced), _np(eu), rtol=0.0, atol=1e-12)
    assert not np.allclose(_np(forced), _np(unforced))


def test_tendency_argument_errors(torch_gpu):
    """A wrong argument count raises TypeError and a wrongly shaped field ValueError, before any
    launch."""
    from my_swampe_amd import modEuler_tdiff as mod

    s = _setup(2, 4, 8)
    args = list(_args(s, True, True))
    with pytest.raises(TypeError):
        mod.phi_timestep(*args[:-1])
    args[0] = np.zeros((4, 2))
    with pytest.raises(ValueError):
        mod.phi_timestep(*args)


@pytest.mark.parametrize("expflag", [False, True])
def test_tstepping_vs_oracle(expflag, torch_gpu):
    """time_stepping.tstepping at T42 (random Fourier fields, forced, diffused): the 13 arrays
    against the oracle's scheme updates, their inverse transforms and the oracle's wind inversion
    (spectra within 1e-12, grids and Fourier coefficients within 1e-11 of each array's largest
    value)."""
    from my_swampe_amd import initial_conditions as ic
    from my_swampe_amd import time_stepping as ts

    print(f"[gpu-case] tstepping expflag={expflag}", flush=True)
    s = _setup(42, 64, 128, seed=13, sigma_K6=1.0e27)
    a = s["arrays"]
    M, N, I, J = s["M"], s["N"], s["I"], s["J"]
    fmn = ic.coriolismn(M, 7.2921159e-5)
    tcmn, marr = ts.tstepcoeffmn(M, N, A), ts.marray(M, N)
    got = ts.tstepping(a["etam0"], a["etam1"], a["deltam0"], a["deltam1"], a["Phim0"], a["Phim1"], I, J, M, N, a["Am"],
                       a["Bm"], a["Cm"], a["Dm"], a["Em"], a["Fm"], a["Gm"], a["Um"], a["Vm"], fmn, s["Pmn"], s["Hmn"],
                       s["Pmnw"], s["Hmnw"], s["tstepcoeff1"], s["tstepcoeff2"], tcmn, marr, s["mJarray"],
                       s["narray"], a["PhiFm"], DT, A, PHIBAR, TAURAD, TAUDRAG, True, True, expflag, s["sigma"],
                       s["sigmaPhi"], None, 2)
    assert len(got) == 13
    eta_mn, delta_mn, Phi_mn = _oracle("explicit" if expflag else "modeuler", s, True, True)
    ref = []
    for spec in (eta_mn, delta_mn, Phi_mn):
        Xm = orc.invrs_leg(spec, I, J, M, N, s["Pmn"])
        ref += [spec, orc.invrs_fft(Xm, I), Xm[:, : M + 1]]
    ref += list(orc.invrs_uv(delta_mn, eta_mn, fmn, I, J, M, N, s["Pmn"], s["Hmn"], tcmn, marr))
    names = ("etamn", "eta", "etam", "deltamn", "delta", "deltam", "Phimn", "Phi", "Phim", "U", "V", "Um", "Vm")
    for name, g, r in zip(names, got, ref):
        g = _np(g)
        assert g.shape == r.shape, name
        tol = 1e-12 if name.endswith("mn") else 1e-11
        assert float(np.max(np.abs(g - r))) <= tol * float(np.max(np.abs(r))), name


@pytest.mark.parametrize("relax,forcflag,diffflag", [(True, True, True), (False, True, True), (False, False, False)])
@pytest.mark.parametrize("size", ["small", "T42"])
def test_semi_implicit_step_vs_oracle(relax, forcflag, diffflag, size, torch_gpu):
    """semi_implicit_tdiff.si_timestep (and time_stepping.tstepping_semi_implicit, the same call)
    against the oracle's semi-implicit update (semi_implicit_tdiff.py:79-264; si_alpha = 0.5,
    lagged momentum remainders, exponential hyperdiffusion): spectra within 1e-12, grids, Fourier
    coefficients and winds within 1e-11 of each array's largest value."""
    import torch

    from my_swampe_amd import filters
    from my_swampe_amd import initial_conditions as ic
    from my_swampe_amd import time_stepping as ts
    from my_swampe_amd.semi_implicit_tdiff import si_timestep

    print(f"[gpu-case] semi-implicit {size} relax={relax} forc={forcflag} diff={diffflag}", flush=True)
    s = _setup(2, 4, 8) if size == "small" else _setup(42, 64, 128, seed=17)
    M, N, I, J = s["M"], s["N"], s["I"], s["J"]
    a = s["arrays"]
    rng = np.random.default_rng(23)
    ru, rv = (rng.standard_normal((J, M + 1)) + 1j * rng.standard_normal((J, M + 1)) for _ in range(2))
    fmn = ic.coriolismn(M, 7.2921159e-5)
    tcmn, marr = ts.tstepcoeffmn(M, N, A), ts.marray(M, N)
    sig_e, sigP_e = filters.sigma6_exponential(M, N, 1.0e27, A, DT), filters.sigma6Phi_exponential(M, N, 1.0e27, A, DT)
    args = (ru, rv, a["etam0"], a["etam1"], a["deltam0"], a["deltam1"], a["Phim0"], a["Phim1"], I, J, M, N, a["Am"],
            a["Bm"], a["Cm"], a["Dm"], a["Em"], a["Fm"], a["Gm"], a["Um"], a["Vm"], fmn, s["Pmn"], s["Hmn"], s["Pmnw"],
            s["Hmnw"], s["tstepcoeff1"], s["tstepcoeff2"], tcmn, marr, s["mJarray"], s["narray"], a["PhiFm"], DT, A,
            PHIBAR, TAURAD, TAUDRAG, forcflag, diffflag, relax, sig_e, sigP_e, 0.5, None, 2)
    got = si_timestep(*args)
    again = ts.tstepping_semi_implicit(*args)
    assert len(got) == 13 and all(torch.equal(x, y) for x, y in zip(got, again))
    a_ = s["arrays"]
    state = SimpleNamespace(Am_curr=a_["Am"], Bm_curr=a_["Bm"], Cm_curr=a_["Cm"], Dm_curr=a_["Dm"], Em_curr=a_["Em"],
                            Fm_curr=a_["Fm"], Gm_curr=a_["Gm"], PhiFm_curr=a_["PhiFm"], Rum_lag=ru, Rvm_lag=rv,
                            etam_curr=a_["etam1"], deltam_curr=a_["deltam1"], Phim_curr=a_["Phim1"],
                            etam_prev=a_["etam0"], deltam_prev=a_["deltam0"], Phim_prev=a_["Phim0"])
    static = SimpleNamespace(Pmnw=s["Pmnw"], Hmnw=s["Hmnw"], mJarray=s["mJarray"], narray=s["narray"], fmn=fmn, dt=DT,
                             a=A, Phibar=PHIBAR, taurad=TAURAD, taudrag=TAUDRAG, tstepcoeff=s["tstepcoeff1"],
                             sigma_exp=sig_e, sigmaPhi_exp=sigP_e)
    eta_mn, delta_mn, Phi_mn = orc._semi_implicit(state, static, forcflag, diffflag, relax, 0.5)
    ref = []
    for spec in (eta_mn, delta_mn, Phi_mn):
        Xm = orc.invrs_leg(spec, I, J, M, N, s["Pmn"])
        ref += [spec, orc.invrs_fft(Xm, I), Xm[:, : M + 1]]
    ref += list(orc.invrs_uv(delta_mn, eta_mn, fmn, I, J, M, N, s["Pmn"], s["Hmn"], tcmn, marr))
    names = ("etamn", "eta", "etam", "deltamn", "delta", "deltam", "Phimn", "Phi", "Phim", "U", "V", "Um", "Vm")
    for name, g, r in zip(names, got, ref):
        g = _np(g)
        assert g.shape == r.shape, name
        tol = 1e-12 if name.endswith("mn") else 1e-11
        assert float(np.max(np.abs(g - r))) <= tol * float(np.max(np.abs(r))), name
