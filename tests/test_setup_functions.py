"""The reference's setup functions on the host (no GPU): initial_conditions.py (spectral_params,
test1_init, state_var_init, velocity_init, ABCDE_init, coriolismn), the coefficient arrays and
RMS wind of time_stepping.py:491-576, filters.py and forcing.py, against the oracle (pinned by the
upstream fixtures, whose runs start from and step with these fields) bit for bit, and against
what the mathematics fixes."""

from __future__ import annotations

import numpy as np
import pytest

from oracle import swampe_oracle as orc


@pytest.mark.parametrize("M", [42, 63, 106])
def test_spectral_params(M):
    from my_swampe_amd import initial_conditions as ic

    got, ref = ic.spectral_params(M), orc.spectral_params(M)
    assert got[:4] == tuple(ref[:4]) and got[0] == M
    for g, r in zip(got[4:], ref[4:]):
        assert np.array_equal(g, np.asarray(r))


def test_spectral_params_refuses_other_truncations():
    from my_swampe_amd import initial_conditions as ic

    with pytest.raises(ValueError, match="Unsupported value of M=50. Only 42, 63, and 106 are supported."):
        ic.spectral_params(50)


@pytest.mark.parametrize("M,test", [(42, None), (42, 1), (63, 1), (63, 2)])
def test_initial_fields_match_the_oracle(M, test):
    """state_var_init, velocity_init and ABCDE_init equal the oracle's fields bit for bit; the two
    levels are the same arrays, delta is zero, the cosine bell lies in [0, Phibar] and is zero
    beyond radius a / 3, and E = (U^2 + V^2) / (2 (1 - mu^2))."""
    from my_swampe_amd import initial_conditions as ic

    N, I, J, dt, lam, mus, w = ic.spectral_params(M)
    a, omega, a1, Phibar = 6.37122e6, 7.2921159e-5, np.pi / 4, 4.0e4
    t1 = ic.test1_init(a, omega, a1)
    assert all(float(x) == float(y) for x, y in zip(t1, orc.test1_init(a, omega, a1)))
    SU0, sina, cosa, etaamp, Phiamp = t1
    args = () if test is None else (a, sina, cosa, Phibar, Phiamp)
    eta0, eta1, delta0, delta1, Phi0, Phi1 = ic.state_var_init(I, J, mus, lam, test, etaamp, *args)
    oe, od, op = orc.state_var_init(I, J, mus, lam, test, etaamp, *args)
    for g, r in ((eta0, oe), (delta0, od), (Phi0, op)):
        assert g.shape == (J, I) and np.array_equal(g, r)
    assert np.array_equal(eta1, eta0) and np.array_equal(Phi1, Phi0) and not np.any(delta0) and not np.any(delta1)
    if test == 1:
        assert Phi0.min() >= 0.0 and Phi0.max() <= Phibar
        cosd = np.sqrt(1.0 - mus[:, None] ** 2) * np.cos(lam[None, :] - 1.5 * np.pi)
        assert not np.any(Phi0[a * np.arccos(np.clip(cosd, -1, 1)) > a / 3.0 * (1 + 1e-12)])
    U, V = ic.velocity_init(I, J, SU0, cosa, sina, mus, lam, test)
    oU, oV = orc.velocity_init(I, J, SU0, cosa, sina, mus, lam, test)
    assert np.array_equal(U, oU) and np.array_equal(V, oV)
    A, B, C, D, E = ic.ABCDE_init(U, V, eta0, Phi0, mus, I, J)
    for g, r in zip((A, B, C, D, E), orc.abcde(U, V, eta0, Phi0, mus)):
        assert np.array_equal(g, r)
    np.testing.assert_allclose(E, (U ** 2 + V ** 2) / (2.0 * (1.0 - mus[:, None] ** 2)), rtol=1e-15, atol=0)
    with pytest.raises(ValueError):
        ic.state_var_init(I, J, mus, lam, 1, etaamp, a, sina)


def test_coriolis_closed_form():
    """f = 2 omega mu has the single coefficient omega / sqrt(0.375) at (m, n) = (0, 1)."""
    from my_swampe_amd import initial_conditions as ic

    f = ic.coriolismn(42, 3.2e-5)
    assert f.shape == (43, 43) and f[0, 1] == 3.2e-5 / np.sqrt(0.375)
    f[0, 1] = 0.0
    assert not np.any(f)
    assert np.array_equal(ic.coriolismn(63, 7.2921159e-5), orc.coriolismn(63, 7.2921159e-5))


def test_time_stepping_coefficients_and_rms():
    """The coefficient arrays equal the oracle's bit for bit, with n = 0 zeroed in tstepcoeffmn and
    m, n (n + 1) along the right axes; RMS_winds equals the oracle's for the test-1 winds and
    scales linearly with the winds."""
    from my_swampe_amd import initial_conditions as ic
    from my_swampe_amd import time_stepping as ts

    N, I, J, dt, lam, mus, w = ic.spectral_params(42)
    a = 8.2e7
    for name, args in (("tstepcoeffmn", (42, 42, a)), ("tstepcoeff2", (J, 42, 120.0, a)), ("narray", (42, 42)),
                       ("mJarray", (J, 42)), ("marray", (42, 42)), ("tstepcoeff", (J, 42, 120.0, mus, a))):
        got = getattr(ts, name)(*args)
        assert np.array_equal(got, np.asarray(getattr(orc, name)(*args))), name
    c = ts.tstepcoeffmn(42, 42, a)
    assert not np.any(c[:, 0]) and c[3, 5] == a / 30.0
    assert ts.narray(42, 42)[7, 4] == 20.0 and ts.marray(42, 42)[7, 4] == 7.0 and ts.mJarray(J, 42)[5, 9] == 9.0
    SU0, sina, cosa, etaamp, Phiamp = ic.test1_init(6.37122e6, 7.2921159e-5, np.pi / 4)
    U, V = ic.velocity_init(I, J, SU0, cosa, sina, mus, lam, 1)
    r = ts.RMS_winds(6.37122e6, I, J, lam, mus, U, V)
    assert r == orc.rms_winds(6.37122e6, I, J, lam, mus, U, V) and r > 0
    np.testing.assert_allclose(ts.RMS_winds(6.37122e6, I, J, lam, mus, 3.0 * U, 3.0 * V), 3.0 * r, rtol=1e-14)


@pytest.mark.parametrize("M,a,dt,K6", [(42, 8.2e7, 120.0, 1.24e33), (106, 6.37122e6, 60.0, 1.0e30)])
def test_filters(M, a, dt, K6):
    """The sixth-order factors equal the oracle's bit for bit; every implicit factor is exactly 1 at
    its neutral degree (n = 1 with the offset, n = 0 without), decreases with n beyond it and lies in
    (0, 1] there, and diffusion multiplies."""
    from my_swampe_amd import filters as fl

    for name in ("sigma6", "sigma6Phi", "sigma6_exponential", "sigma6Phi_exponential"):
        got = getattr(fl, name)(M, M, K6, a, dt)
        assert got.shape == (M + 1, M + 1) and np.array_equal(got, getattr(orc, name)(M, M, K6, a, dt)), name
    K4 = K6 / a ** 2
    for name, neutral, args in (("sigma6", 1, (K6,)), ("sigma6Phi", 0, (K6,)), ("sigma", 1, (K4,)),
                                ("sigmaPhi", 0, (K4,))):
        s = getattr(fl, name)(M, M, *args, a, dt)[0]
        assert s[neutral] == 1.0, name
        assert np.all(np.diff(s[neutral:]) < 0) and np.all(s[neutral + 1:] > 0), name
    x = np.arange(6.0).reshape(2, 3)
    assert np.array_equal(fl.diffusion(x, 0.5), 0.5 * x)


def test_forcing():
    """Phieqfun, Qfun and Rfun equal the oracle's bit for bit; the day side is strict (the nodes at
    lambda = -pi/2 and pi/2 carry Phibar), negative Q drives no wind forcing, taudrag = -1 turns the
    drag off (also per member in an array), and a zero Phi + Phibar takes the smallest normal
    double instead of dividing by zero."""
    from my_swampe_amd import forcing as fo
    from my_swampe_amd import initial_conditions as ic

    N, I, J, dt, lam, mus, w = ic.spectral_params(42)
    Phibar, DPhieq = 3.0e5, 1.0e6
    pe = fo.Phieqfun(Phibar, DPhieq, lam, mus, I, J, 9.8)
    assert np.array_equal(pe, orc.phieqfun(Phibar, DPhieq, lam, mus))
    edge = np.isclose(np.abs(lam), np.pi / 2, rtol=0, atol=1e-14)
    assert edge.sum() == 2 and np.all(pe[:, edge] == Phibar) and np.all(pe[:, np.abs(lam) > np.pi / 2] == Phibar)
    rng = np.random.default_rng(4)
    Phi, U, V = (rng.standard_normal((J, I)) * s for s in (1e4, 10.0, 10.0))
    Q = fo.Qfun(pe, Phi, Phibar, 36000.0)
    assert np.array_equal(Q, orc.qfun(pe, Phi, Phibar, 36000.0))
    for td in (21600.0, -1.0):
        for g, r in zip(fo.Rfun(U, V, Q, Phi, Phibar, td), orc.rfun(U, V, Q, Phi, Phibar, td)):
            assert np.array_equal(g, r)
    F, G = fo.Rfun(U, V, Q, Phi, Phibar, -1.0)
    assert not np.any(F[Q < 0]) and not np.any(G[Q < 0])
    tds = np.array([21600.0, -1.0])[:, None, None]
    Fb, Gb = fo.Rfun(U[None], V[None], Q[None], Phi[None], Phibar, tds)
    assert np.array_equal(Fb[0], fo.Rfun(U, V, Q, Phi, Phibar, 21600.0)[0]) and np.array_equal(Fb[1], F)
    Fz, _ = fo.Rfun(np.ones((1, 1)), np.ones((1, 1)), np.ones((1, 1)), np.full((1, 1), -Phibar), Phibar, -1.0)
    assert Fz[0, 0] == -1.0 / np.finfo(np.float64).tiny
