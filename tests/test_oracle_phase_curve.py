"""Pins of the oracle's phase-curve observable and its vector-Jacobian product (CPU).

The projection follows unit_tests/test_parity_reference_regression.py:67-82; the fixtures' phase
curves pin it on real model output, and the checks below pin what the fixtures leave open
(the un-normalized scale, the phase convention, the derivative).
"""

from __future__ import annotations

import numpy as np
import pytest

from conftest import FIXTURE_CASES, GOLDEN
from oracle import swampe_oracle as orc


@pytest.mark.parametrize("case", sorted(FIXTURE_CASES))
def test_fixture_phase_curve_from_fixture_phi(case):
    ref = np.load(GOLDEN / f"{case}.npz")
    curve = orc.phase_curve(ref["final_Phi"], ref["phase_angles"])
    np.testing.assert_allclose(curve, ref["phase_curve"], rtol=0, atol=1e-12)


@pytest.mark.parametrize("J,I", [(64, 128), (96, 192), (160, 320)])
def test_uniform_grid_closed_form(J, I):
    # Sum over a half circle of grid-aligned cosines: sum_{|i| <= I/4} cos(2 pi i / I) = cot(pi / I).
    phases = np.linspace(-np.pi, np.pi, I, endpoint=False)[:: I // 8]
    _, raw = orc.phase_curve(np.ones((J, I)), phases, raw=True)
    mus = np.polynomial.legendre.leggauss(J)[0]
    want = (1.0 / np.tan(np.pi / I)) * np.sum(np.sqrt(1.0 - mus * mus))
    np.testing.assert_allclose(raw, want, rtol=1e-13)


def test_longitude_roll_shifts_the_phase():
    rng = np.random.default_rng(3)
    J, I, s = 64, 128, 5
    phi = rng.standard_normal((J, I))
    phases = rng.uniform(-np.pi, np.pi, 9)
    a = orc.phase_curve(phi, phases, normalize=False)
    b = orc.phase_curve(np.roll(phi, s, axis=1), phases + 2 * np.pi * s / I, normalize=False)
    np.testing.assert_allclose(b, a, rtol=1e-12, atol=1e-12 * np.abs(a).max())


def test_subsolar_hemisphere_only():
    # A grid that is zero where the phase's visibility is positive gives zero flux.
    J, I = 64, 128
    lam = np.linspace(-np.pi, np.pi, I, endpoint=False)
    phi = np.where(np.cos(lam) > 0, 0.0, 1.0)[None, :].repeat(J, 0)
    _, raw = orc.phase_curve(phi, np.array([0.0]), raw=True)
    assert raw[0] == 0.0


@pytest.mark.parametrize("normalize", [True, False])
@pytest.mark.parametrize("K", [1, 6])
def test_vjp_matches_central_differences(normalize, K):
    rng = np.random.default_rng(11 + K)
    J, I = 32, 64
    phi = 1.0 + 0.3 * rng.standard_normal((J, I))
    phases = rng.uniform(-np.pi, np.pi, K)
    g = rng.standard_normal(K)
    d = rng.standard_normal((J, I))
    _, raw = orc.phase_curve(phi, phases, normalize=normalize, raw=True)
    gphi = orc.phase_curve_vjp(g, raw, J, I, phases, normalize=normalize)
    h = 1e-5
    fp = orc.phase_curve(phi + h * d, phases, normalize=normalize)
    fm = orc.phase_curve(phi - h * d, phases, normalize=normalize)
    fd = float(np.dot(g, (fp - fm) / (2 * h)))
    an = float(np.sum(gphi * d))
    assert abs(fd - an) <= 1e-7 * max(1.0, abs(an)), (fd, an)


def test_vjp_normalization_term_sign():
    # A plausible mistake (dropping the normalization term, or flipping its sign) changes the
    # cotangent along phi itself: the normalized curve is invariant to scaling phi, so the
    # directional derivative along phi is zero.
    rng = np.random.default_rng(5)
    J, I, K = 32, 64, 8
    phi = 1.0 + 0.2 * rng.standard_normal((J, I))
    phases = np.linspace(-np.pi, np.pi, K, endpoint=False)
    g = rng.standard_normal(K)
    _, raw = orc.phase_curve(phi, phases, raw=True)
    gphi = orc.phase_curve_vjp(g, raw, J, I, phases)
    assert abs(float(np.sum(gphi * phi))) <= 1e-12 * float(np.abs(gphi).sum() * np.abs(phi).max())
