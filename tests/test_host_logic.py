"""Host-side logic without a GPU: member splitting, step-index runs, the rollout checkpoint policy."""

from __future__ import annotations

import math

import pytest

from my_swampe_amd.autograd import RolloutConfig
from my_swampe_amd.model import _runs
from my_swampe_amd.plan import split_bounds


@pytest.mark.parametrize("B,parts", [(256, 2), (40, 3), (7, 4), (5, 1), (4, 4)])
def test_split_bounds_cover_members_evenly(B, parts):
    b = split_bounds(B, parts)
    sizes = [b[k + 1] - b[k] for k in range(parts)]
    assert b[0] == 0 and b[-1] == B and len(sizes) == parts
    assert min(sizes) >= 1 and max(sizes) - min(sizes) <= 1


@pytest.mark.parametrize("B,parts", [(3, 5), (2, 0), (8, 5)])
def test_split_bounds_rejects_bad_counts(B, parts):
    with pytest.raises(ValueError):
        split_bounds(B, parts)


def test_runs_of_step_indices():
    assert _runs([2, 3, 4, 7, 8, 10]) == [(2, 3), (7, 2), (10, 1)]
    assert _runs([]) == []
    assert _runs([5]) == [(5, 1)]
    assert _runs([3, 2]) == [(3, 1), (2, 1)]  # not ascending: one step each


def test_checkpoint_policy():
    base = dict(M=42, dt=120.0, a=8.2e7, omega=3.2e-5, Phibar=3e5)
    # 100 steps of 256 T42 members (22.7 MB per state) fit the 32 GiB budget: every state kept.
    assert RolloutConfig(nsteps=100, members=256, **base).every() == 1
    # 5000 steps of 64 T106 members (35 MB per state) do not: every ceil(sqrt(N)) steps.
    assert RolloutConfig(nsteps=5000, members=64, **{**base, "M": 106}).every() == math.ceil(math.sqrt(5000))
    # An explicit interval wins, and a zero budget always recomputes.
    assert RolloutConfig(nsteps=100, members=256, checkpoint_every=7, **base).every() == 7
    assert RolloutConfig(nsteps=100, members=1, checkpoint_budget=0, **base).every() == 10
    assert RolloutConfig(nsteps=100, members=256, **base).state_bytes() == 256 * 3 * 43 * 43 * 16
