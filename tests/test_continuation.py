"""Restart files of the legacy driver (src/my_swampe/continuation.py): file layout, round trip,
the restricted reader, and the timestamp <-> step conversions (CPU; the resume itself is
tests/test_gpu_continuation.py)."""

from __future__ import annotations

import pickle

import numpy as np
import pytest

from my_swampe_amd import continuation as cont


def test_save_load_round_trip(tmp_path):
    rng = np.random.default_rng(0)
    f = [rng.standard_normal((64, 128)) for _ in range(5)]
    spin, geo = rng.standard_normal((10, 2)), rng.standard_normal((10, 2))
    cont.save_data(36, *f, spin, geo, custompath=str(tmp_path))
    names = sorted(p.name for p in tmp_path.iterdir())
    assert names == sorted(["eta-36", "delta-36", "Phi-36", "U-36", "V-36", "spinup-winds", "spinup-geopot"])
    got = cont.load_data(36, custompath=str(tmp_path))
    for a, b in zip(got, f):
        np.testing.assert_array_equal(a, b)
    np.testing.assert_array_equal(cont.read_pickle("spinup-geopot", custompath=str(tmp_path)), geo)
    # Files written by plain pickle (the reference's writer) read back the same.
    with (tmp_path / "eta-99").open("wb") as fh:
        pickle.dump(f[0], fh, protocol=pickle.HIGHEST_PROTOCOL)
    np.testing.assert_array_equal(cont.read_pickle("eta-99", custompath=str(tmp_path)), f[0])


class _NotAnArray:
    pass


def test_restart_reader_refuses_non_array_objects(tmp_path):
    with (tmp_path / "eta-1").open("wb") as fh:
        pickle.dump({"x": _NotAnArray()}, fh)
    with pytest.raises(pickle.UnpicklingError, match="numpy arrays only"):
        cont.read_pickle("eta-1", custompath=str(tmp_path))


@pytest.mark.parametrize("units,t,dt,ts", [("hours", 100, 240.0, "6"), ("minutes", 100, 240.0, "400"),
                                           ("seconds", 7, 240.0, "1680"), ("steps", 37, 240.0, "37"),
                                           ("days", 1000, 240.0, "2")])
def test_timestamp_conversions(units, t, dt, ts):
    """continuation.py:67-132: timestamp = int(dt t / unit), t = int(timestamp unit / dt)."""
    assert cont.compute_timestamp(units, t, dt) == ts
    back = cont.compute_t_from_timestamp(units, float(ts), dt)
    assert back <= t and cont.compute_timestamp(units, back, dt) == ts


def test_timestamp_units_are_checked():
    with pytest.raises(ValueError, match="Cannot parse units"):
        cont.compute_timestamp("weeks", 1, 1.0)
    with pytest.raises(ValueError, match="Cannot parse units"):
        cont.compute_t_from_timestamp("weeks", 1.0, 1.0)
