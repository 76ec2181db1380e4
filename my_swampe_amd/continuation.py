"""Restart files of the legacy driver: save and load the physical fields between runs.

Same on-disk layout and function signatures as the reference (src/my_swampe/continuation.py):
one pickle per field named ``<field>-<timestamp>`` (eta, delta, Phi, U, V; float64 (J, I)
arrays) plus ``spinup-winds`` / ``spinup-geopot`` under ``custompath`` (default ``data/``), and
the timestamp <-> step conversions of the ``timeunits`` argument.

Reading uses a restricted unpickler that rebuilds numpy arrays and scalars only: a restart
file can therefore hold nothing but array data (any other object in it raises
``pickle.UnpicklingError``), where the reference unpickles arbitrary objects.
"""

from __future__ import annotations

import io
import pickle
from pathlib import Path
from typing import Any, Optional, Tuple, Union

import numpy as np

# Globals a pickled numpy array (or numpy scalar) refers to, by module and name.
_ARRAY_GLOBALS = {
    ("numpy", "ndarray"),
    ("numpy", "dtype"),
    ("numpy.core.multiarray", "_reconstruct"),
    ("numpy._core.multiarray", "_reconstruct"),
    ("numpy.core.multiarray", "scalar"),
    ("numpy._core.multiarray", "scalar"),
    ("numpy.core.numeric", "_frombuffer"),
    ("numpy._core.numeric", "_frombuffer"),
    ("builtins", "complex"),
}


class _ArrayUnpickler(pickle.Unpickler):
    def find_class(self, module, name):
        if (module, name) in _ARRAY_GLOBALS:
            return super().find_class(module, name)
        raise pickle.UnpicklingError(f"restart files hold numpy arrays only; refusing {module}.{name}")


def _base(custompath: Optional[str]) -> Path:
    return Path("data") if custompath is None else Path(custompath)


def write_pickle(filename: str, data: Any, custompath: Optional[str] = None) -> None:
    """Write ``data`` (converted to a numpy array) to ``custompath/filename`` (continuation.py:24-43)."""
    base = _base(custompath)
    base.mkdir(parents=True, exist_ok=True)
    if hasattr(data, "detach"):  # torch tensors (the drop-in's device arrays)
        data = data.detach().cpu().numpy()
    with (base / filename).open("wb") as f:
        pickle.dump(np.asarray(data), f, protocol=pickle.HIGHEST_PROTOCOL)


def read_pickle(filename: str, custompath: Optional[str] = None) -> Any:
    """Read ``custompath/filename`` (continuation.py:46-64), numpy arrays only."""
    with (_base(custompath) / filename).open("rb") as f:
        return _ArrayUnpickler(io.BytesIO(f.read())).load()


def _unit_check(units: str) -> None:
    if units not in ("hours", "minutes", "seconds", "steps", "days"):
        raise ValueError("Cannot parse units. Acceptable units are: hours, minutes, seconds, steps, days.")


def compute_timestamp(units: str, t: int, dt: float) -> str:
    """Timestamp token of step t in ``units`` (continuation.py:67-98): int(dt t / unit seconds)."""
    _unit_check(units)
    if units == "steps":
        return str(int(t))
    if units == "seconds":
        return str(int(dt * t))
    # The reference divides by the literal unit length: dt * t / 3600 etc.
    div = {"hours": 3600, "minutes": 60, "days": 3600 * 24}[units]
    return str(int(dt * t / div))


def compute_t_from_timestamp(units: str, timestamp: float, dt: float) -> int:
    """Step index of a timestamp (continuation.py:101-132): int(timestamp * unit seconds / dt)."""
    _unit_check(units)
    if units == "steps":
        return int(timestamp)
    if units == "seconds":
        return int(timestamp / dt)
    if units == "days":
        return int(timestamp * 3600 * 24 / dt)  # the reference's evaluation order
    return int(timestamp * (3600 if units == "hours" else 60) / dt)


def save_data(timestamp: Union[int, str], etadata: Any, deltadata: Any, Phidata: Any, U: Any, V: Any,
              spinupdata: Any, geopotdata: Any, custompath: Optional[str] = None) -> None:
    """Fields of one timestamp and the diagnostics series (continuation.py:135-179)."""
    ts = str(timestamp)
    for name, arr in (("eta", etadata), ("delta", deltadata), ("Phi", Phidata), ("U", U), ("V", V)):
        write_pickle(f"{name}-{ts}", arr, custompath=custompath)
    write_pickle("spinup-winds", spinupdata, custompath=custompath)
    write_pickle("spinup-geopot", geopotdata, custompath=custompath)


def load_data(timestamp: Union[int, str], custompath: Optional[str] = None
              ) -> Tuple[np.ndarray, np.ndarray, np.ndarray, np.ndarray, np.ndarray]:
    """(eta, delta, Phi, U, V) of one timestamp (continuation.py:182-211)."""
    ts = str(timestamp)
    return tuple(np.asarray(read_pickle(f"{name}-{ts}", custompath=custompath))
                 for name in ("eta", "delta", "Phi", "U", "V"))
