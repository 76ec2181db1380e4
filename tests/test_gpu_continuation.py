"""Continuation on the HIP path: unit_tests/test_continuation_roundtrip.py:73-160.

A run saved with continuation.save_data and resumed with contflag / contTime must reproduce a
fresh run started from the same single-level physical fields with starttime at the saved step
(the restart re-derives winds and the spectral state, as in the reference), at the reference's
parity atols; the resumed run is also held to the oracle's explicit-IC restart at the same atols.
run_model's saveflag writes the files of the steps with t % savefreq == 0.
"""

from __future__ import annotations

import numpy as np
import pytest

from conftest import ATOL
from oracle import swampe_oracle as orc

pytestmark = pytest.mark.gpu

BASE = dict(M=42, dt=600.0, Phibar=3.0e5, omega=7.292e-5, a=6.37122e6, test=None, g=9.8, forcflag=True,
            taurad=86400.0, taudrag=86400.0, DPhieq=4.0e6, diffflag=True, modalflag=True, alpha=0.01,
            expflag=False, K6=1.24e33)
N = 4


def _np(t):
    return t.detach().cpu().numpy()


def test_continuation_resume_matches_explicit_ic_restart(tmp_path, torch_gpu):
    from my_swampe_amd.continuation import compute_timestamp, save_data
    from my_swampe_amd.model import run_model_scan, run_model_scan_final

    print("[gpu-case] continuation resume", flush=True)
    first = run_model_scan(**BASE, tmax=2 + N, diagnostics=False)
    last1 = first["last_state"]
    eta_s, delta_s, Phi_s = _np(last1.eta_curr), _np(last1.delta_curr), _np(last1.Phi_curr)
    last_step = int(_np(first["t_seq"])[-1])
    ts = compute_timestamp("seconds", last_step, BASE["dt"])
    save_data(ts, eta_s, delta_s, Phi_s, _np(last1.U_curr), _np(last1.V_curr), np.zeros((2 + N, 2)),
              np.zeros((2 + N, 2)), custompath=str(tmp_path))
    resumed = run_model_scan_final(**BASE, tmax=2 + 2 * N, contflag=True, custompath=str(tmp_path), contTime=ts,
                                   timeunits="seconds", diagnostics=False)["last_state"]
    explicit = run_model_scan_final(**BASE, tmax=2 + 2 * N, eta0_init=eta_s, delta0_init=delta_s, Phi0_init=Phi_s,
                                    starttime=last_step, diagnostics=False)["last_state"]
    ref = orc.run_model_scan(**BASE, tmax=2 + 2 * N, eta0_init=eta_s, delta0_init=delta_s, Phi0_init=Phi_s,
                             starttime=last_step, diagnostics=False, return_history=False)["last_state"]
    for f in ("eta", "delta", "Phi", "U", "V"):
        a = _np(getattr(resumed, f + "_curr"))
        np.testing.assert_array_equal(a, _np(getattr(explicit, f + "_curr")), err_msg=f)
        np.testing.assert_allclose(a, getattr(ref, f + "_curr"), rtol=0, atol=ATOL[f], err_msg=f)


def test_run_model_saveflag_writes_restart_files(tmp_path, torch_gpu):
    from my_swampe_amd.continuation import load_data, read_pickle
    from my_swampe_amd.model import run_model

    print("[gpu-case] saveflag", flush=True)
    res = run_model(BASE["M"], BASE["dt"], 12, BASE["Phibar"], BASE["omega"], BASE["a"], plotflag=False,
                    saveflag=True, savefreq=5, custompath=str(tmp_path), timeunits="steps", verbose=False)
    names = sorted(p.name for p in tmp_path.iterdir())
    assert names == sorted([f"{f}-{t}" for f in ("eta", "delta", "Phi", "U", "V") for t in (5, 10)]
                           + ["spinup-winds", "spinup-geopot"])
    eta10 = load_data(10, custompath=str(tmp_path))[0]
    assert eta10.shape == (64, 128) and np.isfinite(eta10).all()
    np.testing.assert_array_equal(read_pickle("spinup-winds", custompath=str(tmp_path)), res["spinup"])
