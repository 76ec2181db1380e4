"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer (CPU): the FP64 geometry
(geometry.cpp: Gauss-Legendre nodes and weights, the associated Legendre basis and its
reference normalization, the longitudes) built with g++ and run through tests/host/geometry_check.cpp."""

from __future__ import annotations

import os
import shutil
import subprocess

import pytest

from conftest import ROOT


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ is not available")
def test_geometry_under_address_and_undefined_behavior_sanitizers(tmp_path):
    exe = tmp_path / "geometry_check"
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
           "-fno-omit-frame-pointer", str(ROOT / "tests" / "host" / "geometry_check.cpp"),
           str(ROOT / "my_swampe_amd" / "csrc" / "geometry.cpp"), "-o", str(exe)]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")
    run = subprocess.run([str(exe)], capture_output=True, text=True, env=env, timeout=120)
    if "does not come first in initial library list" in run.stderr:
        pytest.skip("the ASan runtime cannot be first in this process environment")
    assert run.returncode == 0, run.stdout + run.stderr
    assert "ERROR: AddressSanitizer" not in run.stderr and "runtime error" not in run.stderr, run.stderr
    for M in (42, 63, 106):
        assert f"M={M} " in run.stdout
