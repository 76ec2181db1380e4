"""Build libswampe_hip.so (gfx950) in-tree with hipcc.

    python -m my_swampe_amd.build        # or my_swampe_amd.build.build()

The library is a plain C-ABI shared object (include/swampe_capi.h); the
Python host module loads it with ctypes.
"""

from __future__ import annotations

import os
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
LIB_DIR = PKG / "_lib"
LIB = LIB_DIR / "libswampe_hip.so"

HIP_SOURCES = ["swx_kernels.hip", "swx_mfma.hip", "swx_bf16x3.hip", "swx_adjoint.hip", "swx_adjoint_mfma.hip",
               "swx_adjoint_leg_mfma.hip", "swx_observe.hip", "swx_transforms.hip", "swx_peaks.hip"]
HOST_SOURCES = ["swx_plan.cpp", "geometry.cpp"]

# Packed FP32 VALU forms are disabled for all device code.
HIPCC_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "--offload-arch=gfx950",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def hipcc() -> str:
    rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
    cand = Path(rocm) / "bin" / "hipcc"
    return str(cand) if cand.exists() else "hipcc"


def build(verbose: bool = True, jobs: int = 0) -> Path:
    """Compile every source (in parallel, `jobs` at a time; 0 = one per CPU up to 8) and link the
    shared library; returns its path."""
    from concurrent.futures import ThreadPoolExecutor

    LIB_DIR.mkdir(exist_ok=True)
    cmds, objs = [], []
    for src in HIP_SOURCES + HOST_SOURCES:
        path = CSRC / src
        if not path.exists():
            continue
        obj = LIB_DIR / (path.stem + ".o")
        lang = ["-x", "hip"] if src.endswith(".hip") or src == "swx_plan.cpp" else []
        cmds.append([hipcc()] + HIPCC_FLAGS + ["-I", str(ROOT / "include"), "-c"] + lang + [str(path), "-o", str(obj)])
        objs.append(str(obj))
    if verbose:
        for cmd in cmds:
            print(" ".join(cmd), flush=True)
    workers = jobs if jobs > 0 else max(1, min(8, os.cpu_count() or 1))
    with ThreadPoolExecutor(max_workers=workers) as pool:
        for r in pool.map(lambda c: subprocess.run(c, capture_output=not verbose), cmds):
            if r.returncode != 0:
                raise subprocess.CalledProcessError(r.returncode, r.args, r.stdout, r.stderr)
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", "-o", str(tmp)] + objs
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    for o in objs:
        try:
            os.remove(o)
        except OSError:
            pass
    return LIB


if __name__ == "__main__":
    build(verbose="-q" not in sys.argv)
