"""Build the gfx950 HIP library in-tree (``unitywindsimulation_amd/libuws_hip.so``).

The library is plain hipcc output with a C ABI (``include/uws.h``); it links only
the HIP runtime.  ``-ffp-contract=off`` keeps every float operation rounded as
written (the kernels restate the reference's operation order), and
``-packed-fp32-ops`` is disabled for all device code.
"""
from __future__ import annotations

import os
import shutil
import subprocess
from pathlib import Path

PKG_DIR = Path(__file__).resolve().parent
REPO_DIR = PKG_DIR.parent
CSRC = PKG_DIR / "csrc"
LIB_PATH = PKG_DIR / "libuws_hip.so"

SOURCES = ["uws_kernels.hip", "uws_adjoint.hip", "uws_api.cpp", "uws_batch.cpp", "uws_adjoint.cpp"]

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-command-line-argument",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found: the uws HIP library cannot be built")


def build_library(verbose: bool = False) -> Path:
    """Compile every HIP source into the in-tree shared library (always rebuilds)."""
    tmp = LIB_PATH.with_suffix(".so.tmp")
    cmd = [hipcc(), *HIPCC_FLAGS, "-I", str(REPO_DIR / "include"), "-o", str(tmp)]
    cmd += [str(CSRC / s) for s in SOURCES]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed ({res.returncode}):\n{res.stderr[-8000:]}")
    if verbose and res.stderr:
        print(res.stderr)
    os.replace(tmp, LIB_PATH)
    return LIB_PATH


if __name__ == "__main__":
    print(build_library(verbose=True))
