"""The step driver's launch sequence (which kernels, on which buffers) for
every combination of single, paired and tripled sweeps: a host program built
from the driver templates with a recording launcher (tests/cpp/step_driver_test.cpp,
hipcc for the host only; nothing runs on a GPU)."""
from __future__ import annotations

import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
SRC = ROOT / "tests" / "cpp" / "step_driver_test.cpp"
EXE = ROOT / "tests" / "cpp" / "step_driver_test"
CSRC = ROOT / "unitywindsimulation_amd" / "csrc"
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"


@pytest.mark.skipif(not Path(HIPCC).exists(), reason="hipcc is needed to build the host program")
def test_step_driver_launch_sequence():
    subprocess.run([HIPCC, "-std=c++17", "-O1", f"-I{CSRC}", str(SRC), "-o", str(EXE)], check=True,
                   capture_output=True, text=True, timeout=600)
    res = subprocess.run([str(EXE)], capture_output=True, text=True, timeout=120)
    print(res.stdout, res.stderr)
    assert res.returncode == 0, res.stdout
    assert "cases passed" in res.stdout
