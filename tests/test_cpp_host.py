"""The C++ host mirror of IFvmSolver (FvmSolverHip.hpp) compiles against the
C ABI and drives a few steps on the GPU."""
from __future__ import annotations

import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
SRC = ROOT / "tests" / "cpp" / "host_api_demo.cpp"
EXE = ROOT / "tests" / "cpp" / "host_api_demo"
LIBDIR = ROOT / "unitywindsimulation_amd"


def build_demo() -> Path:
    cmd = ["g++", "-std=c++17", "-O1", "-Wall", "-Wextra", str(SRC), "-o", str(EXE), f"-L{LIBDIR}", "-luws_hip",
           f"-Wl,-rpath,{LIBDIR}"]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    return EXE


def test_host_mirror_compiles():
    assert build_demo().exists()


@pytest.mark.gpu
def test_host_mirror_runs(gpu):
    exe = EXE if EXE.exists() else build_demo()
    res = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    print(res.stdout, res.stderr)
    assert res.returncode == 0
