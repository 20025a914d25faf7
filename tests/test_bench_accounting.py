"""The byte accounting the bench line reports, against the per-kernel byte
table it is built from: the per-step algorithmic bytes of the reference's
dispatch sequence (every sweep a pass over its arrays) and the compulsory
bytes of the launches the solver runs (each array once per launch).  CPU
only: the functions are plain arithmetic."""
from __future__ import annotations

import importlib.util
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent


@pytest.fixture(scope="module")
def bench():
    spec = importlib.util.spec_from_file_location("bench_module", ROOT / "bench.py")
    mod = importlib.util.module_from_spec(spec)
    sys.modules["bench_module"] = mod
    spec.loader.exec_module(mod)
    return mod


def test_step_bytes_per_kernel_table(bench):
    # wall function 25, eddy viscosity 17, LES term 41, momentum precompute 69, Rhie-Chow precompute 37,
    # p += p' 13, velocity correction 53, zero p' / copy u 8: 263 fixed; 77 per momentum sweep, 29 per
    # pressure sweep, per PISO corrector 37 + 13 + 65 + 8 and 33 per corrector sweep
    assert bench.bytes_per_cell_step(50, 50, 2) == 9109
    assert bench.bytes_per_cell_step(0, 0, 0) == 263
    assert bench.bytes_per_cell_step(1, 0, 0) - bench.bytes_per_cell_step(0, 0, 0) == 77
    assert bench.bytes_per_cell_step(0, 1, 0) - bench.bytes_per_cell_step(0, 0, 0) == 29
    assert bench.bytes_per_cell_step(0, 0, 1) - bench.bytes_per_cell_step(0, 0, 0) == 123
    # without LES: no wall function, eddy viscosity or LES term
    assert bench.bytes_per_cell_step(50, 50, 2, les=False) == 9109 - (25 + 17 + 41)


def test_compulsory_bytes_of_the_launch_plans(bench):
    # one-shot kernels 78 + 146, 25 momentum pairs of 73, Rhie-Chow 37, the velocity correction 53 with
    # p += p' in its launch (+ 8: p read and written), per corrector A_od u' 37, right-hand side 4 and
    # correction 78; a solve of 50 sweeps is 16 triples and one pair (17 launches of 28) or 12 four-sweep
    # launches and one pair (13 of 28)
    fixed = 78 + 146 + 25 * 73 + 37 + 53 + 8 + 2 * (37 + 4 + 78)
    assert bench.compulsory_bytes_per_cell_step(50, 50, 2, pres_depth=3) == fixed + 3 * 17 * 28 == 3813
    assert bench.compulsory_bytes_per_cell_step(50, 50, 2, pres_depth=4) == fixed + 3 * 13 * 28 == 3477
    # with the pressure normalisation p += p' is a launch of its own (13)
    assert (bench.compulsory_bytes_per_cell_step(50, 50, 2, pres_depth=4, normalise=True)
            - bench.compulsory_bytes_per_cell_step(50, 50, 2, pres_depth=4)) == 13 - 8
    # single sweeps everywhere: every sweep one pass of its arrays
    single = bench.compulsory_bytes_per_cell_step(50, 50, 2, pres_depth=1, vel_depth=1)
    assert single == 78 + 146 + 50 * 77 + 37 + 53 + 8 + 2 * (37 + 4 + 78) + 3 * 50 * 29
    # a leftover sweep of a solve: a single launch (29), two or three: one multi-sweep launch (28)
    for P, depth, launches28, singles in ((49, 4, 12, 1), (51, 4, 13, 0), (50, 3, 17, 0), (49, 3, 16, 1)):
        got = bench.compulsory_bytes_per_cell_step(50, P, 0, pres_depth=depth)
        assert got == 78 + 146 + 25 * 73 + 37 + 53 + 8 + launches28 * 28 + singles * 29, (P, depth)
    # an odd momentum count: the leftover sweep is a single launch (77)
    assert (bench.compulsory_bytes_per_cell_step(51, 50, 2, pres_depth=4)
            - bench.compulsory_bytes_per_cell_step(50, 50, 2, pres_depth=4)) == 77


def test_more_than_one_gpu_is_refused():
    # the solver step runs on one GPU; asking for more exits with status 2
    # and a message, before any device work
    import subprocess

    res = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                         timeout=300)
    assert res.returncode == 2
    assert "one GPU" in res.stderr and res.stdout == ""


def test_counter_measured_step_bytes(bench):
    # urban plan (V = P = 50, C = 2, four-sweep pressure launches, momentum pairs), summed by hand from the
    # per-launch counter table: 66.9 + 165.2 (wall function, momentum precompute) + 25 x 84.6 (pairs)
    # + 59.3 + 61.5 (Rhie-Chow precompute, velocity correction with p += p') + 3 solves of
    # (30.5 + 11 x 36.2 + 32.2) + 2 x (38.9 + 46.2 + 73.5) = 4167.8
    assert bench.measured_bytes_per_cell_step("urban", 50, 50, 2, True, 4, 2) == 4167.8
    # no profile of other workloads or of the laminar plan
    assert bench.measured_bytes_per_cell_step("isolated_building", 50, 50, 2) is None
    assert bench.measured_bytes_per_cell_step("urban", 50, 50, 2, les=False) is None
