"""The CPU oracle against its frozen regression vectors
(tests/golden/oracle_tiny.npz, written by tests/golden/make_oracle_fixtures.py):
bit for bit in FP32 and FP64, so any change of the oracle's arithmetic is
seen.  The workloads and the seeded start are checked against the stored
flags and initial velocity first."""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np
import pytest

sys.path.insert(0, str(Path(__file__).resolve().parent / "golden"))

import make_oracle_fixtures as G  # noqa: E402

from unitywindsimulation_amd.geometry import noise_velocity  # noqa: E402

FIX = np.load(G.OUT, allow_pickle=False)


@pytest.mark.parametrize("name", sorted(G.CASES))
def test_workload_and_start_unchanged(name):
    w = G.case(name)
    np.testing.assert_array_equal(np.asarray(w.flags, dtype=np.int32).reshape(-1), FIX[f"{name}_flags"])
    np.testing.assert_array_equal(noise_velocity(w.flags, w.noise_seed, w.noise_amp), FIX[f"{name}_u0"])


@pytest.mark.parametrize("name", sorted(G.CASES))
@pytest.mark.parametrize("prec,tag,steps", [(0, "fp32", 1), (0, "fp32", 3), (1, "fp64", 3)])
def test_oracle_matches_frozen_vectors(name, prec, tag, steps):
    u, p = G.run(G.case(name), prec, steps)
    dt = np.float32 if prec == 0 else np.float64
    u = np.asarray(u, dtype=dt).reshape(-1, 3)
    p = np.asarray(p, dtype=dt).reshape(-1)
    ref_u, ref_p = FIX[f"{name}_{tag}_{steps}_u"], FIX[f"{name}_{tag}_{steps}_p"]
    assert np.array_equal(u.view(np.uint8), ref_u.view(np.uint8)), (name, tag, steps)
    assert np.array_equal(p.view(np.uint8), ref_p.view(np.uint8)), (name, tag, steps)
