"""The HIP path against the oracle's frozen vectors (tests/golden/oracle_tiny.npz),
without running the oracle: FP32 within 1e-5 relative L2 after one and three
steps, FP64 within 1e-12 after three."""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np
import pytest

sys.path.insert(0, str(Path(__file__).resolve().parent / "golden"))

import make_oracle_fixtures as G  # noqa: E402

from unitywindsimulation_amd import FvmSolverHip  # noqa: E402
from unitywindsimulation_amd import _native as N  # noqa: E402

pytestmark = pytest.mark.gpu
FIX = np.load(G.OUT, allow_pickle=False)


def rel(a, b):
    a = np.asarray(a, dtype=np.float64).reshape(-1)
    b = np.asarray(b, dtype=np.float64).reshape(-1)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("name", sorted(G.CASES))
@pytest.mark.parametrize("prec,tag,steps,tol", [(0, "fp32", 1, 1e-5), (0, "fp32", 3, 1e-5), (1, "fp64", 3, 1e-12)])
def test_hip_matches_frozen_oracle_vectors(gpu, name, prec, tag, steps, tol):
    w = G.case(name)
    if prec == 1:
        w.cf.precision = N.PREC_FP64
    with FvmSolverHip(w.cf) as s:
        s.InitFlags(FIX[f"{name}_flags"])
        s.set_velocity(FIX[f"{name}_u0"])
        s.step(steps)
        if prec == 1:
            u, p = s.field64(N.FIELD_VEL), s.field64(N.FIELD_PRES)
        else:
            u, p = s.velocity(3), s.pressure()
    eu = rel(u, FIX[f"{name}_{tag}_{steps}_u"])
    ep = rel(p, FIX[f"{name}_{tag}_{steps}_p"])
    print(f"{name} {tag} {steps} steps: rel-L2 u {eu:.3e} p {ep:.3e}", flush=True)
    assert eu <= tol and ep <= tol, (eu, ep)
