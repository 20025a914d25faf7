"""Writes tests/golden/oracle_tiny.npz: regression vectors of the CPU oracle.

The oracle restates the reference's kernels; its functions are pinned by
closed forms, assembled systems and planted mistakes (tests/test_oracle*.py).
These vectors freeze its outputs on two small cases, so that any later change
of its arithmetic shows up as a failure of tests/test_golden_oracle.py (the
oracle against these vectors, bit for bit) and the HIP path can be checked
against them without the oracle (tests/test_gpu_golden.py).

Cases: the tiny 16 x 8 x 12 workload with two buildings (LES, PISO, rough
wall, V = 6, P = 8) and the same without LES and with SIMPLE; the seeded
noise start of the workload; FP32 after 1 and 3 steps, FP64 after 3 steps;
velocity [cell][3] and pressure [cell].  Run from the repository root:

    python tests/golden/make_oracle_fixtures.py
"""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from oracle.oracle import Oracle  # noqa: E402
from unitywindsimulation_amd import workloads  # noqa: E402
from unitywindsimulation_amd.geometry import noise_velocity  # noqa: E402

OUT = Path(__file__).resolve().parent / "oracle_tiny.npz"

CASES = {"les_piso": dict(), "laminar_simple": dict(les=False, piso=False)}


def case(name):
    return workloads.tiny(**CASES[name])


def run(w, precision, steps):
    o = Oracle(w.cf, precision)
    o.init_like_solver()
    o.init_flags(w.flags)
    o.set_velocity(noise_velocity(w.flags, w.noise_seed, w.noise_amp))
    o.step(steps)
    u, p = o.velocity(), o.pressure()
    o.close()
    return u, p


def main() -> int:
    arrays = {}
    for name in CASES:
        w = case(name)
        arrays[f"{name}_flags"] = np.asarray(w.flags, dtype=np.int32).reshape(-1)
        arrays[f"{name}_u0"] = noise_velocity(w.flags, w.noise_seed, w.noise_amp)
        for prec, tag, steps in ((0, "fp32", 1), (0, "fp32", 3), (1, "fp64", 3)):
            u, p = run(case(name), prec, steps)
            dt = np.float32 if prec == 0 else np.float64
            arrays[f"{name}_{tag}_{steps}_u"] = np.asarray(u, dtype=dt).reshape(-1, 3)
            arrays[f"{name}_{tag}_{steps}_p"] = np.asarray(p, dtype=dt).reshape(-1)
    np.savez_compressed(OUT, **arrays)
    print(f"wrote {OUT} ({OUT.stat().st_size} bytes, {len(arrays)} arrays)")
    return 0


if __name__ == "__main__":
    sys.exit(main())
