"""Shared helpers: a HIP solver and the CPU oracle set to the same state."""
from __future__ import annotations

import numpy as np

from oracle.oracle import Oracle
from unitywindsimulation_amd import _native as N
from unitywindsimulation_amd.geometry import noise_velocity


def rel_l2(a, b) -> float:
    a = np.asarray(a, dtype=np.float64).reshape(-1)
    b = np.asarray(b, dtype=np.float64).reshape(-1)
    nb = np.linalg.norm(b)
    d = np.linalg.norm(a - b)
    return float(d / nb) if nb > 0 else float(d)


def oracle_for(w, precision=0, u0=None, p0=None):
    o = Oracle(w.cf, precision)
    o.init_like_solver(face_positions=getattr(w, "face_positions", None))
    o.init_flags(w.flags)
    o.set_velocity(noise_velocity(w.flags, w.noise_seed, w.noise_amp) if u0 is None else u0)
    if p0 is not None:
        o.set_pressure(p0)
    return o


def solver_for(w, u0=None, p0=None):
    from unitywindsimulation_amd import FvmSolverHip

    s = FvmSolverHip(w.cf, face_positions=getattr(w, "face_positions", None))
    s.InitFlags(w.flags)
    s.set_velocity(noise_velocity(w.flags, w.noise_seed, w.noise_amp) if u0 is None else u0)
    if p0 is not None:
        s.set_pressure(p0)
    return s


def seeded_pressure(w, seed=11, amp=0.5):
    rng = np.random.default_rng(seed)
    p = (rng.random(w.ncells, dtype=np.float32) - np.float32(0.5)) * np.float32(2 * amp)
    p[w.flags != 0] = 0.0
    return p
