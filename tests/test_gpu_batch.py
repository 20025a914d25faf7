"""The ensemble (one launch per kernel for many uneven cases) against the
same cases stepped one by one: bit-for-bit equal for every launch shape."""
from __future__ import annotations

import numpy as np
import pytest

from unitywindsimulation_amd import FvmSolverHip, workloads
from unitywindsimulation_amd.geometry import noise_velocity
from unitywindsimulation_amd.solver import FvmEnsembleHip

pytestmark = pytest.mark.gpu


def small_cases(vel_iters=6):
    ws = [workloads.tiny(grid=(16, 8, 12), vel_iters=vel_iters),
          workloads.tiny(grid=(40, 12, 20), seed=3, vel_iters=vel_iters),
          workloads.tiny(grid=(70, 10, 9), seed=5, vel_iters=vel_iters)]
    return ws


def singles(ws, steps):
    out = []
    for w in ws:
        with FvmSolverHip(w.cf) as s:
            s.InitFlags(w.flags)
            s.set_velocity(noise_velocity(w.flags, w.noise_seed, w.noise_amp))
            s.step(steps)
            out.append((s.velocity(3), s.pressure()))
    return out


def ensemble_run(ws, steps, ty=None, persistent=False, autotune=False, group_bytes=0):
    with FvmEnsembleHip([w.cf for w in ws]) as e:
        for c, w in enumerate(ws):
            e.init_flags(c, w.flags)
            e.set_velocity(c, noise_velocity(w.flags, w.noise_seed, w.noise_amp))
        if autotune:
            e.autotune(2)
        if ty is not None:
            e.set_launch(ty, persistent, group_bytes)
        e.step(steps)
        return [(e.velocity(c, 3), e.pressure(c)) for c in range(len(ws))]


@pytest.mark.parametrize("ty,persistent,group_bytes,pack", [(4, False, 0, None), (2, False, 0, None),
                                                            (1, False, 0, None), (4, True, 0, None),
                                                            (1, True, 0, None), (4, False, 1, None),
                                                            (2, True, 400_000, None), (4, False, -1, None),
                                                            (4, False, -1, 2), (1, True, -1, 1)])
def test_ensemble_matches_single_cases_bitwise(gpu, monkeypatch, ty, persistent, group_bytes, pack):
    """pack: workgroups per packed group (UWS_PACK_BLOCKS), small enough
    that the packed mode (group_bytes < 0) forms several groups of these
    small cases"""
    if pack is not None:
        monkeypatch.setenv("UWS_PACK_BLOCKS", str(pack))
    ws = small_cases()
    ref = singles(ws, 4)
    got = ensemble_run(ws, 4, ty, persistent, group_bytes=group_bytes)
    for (ug, pg), (ur, pr) in zip(got, ref):
        np.testing.assert_array_equal(ug, ur)
        np.testing.assert_array_equal(pg, pr)


@pytest.mark.parametrize("ty,persistent,group_bytes,pack", [(4, False, 0, None), (2, True, 0, None),
                                                            (4, False, -1, 3)])
def test_six_case_groups_match_single_cases_bitwise(gpu, monkeypatch, ty, persistent, group_bytes, pack):
    """Six cases: one group of six reads the cases' field tables from device
    memory; packed groups of at most four cases pass them as kernel
    arguments.  Both equal the single cases bit for bit."""
    if pack is not None:
        monkeypatch.setenv("UWS_PACK_BLOCKS", str(pack))
    ws = small_cases() + [workloads.tiny(grid=(24, 9, 14), seed=9), workloads.tiny(grid=(65, 8, 10), seed=4),
                          workloads.tiny(grid=(16, 8, 12), seed=12)]
    ref = singles(ws, 3)
    with FvmEnsembleHip([w.cf for w in ws]) as e:
        for c, w in enumerate(ws):
            e.init_flags(c, w.flags)
            e.set_velocity(c, noise_velocity(w.flags, w.noise_seed, w.noise_amp))
        e.set_launch(ty, persistent, group_bytes)
        groups = e.num_groups()
        e.step(3)
        got = [(e.velocity(c, 3), e.pressure(c)) for c in range(len(ws))]
    print(f"  six cases, launch {ty}/{persistent}/{group_bytes}/{pack}: {groups} groups", flush=True)
    assert groups == 1 if group_bytes == 0 else groups >= 2
    for (ug, pg), (ur, pr) in zip(got, ref):
        np.testing.assert_array_equal(ug, ur)
        np.testing.assert_array_equal(pg, pr)


def test_autotune_leaves_state_unchanged(gpu):
    ws = small_cases()
    ref = singles(ws, 3)
    got = ensemble_run(ws, 3, autotune=True)
    for (ug, pg), (ur, pr) in zip(got, ref):
        np.testing.assert_array_equal(ug, ur)
        np.testing.assert_array_equal(pg, pr)


def test_ensemble_rejects_mixed_schedules(gpu):
    a, b = workloads.tiny(), workloads.tiny()
    b.cf.presMaxNumIter = 9
    with pytest.raises(Exception):
        FvmEnsembleHip([a.cf, b.cf])


@pytest.mark.parametrize("group_bytes,vel_iters,zchunks", [(0, 6, None), (1, 6, None), (0, 4, None), (1, 5, None),
                                                             (0, 5, (3, 1)), (-1, 6, None), (-1, 5, (3, 1))])
def test_ensemble_two_sweep_kernel_matches_single_cases(gpu, monkeypatch, group_bytes, vel_iters, zchunks):
    """With the multi-sweep kernels forced at every size (momentum pairs,
    pressure pairs and triples), the ensemble's launches over every case's
    z-marching blocks equal the single cases.  Four momentum sweeps run as
    two pairs, five as two pairs and a single sweep; either way the step ends
    in the buffer single sweeps would (the groups of an ensemble share buffer
    roles).  zchunks: z planes per workgroup of the two- and three-sweep kernels, for
    every case (chunks that start at z = 1 and 2)."""
    monkeypatch.setenv("UWS_SWEEP_FUSION", "2")
    if zchunks is not None:
        monkeypatch.setenv("UWS_ZCHUNK", str(zchunks[0]))
        monkeypatch.setenv("UWS_ZCHUNK3", str(zchunks[1]))
    ws = small_cases(vel_iters)
    ref = singles(ws, 3)
    got = ensemble_run(ws, 3, 4, False, group_bytes=group_bytes)
    for (ug, pg), (ur, pr) in zip(got, ref):
        np.testing.assert_array_equal(ug, ur)
        np.testing.assert_array_equal(pg, pr)
    # and the one-sweep kernels agree with both
    monkeypatch.setenv("UWS_SWEEP_FUSION", "0")
    for (ug, pg), (ur, pr) in zip(singles(ws, 3), ref):
        np.testing.assert_array_equal(ug, ur)
        np.testing.assert_array_equal(pg, pr)


@pytest.mark.parametrize("group_bytes,zchunk4", [(0, None), (-1, None), (0, 3), (1, 1)])
def test_ensemble_four_sweep_kernel_matches_single_sweeps(gpu, monkeypatch, group_bytes, zchunk4):
    """With four-sweep pressure launches forced (UWS_SWEEP_DEPTH=4: every
    solve of P = 8 sweeps runs two of them), the ensemble's launches over
    every case's z-marching blocks equal the same cases stepped with one-sweep
    kernels only, bit for bit.  zchunk4: z planes per four-sweep workgroup."""
    ws = small_cases()
    monkeypatch.setenv("UWS_SWEEP_FUSION", "0")
    ref = singles(ws, 3)
    monkeypatch.delenv("UWS_SWEEP_FUSION")
    monkeypatch.setenv("UWS_SWEEP_DEPTH", "4")
    if zchunk4 is not None:
        monkeypatch.setenv("UWS_ZCHUNK4", str(zchunk4))
    got = ensemble_run(ws, 3, 4, False, group_bytes=group_bytes)
    for (ug, pg), (ur, pr) in zip(got, ref):
        np.testing.assert_array_equal(ug, ur)
        np.testing.assert_array_equal(pg, pr)
