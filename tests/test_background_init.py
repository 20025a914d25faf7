"""Warm start from a background flow (InitVelPresFieldsFromBackGroundFlow) and
the FlagExporter mask format: the CPU restatement pinned by closed forms, and
(GPU) the kernel against it."""
from __future__ import annotations

import numpy as np
import pytest

from oracle.oracle import background_sample
from unitywindsimulation_amd import FvmSolverHip
from unitywindsimulation_amd import _native as N
from unitywindsimulation_amd.config import FluidSimConfig
from unitywindsimulation_amd.geometry import read_flag_file, write_flag_file


def _cfg(grid, dx=1.0, pos=(0.0, 0.0, 0.0), precision=0):
    X, Y, Z = grid
    cf = FluidSimConfig(physDomainSize=(X * dx, Y * dx, Z * dx), gridResX=X, dt=0.05, mu=1.5e-5,
                        physFieldPos=pos, precision=precision)
    cf.Init()
    assert cf.gridRes == tuple(grid)
    return cf


def test_restatement_identity_copies_the_field():
    """Same grid, same position, no rotation: with odd sizes every cell samples
    its own texel centre.  With an even size the cell position (i - N/2) dx
    sits half a texel below the texel centre (i + 1/2) / N: the reference's
    warm start averages each cell with its lower neighbour along that axis
    (clamped at 0)."""
    g = (9, 7, 5)
    rng = np.random.default_rng(0)
    v = rng.standard_normal((5, 7, 9, 4)).astype(np.float32)
    p = rng.standard_normal((5, 7, 9)).astype(np.float32)
    u, q = background_sample(g, 1.0, (3.0, 1.0, -2.0), 0.0, g, 1.0, (3.0, 1.0, -2.0), v, p, dtype=np.float64)
    np.testing.assert_allclose(u, v.reshape(-1, 4)[:, :3], rtol=0, atol=1e-12)
    np.testing.assert_allclose(q, p.reshape(-1), rtol=0, atol=1e-12)
    # in float32 the texel coordinate (i - N/2) / N + 1/2 rounds: weights of ~1e-7 remain
    u32, _ = background_sample(g, 1.0, (3.0, 1.0, -2.0), 0.0, g, 1.0, (3.0, 1.0, -2.0), v, p)
    np.testing.assert_allclose(u32, v.reshape(-1, 4)[:, :3], rtol=0, atol=1e-5)
    ge = (8, 7, 5)  # even x
    pe = rng.standard_normal((5, 7, 8)).astype(np.float64)
    _, qe = background_sample(ge, 1.0, (0.0, 0.0, 0.0), 0.0, ge, 1.0, (0.0, 0.0, 0.0), np.zeros((5, 7, 8, 3)), pe,
                              dtype=np.float64)
    lower = np.concatenate([pe[:, :, :1], pe[:, :, :-1]], axis=2)
    np.testing.assert_allclose(qe, (0.5 * (pe + lower)).reshape(-1), rtol=0, atol=1e-15)


def test_restatement_reproduces_linear_fields_inside():
    """Trilinear filtering reproduces a linear field exactly away from the clamped border."""
    bg = (20, 12, 16)
    bdx = 2.0
    zz, yy, xx = np.meshgrid(np.arange(16), np.arange(12), np.arange(20), indexing="ij")
    # world coordinates of background texel centres (pos = 0): ((i + 1/2) - N/2) dx
    wx, wy, wz = ((xx + 0.5) - 10) * bdx, ((yy + 0.5) - 6) * bdx, ((zz + 0.5) - 8) * bdx
    lin = 1.0 + 0.3 * wx - 0.2 * wy + 0.1 * wz
    v = np.stack([lin, 2 * lin, -lin], axis=-1)
    fg = (10, 8, 10)
    for rot in (0.0, 30.0):
        u, q = background_sample(fg, 1.0, (1.0, 0.5, -1.5), rot, bg, bdx, (0.0, 0.0, 0.0), v, lin, dtype=np.float64)
        zf, yf, xf = np.meshgrid(np.arange(10), np.arange(8), np.arange(10), indexing="ij")
        a = float(np.float32(rot) * np.float32(3.1415926 / 180))
        lx, ly, lz = (xf - 5) * 1.0, (yf - 4) * 1.0, (zf - 5) * 1.0
        # world position of the cell, then its coordinates in the background frame (rotated by -a)
        gx = 1.0 + lx * np.cos(a) + lz * np.sin(a)
        gy = 0.5 + ly
        gz = -1.5 + lz * np.cos(a) - lx * np.sin(a)
        bx, by, bz = gx * np.cos(a) - gz * np.sin(a), gy, gz * np.cos(a) + gx * np.sin(a)
        exact = (1.0 + 0.3 * bx - 0.2 * by + 0.1 * bz).reshape(-1)
        np.testing.assert_allclose(q, exact, rtol=0, atol=1e-9)
        np.testing.assert_allclose(u[:, 1], 2 * exact, rtol=0, atol=1e-9)


def test_flag_file_round_trip(tmp_path):
    rng = np.random.default_rng(3)
    flags = (rng.random(5 * 4 * 3) < 0.3).astype(np.int32)
    path = tmp_path / "flags.txt"
    write_flag_file(path, flags)
    lines = path.read_text().split()
    assert [int(s) for s in lines] == list(np.flatnonzero(flags))
    np.testing.assert_array_equal(read_flag_file(path, flags.size), flags)
    path.write_text("3\n999\n")
    with pytest.raises(ValueError):
        read_flag_file(path, 60)


@pytest.mark.gpu
@pytest.mark.parametrize("prec,tol", [(N.PREC_FP32, 1e-6), (N.PREC_FP64, 1e-12)])
def test_gpu_background_init_matches_restatement(gpu, prec, tol):
    bg_grid, fg_grid = (24, 10, 20), (16, 8, 12)
    cbg = _cfg(bg_grid, dx=2.0, pos=(4.0, 3.0, -2.0))
    cfg = _cfg(fg_grid, dx=1.0, pos=(6.0, 4.0, 1.0), precision=prec)
    rng = np.random.default_rng(11)
    v = rng.standard_normal((20, 10, 24, 4)).astype(np.float32)
    p = rng.standard_normal((20, 10, 24)).astype(np.float32)
    with FvmSolverHip(cfg) as s:
        s.rcf.flowFieldOrientation = 37.5
        s.InitVelPresFieldsFromBackGroundFlow(cbg, v, p)
        u = s.field64(N.FIELD_VEL)
        q = s.field64(N.FIELD_PRES)
        assert np.all(s.field64(N.FIELD_VEL_CORR) == 0)
    ur, qr = background_sample(fg_grid, 1.0, (6.0, 4.0, 1.0), 37.5, bg_grid, 2.0, (4.0, 3.0, -2.0), v, p,
                               dtype=np.float32 if prec == N.PREC_FP32 else np.float64)
    assert np.max(np.abs(u - ur)) <= tol * max(1.0, np.max(np.abs(ur)))
    assert np.max(np.abs(q - qr)) <= tol * max(1.0, np.max(np.abs(qr)))
