"""The allocation-free warm start (uws_init_vel_pres_fields_from_background_noalloc)
against the allocating one (uws_init_vel_pres_fields_from_background), through
the C ABI: the same fields bit for bit, on two calls in a row on one handle,
the second with a smaller background (the staging array still holds the
first call's larger arrays), for the FP32, FP64 and complex-step solvers,
with and without a background pressure; and an input beyond the staging
array's size is refused with UWS_ERR_TOO_LARGE."""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

from unitywindsimulation_amd import FvmSolverHip
from unitywindsimulation_amd import _native as N
from unitywindsimulation_amd.config import FluidSimConfig

pytestmark = pytest.mark.gpu

FIELDS = (N.FIELD_VEL, N.FIELD_PRES, N.FIELD_VEL_CORR, N.FIELD_PRES_CORR)


def _cfg(grid, dx=1.0, pos=(0.0, 0.0, 0.0), precision=0):
    X, Y, Z = grid
    cf = FluidSimConfig(physDomainSize=(X * dx, Y * dx, Z * dx), gridResX=X, dt=0.05, mu=1.5e-5,
                        physFieldPos=pos, precision=precision)
    cf.Init()
    assert cf.gridRes == tuple(grid)
    return cf


def _background(grid, dx, pos, ncomp, seed, with_pres):
    X, Y, Z = grid
    rng = np.random.default_rng(seed)
    v = np.ascontiguousarray(rng.standard_normal((Z * Y * X, ncomp)).astype(np.float32))
    p = np.ascontiguousarray(rng.standard_normal(Z * Y * X).astype(np.float32)) if with_pres else None
    bg = N.uws_background()
    bg.grid_res[:] = grid
    bg.dx = dx
    bg.phys_field_pos[:] = pos
    return bg, v, p


def _call(lib, name, h, pos, rot, bg, v, p):
    fpos = (ctypes.c_float * 3)(*pos)
    return getattr(lib, name)(h, ctypes.cast(fpos, ctypes.c_void_p), rot, ctypes.byref(bg), N.ptr(v), v.shape[1],
                              N.ptr(p) if p is not None else None)


def _state(s):
    out = [s.field64(f) for f in FIELDS]
    # every velocity buffer got the sample: one step reads them all
    s.Step()
    out += [s.field64(N.FIELD_VEL), s.field64(N.FIELD_PRES)]
    return out


@pytest.mark.parametrize("prec", [N.PREC_FP32, N.PREC_FP64, N.PREC_CSTEP])
@pytest.mark.parametrize("with_pres", [True, False])
def test_noalloc_warm_start_equals_allocating_one(gpu, prec, with_pres):
    lib = N.lib()
    fg_grid, fg_pos = (16, 8, 12), (6.0, 4.0, 1.0)
    calls = [  # (background grid, dx, position, ncomp, seed, orientation): the second is smaller
        ((40, 16, 36), 2.0, (4.0, 3.0, -2.0), 4, 11, 37.5),
        ((12, 6, 10), 3.0, (7.0, 2.0, 2.0), 3, 12, -20.0),
    ]
    cfg = _cfg(fg_grid, pos=fg_pos, precision=prec)
    cfg.velMaxNumIter, cfg.presMaxNumIter = 4, 6
    with FvmSolverHip(cfg) as a, FvmSolverHip(cfg) as b:
        for k, (grid, dx, pos, ncomp, seed, rot) in enumerate(calls):
            bg, v, p = _background(grid, dx, pos, ncomp, seed, with_pres)
            print(f"call {k}: precision {prec}, background {grid} ncomp {ncomp} pressure {with_pres}", flush=True)
            assert _call(lib, "uws_init_vel_pres_fields_from_background_noalloc", a._h, fg_pos, rot, bg, v, p) == 0
            assert _call(lib, "uws_init_vel_pres_fields_from_background", b._h, fg_pos, rot, bg, v, p) == 0
            sa, sb = _state(a), _state(b)
            assert np.any(sa[0] != 0)
            for x, y in zip(sa, sb):
                assert np.array_equal(x.view(np.uint64), y.view(np.uint64)), k


def test_noalloc_warm_start_refuses_oversize(gpu):
    lib = N.lib()
    cfg = _cfg((16, 8, 12))
    with FvmSolverHip(cfg) as s:
        u0 = s.velocity(3).copy()
        # 4 velocity components + pressure of 4 bytes: 20 B per background cell
        cells = N.BG_STAGING_MAX_BYTES // 20 + 1
        grid = (cells, 1, 1)
        bg = N.uws_background()
        bg.grid_res[:] = grid
        bg.dx = 1.0
        v = np.zeros((cells, 4), dtype=np.float32)
        p = np.zeros(cells, dtype=np.float32)
        rc = _call(lib, "uws_init_vel_pres_fields_from_background_noalloc", s._h, (0.0, 0.0, 0.0), 0.0, bg, v, p)
        assert rc == N.UWS_ERR_TOO_LARGE
        assert b"staging" in lib.uws_strerror(rc)
        np.testing.assert_array_equal(s.velocity(3), u0)
        # the same cells without the pressure fit
        rc = _call(lib, "uws_init_vel_pres_fields_from_background_noalloc", s._h, (0.0, 0.0, 0.0), 0.0, bg, v, None)
        assert rc == 0
