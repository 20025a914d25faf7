"""IFvmSolver members on a live HIP solver: ChangePhysDomainSize
(FvmSolverGpu.cs:1034-1046) and the netCDF export of live frames
(NcFlowFieldExporter.cs:295-339)."""
from __future__ import annotations

import copy

import numpy as np
import pytest

from unitywindsimulation_amd import FvmSolverHip, workloads
from unitywindsimulation_amd import _native as N
from unitywindsimulation_amd.export import NcFlowFieldExporter
from unitywindsimulation_amd.geometry import noise_velocity, rasterise

pytestmark = pytest.mark.gpu


def _bits_equal(a, b):
    return np.array_equal(np.asarray(a).view(np.uint32), np.asarray(b).view(np.uint32))


def test_change_phys_domain_size_same_grid(gpu):
    """A new physDomainSize with the same gridRes (dx 1 -> 1.5): the fields
    are re-initialised (InitVelPresFields), flags and boundary textures kept,
    the clock kept; the next steps equal those of a new solver made with the
    new configuration, the same mask and the same (kept) textures."""
    w = workloads.tiny()
    with FvmSolverHip(w.cf) as a:
        a.InitFlags(w.flags)
        a.set_velocity(noise_velocity(w.flags, w.noise_seed, w.noise_amp))
        a.step(2)
        bnd = [a.bnd_field(f) for f in range(6)]
        X, Y, Z = w.grid
        a.cf.physDomainSize = (1.5 * X, 1.5 * Y, 1.5 * Z)
        a.ChangePhysDomainSize()
        assert a.cf.gridRes == (X, Y, Z) and abs(a.cf.dx - 1.5) < 1e-7
        np.testing.assert_array_equal(a.GetFlagField().reshape(-1), (w.flags > 0).astype(np.int32))
        for f in range(6):
            np.testing.assert_array_equal(a.bnd_field(f), bnd[f])
        assert np.all(a.velocity(3) == 0) and np.all(a.pressure() == 0)
        step = np.zeros(1, dtype=np.int64)
        a.step(2)
        cf_b = copy.deepcopy(a.cf)
        with FvmSolverHip(cf_b) as b:
            b.InitFlags(w.flags)
            for f in range(6):
                b.set_bnd_field(f, bnd[f])
            b.step(2)
            assert _bits_equal(a.velocity(3), b.velocity(3)) and _bits_equal(a.pressure(), b.pressure())
        import ctypes
        t = ctypes.c_double()
        assert a._lib.uws_get_time(a._h, step.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)), ctypes.byref(t)) == 0
        assert step[0] == 4


def test_change_phys_domain_size_new_grid(gpu):
    """A longer domain (16 -> 24 cells in x): the handle's fields are
    reallocated; with the new mask the steps equal those of a new solver."""
    w = workloads.tiny()
    with FvmSolverHip(w.cf) as a:
        a.InitFlags(w.flags)
        a.step(1)
        a.cf.gridResX = 24
        a.cf.physDomainSize = (24.0, 8.0, 12.0)
        a.ChangePhysDomainSize()
        assert a.cf.gridRes == (24, 8, 12)
        assert np.all(a.GetFlagField() == 0)
        flags = rasterise((24, 8, 12), w.boxes)
        a.InitFlags(flags)
        u0 = noise_velocity(flags, 5, 0.05)
        a.set_velocity(u0)
        a.step(2)
        with FvmSolverHip(copy.deepcopy(a.cf)) as b:
            b.InitFlags(flags)
            b.set_velocity(u0)
            b.step(2)
            assert _bits_equal(a.velocity(3), b.velocity(3)) and _bits_equal(a.pressure(), b.pressure())
    # precision changes are refused
    with FvmSolverHip(workloads.tiny().cf) as a:
        cf = copy.deepcopy(a.cf)
        cf.precision = N.PREC_FP64
        cf.Init()
        rc = a._lib.uws_change_phys_domain_size(a._h, cf.to_c())
        assert rc == N.UWS_ERR_CONFIG


def test_export_live_frames(gpu, tmp_path):
    """Three frames of a stepping HIP solver through the exporter (writer
    thread), read back with scipy's netCDF-3 reader: each equals
    GetVelField / GetPresField at the time it was enqueued."""
    scipy_io = pytest.importorskip("scipy.io")
    w = workloads.tiny()
    X, Y, Z = w.grid
    path = tmp_path / "live.nc"
    frames = []
    with FvmSolverHip(w.cf) as s, NcFlowFieldExporter(path, s.cf, threaded=True) as ex:
        s.InitFlags(w.flags)
        s.set_velocity(noise_velocity(w.flags, w.noise_seed, w.noise_amp))
        for _ in range(3):
            s.step(2)
            ex.enqueue_field(s)
            frames.append((s.cf.currPhyTime, s.GetVelField().reshape(Z, Y, X, 4), s.GetPresField().reshape(Z, Y, X)))
    f = scipy_io.netcdf_file(str(path), "r", mmap=False)
    try:
        assert f.variables["velField"].shape == (3, Z, Y, X, 4)
        for k, (t, u, p) in enumerate(frames):
            assert f.variables["time"][k] == np.float32(t)
            np.testing.assert_array_equal(f.variables["velField"][k], u)
            np.testing.assert_array_equal(f.variables["presField"][k], p)
        assert np.any(f.variables["velField"][2] != f.variables["velField"][0])
    finally:
        f.close()
