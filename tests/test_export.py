"""netCDF export in the reference exporter's layout, read back with scipy's
independent netCDF-3 reader."""
from __future__ import annotations

import numpy as np
import pytest

from unitywindsimulation_amd.config import FluidSimConfig
from unitywindsimulation_amd.export import NcFlowFieldExporter

scipy_io = pytest.importorskip("scipy.io")


class _FakeSolver:
    """Duck-typed solver: a config and velocity/pressure getters."""

    def __init__(self, cf, seed):
        self.cf = cf
        X, Y, Z = cf.gridRes
        rng = np.random.default_rng(seed)
        self.u = rng.standard_normal((Z, Y, X, 4)).astype(np.float32)
        self.u[..., 3] = 0
        self.p = rng.standard_normal((Z, Y, X)).astype(np.float32)

    def velocity(self, ncomp=3):
        return self.u[..., :ncomp].copy()

    def pressure(self):
        return self.p.copy()


@pytest.mark.parametrize("threaded", [False, True])
def test_nc_export_layout_and_records(tmp_path, threaded):
    cf = FluidSimConfig(physDomainSize=(10.0, 4.0, 6.0), gridResX=5, dt=0.25)
    cf.Init()
    X, Y, Z = cf.gridRes
    path = tmp_path / "flow.nc"
    frames = []
    with NcFlowFieldExporter(path, cf, threaded=threaded) as ex:
        for k in range(3):
            cf._advance(k + 1, 0.25 * (k + 1))
            s = _FakeSolver(cf, k)
            frames.append(s)
            ex.enqueue_field(s)
    f = scipy_io.netcdf_file(str(path), "r", mmap=False)
    try:
        assert f.dimensions["time"] is None and f.dimensions["velX"] == X and f.dimensions["velVec"] == 4
        assert f.dimensions["presZ"] == Z
        ga = f._attributes
        assert ga["gridLayoutType"] == b"Collocated" and float(ga["dt"]) == pytest.approx(0.25)
        np.testing.assert_array_equal(f.variables["velX"][:], (np.arange(X) + 0.5) * 2.0)
        np.testing.assert_array_equal(f.variables["velVec"][:], [0, 1, 2, 3])
        assert f.variables["velField"].dimensions == ("time", "velZ", "velY", "velX", "velVec")
        assert f.variables["presField"].dimensions == ("time", "presZ", "presY", "presX")
        assert f.variables["velField"].shape == (3, Z, Y, X, 4)
        np.testing.assert_array_equal(f.variables["time"][:], [0.25, 0.5, 0.75])
        for k, s in enumerate(frames):
            np.testing.assert_array_equal(f.variables["velField"][k], s.u)
            np.testing.assert_array_equal(f.variables["presField"][k], s.p)
    finally:
        f.close()
