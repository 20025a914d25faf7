"""Host side of the stretched grid (no GPU): InitMeshBuffers' face
coordinates and the configuration checks of the C ABI."""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

from unitywindsimulation_amd import _native as N
from unitywindsimulation_amd.config import FluidSimConfig, FvmSolverType, GridType
from unitywindsimulation_amd.solver import MESH_ONE_SIDE, MESH_RATIO, MESH_UNIFORM_SIZE, stretched_face_positions


def test_reference_mesh_closed_form():
    """100 cells, dx = 1: 25 cells growing by 1.08 away from a 50-cell
    uniform block on each side.  Sizes are dx * 1.08^k (k = distance from the
    block), mirrored about the block; the total length is the geometric sum."""
    f = stretched_face_positions(100, 1.0, MESH_UNIFORM_SIZE[0], MESH_ONE_SIDE[0], MESH_RATIO).astype(np.float64)
    assert f[0] == 0.0 and np.all(np.diff(f) > 0)
    h = np.diff(f)
    k = np.concatenate([np.arange(24, -1, -1), np.zeros(50), np.arange(25)])
    np.testing.assert_allclose(h, 1.08 ** k, rtol=2e-5)
    np.testing.assert_allclose(h[25:75], 1.0, rtol=1e-5)
    np.testing.assert_allclose(h, h[::-1], rtol=2e-5)
    total = 50 + 2 * (1.08 ** 25 - 1) / 0.08
    assert f[-1] == pytest.approx(total, rel=1e-5)
    # the y axis of the reference mesh: 50 uniform cells, no growth
    fy = stretched_face_positions(50, 1.0, MESH_UNIFORM_SIZE[1], MESH_ONE_SIDE[1], MESH_RATIO)
    np.testing.assert_array_equal(fy, np.arange(51, dtype=np.float32))


@pytest.mark.parametrize("n,uniform,one_side", [(60, 50, 25), (60, 50, 0), (49, 50, 0), (110, 50, 25)])
def test_mesh_generator_rejects_tables_the_reference_overruns(n, uniform, one_side):
    """The reference fills the trailing side from the leading side's size
    table and the uniform block for `uniform` cells: fewer cells than
    uniform + one_side, or a trailing side longer than the leading one, index
    outside its arrays."""
    with pytest.raises(ValueError):
        stretched_face_positions(n, 1.0, uniform, one_side, 1.08)


def _create_status(cf: FluidSimConfig) -> int:
    c = cf.to_c()
    h = ctypes.c_void_p()
    rc = N.lib().uws_create(ctypes.byref(c), ctypes.byref(h))
    assert rc != N.UWS_OK  # the configurations below are all rejected before any device work
    return rc


def test_stretched_grid_configuration_checks():
    cf = FluidSimConfig(gridResX=16, physDomainSize=(16.0, 8.0, 8.0), gridType=GridType.CollNonUniform,
                        fvmSolverType=FvmSolverType.PISO)
    cf.Init()
    # no PISO corrector on the stretched grid (InitPisoShader throws)
    assert _create_status(cf) == N.UWS_ERR_UNSUPPORTED
    c = cf.to_c()
    c.fvm_solver_type = N.SIMPLE
    c.grid_type = N.GRID_STAGGERED
    h = ctypes.c_void_p()
    assert N.lib().uws_create(ctypes.byref(c), ctypes.byref(h)) == N.UWS_ERR_UNSUPPORTED
    c.grid_type = 7
    assert N.lib().uws_create(ctypes.byref(c), ctypes.byref(h)) == N.UWS_ERR_CONFIG
    with pytest.raises(NotImplementedError):
        FluidSimConfig(gridType=GridType.Staggered).Init()
