"""MI355X-native finite-volume incompressible wind solver step.

Drop-in for UnityWindSimulation's ``FvmSolverGpu`` (the ``IFvmSolver`` plugin):
the numerics run as hand-written gfx950 HIP kernels in ``libuws_hip.so``
behind the C ABI of ``include/uws.h``; this package is the host-side mirror
of the reference's ``FluidSimConfig`` / ``RuntimeConfig`` / ``IFvmSolver``.
"""
from .config import (FluidSimConfig, FvmSolverType, GridType, InflowType, PresBndCond, RuntimeConfig, SolidType,
                     TurbulenceModel, VelBndCond, WallModel, load_asset)
from .solver import FvmSolverHip, IFvmSolver

__all__ = [
    "FluidSimConfig", "RuntimeConfig", "IFvmSolver", "FvmSolverHip", "load_asset", "VelBndCond", "PresBndCond",
    "FvmSolverType", "TurbulenceModel", "InflowType", "SolidType", "GridType", "WallModel",
]
