"""Host mirror of the reference's configuration classes.

``FluidSimConfig`` keeps the field names, defaults and enum integers of
``Assets/Code/FvmSolver/FluidSimConfig.cs:9-370`` (the fields the solver step
reads) and ``Init()`` derives dx, ds, dv, gridRes, nu, D, DInv in float32 as the
C# code does.  ``RuntimeConfig`` mirrors ``Assets/Code/RuntimeConfig/RuntimeConfig.cs``.
``load_asset`` reads the reference's Unity YAML ``.asset`` config files.
"""
from __future__ import annotations

import enum
import re
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np

from . import _native as N


class GridType(enum.IntEnum):          # FluidSimConfig.cs:9-14
    Collocated = 0
    CollNonUniform = 1
    Staggered = 2


class VelBndCond(enum.IntEnum):        # FluidSimConfig.cs:16-21
    FixedValue = 0
    ZeroGrad = 1
    Symmetry = 2


class PresBndCond(enum.IntEnum):       # FluidSimConfig.cs:23-28
    FixedValue = 0
    ZeroGrad = 1
    Symmetry = 2


class FvmSolverType(enum.IntEnum):     # FluidSimConfig.cs:30-34
    SIMPLE = 0
    PISO = 1


class SolidType(enum.IntEnum):         # FluidSimConfig.cs:42-47
    NoSolid = 0
    Box = 1
    Model = 2


class TurbulenceModel(enum.IntEnum):   # FluidSimConfig.cs:73-77
    None_ = 0
    Smagorinsky = 1


class InflowType(enum.IntEnum):        # FluidSimConfig.cs:86-90 (+ PowerLaw extension)
    Constant = 0
    Log = 1
    PowerLaw = 2


class WallModel(enum.IntEnum):         # extension
    SmoothLogLaw = 0
    RoughLogLaw = 1


f32 = np.float32


def _round_to_int(v: float) -> int:
    """Mathf.RoundToInt: round half to even (System.Math.Round)."""
    return int(round(float(v)))


@dataclass
class FluidSimConfig:
    """Fields of FluidSimConfig (FluidSimConfig.cs:113-178) the solver step reads."""

    physDomainSize: tuple = (10.0, 5.0, 10.0)
    gridResX: int = 100
    dt: float = 0.1
    mu: float = 0.005
    den: float = 1.0
    externalForce: tuple = (0.0, 0.0, 0.0)
    velBndCondX0: VelBndCond = VelBndCond.FixedValue
    velBndCondXn: VelBndCond = VelBndCond.FixedValue
    velBndCondY0: VelBndCond = VelBndCond.FixedValue
    velBndCondYn: VelBndCond = VelBndCond.FixedValue
    velBndCondZ0: VelBndCond = VelBndCond.FixedValue
    velBndCondZn: VelBndCond = VelBndCond.FixedValue
    presBndCondX0: PresBndCond = PresBndCond.ZeroGrad
    presBndCondXn: PresBndCond = PresBndCond.ZeroGrad
    presBndCondY0: PresBndCond = PresBndCond.ZeroGrad
    presBndCondYn: PresBndCond = PresBndCond.ZeroGrad
    presBndCondZ0: PresBndCond = PresBndCond.ZeroGrad
    presBndCondZn: PresBndCond = PresBndCond.ZeroGrad
    velX0: tuple = (0.0, 0.0, 0.0)
    velZn: tuple = (0.0, 0.0, 0.0)
    physFieldPos: tuple = (0.0, 0.0, 0.0)   # world position of the grid centre (FluidSimConfig.cs:256)
    solidType: SolidType = SolidType.NoSolid
    fvmSolverType: FvmSolverType = FvmSolverType.SIMPLE
    PISONumCorrectors: int = 2
    turbulenceModel: TurbulenceModel = TurbulenceModel.None_
    smagorinskyConstant: float = 0.15
    inflowType: InflowType = InflowType.Constant
    gridType: GridType = GridType.Collocated
    calResidual: bool = False
    velResidualCheckInterval: int = 10
    velTolerance: float = 1e-15
    velMaxNumIter: int = 50
    presResidualCheckInterval: int = 20
    presTolerance: float = 1e-3
    presMaxNumIter: int = 400
    # extensions (no reference field)
    wallSlots: int = 1               # 1: single-channel wall-viscosity texture as the reference runs
    wallModel: WallModel = WallModel.SmoothLogLaw
    wallZ0: float = 0.01
    inflowRefHeight: float = 16.0    # power-law inlet: u = U_ref (y/h_ref)^alpha
    inflowAlpha: float = 0.25
    precision: int = 0               # 0 FP32, 1 FP64, 2 FP32 + fp16 corrections, 3 FP32 + bf16 corrections
    refineRounds: int = 3            # low-precision modes: defect-correction rounds per pressure solve
    refineSweeps: int = 20           # low-precision modes: sweeps per round
    computeNorms: bool = False       # FP64 residual / divergence norms every step
    # derived by Init()
    _derived: dict = field(default_factory=dict, repr=False)

    # ---- FluidSimConfig.Init (FluidSimConfig.cs:322-346) ----
    def Init(self) -> "FluidSimConfig":
        if self.gridType == GridType.Staggered:
            raise NotImplementedError("GridType.Staggered is not built (collocated and stretched grids only)")
        dx = f32(self.physDomainSize[0]) / f32(self.gridResX)
        ds = dx * dx
        dv = dx * dx * dx
        gy = _round_to_int(f32(self.physDomainSize[1]) / dx)
        gz = _round_to_int(f32(self.physDomainSize[2]) / dx)
        nu = f32(self.mu) / f32(self.den)
        D = dv / f32(self.dt) + f32(6) * nu * ds / dx
        self._derived = dict(dx=dx, ds=ds, dv=dv, gridRes=(int(self.gridResX), gy, gz), nu=nu, D=D,
                             DInv=f32(1) / D, currPhyTime=0.0, currSimStep=0)
        return self

    def _d(self, k):
        if not self._derived:
            self.Init()
        return self._derived[k]

    dx = property(lambda self: self._d("dx"))
    ds = property(lambda self: self._d("ds"))
    dv = property(lambda self: self._d("dv"))
    gridRes = property(lambda self: self._d("gridRes"))
    nu = property(lambda self: self._d("nu"))

    @property
    def currPhyTime(self) -> float:
        return self._d("currPhyTime")

    @property
    def currSimStep(self) -> int:
        return self._d("currSimStep")

    def _advance(self, steps: int, t: float) -> None:
        self._derived["currSimStep"] = steps
        self._derived["currPhyTime"] = t

    def to_c(self) -> N.uws_config:
        """The POD configuration that crosses the C ABI."""
        c = N.uws_config()
        gx, gy, gz = self.gridRes
        c.grid_res[:] = (gx, gy, gz)
        c.dx = float(self.dx)
        c.dt = float(self.dt)
        c.mu = float(self.mu)
        c.den = float(self.den)
        c.external_force[:] = tuple(float(v) for v in self.externalForce)
        c.vel_bnd[:] = [int(v) for v in (self.velBndCondX0, self.velBndCondXn, self.velBndCondY0,
                                          self.velBndCondYn, self.velBndCondZ0, self.velBndCondZn)]
        c.pres_bnd[:] = [int(v) for v in (self.presBndCondX0, self.presBndCondXn, self.presBndCondY0,
                                           self.presBndCondYn, self.presBndCondZ0, self.presBndCondZn)]
        c.fvm_solver_type = int(self.fvmSolverType)
        c.piso_num_correctors = int(self.PISONumCorrectors)
        c.turbulence_model = int(self.turbulenceModel)
        c.smagorinsky_constant = float(self.smagorinskyConstant)
        c.vel_max_num_iter = int(self.velMaxNumIter)
        c.pres_max_num_iter = int(self.presMaxNumIter)
        c.cal_residual = int(bool(self.calResidual))
        c.vel_residual_check_interval = int(self.velResidualCheckInterval)
        c.vel_tolerance = float(self.velTolerance)
        c.pres_residual_check_interval = int(self.presResidualCheckInterval)
        c.pres_tolerance = float(self.presTolerance)
        c.wall_slots = int(self.wallSlots)
        c.wall_model = int(self.wallModel)
        c.wall_z0 = float(self.wallZ0)
        c.precision = int(self.precision)
        c.refine_rounds = int(self.refineRounds)
        c.refine_sweeps = int(self.refineSweeps)
        c.compute_norms = int(bool(self.computeNorms))
        c.grid_type = int(self.gridType)
        return c


@dataclass
class RuntimeConfig:
    """RuntimeConfig (RuntimeConfig.cs:6-17): windSpeed starts as cf.velX0.x."""

    flowFieldOrientation: float = 0.0
    windSpeed: float = 10.0
    llhCoord: tuple = (0.0, 0.0, 0.0)

    @classmethod
    def from_config(cls, cf: FluidSimConfig, llh=(0.0, 0.0, 0.0)) -> "RuntimeConfig":
        return cls(windSpeed=float(cf.velX0[0]), llhCoord=tuple(llh))


_ENUM_FIELDS = {
    "velBndCondX0": VelBndCond, "velBndCondXn": VelBndCond, "velBndCondY0": VelBndCond,
    "velBndCondYn": VelBndCond, "velBndCondZ0": VelBndCond, "velBndCondZn": VelBndCond,
    "presBndCondX0": PresBndCond, "presBndCondXn": PresBndCond, "presBndCondY0": PresBndCond,
    "presBndCondYn": PresBndCond, "presBndCondZ0": PresBndCond, "presBndCondZn": PresBndCond,
    "solidType": SolidType, "fvmSolverType": FvmSolverType, "turbulenceModel": TurbulenceModel,
    "inflowType": InflowType, "gridType": GridType,
}
_VEC_FIELDS = ("physDomainSize", "externalForce", "velX0", "velZn", "physFieldPos")
_INT_FIELDS = ("gridResX", "PISONumCorrectors", "velResidualCheckInterval", "velMaxNumIter",
               "presResidualCheckInterval", "presMaxNumIter")
_FLOAT_FIELDS = ("dt", "mu", "den", "smagorinskyConstant", "velTolerance", "presTolerance")


def load_asset(path) -> FluidSimConfig:
    """Read a Unity ``FluidSimConfig`` ``.asset`` file (YAML with Unity tags).

    Only ``key: scalar`` lines and the ``x/y/z`` children of vector fields are
    interpreted; fields the solver step does not read are ignored.
    """
    text = Path(path).read_text()
    scalars: dict[str, str] = {}
    vectors: dict[str, dict[str, float]] = {}
    current = None
    for line in text.splitlines():
        m = re.match(r"^  (\w+):\s*(.*)$", line)
        if m:
            key, val = m.group(1), m.group(2).strip()
            current = key if val == "" else None
            if val.startswith("{"):
                vectors[key] = {k: float(v) for k, v in re.findall(r"(\w+):\s*([-+0-9.eE]+)", val)}
            elif val != "":
                scalars[key] = val
            continue
        m = re.match(r"^    ([xyz]):\s*([-+0-9.eE]+)\s*$", line)
        if m and current:
            vectors.setdefault(current, {})[m.group(1)] = float(m.group(2))
    cf = FluidSimConfig()
    for k in _VEC_FIELDS:
        if k in vectors:
            v = vectors[k]
            setattr(cf, k, (v.get("x", 0.0), v.get("y", 0.0), v.get("z", 0.0)))
    for k in _INT_FIELDS:
        if k in scalars:
            setattr(cf, k, int(scalars[k]))
    for k in _FLOAT_FIELDS:
        if k in scalars:
            setattr(cf, k, float(scalars[k]))
    for k, en in _ENUM_FIELDS.items():
        if k in scalars:
            setattr(cf, k, en(int(scalars[k])))
    if "calResidual" in scalars:
        cf.calResidual = bool(int(scalars["calResidual"]))
    return cf
