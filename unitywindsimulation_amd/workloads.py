"""Synthetic wind-engineering workloads (grids, buildings, inflow, physics).

Axes are the reference's: x streamwise, y up, z spanwise; a grid written
``length x span x height`` in the workload descriptions is ``gridRes =
(length, height, span)`` here.

Common physics (all workloads): dx = 1 m, air (mu = 1.5e-5, den = 1), inlet
X0 fixed-value power law u = 5 (y/16)^0.25 m/s at the cell-centre height
y = (j + 1/2) dx, zero-gradient outlet Xn with fixed (zero) pressure,
ground Y0 a no-slip wall, free-slip (symmetry) top and sides, pressure zero
gradient elsewhere (the boundary set of the reference's AijBuilding0 case,
Code/Config/AijBuilding0.asset:35-46).  Smagorinsky LES with Cs = 0.1, the
rough-wall log law (z0 = 0.01 m) on the ground and all building faces (all
three wall-viscosity slots), PISO with 2 correctors, 50 momentum and 50
pressure Jacobi sweeps, dt from CFL 0.5 on the peak cell speed, taken as
PEAK_TO_INLET = 4 times the largest inlet speed (the flow between and over
the buildings reaches 20-31 m/s for a 7-8.4 m/s inlet peak).  The reference's
step applies the full SIMPLE pressure correction without under-relaxation to
a centrally differenced momentum equation, which is only stable for small
convective Courant numbers: with CFL 0.5 on the inlet speed (dt four times
larger) the urban case diverges near step 140, with dt/2 near step 820.  The
initial velocity is U(-0.05, 0.05) m/s noise per component in fluid cells.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from .config import (FluidSimConfig, FvmSolverType, GridType, InflowType, PresBndCond, SolidType, TurbulenceModel,
                     VelBndCond, WallModel)
from .geometry import Box, SplitMix64, lattice_layout, rasterise

U_REF, H_REF, ALPHA = 5.0, 16.0, 0.25
PEAK_TO_INLET = 4.0


@dataclass
class Workload:
    name: str
    cf: FluidSimConfig
    flags: np.ndarray            # int32 [X*Y*Z]
    noise_seed: int
    noise_amp: float = 0.05
    steps: int = 200
    mean_last: int = 0
    boxes: list = field(default_factory=list)
    face_positions: list | None = None  # stretched grid: facePosX/Y/Z (float32)

    @property
    def grid(self):
        return self.cf.gridRes

    @property
    def ncells(self) -> int:
        X, Y, Z = self.grid
        return X * Y * Z


def power_law_umax(ny: int, dx: float = 1.0) -> float:
    return U_REF * ((ny - 0.5) * dx / H_REF) ** ALPHA


def base_config(grid, vel_iters=50, pres_iters=50, correctors=2) -> FluidSimConfig:
    X, Y, Z = grid
    dt = 0.5 * 1.0 / (PEAK_TO_INLET * power_law_umax(Y))
    cf = FluidSimConfig(
        physDomainSize=(float(X), float(Y), float(Z)), gridResX=X, dt=dt, mu=1.5e-5, den=1.0,
        velBndCondX0=VelBndCond.FixedValue, velBndCondXn=VelBndCond.ZeroGrad, velBndCondY0=VelBndCond.FixedValue,
        velBndCondYn=VelBndCond.Symmetry, velBndCondZ0=VelBndCond.Symmetry, velBndCondZn=VelBndCond.Symmetry,
        presBndCondX0=PresBndCond.ZeroGrad, presBndCondXn=PresBndCond.FixedValue,
        presBndCondY0=PresBndCond.ZeroGrad, presBndCondYn=PresBndCond.Symmetry,
        presBndCondZ0=PresBndCond.Symmetry, presBndCondZn=PresBndCond.Symmetry,
        velX0=(U_REF, 0.0, 0.0), fvmSolverType=FvmSolverType.PISO, PISONumCorrectors=correctors,
        turbulenceModel=TurbulenceModel.Smagorinsky, smagorinskyConstant=0.1, inflowType=InflowType.PowerLaw,
        velMaxNumIter=vel_iters, presMaxNumIter=pres_iters, wallSlots=3, wallModel=WallModel.RoughLogLaw,
        wallZ0=0.01, inflowRefHeight=H_REF, inflowAlpha=ALPHA,
    )
    cf.Init()
    assert cf.gridRes == tuple(grid), (cf.gridRes, grid)
    return cf


def isolated_building(grid=(128, 64, 64), b=8) -> Workload:
    """A b x b x 2b square prism in a 128 x 64 x 64 grid (the small plain case):
    x in [32, 40), y in [0, 16), z in [28, 36) for b = 8 (a quarter of the
    length downstream of the inlet, centred spanwise)."""
    X, Y, Z = grid
    x0 = X // 4
    z0 = Z // 2 - b // 2
    boxes = [Box(x0, x0 + b, z0, z0 + b, 2 * b)]
    cf = base_config(grid)
    return Workload(f"isolated_building_{X}x{Y}x{Z}", cf, rasterise(grid, boxes), noise_seed=0, steps=200,
                    boxes=boxes)


def urban(grid=(256, 96, 128), n=40, lattice=(8, 5), seed=1, height=(4, 24), steps=1000, mean_last=500) -> Workload:
    """Urban case: n seeded rectangular buildings on a jittered lattice
    (footprints U{4..12} per side, heights U{height}), 256 x 128 span x 96 high."""
    boxes = lattice_layout(grid, n, lattice, seed, foot=(4, 12), height=height)
    cf = base_config(grid)
    X, Y, Z = grid
    return Workload(f"urban_{n}bldg_{X}x{Y}x{Z}", cf, rasterise(grid, boxes), noise_seed=seed, steps=steps,
                    mean_last=mean_last, boxes=boxes)


def tiny(grid=(16, 8, 12), seed=7, les=True, piso=True, vel_iters=6, pres_iters=8) -> Workload:
    """A few-thousand-cell case with two buildings, for oracle-sized checks."""
    X, Y, Z = grid
    boxes = [Box(4, 7, 3, 6, 4), Box(9, 11, 7, 10, 6)]
    cf = base_config(grid, vel_iters=vel_iters, pres_iters=pres_iters)
    if not les:
        cf.turbulenceModel = TurbulenceModel.None_
    if not piso:
        cf.fvmSolverType = FvmSolverType.SIMPLE
    cf.Init()
    return Workload(f"tiny_{X}x{Y}x{Z}", cf, rasterise(grid, boxes), noise_seed=seed, steps=4, boxes=boxes)


ENSEMBLE_GRIDS = [(64, 64, 64), (128, 64, 64), (160, 64, 96), (192, 64, 96)]


def ensemble(n_cases=24, seed=2, grids=ENSEMBLE_GRIDS, buildings=(5, 30), steps=500,
             region=((0.2, 0.5), (0.2, 0.8)), centre_x=0.35) -> list:
    """The uneven ensemble: each case draws its grid from `grids` (x, height,
    span: 64^3, 128 x 64 span x 64 high, 160 x 96 x 64, 192 x 96 x 64), a
    building count in `buildings` and an inflow direction 15 deg * U{0..23};
    the inflow always enters at X0, so the building layout is rotated into the
    solver frame by the negative inflow angle about (centre_x * X, Z / 2).
    The layout lattice covers `region` (fractions of x and z) around that
    centre, so every rotation keeps the buildings in the upstream two thirds
    of the domain: a wake that reaches the zero-gradient outlet draws
    backflow through it, which the reference's outlet condition lets grow."""
    rng = SplitMix64(seed)
    out = []
    for c in range(n_cases):
        grid = grids[rng.randint(0, len(grids) - 1)]
        nb = rng.randint(buildings[0], buildings[1])
        angle = 15.0 * rng.randint(0, 23)
        bseed = rng.next()
        X, Y, Z = grid
        lat = (6, 5)
        boxes = lattice_layout(grid, nb, lat, bseed, foot=(4, 12), height=(4, 24), region=region)
        cf = base_config(grid)
        w = Workload(f"ens{c:02d}_{X}x{Y}x{Z}_{nb}bldg_{int(angle)}deg", cf,
                     rasterise(grid, boxes, angle, centre=(centre_x * X, Z / 2.0)),
                     noise_seed=bseed & 0xFFFFFFFF, steps=steps, boxes=boxes)
        w.angle = angle
        out.append(w)
    return out


def precision(grid=(384, 128, 256), n=120, lattice=(15, 10), seed=3, height=(4, 40), steps=500, mean_last=250,
              rounds=3, sweeps=20) -> Workload:
    """The precision case: 384 x 256 span x 128 high with n seeded buildings
    of heights U{4..40}; every pressure solve is rounds x sweeps Jacobi
    sweeps (presMaxNumIter = rounds * sweeps for the plain FP32/FP64 runs,
    refineRounds / refineSweeps for the low-precision modes)."""
    boxes = lattice_layout(grid, n, lattice, seed, foot=(4, 12), height=height)
    cf = base_config(grid, pres_iters=rounds * sweeps)
    cf.refineRounds, cf.refineSweeps = rounds, sweeps
    X, Y, Z = grid
    return Workload(f"precision_{n}bldg_{X}x{Y}x{Z}", cf, rasterise(grid, boxes), noise_seed=seed, steps=steps,
                    mean_last=mean_last, boxes=boxes)


@dataclass
class DiffWorkload:
    """The differentiation case: a workload plus its objective and parameters."""
    base: Workload
    nsteps: int
    ckpt_every: int
    region: tuple          # (x0, z0, width, depth) of the objective's region
    delta: np.ndarray      # per-height-level inlet perturbations [m/s]
    u_ref: float = U_REF
    alpha: float = ALPHA
    cs: float = 0.1


def normal_pairs(rng: SplitMix64, n: int) -> np.ndarray:
    """n standard normal draws from SplitMix64 uniforms (Box-Muller, cosine branch)."""
    out = np.empty(n)
    for i in range(n):
        u1 = max(rng.uniform(), 1.0 / 16777216.0)
        u2 = rng.uniform()
        out[i] = math.sqrt(-2.0 * math.log(u1)) * math.cos(2.0 * math.pi * u2)
    return out


def differentiation(grid=(128, 64, 96), n=20, lattice=(5, 4), seed=4, nsteps=64, ckpt_every=8,
                    region_size=32) -> DiffWorkload:
    """128 x 96 span x 64 high with n seeded buildings; J = mean |u|^2 on the
    2 m layer (average of the cell layers j = 1, 2) over a seeded
    region_size^2 (x, z) region; parameters U_ref, alpha, Cs and one inlet
    perturbation per height level, N(0, 0.05 U_ref)."""
    boxes = lattice_layout(grid, n, lattice, seed, foot=(4, 12), height=(4, 24))
    cf = base_config(grid)
    X, Y, Z = grid
    w = Workload(f"differentiation_{n}bldg_{X}x{Y}x{Z}", cf, rasterise(grid, boxes), noise_seed=seed,
                 steps=nsteps, boxes=boxes)
    rng = SplitMix64(seed ^ 0xD1FF)
    rs = min(region_size, X, Z)
    region = (rng.randint(0, X - rs), rng.randint(0, Z - rs), rs, rs)
    delta = 0.05 * U_REF * normal_pairs(rng, Y)
    return DiffWorkload(w, nsteps, ckpt_every, region, delta)


def growing_faces(n: int, dx: float, ratio: float) -> np.ndarray:
    """n + 1 face coordinates (float32) of cells dx, dx*ratio, dx*ratio^2, ...
    from 0 upwards: a ground-refined vertical axis."""
    sizes = np.empty(n, dtype=np.float32)
    sizes[0] = np.float32(dx)
    for i in range(1, n):
        sizes[i] = sizes[i - 1] * np.float32(ratio)
    out = np.zeros(n + 1, dtype=np.float32)
    for i in range(n):
        out[i + 1] = out[i] + sizes[i]
    return out


def tiny_stretched(grid=(16, 10, 12), seed=7, les=True, vel_iters=6, pres_iters=8, ratio=1.12) -> Workload:
    """tiny() on a stretched grid (GridType.CollNonUniform, SIMPLE: the
    stretched grid has no PISO corrector).  x and z use the reference's
    InitMeshBuffers layout scaled down (a uniform core with cells growing by
    `ratio` on both sides), y grows upward from the ground by `ratio`."""
    from .solver import stretched_face_positions

    X, Y, Z = grid
    w = tiny(grid=grid, seed=seed, les=les, piso=False, vel_iters=vel_iters, pres_iters=pres_iters)
    cf = w.cf
    cf.gridType = GridType.CollNonUniform
    cf.Init()
    fx = stretched_face_positions(X, 1.0, X - 2 * (X // 3), X // 3, ratio)
    fy = growing_faces(Y, 1.0, ratio)
    fz = stretched_face_positions(Z, 1.0, Z - 2 * (Z // 3), Z // 3, ratio)
    w.name = f"tiny_stretched_{X}x{Y}x{Z}"
    w.face_positions = [fx, fy, fz]
    return w


def wind_tunnel_box(vel_iters=800, pres_iters=2000, steps=20) -> Workload:
    """The reference's stretched-grid case, Code/Config/WindTunnelBoxConfig.asset:
    100 x 50 x 100 cells, dx = 1 m (physDomainSize 100 x 50 x 100), dt = 0.02 s,
    mu = 1, den = 1, uniform inflow 10 m/s at X0, zero-gradient outlet with
    fixed pressure, symmetry elsewhere, the hard-coded box x in [20, 24),
    y in [0, 8), z in [18, 22), laminar SIMPLE with 800 momentum and 2000
    pressure sweeps, on InitMeshBuffers' mesh (a 50-cell uniform core with 25
    cells growing by 1.08 on both sides along x and z, FvmSolverGpu.cs:682-788).
    The asset's initial field is the constant inflow velocity."""
    from .solver import reference_mesh

    cf = FluidSimConfig(
        physDomainSize=(100.0, 50.0, 100.0), gridResX=100, dt=0.02, mu=1.0, den=1.0,
        velBndCondX0=VelBndCond.FixedValue, velBndCondXn=VelBndCond.ZeroGrad, velBndCondY0=VelBndCond.Symmetry,
        velBndCondYn=VelBndCond.Symmetry, velBndCondZ0=VelBndCond.Symmetry, velBndCondZn=VelBndCond.Symmetry,
        presBndCondX0=PresBndCond.ZeroGrad, presBndCondXn=PresBndCond.FixedValue,
        presBndCondY0=PresBndCond.Symmetry, presBndCondYn=PresBndCond.Symmetry,
        presBndCondZ0=PresBndCond.Symmetry, presBndCondZn=PresBndCond.Symmetry,
        velX0=(10.0, 0.0, 0.0), solidType=SolidType.Box, fvmSolverType=FvmSolverType.SIMPLE,
        turbulenceModel=TurbulenceModel.None_, inflowType=InflowType.Constant, gridType=GridType.CollNonUniform,
        velMaxNumIter=vel_iters, presMaxNumIter=pres_iters,
    )
    cf.Init()
    X, Y, Z = cf.gridRes
    flags = np.zeros((Z, Y, X), dtype=np.int32)
    flags[18:22, 0:8, 20:24] = 1
    return Workload(f"wind_tunnel_box_stretched_{X}x{Y}x{Z}", cf, flags.reshape(-1), noise_seed=0, noise_amp=0.0,
                    steps=steps, boxes=[Box(20, 24, 18, 22, 8)], face_positions=reference_mesh(cf))


WORKLOADS = {
    "isolated_building": isolated_building,
    "urban": urban,
    "tiny": tiny,
    "tiny_stretched": tiny_stretched,
    "wind_tunnel_box": wind_tunnel_box,
}
