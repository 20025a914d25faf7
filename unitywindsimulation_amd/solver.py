"""Host-side mirror of the reference's solver plugin interface.

``IFvmSolver`` restates ``Assets/Code/FvmSolver/IFvmSolver.cs:3-24``;
``FvmSolverHip`` implements it over the C ABI of ``libuws_hip.so`` the way
``FvmSolverGpu`` (``Assets/Code/FvmSolver/FvmSolverGpu.cs``) implements it over
Unity compute shaders.  Fields are returned as numpy arrays: velocity
``[z][y][x][4]`` (the reference's float4 texel, w = 0), pressure ``[z][y][x]``,
flags ``[z][y][x]`` int32.
"""
from __future__ import annotations

import abc
import ctypes
from ctypes import byref, c_double, c_int64, c_void_p

import numpy as np

from . import _native as N
from .config import FluidSimConfig, GridType, InflowType, RuntimeConfig

# InitMeshBuffers' hard-coded stretched mesh (FvmSolverGpu.cs:692-697): a
# uniform block of 50 cells per axis, 25 growing cells before it along x and
# z, none along y, growth ratio 1.08
MESH_UNIFORM_SIZE = (50, 50, 50)
MESH_ONE_SIDE = (25, 0, 25)
MESH_RATIO = 1.08


def stretched_face_positions(n: int, dx: float, uniform: int, one_side: int, ratio: float) -> np.ndarray:
    """One axis of InitMeshBuffers (FvmSolverGpu.cs:682-788): n + 1 face
    coordinates (float32).  Raises ValueError where the reference's generator
    indexes outside its tables."""
    out = np.zeros(n + 1, dtype=np.float32)
    rc = N.lib().uws_stretched_face_positions(int(n), float(dx), int(uniform), int(one_side), float(ratio),
                                              N.ptr(out))
    if rc == N.UWS_ERR_CONFIG:
        raise ValueError(f"stretched mesh: {n} cells do not fit a uniform block of {uniform} "
                         f"with {one_side} growing cells on the leading side")
    N.check(rc, "uws_stretched_face_positions")
    return out


def reference_mesh(cf: FluidSimConfig):
    """The face coordinates FvmSolverGpu builds for GridType.CollNonUniform."""
    return [stretched_face_positions(n, float(cf.dx), u, o, MESH_RATIO)
            for n, u, o in zip(cf.gridRes, MESH_UNIFORM_SIZE, MESH_ONE_SIDE)]


class IFvmSolver(abc.ABC):
    """IFvmSolver (IFvmSolver.cs:3-24)."""

    @abc.abstractmethod
    def Step(self) -> bool: ...

    @abc.abstractmethod
    def InitFlags(self, flags) -> None: ...

    @abc.abstractmethod
    def InitVelPresFields(self) -> None: ...

    @abc.abstractmethod
    def InitVelPresFieldsFromBackGroundFlow(self, cfBackGround, velTexBackGround, presTexBackGround) -> None: ...

    @abc.abstractmethod
    def SetFixedValueVelBndCond(self) -> None: ...

    @abc.abstractmethod
    def ChangePhysFieldPos(self) -> None: ...

    @abc.abstractmethod
    def ChangePhysDomainSize(self) -> None: ...

    @abc.abstractmethod
    def GetVelField(self): ...

    @abc.abstractmethod
    def GetPresField(self): ...

    @abc.abstractmethod
    def GetFlagField(self): ...


class FvmSolverHip(IFvmSolver):
    """FvmSolverGpu's behaviour on MI355X through ``libuws_hip.so``.

    Constructor = ``FvmSolverGpu(FluidSimConfig, RuntimeConfig)`` (:173-222):
    fields allocated, on a stretched grid the face coordinates set
    (``InitMeshBuffers``, the reference's hard-coded mesh unless
    ``face_positions`` gives three arrays), flags all fluid, fields
    initialised by ``InitVelPresFields``, boundary textures zeroed then filled
    by ``SetFixedValueVelBndCond``; ``SolidType.Box`` applies the hard-coded box.
    """

    def __init__(self, cf: FluidSimConfig, rcf: RuntimeConfig | None = None, face_positions=None):
        self.cf = cf
        cf.Init()
        self.rcf = rcf or RuntimeConfig.from_config(cf)
        self._lib = N.lib()
        self._cfg = cf.to_c()
        h = c_void_p()
        N.check(self._lib.uws_create(byref(self._cfg), byref(h)), "uws_create")
        self._h = h
        X, Y, Z = cf.gridRes
        self.shape = (Z, Y, X)
        self.n = X * Y * Z
        if int(cf.gridType) == int(GridType.CollNonUniform):
            self.set_face_positions(*(face_positions if face_positions is not None else reference_mesh(cf)))
        elif face_positions is not None:
            raise ValueError("face positions need GridType.CollNonUniform")
        self.InitVelPresFields()
        if int(cf.solidType) == 1:  # SolidType.Box (FvmSolverGpu.cs:216-217)
            N.check(self._lib.uws_init_box(self._h), "uws_init_box")
        self.SetFixedValueVelBndCond()

    # ---- lifetime ----
    def close(self) -> None:
        if getattr(self, "_h", None):
            self._lib.uws_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- stretched grid ----
    def set_face_positions(self, fx, fy, fz) -> None:
        """facePosX/Y/Z: X+1, Y+1, Z+1 strictly increasing coordinates."""
        Z, Y, X = self.shape
        a = [np.ascontiguousarray(np.asarray(v, dtype=np.float32).reshape(-1)) for v in (fx, fy, fz)]
        if [v.size for v in a] != [X + 1, Y + 1, Z + 1]:
            raise ValueError("face positions need X+1, Y+1, Z+1 entries")
        N.check(self._lib.uws_set_face_positions(self._h, *[N.ptr(v) for v in a]), "uws_set_face_positions")

    def face_positions(self):
        Z, Y, X = self.shape
        a = [np.zeros(n + 1, dtype=np.float32) for n in (X, Y, Z)]
        N.check(self._lib.uws_get_face_positions(self._h, *[N.ptr(v) for v in a]), "uws_get_face_positions")
        return a

    # ---- IFvmSolver ----
    def Step(self) -> bool:
        """One whole step; returns readyForRender (always True: no time slicing)."""
        self.step(1)
        return True

    def InitFlags(self, flags) -> None:
        a = np.ascontiguousarray(np.asarray(flags, dtype=np.int32).reshape(-1))
        N.check(self._lib.uws_init_flags(self._h, N.ptr(a), a.size), "uws_init_flags")

    def InitVelPresFields(self) -> None:
        it = int(self.cf.inflowType)
        # power-law inflow has no reference init kernel; the field starts at rest
        if it == int(InflowType.PowerLaw):
            N.check(self._lib.uws_init_vel_pres_fields(self._h, N.INFLOW_CONSTANT, 0.0), "uws_init_vel_pres_fields")
        else:
            N.check(self._lib.uws_init_vel_pres_fields(self._h, it, float(self.rcf.windSpeed)),
                    "uws_init_vel_pres_fields")

    def InitVelPresFieldsFromBackGroundFlow(self, cfBackGround, velTexBackGround, presTexBackGround) -> None:
        """Warm start from a background flow (FvmSolverGpu.cs:898-933): cfBackGround
        gives its grid, dx and physFieldPos, velTexBackGround its velocity as
        [z][y][x][3|4] (what GetVelField returns), presTexBackGround its pressure
        [z][y][x] (or None); this solver's physFieldPos and the runtime flow-field
        orientation place the grid (uws_init_vel_pres_fields_from_background)."""
        cfBackGround.Init()
        bX, bY, bZ = cfBackGround.gridRes
        nb = bX * bY * bZ
        v = np.ascontiguousarray(np.asarray(velTexBackGround, dtype=np.float32).reshape(nb, -1))
        if v.shape[1] not in (3, 4):
            raise ValueError("background velocity must have 3 or 4 components per cell")
        bg = N.uws_background()
        bg.grid_res[:] = (bX, bY, bZ)
        bg.dx = float(cfBackGround.dx)
        bg.phys_field_pos[:] = tuple(float(x) for x in cfBackGround.physFieldPos)
        pos = (ctypes.c_float * 3)(*[float(x) for x in self.cf.physFieldPos])
        pp = None
        if presTexBackGround is not None:
            pr = np.ascontiguousarray(np.asarray(presTexBackGround, dtype=np.float32).reshape(nb))
            pp = N.ptr(pr)
        N.check(self._lib.uws_init_vel_pres_fields_from_background(
            self._h, ctypes.cast(pos, c_void_p), float(self.rcf.flowFieldOrientation), byref(bg), N.ptr(v),
            v.shape[1], pp), "uws_init_vel_pres_fields_from_background")

    def SetFixedValueVelBndCond(self) -> None:
        vz = (ctypes.c_float * 3)(*[float(v) for v in self.cf.velZn])
        N.check(self._lib.uws_set_fixed_value_vel_bnd_cond(
            self._h, int(self.cf.inflowType), float(self.rcf.windSpeed), ctypes.cast(vz, c_void_p),
            float(self.cf.inflowRefHeight), float(self.cf.inflowAlpha)), "uws_set_fixed_value_vel_bnd_cond")

    def ChangePhysFieldPos(self) -> None:
        self.InitVelPresFields()

    def ChangePhysDomainSize(self) -> None:
        """FvmSolverGpu.ChangePhysDomainSize (FvmSolverGpu.cs:1034-1046): take
        the configuration's re-derived gridRes and dx (cf.Init()) and
        re-initialise the fields (InitVelPresFields).  The handle's fields are
        reallocated for the new grid (uws_change_phys_domain_size); flags and
        boundary textures survive where gridRes is unchanged.  After a grid
        change the mask is all fluid (the controller re-voxelizes and calls
        InitFlags, FvmSolverController.cs:869-874; SolidType.Box re-applies
        the hard-coded box), the stretched mesh is rebuilt as in the
        constructor, and the inflow texture is refilled."""
        old_grid = self.cf.gridRes
        self.cf.Init()
        self._cfg = self.cf.to_c()
        N.check(self._lib.uws_change_phys_domain_size(self._h, byref(self._cfg)), "uws_change_phys_domain_size")
        X, Y, Z = self.cf.gridRes
        self.shape = (Z, Y, X)
        self.n = X * Y * Z
        changed = tuple(self.cf.gridRes) != tuple(old_grid)
        if int(self.cf.gridType) == int(GridType.CollNonUniform):
            self.set_face_positions(*reference_mesh(self.cf))
        self.InitVelPresFields()
        if changed:
            if int(self.cf.solidType) == 1:
                N.check(self._lib.uws_init_box(self._h), "uws_init_box")
            self.SetFixedValueVelBndCond()

    def GetVelField(self) -> np.ndarray:
        return self.velocity(4)

    def GetPresField(self) -> np.ndarray:
        return self.pressure()

    def GetFlagField(self) -> np.ndarray:
        out = np.empty(self.n, dtype=np.int32)
        N.check(self._lib.uws_get_flags(self._h, N.ptr(out), self.n), "uws_get_flags")
        return out.reshape(self.shape)

    # ---- extensions ----
    def step(self, nsteps: int) -> None:
        N.check(self._lib.uws_step(self._h, int(nsteps)), "uws_step")
        s, t = c_int64(), c_double()
        self._lib.uws_get_time(self._h, byref(s), byref(t))
        self.cf._advance(int(s.value), float(t.value))

    def synchronize(self) -> None:
        N.check(self._lib.uws_synchronize(self._h), "uws_synchronize")

    def init_noise(self, seed: int, amp: float) -> None:
        N.check(self._lib.uws_init_noise(self._h, int(seed), float(amp)), "uws_init_noise")

    def velocity(self, ncomp: int = 3) -> np.ndarray:
        out = np.empty((self.n, ncomp), dtype=np.float32)
        N.check(self._lib.uws_get_velocity(self._h, N.ptr(out), ncomp, self.n), "uws_get_velocity")
        return out.reshape(*self.shape, ncomp)

    def set_velocity(self, u) -> None:
        a = np.ascontiguousarray(np.asarray(u, dtype=np.float32).reshape(self.n, -1))
        N.check(self._lib.uws_set_velocity(self._h, N.ptr(a), a.shape[1], self.n), "uws_set_velocity")

    def pressure(self) -> np.ndarray:
        out = np.empty(self.n, dtype=np.float32)
        N.check(self._lib.uws_get_pressure(self._h, N.ptr(out), self.n), "uws_get_pressure")
        return out.reshape(self.shape)

    def set_pressure(self, p) -> None:
        a = np.ascontiguousarray(np.asarray(p, dtype=np.float32).reshape(-1))
        N.check(self._lib.uws_set_pressure(self._h, N.ptr(a), self.n), "uws_set_pressure")

    def field(self, fid: int) -> np.ndarray:
        nc = N.FIELD_NCOMP[fid]
        out = np.empty(self.n * nc, dtype=np.float32)
        N.check(self._lib.uws_get_field(self._h, fid, N.ptr(out), out.size), f"uws_get_field({fid})")
        return out.reshape(self.n, nc) if nc == 3 else out

    def set_field(self, fid: int, a) -> None:
        a = np.ascontiguousarray(np.asarray(a, dtype=np.float32).reshape(-1))
        N.check(self._lib.uws_set_field(self._h, fid, N.ptr(a), a.size), f"uws_set_field({fid})")

    def bnd_field(self, face: int) -> np.ndarray:
        X, Y, Z = self.cf.gridRes
        n = (Y if face < 2 else X) * (Z if face < 4 else Y)
        out = np.empty((n, 3), dtype=np.float32)
        N.check(self._lib.uws_get_bnd_field(self._h, face, N.ptr(out), n), "uws_get_bnd_field")
        return out

    def set_bnd_field(self, face: int, xyz) -> None:
        a = np.ascontiguousarray(np.asarray(xyz, dtype=np.float32).reshape(-1, 3))
        N.check(self._lib.uws_set_bnd_field(self._h, face, N.ptr(a), a.shape[0]), "uws_set_bnd_field")

    def mean_reset(self) -> None:
        N.check(self._lib.uws_mean_reset(self._h), "uws_mean_reset")

    def mean_accumulate(self) -> None:
        """Add the current velocity to the FP64 time sum on the device."""
        N.check(self._lib.uws_mean_accumulate(self._h), "uws_mean_accumulate")

    def mean_velocity(self):
        """(time mean of u as (N, 3) float64, number of accumulated steps)."""
        out = np.empty((self.n, 3), dtype=np.float64)
        cnt = c_int64()
        N.check(self._lib.uws_get_mean_velocity(self._h, N.ptr(out), out.size, byref(cnt)), "uws_get_mean_velocity")
        return out, int(cnt.value)

    def bnd_size(self, face: int) -> int:
        X, Y, Z = self.cf.gridRes
        return (Y if face < 2 else X) * (Z if face < 4 else Y)

    def set_bnd_field64(self, face: int, xyz) -> None:
        """Boundary texture upload in double (exact for FP64 and complex-step runs)."""
        a = np.ascontiguousarray(np.asarray(xyz, dtype=np.float64).reshape(-1, 3))
        N.check(self._lib.uws_set_bnd_field_f64(self._h, face, N.ptr(a), a.shape[0]), "uws_set_bnd_field_f64")

    def set_smagorinsky_constant(self, cs: float) -> None:
        """smagorinskyConstant of the live solver, used from the next step on."""
        N.check(self._lib.uws_set_smagorinsky_constant(self._h, float(cs)), "uws_set_smagorinsky_constant")
        self.cf.smagorinskyConstant = float(cs)

    def _cstep_len(self, what: int, face: int) -> int:
        if what == N.CSTEP_VEL:
            return 3 * self.n
        if what == N.CSTEP_PRES:
            return self.n
        if what == N.CSTEP_BND:
            return 3 * self.bnd_size(face)
        return 1

    def cstep_set_imag(self, what: int, im, face: int = 0) -> None:
        """Imaginary parts of a complex-step run (precision PREC_CSTEP)."""
        a = np.ascontiguousarray(np.asarray(im, dtype=np.float64).reshape(-1))
        if a.size != self._cstep_len(what, face):
            raise ValueError(f"expected {self._cstep_len(what, face)} values, got {a.size}")
        N.check(self._lib.uws_cstep_set_imag(self._h, what, face, N.ptr(a), a.size), "uws_cstep_set_imag")

    def cstep_get_imag(self, what: int, face: int = 0) -> np.ndarray:
        out = np.empty(self._cstep_len(what, face), dtype=np.float64)
        N.check(self._lib.uws_cstep_get_imag(self._h, what, face, N.ptr(out), out.size), "uws_cstep_get_imag")
        return out.reshape(-1, 3) if what in (N.CSTEP_VEL, N.CSTEP_BND) else out

    def run_stage(self, stage: int) -> None:
        N.check(self._lib.uws_run_stage(self._h, int(stage)), f"uws_run_stage({stage})")

    def time_kernel(self, kernel: int, reps: int = 20) -> float:
        """Average milliseconds per launch of a hot-loop kernel (HIP events on the solver's stream)."""
        ms = c_double()
        N.check(self._lib.uws_time_kernel(self._h, int(kernel), int(reps), byref(ms)), "uws_time_kernel")
        return float(ms.value)

    def norms(self, strict: bool = False) -> np.ndarray:
        """FP64 norms of the last step (needs ``computeNorms``): see ``uws_get_norms``.
        A step that left a non-finite value returns the norms all the same
        (``out[NORM_NONFINITE]`` counts the cells); ``strict`` raises instead."""
        out = np.zeros(N.NORM_COUNT, dtype=np.float64)
        rc = self._lib.uws_get_norms(self._h, N.ptr(out), N.NORM_COUNT)
        if rc != N.UWS_ERR_NONFINITE or strict:
            N.check(rc, "uws_get_norms")
        return out

    def sweep_counts(self):
        """(momentum, SIMPLE pressure) Jacobi sweeps the last step ran (uws_get_sweep_counts)."""
        v, p = ctypes.c_int32(), ctypes.c_int32()
        N.check(self._lib.uws_get_sweep_counts(self._h, byref(v), byref(p)), "uws_get_sweep_counts")
        return v.value, p.value

    def launch_plan(self):
        """(pressure, momentum) sweeps per launch the step's multi-sweep
        policy picks for this handle (uws_get_launch_plan)."""
        pr, ve = ctypes.c_int32(), ctypes.c_int32()
        N.check(self._lib.uws_get_launch_plan(self._h, byref(pr), byref(ve)), "uws_get_launch_plan")
        return pr.value, ve.value

    def field64(self, fid: int) -> np.ndarray:
        nc = N.FIELD_NCOMP[fid]
        out = np.empty(self.n * nc, dtype=np.float64)
        N.check(self._lib.uws_get_field_f64(self._h, fid, N.ptr(out), out.size), f"uws_get_field_f64({fid})")
        return out.reshape(self.n, nc) if nc == 3 else out


class FvmEnsembleHip:
    """Several uneven cases stepped together through ``uws_batch_*``: one
    launch per kernel for the whole ensemble, one HIP graph per step.  Case
    ``c`` behaves exactly like ``FvmSolverHip(cfs[c])``."""

    def __init__(self, cfs):
        self._lib = N.lib()
        self.cfs = list(cfs)
        for cf in self.cfs:
            cf.Init()
        arr = (N.uws_config * len(self.cfs))(*[cf.to_c() for cf in self.cfs])
        h = c_void_p()
        N.check(self._lib.uws_batch_create(ctypes.cast(arr, c_void_p), len(self.cfs), byref(h)), "uws_batch_create")
        self._h = h
        self.shapes = [(cf.gridRes[2], cf.gridRes[1], cf.gridRes[0]) for cf in self.cfs]
        for c, cf in enumerate(self.cfs):
            vz = (ctypes.c_float * 3)(*[float(v) for v in cf.velZn])
            N.check(self._lib.uws_batch_set_fixed_value_vel_bnd_cond(
                self._h, c, int(cf.inflowType), float(cf.velX0[0]), ctypes.cast(vz, c_void_p),
                float(cf.inflowRefHeight), float(cf.inflowAlpha)), "uws_batch_set_fixed_value_vel_bnd_cond")

    def close(self) -> None:
        if getattr(self, "_h", None):
            self._lib.uws_batch_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def ncells(self, c: int) -> int:
        z, y, x = self.shapes[c]
        return x * y * z

    def init_flags(self, c: int, flags) -> None:
        a = np.ascontiguousarray(np.asarray(flags, dtype=np.int32).reshape(-1))
        N.check(self._lib.uws_batch_init_flags(self._h, c, N.ptr(a), a.size), "uws_batch_init_flags")

    def init_noise(self, c: int, seed: int, amp: float) -> None:
        N.check(self._lib.uws_batch_init_noise(self._h, c, int(seed), float(amp)), "uws_batch_init_noise")

    def set_velocity(self, c: int, u) -> None:
        a = np.ascontiguousarray(np.asarray(u, dtype=np.float32).reshape(self.ncells(c), -1))
        N.check(self._lib.uws_batch_set_velocity(self._h, c, N.ptr(a), a.shape[1], a.shape[0]),
                "uws_batch_set_velocity")

    def velocity(self, c: int, ncomp: int = 3) -> np.ndarray:
        out = np.empty((self.ncells(c), ncomp), dtype=np.float32)
        N.check(self._lib.uws_batch_get_velocity(self._h, c, N.ptr(out), ncomp, out.shape[0]),
                "uws_batch_get_velocity")
        return out.reshape(*self.shapes[c], ncomp)

    def pressure(self, c: int) -> np.ndarray:
        out = np.empty(self.ncells(c), dtype=np.float32)
        N.check(self._lib.uws_batch_get_pressure(self._h, c, N.ptr(out), out.size), "uws_batch_get_pressure")
        return out.reshape(self.shapes[c])

    def step(self, nsteps: int) -> None:
        N.check(self._lib.uws_batch_step(self._h, int(nsteps)), "uws_batch_step")

    def synchronize(self) -> None:
        N.check(self._lib.uws_batch_synchronize(self._h), "uws_batch_synchronize")

    def set_launch(self, ty: int, persistent: bool, group_bytes: int = 0) -> None:
        N.check(self._lib.uws_batch_set_launch(self._h, int(ty), int(bool(persistent)), int(group_bytes)),
                "uws_batch_set_launch")

    def num_groups(self) -> int:
        n = ctypes.c_int32()
        N.check(self._lib.uws_batch_num_groups(self._h, byref(n)), "uws_batch_num_groups")
        return int(n.value)

    def autotune(self, steps: int = 3):
        """Pick the fastest launch shape; returns (tile rows, persistent, group bytes, ms per step)."""
        best = (ctypes.c_int64 * 3)()
        ms = c_double()
        N.check(self._lib.uws_batch_autotune(self._h, int(steps), ctypes.cast(best, c_void_p), byref(ms)),
                "uws_batch_autotune")
        return int(best[0]), bool(best[1]), int(best[2]), float(ms.value)
