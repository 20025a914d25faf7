"""ctypes wrapper of the CPU oracle (``oracle/liboracle_uws.so``).

Test infrastructure: only the tests, the smoke check and the bench's
CPU-baseline leg use it, always as the reference the HIP results are
compared with (or, in the CPU-baseline leg, timed beside).
"""
from __future__ import annotations

import ctypes
import subprocess
from ctypes import POINTER, c_double, c_float, c_int, c_int64, c_void_p
from pathlib import Path

import numpy as np

from unitywindsimulation_amd import _native as N

ORACLE_DIR = Path(__file__).resolve().parent
LIB = ORACLE_DIR / "liboracle_uws.so"

_SIG = {
    "uwso_num_threads": (c_int, []),
    "uwso_create": (c_void_p, [POINTER(N.uws_config), c_int]),
    "uwso_destroy": (None, [c_void_p]),
    "uwso_init_flags": (None, [c_void_p, c_void_p]),
    "uwso_init_box": (None, [c_void_p]),
    "uwso_init_vel_pres_fields": (None, [c_void_p, c_int, c_float]),
    "uwso_set_fixed_value_vel_bnd_cond": (None, [c_void_p, c_int, c_float, c_void_p, c_float, c_float]),
    "uwso_step": (None, [c_void_p, c_int]),
    "uwso_run_stage": (None, [c_void_p, c_int]),
    "uwso_get_field": (c_int, [c_void_p, c_int, c_void_p]),
    "uwso_set_field": (c_int, [c_void_p, c_int, c_void_p]),
    "uwso_get_flags": (None, [c_void_p, c_void_p]),
    "uwso_set_bnd_field": (c_int, [c_void_p, c_int, c_void_p]),
    "uwso_get_bnd_field": (c_int, [c_void_p, c_int, c_void_p]),
    "uwso_log_law_wall_func": (c_double, [c_void_p, c_double]),
    "uwso_residuals": (None, [c_void_p, c_void_p]),
    "uwso_sweep_counts": (None, [c_void_p, c_void_p]),
    "uwso_time": (None, [c_void_p, POINTER(c_int64), POINTER(c_double)]),
    "uwso_set_face_positions": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p]),
    "uwso_get_face_flux": (c_int, [c_void_p, c_void_p]),
    "uwso_set_face_flux": (c_int, [c_void_p, c_void_p]),
    "uwso_run_kernel": (c_int, [c_void_p, c_int]),
}
_lib = None
_variants = {}

# single kernels and extra field ids of the pinning tests (uws_oracle.cpp)
KERNEL_VEL_SWEEP, KERNEL_PISO_AODU, KERNEL_PISO_PRES_SWEEP, KERNEL_PISO_VEL_CORRECT = 0, 1, 2, 3
FIELD_VEL_ITER, FIELD_PC_ITER = 100, 101
_NCOMP = {**N.FIELD_NCOMP, FIELD_VEL_ITER: 3, FIELD_PC_ITER: 1}


def build() -> Path:
    # -B: compile on every call, also over an existing library
    subprocess.run(["make", "-s", "-B", "-C", str(ORACLE_DIR)], check=True)
    return LIB


def _load(path):
    d = ctypes.CDLL(str(path))
    for k, (r, a) in _SIG.items():
        fn = getattr(d, k)
        fn.restype = r
        fn.argtypes = a
    return d


def lib(path=None):
    """The oracle library; `path` loads another build of the same source
    (the planted-mistake builds of the pinning tests)."""
    global _lib
    if path is not None:
        key = str(path)
        if key not in _variants:
            _variants[key] = _load(path)
        return _variants[key]
    if _lib is None:
        if not LIB.exists():
            build()
        _lib = _load(LIB)
    return _lib


def build_variant(mutant: int, out_dir) -> Path:
    """The oracle compiled with planted mistake number `mutant` (-DUWSO_MUTANT)."""
    out = Path(out_dir) / f"liboracle_mutant{mutant}.so"
    subprocess.run(["g++", "-O1", "-std=c++17", "-fPIC", "-fopenmp", "-ffp-contract=off", "-fno-fast-math",
                    f"-DUWSO_MUTANT={int(mutant)}", "-shared", "-o", str(out), str(ORACLE_DIR / "uws_oracle.cpp")],
                   check=True)
    return out


class Oracle:
    """The reference step on the CPU; precision 0 = FP32, 1 = FP64."""

    def __init__(self, cf, precision: int = 0, libpath=None):
        self.cf = cf
        cf.Init()
        self._cfg = cf.to_c()
        self.L = lib(libpath)
        self.h = self.L.uwso_create(ctypes.byref(self._cfg), precision)
        X, Y, Z = cf.gridRes
        self.grid = (X, Y, Z)
        self.n = X * Y * Z

    def close(self):
        if self.h:
            self.L.uwso_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # reference-style initialisation, mirroring FvmSolverHip's constructor
    def init_like_solver(self, wind_speed=None, face_positions=None):
        cf = self.cf
        if face_positions is not None:
            self.set_face_positions(*face_positions)
        ws = float(cf.velX0[0] if wind_speed is None else wind_speed)
        it = int(cf.inflowType)
        if it == 2:
            self.L.uwso_init_vel_pres_fields(self.h, 0, 0.0)
        else:
            self.L.uwso_init_vel_pres_fields(self.h, it, ws)
        if int(cf.solidType) == 1:
            self.L.uwso_init_box(self.h)
        vz = (ctypes.c_float * 3)(*[float(v) for v in cf.velZn])
        self.L.uwso_set_fixed_value_vel_bnd_cond(self.h, it, ws, ctypes.cast(vz, c_void_p),
                                                 float(cf.inflowRefHeight), float(cf.inflowAlpha))

    def set_face_positions(self, fx, fy, fz):
        """Stretched grid: face coordinates of each axis (X+1, Y+1, Z+1 values)."""
        X, Y, Z = self.grid
        a = [np.ascontiguousarray(np.asarray(v, dtype=np.float32).reshape(-1)) for v in (fx, fy, fz)]
        assert [v.size for v in a] == [X + 1, Y + 1, Z + 1]
        self._face_pos = a
        assert self.L.uwso_set_face_positions(self.h, *[v.ctypes.data_as(c_void_p) for v in a]) == 0

    def face_flux(self):
        """velFaceFluxField as [Z+1][Y+1][X+1][3]."""
        X, Y, Z = self.grid
        out = np.empty((Z + 1) * (Y + 1) * (X + 1) * 3, dtype=np.float64)
        assert self.L.uwso_get_face_flux(self.h, out.ctypes.data_as(c_void_p)) == 0
        return out.reshape(Z + 1, Y + 1, X + 1, 3)

    def set_face_flux(self, a):
        X, Y, Z = self.grid
        a = np.ascontiguousarray(np.asarray(a, dtype=np.float64).reshape(-1))
        assert a.size == (Z + 1) * (Y + 1) * (X + 1) * 3
        assert self.L.uwso_set_face_flux(self.h, a.ctypes.data_as(c_void_p)) == 0

    def init_flags(self, flags):
        a = np.ascontiguousarray(np.asarray(flags, dtype=np.int32).reshape(-1))
        self.L.uwso_init_flags(self.h, a.ctypes.data_as(c_void_p))

    def step(self, n=1):
        self.L.uwso_step(self.h, int(n))

    def run_stage(self, s):
        self.L.uwso_run_stage(self.h, int(s))

    def field(self, fid):
        nc = _NCOMP[fid]
        out = np.empty(self.n * nc, dtype=np.float64)
        assert self.L.uwso_get_field(self.h, fid, out.ctypes.data_as(c_void_p)) == 0
        return out.reshape(self.n, nc) if nc == 3 else out

    def set_field(self, fid, a):
        a = np.ascontiguousarray(np.asarray(a, dtype=np.float64).reshape(-1))
        assert a.size == self.n * _NCOMP[fid]
        assert self.L.uwso_set_field(self.h, fid, a.ctypes.data_as(c_void_p)) == 0

    def velocity(self):
        return self.field(N.FIELD_VEL)

    def pressure(self):
        return self.field(N.FIELD_PRES)

    def set_velocity(self, u):
        self.set_field(N.FIELD_VEL, u)

    def set_pressure(self, p):
        self.set_field(N.FIELD_PRES, p)

    def flags(self):
        out = np.empty(self.n, dtype=np.int32)
        self.L.uwso_get_flags(self.h, out.ctypes.data_as(c_void_p))
        return out

    def bnd_field(self, face):
        X, Y, Z = self.grid
        n = (Y if face < 2 else X) * (Z if face < 4 else Y)
        out = np.empty(n * 3, dtype=np.float64)
        assert self.L.uwso_get_bnd_field(self.h, face, out.ctypes.data_as(c_void_p)) == 0
        return out.reshape(n, 3)

    def set_bnd_field(self, face, xyz):
        a = np.ascontiguousarray(np.asarray(xyz, dtype=np.float64).reshape(-1))
        assert self.L.uwso_set_bnd_field(self.h, face, a.ctypes.data_as(c_void_p)) == 0

    def wall_func(self, uP):
        return self.L.uwso_log_law_wall_func(self.h, float(uP))

    def residuals(self):
        out = np.zeros(2)
        self.L.uwso_residuals(self.h, out.ctypes.data_as(c_void_p))
        return out

    def run_kernel(self, k: int):
        """One kernel alone (KERNEL_*), for the pinning tests."""
        assert self.L.uwso_run_kernel(self.h, int(k)) == 0

    def sweep_counts(self):
        """(momentum, SIMPLE pressure) sweeps run by the last step."""
        out = np.zeros(2, dtype=np.int32)
        self.L.uwso_sweep_counts(self.h, out.ctypes.data_as(c_void_p))
        return int(out[0]), int(out[1])

    def time(self):
        s, t = c_int64(), c_double()
        self.L.uwso_time(self.h, ctypes.byref(s), ctypes.byref(t))
        return s.value, t.value


def background_sample(fg_grid, fg_dx, fg_pos, rot_deg, bg_grid, bg_dx, bg_pos, bg_vel, bg_pres=None,
                      dtype=np.float32):
    """CPU restatement of CSInitVelPresFieldsFromBGFlowRotate
    (InitShaderColl.compute:184-226; angle formed as in FvmSolverGpu.cs:904,
    flowFieldOrientation * (float)(3.1415926 / 180)): for every foreground
    cell, rotate its offset from the grid centre (integer gridRes / 2) by the
    angle about y, translate to world coordinates, rotate back into the
    background frame and sample the background fields with trilinear
    filtering and clamped addressing (texel centres at (i + 1/2) / N, lerp
    along x, then y, then z).  Arithmetic in `dtype`.  bg_vel: [z][y][x][3|4],
    bg_pres: [z][y][x] or None.  Returns (u [N, 3], p [N])."""
    t = np.dtype(dtype).type
    X, Y, Z = fg_grid
    bX, bY, bZ = bg_grid
    ang = t(np.float32(rot_deg) * np.float32(3.1415926 / 180))
    zz, yy, xx = np.meshgrid(np.arange(Z), np.arange(Y), np.arange(X), indexing="ij")
    dxf = t(fg_dx)
    lx = (xx - X // 2).astype(t) * dxf
    ly = (yy - Y // 2).astype(t) * dxf
    lz = (zz - Z // 2).astype(t) * dxf
    s, c = t(np.sin(ang)), t(np.cos(ang))
    gx = t(fg_pos[0]) + (lx * c + lz * s)
    gy = t(fg_pos[1]) + ly
    gz = t(fg_pos[2]) + (lz * c - lx * s)
    s, c = t(np.sin(-ang)), t(np.cos(-ang))
    rx, ry, rz = gx - t(bg_pos[0]), gy - t(bg_pos[1]), gz - t(bg_pos[2])
    bxp, byp, bzp = rx * c + rz * s, ry, rz * c - rx * s
    half = t(0.5)
    u = bxp / (t(bX) * t(bg_dx)) + half
    v = byp / (t(bY) * t(bg_dx)) + half
    w = bzp / (t(bZ) * t(bg_dx)) + half

    def axis(q, n):
        tq = q * t(n) - half
        f0 = np.floor(tq)
        a = (tq - f0).astype(t)
        i0 = f0.astype(np.int64)
        return np.clip(i0, 0, n - 1), np.clip(i0 + 1, 0, n - 1), a

    x0, x1, ax = axis(u, bX)
    y0, y1, ay = axis(v, bY)
    z0, z1, az = axis(w, bZ)
    one = t(1)

    def samp(a3):
        a3 = np.asarray(a3, dtype=t).reshape(bZ, bY, bX)
        bx_, cx_, by_, cy_, bz_, cz_ = one - ax, ax, one - ay, ay, one - az, az
        c00 = a3[z0, y0, x0] * bx_ + a3[z0, y0, x1] * cx_
        c10 = a3[z0, y1, x0] * bx_ + a3[z0, y1, x1] * cx_
        c01 = a3[z1, y0, x0] * bx_ + a3[z1, y0, x1] * cx_
        c11 = a3[z1, y1, x0] * bx_ + a3[z1, y1, x1] * cx_
        c0 = c00 * by_ + c10 * cy_
        c1 = c01 * by_ + c11 * cy_
        return (c0 * bz_ + c1 * cz_).reshape(-1)

    bv = np.asarray(bg_vel).reshape(bZ * bY * bX, -1)
    uo = np.stack([samp(bv[:, k]) for k in range(3)], axis=1)
    po = samp(bg_pres) if bg_pres is not None else np.zeros(X * Y * Z, dtype=t)
    return uo, po
