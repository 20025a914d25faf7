"""Discrete adjoint of the step against complex-step and central finite
differences.

J is a seeded linear functional of the final state, J = <w_u, u_T> + <w_p, p_T>;
its gradient from one backward pass is checked along seeded directions of
every input the adjoint differentiates: the initial velocity and pressure,
the boundary textures of the fixed-value faces and the Smagorinsky constant.
The complex-step reference Im(J(x + i h v)) / h runs the same step in complex
FP64 on the GPU (precision UWS_PREC_CSTEP, h = 1e-20: no cancellation)."""
from __future__ import annotations

import copy

import numpy as np
import pytest

from unitywindsimulation_amd import _native as N
from unitywindsimulation_amd import workloads
from unitywindsimulation_amd.config import PresBndCond, VelBndCond
from unitywindsimulation_amd.adjoint import FvmAdjoint, fvm_unroll, layer_objective, power_law_bnd

from .helpers import seeded_pressure, solver_for

pytestmark = pytest.mark.gpu

H = 1e-20

VARIANTS = {
    # name: (workload factory, steps, checkpoint interval)
    "piso_les": (lambda: workloads.tiny(), 4, 2),
    "simple_les": (lambda: workloads.tiny(piso=False), 3, 3),
    "piso_laminar": (lambda: workloads.tiny(les=False), 3, 2),
    "no_momentum_sweeps": (lambda: workloads.tiny(vel_iters=0, pres_iters=3), 2, 1),
    # two x tiles, the second 6 columns wide; Y and Z divided by no tile height or z chunk
    "piso_les_ragged": (lambda: workloads.tiny(grid=(70, 9, 11)), 2, 1),
}


def _with_precision(w, prec):
    w2 = copy.deepcopy(w)
    w2.cf.precision = prec
    w2.cf.Init()
    return w2


def _solver(w, prec):
    w2 = _with_precision(w, prec)
    return solver_for(w2, p0=seeded_pressure(w2))


def _functional(n, seed=5):
    rng = np.random.default_rng(seed)
    return rng.standard_normal((n, 3)), rng.standard_normal(n)


def _directions(w, s, seed=9):
    """name -> (cstep target, face, direction array); fluid cells only for the state."""
    rng = np.random.default_rng(seed)
    fluid = (w.flags == 0)
    du = rng.standard_normal((w.ncells, 3)) * fluid[:, None]
    dp = rng.standard_normal(w.ncells) * fluid
    dirs = {"u0": (N.CSTEP_VEL, 0, du), "p0": (N.CSTEP_PRES, 0, dp)}
    for face in (N.X0, N.Y0):
        dirs[f"bnd{face}"] = (N.CSTEP_BND, face, rng.standard_normal((s.bnd_size(face), 3)))
    if int(w.cf.turbulenceModel) == 1:
        dirs["cs"] = (N.CSTEP_CS, 0, np.ones(1))
    return dirs


def _adjoint_dot(adj, name, target, face, d):
    if target == N.CSTEP_VEL:
        g = adj.grad_velocity0()
    elif target == N.CSTEP_PRES:
        g = adj.grad_pressure0()
    elif target == N.CSTEP_BND:
        g = adj.grad_bnd(face)
    else:
        g = np.array([adj.grad_cs()])
    return float(np.sum(g.reshape(d.shape) * d))


def _cstep_dot(w, nsteps, wu, wp, target, face, d):
    with _solver(w, N.PREC_CSTEP) as s:
        s.cstep_set_imag(target, H * d, face)
        s.step(nsteps)
        iu = s.cstep_get_imag(N.CSTEP_VEL)
        ip = s.cstep_get_imag(N.CSTEP_PRES)
    return float((np.sum(wu * iu) + np.sum(wp * ip)) / H)


@pytest.mark.parametrize("variant", list(VARIANTS))
def test_adjoint_matches_complex_step(gpu, variant):
    make, nsteps, K = VARIANTS[variant]
    w = make()
    print(f"[adjoint] {variant}: grid {w.grid} V={w.cf.velMaxNumIter} P={w.cf.presMaxNumIter} "
          f"steps={nsteps} K={K}", flush=True)
    wu, wp = _functional(w.ncells)
    results = {}
    for prec in (N.PREC_FP64, N.PREC_FP32):
        with _solver(w, prec) as s:
            dirs = _directions(w, s)
            with FvmAdjoint(s, nsteps, K) as adj:
                adj.forward()
                adj.backward(wu, wp)
                results[prec] = {k: _adjoint_dot(adj, k, *v) for k, v in dirs.items()}
    for name, (target, face, d) in dirs.items():
        ref = _cstep_dot(w, nsteps, wu, wp, target, face, d)
        e64 = abs(results[N.PREC_FP64][name] - ref) / max(abs(ref), 1e-300)
        e32 = abs(results[N.PREC_FP32][name] - ref) / max(abs(ref), 1e-300)
        print(f"  {name:6s} complex-step {ref:+.10e}  FP64 adjoint rel {e64:.2e}  FP32 adjoint rel {e32:.2e}",
              flush=True)
        assert e64 < 1e-9, (name, ref, results[N.PREC_FP64][name])
        assert e32 < 2e-3, (name, ref, results[N.PREC_FP32][name])


def test_adjoint_matches_central_differences(gpu):
    """FP64 central differences, h = 1e-3 on the scale of each direction."""
    w = workloads.tiny()
    nsteps, K = 3, 2
    wu, wp = _functional(w.ncells, seed=21)

    def J(s):
        return float(np.sum(wu * s.field64(N.FIELD_VEL)) + np.sum(wp * s.field64(N.FIELD_PRES)))

    bnd0 = power_law_bnd(w.grid, 5.0, 0.25, np.zeros(w.grid[1]))
    with _solver(w, N.PREC_FP64) as s:
        s.set_bnd_field64(N.X0, bnd0)
        with FvmAdjoint(s, nsteps, K) as adj:
            adj.forward()
            adj.backward(wu, wp)
            g_bnd = adj.grad_bnd(N.X0)
            g_cs = adj.grad_cs()
    rng = np.random.default_rng(3)
    h = 1e-3
    for trial in range(3):
        d = rng.standard_normal(bnd0.shape)
        d[:, 1:] = 0.0
        dcs = float(rng.standard_normal())
        vals = []
        for sgn in (1.0, -1.0):
            with _solver(w, N.PREC_FP64) as s:
                s.set_bnd_field64(N.X0, bnd0 + sgn * h * d)
                s.set_smagorinsky_constant(0.1 + sgn * h * dcs)
                s.step(nsteps)
                vals.append(J(s))
        fd = (vals[0] - vals[1]) / (2 * h)
        ad = float(np.sum(g_bnd * d)) + g_cs * dcs
        err = abs(fd - ad) / abs(fd)
        print(f"  direction {trial}: central difference {fd:+.10e} adjoint {ad:+.10e} rel {err:.2e}", flush=True)
        assert err < 1e-5


def test_torch_function_power_law_parameters(gpu):
    """dJ/d(U_ref, alpha, Cs, delta) through torch autograd over the C ABI,
    against complex step along a random parameter direction."""
    import torch

    w = workloads.tiny()
    X, Y, Z = w.grid
    nsteps, K = 3, 2
    region = (2, 2, 8, 6)
    rng = np.random.default_rng(8)
    delta0 = 0.05 * 5.0 * rng.standard_normal(Y)
    theta = torch.tensor([5.0, 0.25, 0.1], dtype=torch.float64, requires_grad=True)
    delta = torch.tensor(delta0, dtype=torch.float64, requires_grad=True)
    with _solver(w, N.PREC_FP64) as s:
        with FvmAdjoint(s, nsteps, K) as adj:
            bnd = power_law_bnd(w.grid, theta[0], theta[1], delta, xp=torch)
            u, p = fvm_unroll(adj, bnd, theta[2])
            uu = u.reshape(Z, Y, X, 3)
            x0, z0, wd, dp = region
            ub = 0.5 * (uu[z0:z0 + dp, 1, x0:x0 + wd, :] + uu[z0:z0 + dp, 2, x0:x0 + wd, :])
            J = (ub * ub).sum() / (wd * dp)
            J.backward()
            J_np, _ = layer_objective(u.detach().numpy(), w.grid, region)
            assert abs(J.item() - J_np) < 1e-12 * max(1.0, abs(J_np))
    # complex step along a random direction of (U_ref, alpha, Cs, delta)
    dth = rng.standard_normal(3)
    ddl = rng.standard_normal(Y)
    ad = float(theta.grad.numpy() @ dth + delta.grad.numpy() @ ddl)
    yc = (np.arange(Y) + 0.5) / 16.0
    dcol = dth[0] * yc ** 0.25 + 5.0 * yc ** 0.25 * np.log(yc) * dth[1] + ddl
    dbnd = np.zeros((Y * Z, 3))
    dbnd[:, 0] = np.tile(dcol, Z)
    with _solver(w, N.PREC_CSTEP) as s:
        s.set_bnd_field64(N.X0, power_law_bnd(w.grid, 5.0, 0.25, delta0))
        s.set_smagorinsky_constant(0.1)
        s.cstep_set_imag(N.CSTEP_BND, H * dbnd, N.X0)
        s.cstep_set_imag(N.CSTEP_CS, np.array([H * dth[2]]))
        s.step(nsteps)
        re = s.field64(N.FIELD_VEL)
        im = s.cstep_get_imag(N.CSTEP_VEL)
    x0, z0, wd, dp = region
    r4 = re.reshape(Z, Y, X, 3)
    i4 = im.reshape(Z, Y, X, 3)
    ubr = 0.5 * (r4[z0:z0 + dp, 1, x0:x0 + wd] + r4[z0:z0 + dp, 2, x0:x0 + wd])
    ubi = 0.5 * (i4[z0:z0 + dp, 1, x0:x0 + wd] + i4[z0:z0 + dp, 2, x0:x0 + wd])
    cs_ref = float(np.sum(2.0 * ubr * ubi) / (wd * dp) / H)
    err = abs(ad - cs_ref) / abs(cs_ref)
    print(f"  torch adjoint {ad:+.10e} complex-step {cs_ref:+.10e} rel {err:.2e}", flush=True)
    assert err < 1e-9


def _variant(w, slots=None, model=None, **bnd):
    w2 = copy.deepcopy(w)
    if slots is not None:
        w2.cf.wallSlots = slots
        w2.cf.wallModel = model
    for k, v in bnd.items():
        setattr(w2.cf, k, v)
    w2.cf.Init()
    return w2


REVERSE_VARIANTS = {
    "piso_les_slots3_rough": lambda: workloads.tiny(),
    "piso_les_ragged": lambda: workloads.tiny(grid=(70, 9, 11)),
    "simple_les_slots1_smooth": lambda: _variant(workloads.tiny(piso=False), 1, 0),
    "piso_les_slots1_rough": lambda: _variant(workloads.tiny(), 1, 1),
    "piso_laminar_slots3_smooth": lambda: _variant(workloads.tiny(les=False), 3, 0),
    # fixed-value and fixed-pressure Z0, zero-gradient Yn: every boundary branch
    "simple_laminar_boundaries": lambda: _variant(
        workloads.tiny(les=False, piso=False), velBndCondZ0=VelBndCond.FixedValue,
        presBndCondZ0=PresBndCond.FixedValue, velBndCondYn=VelBndCond.ZeroGrad),
    "piso_les_boundaries": lambda: _variant(
        workloads.tiny(), velBndCondZn=VelBndCond.FixedValue, presBndCondZn=PresBndCond.FixedValue,
        velBndCondYn=VelBndCond.ZeroGrad, velZn=(0.5, -0.1, -0.2)),
}


@pytest.mark.parametrize("variant", list(REVERSE_VARIANTS))
def test_adjoint_hand_reverse_matches_colouring(gpu, variant, monkeypatch):
    """The hand-written reverses of the one-shot kernels (momentum
    coefficients with wall function, eddy viscosity and LES term; Rhie-Chow
    coefficients; SIMPLE and PISO corrections; PISO coefficients and aodu)
    against the colouring path, which differentiates the same forward bodies
    with forward-mode Dual numbers: every gradient the adjoint returns agrees
    to FP64 rounding, on FP64 and FP32 records."""
    w = REVERSE_VARIANTS[variant]()
    nsteps, K = 3, 2
    wu, wp = _functional(w.ncells, seed=13)
    for prec in (N.PREC_FP64, N.PREC_FP32):
        out = {}
        for mode in ("hand", "colour"):
            if mode == "colour":
                monkeypatch.setenv("UWS_ADJOINT_REVERSE", "colour")
            else:
                monkeypatch.delenv("UWS_ADJOINT_REVERSE", raising=False)
            with _solver(w, prec) as s:
                with FvmAdjoint(s, nsteps, K) as adj:
                    adj.forward()
                    adj.backward(wu, wp)
                    g = {"u0": adj.grad_velocity0(), "p0": adj.grad_pressure0()}
                    for face in range(6):
                        g[f"bnd{face}"] = adj.grad_bnd(face)
                    g["cs"] = np.array([adj.grad_cs()])
            out[mode] = g
        for k in out["colour"]:
            a, b = out["hand"][k], out["colour"][k]
            scale = max(float(np.max(np.abs(b))), 1e-300)
            err = float(np.max(np.abs(a - b))) / scale
            print(f"  {variant} prec {prec} {k:6s} max|colour| {scale:.3e}  "
                  f"max|hand - colour| / max|colour| {err:.2e}", flush=True)
            assert err < 1e-11, (prec, k, err)


def test_adjoint_records_independent_of_sweep_fusion(gpu, monkeypatch):
    """The recording step runs its pressure solves with the three-sweep
    kernel, which also stores its first two sweeps' iterates; every record
    must equal what single sweeps give, so the FP32 gradients with the
    multi-sweep kernels and with single sweeps only (UWS_SWEEP_FUSION=0, in
    the forward and the recording alike) are identical."""
    w = workloads.tiny(grid=(20, 14, 17), pres_iters=8)
    nsteps, K = 3, 2
    wu, wp = _functional(w.ncells, seed=17)
    out = []
    for env in ({}, {"UWS_SWEEP_FUSION": "0"}, {"UWS_SWEEP_FUSION": "2", "UWS_ZCHUNK3": "5"}):
        monkeypatch.delenv("UWS_SWEEP_FUSION", raising=False)
        monkeypatch.delenv("UWS_ZCHUNK3", raising=False)
        for k, v in env.items():
            monkeypatch.setenv(k, v)
        with _solver(w, N.PREC_FP32) as s:
            with FvmAdjoint(s, nsteps, K) as adj:
                adj.forward()
                adj.backward(wu, wp)
                out.append((adj.grad_velocity0(), adj.grad_pressure0(), adj.grad_bnd(N.X0), adj.grad_cs()))
    for r in out[1:]:
        for a, b in zip(r, out[0]):
            assert np.array_equal(a, b)
