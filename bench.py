"""Benchmark: PISO solver steps per second on one MI355X, reported as MLUPS.

One "step" is one FvmSolverGpu.Step() (LES, momentum predictor with V
Jacobi sweeps, Rhie-Chow pressure correction with P sweeps, C PISO correctors)
over the whole grid, inputs resident in HBM.  Default workload: the urban
case, 40 seeded buildings on a 256 (x) x 96 (height) x 128 (span) grid, V = P = 50,
C = 2.  Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "MLUPS per PISO step + %HBM peak; batched vs plain cases/s at ≤1e-3 rel L2"
HBM_PEAK_GBS = 8000.0
# HBM bytes per cell per launch read by rocprofv3 FETCH_SIZE (x2, calibrated on a
# 256 MiB copy) + WRITE_SIZE on the urban case: profiles/urban_256x96x128_hbm_traffic.md
MEASURED_TRAFFIC_B_PER_CELL = {"urban": {"k_vel_sweep": 75.9, "k_pres_sweep": 29.7, "k_pres_sweep2": 32.2,
                                         "k_pres_sweep3": 35.2, "k_pres_sweep4": 36.2, "k_vel_sweep2": 84.6}}


def bytes_per_cell_step(V: int, P: int, C: int, les: bool = True) -> int:
    """Algorithmic HBM bytes per cell per step: vectors 12 B, scalars 4 B,
    flag 1 B, each array counted once per read and per write per kernel.
    Without LES the wall function (25 B), eddy viscosity (17 B) and LES term
    (41 B) do not run."""
    return 263 - (0 if les else 83) + 77 * V + 29 * P + C * (123 + 33 * P)


def compulsory_bytes_per_cell_step(V: int, P: int, C: int, les: bool = True, pres_depth: int = 3,
                                   vel_depth: int = 2, normalise: bool = False) -> int:
    """Bytes per cell one step of the solver's launch plan must move, each array once per
    launch: wall function + eddy viscosity with the gradient store (78), the
    momentum precompute with the LES term and the gradient read (146), V / 2
    momentum pairs (73 each, a leftover single sweep 77; vel_depth 1: V single
    sweeps), the Rhie-Chow precompute (37), per pressure solve P / pres_depth
    launches of pres_depth sweeps (28 each; a leftover triple or pair 28, a
    single 29), the velocity correction with p += p' in the same launch (53 +
    8; with the pressure normalisation they are separate launches, 53 + 13)
    and per PISO corrector A_od u' (37), the corrector's right-hand side (4)
    and the correction (78).  3,813 B/cell at V = P = 50, C = 2 with triples,
    3,477 with four-sweep launches."""
    d = max(1, pres_depth)
    solve = (P // d) * 28 + (0 if P % d == 0 else (29 if P % d == 1 else 28)) if d > 1 else 29 * P
    mom = (V // 2) * 73 + (V % 2) * 77 if vel_depth == 2 else 77 * V
    one_shot = (78 + 146) if les else 69
    correct = 53 + 13 if normalise else 53 + 8
    return one_shot + mom + 37 + solve + correct + C * (37 + 4 + solve + 78)


# HBM bytes per cell each kernel moved per launch on the urban case
# (256 x 96 x 128), read from the FETCH_SIZE (scaled by 2.000) and WRITE_SIZE
# counters (profiles/urban_256x96x128_hbm_traffic.md); "_zin": the first
# launch of a pressure solve, which starts from p' = 0; k_vel_correct
# includes the p += p' it applies in the same launch
MEASURED_LAUNCH_B_PER_CELL = {"urban": {
    "k_wall_eddy": 66.9, "k_vel_precompute": 165.2, "k_pres_precompute": 59.3,
    "k_vel_correct": 61.5, "k_piso_aodu": 38.9, "k_piso_precompute": 46.2, "k_piso_correct": 73.5,
    "k_vel_sweep": 75.9, "k_vel_sweep2": 84.6, "k_pres_sweep": 29.7, "k_pres_sweep2": 32.2,
    "k_pres_sweep3": 35.2, "k_pres_sweep4": 36.2, "k_pres_sweep4_zin": 30.5}}


def measured_bytes_per_cell_step(workload: str, V: int, P: int, C: int, les: bool = True, pres_depth: int = 4,
                                 vel_depth: int = 2):
    """Counter-measured HBM bytes per cell of one step: the per-launch
    figures above summed over the launch plan (the launches
    compulsory_bytes_per_cell_step counts), or None where no profile of the
    workload holds them."""
    m = MEASURED_LAUNCH_B_PER_CELL.get(workload)
    if m is None or not les:
        return None
    names = {1: "k_pres_sweep", 2: "k_pres_sweep2", 3: "k_pres_sweep3", 4: "k_pres_sweep4"}
    d = max(1, pres_depth)
    n_main, rem = (P // d, P % d) if d > 1 else (P, 0)
    solve = n_main * m[names[d]] + (0.0 if rem == 0 else m[names[rem]])
    if n_main > 0 and names[d] + "_zin" in m:
        solve += m[names[d] + "_zin"] - m[names[d]]
    mom = (V // 2) * m["k_vel_sweep2"] + (V % 2) * m["k_vel_sweep"] if vel_depth == 2 else V * m["k_vel_sweep"]
    total = (m["k_wall_eddy"] + m["k_vel_precompute"] + mom + m["k_pres_precompute"] + solve + m["k_vel_correct"]
             + C * (m["k_piso_aodu"] + m["k_piso_precompute"] + solve + m["k_piso_correct"]))
    return round(total, 1)


def stream_ceiling(cells: int, nin: int, reps: int = 20) -> float:
    """GB/s of the library's plain streaming kernel over nin + 1 float arrays
    of `cells` values (uws_stream_bandwidth), launched back to back on the
    same arrays as the sweeps are."""
    import ctypes

    from unitywindsimulation_amd import _native as N

    g = ctypes.c_double()
    N.check(N.lib().uws_stream_bandwidth(int(cells), int(nin), int(reps), ctypes.byref(g)), "uws_stream_bandwidth")
    return g.value


def cpu_baseline(budget_s: float = 15.0) -> dict:
    """The CPU oracle (a restatement of the reference kernels, its cell loops
    parallel over the host cores with OpenMP like the reference's Parallel.For
    transcription would be) on the isolated-building case, whole steps until
    ~budget_s of wall time."""
    from oracle.oracle import Oracle, lib as oracle_lib
    from unitywindsimulation_amd import workloads
    from unitywindsimulation_amd.geometry import noise_velocity

    w = workloads.isolated_building()
    o = Oracle(w.cf)
    o.init_like_solver()
    o.init_flags(w.flags)
    o.set_velocity(noise_velocity(w.flags, w.noise_seed, w.noise_amp))
    t0 = time.perf_counter()
    n = 0
    while True:
        o.step(1)
        n += 1
        if time.perf_counter() - t0 >= budget_s:
            break
    dt = time.perf_counter() - t0
    X, Y, Z = w.grid
    cf = w.cf
    C = cf.PISONumCorrectors if int(cf.fvmSolverType) == 1 else 0
    b_step = bytes_per_cell_step(cf.velMaxNumIter, cf.presMaxNumIter, C, int(cf.turbulenceModel) == 1)
    threads = int(oracle_lib().uwso_num_threads())
    cpu_model = "unknown"
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    cpu_model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    import shutil

    dotnet = {t: shutil.which(t) is not None for t in ("dotnet", "mono", "csc")}
    return {"value": round(X * Y * Z * n / dt / 1e6, 4), "unit": "MLUPS", "cores": threads, "kind": "port",
            "sample": f"{n} PISO steps of {w.name} (V=P=50, C=2), FP32 oracle on {threads} OpenMP threads "
                      f"({os.cpu_count()} logical CPUs visible)",
            "stand_in_for": "the C# Parallel.For transcription of the reference kernels, which cannot be built "
                            "without a .NET toolchain (none is installed): the C++ OpenMP oracle, the same "
                            "kernels, order and FP32 arithmetic, runs instead",
            "dotnet_toolchain_on_host": dotnet,
            "ms_per_step": round(dt / n * 1e3, 2), "algorithmic_GBps": round(b_step * X * Y * Z * n / dt / 1e9, 2),
            "cpu_model": cpu_model}


def ensemble_report(steps: int = 50, check_steps: int = 500, launch: str = "auto") -> dict:
    """The 24-case uneven ensemble: cases/s of the batched ensemble (one graph
    per step over all cases) against the same cases stepped one after another,
    each with its own graph per step ("plain") and with one host launch per
    kernel ("eager_plain", how the reference dispatches), and the FP32
    ensemble's velocity error against FP64 runs after check_steps steps (0:
    not checked).  launch: "auto" (the autotuner picks the launch shape),
    "packed" (groups of at most one round of two-sweep workgroups) or "one"
    (one group)."""
    import numpy as np

    from unitywindsimulation_amd import FvmSolverHip, workloads
    from unitywindsimulation_amd import _native as N
    from unitywindsimulation_amd.solver import FvmEnsembleHip

    ws = workloads.ensemble()
    full_steps = ws[0].steps

    def sequential(eager: bool, single_sweeps: bool = False, nsteps: int = steps) -> float:
        """seconds per step of all cases stepped one after another; eager: one host launch per
        kernel, else one captured graph per step and case; single_sweeps: one launch per Jacobi
        sweep too (with eager, the reference's own dispatch pattern)"""
        saved = {k: os.environ.get(k) for k in ("UWS_NO_GRAPH", "UWS_SWEEP_FUSION")}
        if eager:
            os.environ["UWS_NO_GRAPH"] = "1"
        if single_sweeps:
            os.environ["UWS_SWEEP_FUSION"] = "0"
        try:
            total = 0.0
            for w in ws:
                with FvmSolverHip(w.cf) as s:
                    s.InitFlags(w.flags)
                    s.init_noise(w.noise_seed, w.noise_amp)
                    s.step(2)
                    s.synchronize()
                    t0 = time.perf_counter()
                    s.step(nsteps)
                    s.synchronize()
                    total += (time.perf_counter() - t0) / nsteps
            return total
        finally:
            for k, v in saved.items():
                if v is None:
                    os.environ.pop(k, None)
                else:
                    os.environ[k] = v

    t_plain = sequential(eager=False)
    t_eager = sequential(eager=True)
    t_ref = sequential(eager=True, single_sweeps=True, nsteps=max(1, min(steps, 10)))
    with FvmEnsembleHip([w.cf for w in ws]) as e:
        for c, w in enumerate(ws):
            e.init_flags(c, w.flags)
            e.init_noise(c, w.noise_seed, w.noise_amp)
        if launch == "auto":
            ty, pers, gbytes, _ = e.autotune(2)
        else:
            ty, pers, gbytes = 4, False, (-1 if launch == "packed" else 0)
            e.set_launch(ty, pers, gbytes)
        ngroups = e.num_groups()
        e.step(1)
        e.synchronize()
        t0 = time.perf_counter()
        e.step(steps)
        e.synchronize()
        t_batch = (time.perf_counter() - t0) / steps
    # accuracy: FP32 ensemble vs FP64 single runs after check_steps steps
    errs = []
    if check_steps > 0:
        with FvmEnsembleHip([w.cf for w in ws]) as e:
            for c, w in enumerate(ws):
                e.init_flags(c, w.flags)
                e.init_noise(c, w.noise_seed, w.noise_amp)
            e.set_launch(ty, pers, gbytes)
            e.step(check_steps)
            for c, w in enumerate(ws):
                w.cf.precision = N.PREC_FP64
                with FvmSolverHip(w.cf) as d:
                    d.InitFlags(w.flags)
                    d.init_noise(w.noise_seed, w.noise_amp)
                    d.step(check_steps)
                    ref = d.field64(N.FIELD_VEL)
                w.cf.precision = N.PREC_FP32
                got = e.velocity(c, 3).reshape(-1, 3).astype(np.float64)
                errs.append(float(np.linalg.norm(got - ref) / np.linalg.norm(ref)))
    cells = sum(w.ncells for w in ws)
    # bound: every step must move each case's compulsory bytes; no launch shape moves them faster than
    # the plain streaming kernel does on any working set these cases form (one case's arrays, which the
    # last-level cache holds, or all cases' arrays at once)
    cf0 = ws[0].cf
    C0 = cf0.PISONumCorrectors if int(cf0.fvmSolverType) == 1 else 0
    b_comp = compulsory_bytes_per_cell_step(cf0.velMaxNumIter, cf0.presMaxNumIter, C0,
                                            int(cf0.turbulenceModel) == 1)
    sizes = sorted({w.ncells for w in ws})
    ceilings = {f"{n}_cells_nin{k}": stream_ceiling(n, k) for n in sizes for k in (6, 16)}
    ceilings[f"{cells}_cells_nin16"] = stream_ceiling(cells, 16, 5)
    best = max(ceilings.values())
    t_min = b_comp * cells / (best * 1e9)
    bound = {"compulsory_B_per_cell_step": b_comp, "compulsory_GB_per_step": round(b_comp * cells / 1e9, 3),
             "stream_ceilings_GBps": {k: round(v, 1) for k, v in ceilings.items()},
             "min_ms_per_step_at_best_ceiling": round(t_min * 1e3, 3),
             "max_speedup_vs_plain": round(t_plain / t_min, 3),
             "batched_frac_of_bound": round(t_min / t_batch, 4)}
    return {"cases": len(ws), "cells": cells, "steps_per_case": full_steps, "bound": bound,
            "plain_cases_per_s": round(len(ws) / (t_plain * full_steps), 4),
            "batched_cases_per_s": round(len(ws) / (t_batch * full_steps), 4),
            "speedup": round(t_plain / t_batch, 3),
            "eager_plain_cases_per_s": round(len(ws) / (t_eager * full_steps), 4),
            "speedup_vs_eager_plain": round(t_eager / t_batch, 3),
            "plain_ms_per_step_all_cases": round(t_plain * 1e3, 3), "batched_ms_per_step": round(t_batch * 1e3, 3),
            "eager_plain_ms_per_step_all_cases": round(t_eager * 1e3, 3),
            # the same cases dispatched as FvmSolverGpu.Step does: one launch per kernel and per
            # Jacobi sweep, no graph (a reference point; "speedup" above is against the plain
            # solver as this library runs it)
            "reference_dispatch_plain_ms_per_step_all_cases": round(t_ref * 1e3, 3),
            "speedup_vs_reference_dispatch_plain": round(t_ref / t_batch, 3),
            "launch": {"tile_rows": ty, "persistent": pers, "groups": ngroups,
                       "grouping": ("packed: at most one round of two-sweep workgroups per group" if gbytes < 0
                                    else "one group" if gbytes == 0 else f"consecutive cases within {gbytes >> 20} MiB"),
                       "chosen_by": "autotuner" if launch == "auto" else "fixed"},
            "max_rel_l2_u_vs_fp64": max(errs) if errs else None, "compared_after_steps": check_steps}


def precision_report(steps: int = 500, mean_last: int = 250, only=None) -> dict:
    """The precision case (384 x 128 high x 256 span, 120 buildings): the
    same run in FP64, FP32 and FP32 with FP16 / BF16 correction storage
    (3 rounds x 20 sweeps of defect correction per pressure solve; the plain
    runs do the 60 sweeps directly), FP64 norms every step.  Errors of u, p
    after `steps` steps and of the time mean of u over the last `mean_last`
    steps, relative L2 against the FP64 run."""
    import copy

    import numpy as np

    from unitywindsimulation_amd import FvmSolverHip, workloads
    from unitywindsimulation_amd import _native as N

    w = workloads.precision(steps=steps, mean_last=mean_last)
    X, Y, Z = w.grid
    modes = [("fp64", N.PREC_FP64), ("fp32", N.PREC_FP32), ("fp32_fp16", N.PREC_FP32_FP16),
             ("fp32_bf16", N.PREC_FP32_BF16)]
    if only:
        modes = [m for m in modes if m[0] in only]
    out = {"workload": w.name, "grid_xyz": [X, Y, Z], "cells": w.ncells, "buildings": len(w.boxes),
           "steps": steps, "mean_over_last": mean_last, "refine": [w.cf.refineRounds, w.cf.refineSweeps]}
    ref = None

    def rel(a, b):
        return float(np.linalg.norm(a - b) / np.linalg.norm(b))

    for name, prec in modes:
        cf = copy.deepcopy(w.cf)
        cf.precision = prec
        cf.computeNorms = True
        cf.Init()
        with FvmSolverHip(cf) as s:
            s.InitFlags(w.flags)
            s.init_noise(w.noise_seed, w.noise_amp)
            s.synchronize()
            t0 = time.perf_counter()
            s.step(steps - mean_last)
            s.mean_reset()
            for _ in range(mean_last):
                s.step(1)
                s.mean_accumulate()
            s.synchronize()
            t = time.perf_counter() - t0
            u = s.field64(N.FIELD_VEL)
            p = s.field64(N.FIELD_PRES)
            um, cnt = s.mean_velocity()
            nrm = s.norms()
        rec = {"ms_per_step": round(t / steps * 1e3, 3), "finite": bool(np.isfinite(u).all() and np.isfinite(p).all()),
               "norms_last_step": {"vel_residual": float(nrm[N.NORM_VEL_RESIDUAL]),
                                   "pres_residual": float(nrm[N.NORM_PRES_RESIDUAL]),
                                   "divergence": float(nrm[N.NORM_DIVERGENCE]),
                                   "nonfinite_cells": float(nrm[N.NORM_NONFINITE])}}
        if prec in (N.PREC_FP32_FP16, N.PREC_FP32_BF16):
            # the low-precision modes have no plain SIMPLE sweep: the defect of every round stands for it
            del rec["norms_last_step"]["pres_residual"]
            rec["norms_last_step"]["refine_defect_simple_rounds"] = [float(nrm[N.NORM_REFINE0 + r])
                                                                     for r in range(w.cf.refineRounds)]
        if prec == N.PREC_FP64:
            ref = (u, p, um, rec["norms_last_step"])
        elif ref is not None:
            rec["rel_l2_vs_fp64"] = {"u": rel(u, ref[0]), "p": rel(p, ref[1]), "u_time_mean": rel(um, ref[2])}
            # the momentum residual of the last sweep is rounding-limited (1e-15 in FP64), so only the
            # divergence, and for the plain modes the pressure residual, are compared
            keys = ("divergence",) if prec in (N.PREC_FP32_FP16, N.PREC_FP32_BF16) else ("divergence",
                                                                                           "pres_residual")
            rec["norm_rel_err_vs_fp64"] = {k: abs(rec["norms_last_step"][k] - ref[3][k]) / abs(ref[3][k])
                                           for k in keys if ref[3][k] != 0}
        out[name] = rec
        print(f"[precision] {name}: {rec['ms_per_step']} ms/step {rec.get('rel_l2_vs_fp64', '')}", file=sys.stderr,
              flush=True)
    return out


def adjoint_report(ndir: int = 16, h_fd=(1e-3, 1e-6)) -> dict:
    """The differentiation case: 64 steps with a checkpoint every 8, J = mean
    |u|^2 on the 2 m layer (average of cell layers 1 and 2) of a seeded 32 x 32
    region; gradient w.r.t. U_ref, alpha, Cs and the 64 per-level inlet
    perturbations through the torch autograd function on FP32 and FP64
    handles.  Checked against FP64 complex step (Im J(theta + i h v) / h,
    h = 1e-20) along the three unit directions of U_ref, alpha, Cs and ndir
    random unit directions, and against FP64 central differences along the
    same ndir random directions for each step h in h_fd (the truncation error
    of the 64-step objective is O(h^2) with a large constant: h = 1e-3 is
    reported as specified, the smaller h shows the convergence)."""
    import copy

    import numpy as np
    import torch

    from unitywindsimulation_amd import FvmSolverHip, workloads
    from unitywindsimulation_amd import _native as N
    from unitywindsimulation_amd.adjoint import FvmAdjoint, fvm_unroll, layer_objective, power_law_bnd

    d = workloads.differentiation()
    w = d.base
    grid = w.grid
    X, Y, Z = grid
    theta0 = np.concatenate([[d.u_ref, d.alpha, d.cs], d.delta])
    x0, z0, rw, rd = d.region
    yc = (np.arange(Y) + 0.5) * float(w.cf.dx) / float(w.cf.inflowRefHeight)

    def solver(prec, theta=theta0):
        cf = copy.deepcopy(w.cf)
        cf.precision = prec
        cf.Init()
        s = FvmSolverHip(cf)
        s.InitFlags(w.flags)
        s.init_noise(w.noise_seed, w.noise_amp)
        s.set_bnd_field64(N.X0, power_law_bnd(grid, theta[0], theta[1], theta[3:]))
        s.set_smagorinsky_constant(theta[2])
        return s

    out = {"workload": w.name, "grid_xyz": [X, Y, Z], "cells": w.ncells, "buildings": len(w.boxes),
           "steps": d.nsteps, "checkpoint_every": d.ckpt_every, "region_x0_z0_w_d": list(d.region),
           "objective": "mean |u|^2 over the region at y = 2 m (cell layers 1, 2 averaged)",
           "parameters": "U_ref, alpha, Cs, 64 per-level inlet perturbations"}
    grads = {}
    for prec, name in ((N.PREC_FP32, "fp32"), (N.PREC_FP64, "fp64")):
        with solver(prec) as s, FvmAdjoint(s, d.nsteps, d.ckpt_every) as adj:
            th = torch.tensor(theta0, dtype=torch.float64, requires_grad=True)
            bnd = power_law_bnd(grid, th[0], th[1], th[3:], xp=torch)
            t0 = time.perf_counter()
            u, _ = fvm_unroll(adj, bnd, th[2])
            t1 = time.perf_counter()
            uu = u.reshape(Z, Y, X, 3)
            ub = 0.5 * (uu[z0:z0 + rd, 1, x0:x0 + rw] + uu[z0:z0 + rd, 2, x0:x0 + rw])
            J = (ub * ub).sum() / (rw * rd)
            J.backward()
            t2 = time.perf_counter()
            grads[name] = th.grad.numpy().copy()
            out[name] = {"J": J.item(), "forward_s": round(t1 - t0, 3), "backward_s": round(t2 - t1, 3),
                         "device_GiB": round(adj.memory_bytes / 2 ** 30, 3),
                         "dJ_dUref": float(grads[name][0]), "dJ_dalpha": float(grads[name][1]),
                         "dJ_dCs": float(grads[name][2])}
        print(f"[adjoint] {name}: J={out[name]['J']:.6e} forward {out[name]['forward_s']} s "
              f"backward {out[name]['backward_s']} s", file=sys.stderr, flush=True)

    def dbnd(v):
        col = v[0] * yc ** theta0[1] + theta0[0] * yc ** theta0[1] * np.log(yc) * v[1] + v[3:]
        b = np.zeros((Z * Y, 3))
        b[:, 0] = np.tile(col, Z)
        return b

    H = 1e-20

    def cstep(v):
        with solver(N.PREC_CSTEP) as s:
            s.cstep_set_imag(N.CSTEP_BND, H * dbnd(v), N.X0)
            s.cstep_set_imag(N.CSTEP_CS, np.array([H * v[2]]))
            s.step(d.nsteps)
            re = s.field64(N.FIELD_VEL).reshape(Z, Y, X, 3)
            im = s.cstep_get_imag(N.CSTEP_VEL).reshape(Z, Y, X, 3)
        ubr = 0.5 * (re[z0:z0 + rd, 1, x0:x0 + rw] + re[z0:z0 + rd, 2, x0:x0 + rw])
        ubi = 0.5 * (im[z0:z0 + rd, 1, x0:x0 + rw] + im[z0:z0 + rd, 2, x0:x0 + rw])
        return float(np.sum(2.0 * ubr * ubi) / (rw * rd) / H)

    def J64(theta):
        with solver(N.PREC_FP64, theta) as s:
            s.step(d.nsteps)
            return layer_objective(s.field64(N.FIELD_VEL), grid, d.region)[0]

    rng = np.random.default_rng(4)
    rand = [v / np.linalg.norm(v) for v in rng.standard_normal((ndir, theta0.size))]
    dirs = [np.eye(theta0.size)[i] for i in range(3)] + rand
    t0 = time.perf_counter()
    ref = np.array([cstep(v) for v in dirs])
    t_cs = time.perf_counter() - t0
    scale = float(np.sqrt(np.mean(ref ** 2)))
    for name in ("fp64", "fp32"):
        ad = np.array([grads[name] @ v for v in dirs])
        out[name]["vs_complex_step_max_rel"] = float(np.max(np.abs(ad - ref) / np.abs(ref)))
        out[name]["vs_complex_step_max_abs_over_rms"] = float(np.max(np.abs(ad - ref)) / scale)
    out["complex_step"] = {"directions": len(dirs), "seconds": round(t_cs, 2),
                           "unit_UrefAlphaCs": [float(x) for x in ref[:3]]}
    out["central_fd"] = []
    for h in h_fd:
        t0 = time.perf_counter()
        fd = np.array([(J64(theta0 + h * v) - J64(theta0 - h * v)) / (2 * h) for v in rand])
        t_fd = time.perf_counter() - t0
        rec = {"h": h, "directions": len(rand), "seconds": round(t_fd, 2),
               "fd_vs_complex_step_max_rel": float(np.max(np.abs(fd - ref[3:]) / np.abs(ref[3:])))}
        for name in ("fp64", "fp32"):
            ad = np.array([grads[name] @ v for v in rand])
            rec[f"{name}_adjoint_vs_fd_max_rel"] = float(np.max(np.abs(ad - fd) / np.abs(fd)))
        out["central_fd"].append(rec)
    return out


def stretched_report(steps: int = 5, warmup: int = 2) -> dict:
    """The reference's stretched-grid case: WindTunnelBoxConfig (100 x 50 x 100,
    laminar SIMPLE, 800 momentum and 2000 pressure sweeps per step) on
    InitMeshBuffers' mesh (25 cells growing by 1.08 on both sides of a
    50-cell uniform block along x and z)."""
    import torch

    from unitywindsimulation_amd import FvmSolverHip, workloads
    from unitywindsimulation_amd import _native as N

    w = workloads.wind_tunnel_box()
    cf = w.cf
    X, Y, Z = w.grid
    n = X * Y * Z
    V, P = cf.velMaxNumIter, cf.presMaxNumIter
    s = FvmSolverHip(cf, face_positions=w.face_positions)
    s.InitFlags(w.flags)
    s.step(warmup)
    s.synchronize()
    t0 = time.perf_counter()
    s.step(steps)
    s.synchronize()
    ms = (time.perf_counter() - t0) / steps * 1e3
    b = bytes_per_cell_step(V, P, 0, les=False)
    t_vel = s.time_kernel(N.KERNEL_VEL_SWEEP, 20)
    t_vel2 = s.time_kernel(N.KERNEL_VEL_SWEEP2, 20)
    t_pres = s.time_kernel(N.KERNEL_PRES_SWEEP, 20)
    t_pres3 = s.time_kernel(N.KERNEL_PRES_SWEEP3, 20)
    u = s.velocity(3)
    s.close()
    return {"workload": w.name, "grid_xyz": [X, Y, Z], "cells": n, "vel_sweeps": V, "pres_sweeps": P,
            "steps": steps, "ms_per_step": round(ms, 3), "mlups": round(n / (ms * 1e-3) / 1e6, 2),
            "bytes_per_cell_step": b, "hbm_frac": round(b * n / (ms * 1e-3) / (HBM_PEAK_GBS * 1e9), 4),
            "kernel_ms": {"k_vel_sweep (stretched)": round(t_vel, 5), "k_vel_sweep2 (stretched)": round(t_vel2, 5),
                          "k_pres_sweep": round(t_pres, 5), "k_pres_sweep3": round(t_pres3, 5)},
            "domain_m": [float(v[-1]) for v in w.face_positions],
            "finite": bool(torch.isfinite(torch.from_numpy(u)).all()),
            "max_abs_u": float(abs(u).max())}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--workload", default="urban")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    ap.add_argument("--no-ensemble", action="store_true", help="skip the 24-case ensemble leg")
    ap.add_argument("--ensemble-steps", type=int, default=50, help="timed steps of the ensemble legs")
    ap.add_argument("--ensemble-check-steps", type=int, default=500,
                    help="steps after which the ensemble is compared with FP64 runs (0: no comparison)")
    ap.add_argument("--ensemble-launch", default="auto", choices=("auto", "packed", "one"),
                    help="launch grouping of the batched ensemble (auto: the autotuner's choice)")
    ap.add_argument("--no-adjoint", action="store_true",
                    help="skip the differentiation leg (adjoint vs complex step and central differences)")
    ap.add_argument("--adjoint-dirs", type=int, default=16)
    ap.add_argument("--no-precision", action="store_true",
                    help="skip the precision leg (FP32 / FP16 / BF16 correction storage vs an FP64 run, 500 steps)")
    ap.add_argument("--precision-steps", type=int, default=500)
    ap.add_argument("--precision-modes", default="",
                    help="comma list of the precision leg's modes (fp64, fp32, fp32_fp16, fp32_bf16); default all; "
                         "errors against FP64 need fp64 in the list")
    ap.add_argument("--no-stretched", action="store_true",
                    help="skip the stretched-grid leg (the reference's wind-tunnel box on its stretched mesh)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="also run a plain 256 MiB device copy (reference point for HBM counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: the solver step runs on a single MI355X and the library has no "
              f"multi-GPU path (asked for --gpus {args.gpus})", file=sys.stderr)
        return 2

    import torch

    from unitywindsimulation_amd import FvmSolverHip, workloads
    from unitywindsimulation_amd import _native as N

    torch.cuda.set_device(0)
    w = workloads.WORKLOADS[args.workload]()
    cf = w.cf
    X, Y, Z = w.grid
    ncell = X * Y * Z
    V, P = cf.velMaxNumIter, cf.presMaxNumIter
    C = cf.PISONumCorrectors if int(cf.fvmSolverType) == 1 else 0
    les = int(cf.turbulenceModel) == 1
    s = FvmSolverHip(cf, face_positions=w.face_positions)
    s.InitFlags(w.flags)
    s.init_noise(w.noise_seed, w.noise_amp)
    s.step(args.warmup)
    s.synchronize()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    s.step(args.steps)
    s.synchronize()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    ms_step = dt / args.steps * 1e3
    mlups = ncell / (ms_step * 1e-3) / 1e6
    b_step = bytes_per_cell_step(V, P, C, les)
    hbm_frac = b_step * ncell / (ms_step * 1e-3) / (HBM_PEAK_GBS * 1e9)

    # dominant kernel: the Jacobi sweep kernel with the larger share of step time.  The step runs the
    # sweeps in groups (temporal blocking: k_pres_sweep4 or k_pres_sweep3, four or three pressure sweeps
    # per launch as the handle's launch plan says, the rest in shorter launches; FP32 momentum sweeps as
    # k_vel_sweep2 pairs where the plan has them).  A launch's algorithmic bytes are what it must move,
    # each array once: pressure 28 B/cell (c_E, c_N, c_U, 1/Dp, rhs, p' in, p' out), momentum 73 B/cell
    # (S, G, 1/D, b: 3 planes each, the flag, u in, u out); separate sweeps would move 29 and 77 B/cell.
    pres_depth, vel_depth = s.launch_plan()
    t_vel = s.time_kernel(N.KERNEL_VEL_SWEEP, 20)
    t_pres = s.time_kernel(N.KERNEL_PRES_SWEEP, 20)
    t_pres2 = s.time_kernel(N.KERNEL_PRES_SWEEP2, 20)
    t_vel2 = s.time_kernel(N.KERNEL_VEL_SWEEP2, 20)
    t_pres3 = s.time_kernel(N.KERNEL_PRES_SWEEP3, 20)
    t_pres4 = s.time_kernel(N.KERNEL_PRES_SWEEP4, 20)
    t_pd = {1: t_pres, 2: t_pres2, 3: t_pres3, 4: t_pres4}[pres_depth]
    share_vel = (V // 2) * t_vel2 if vel_depth == 2 else V * t_vel
    share_pres = (1 + C) * (P // pres_depth) * t_pd
    if share_vel >= share_pres:
        if vel_depth == 2:
            kname, kms, kbytes, kbytes_sweeps = "k_vel_sweep2", t_vel2, 73 * ncell, 2 * 77 * ncell
        else:
            kname, kms, kbytes, kbytes_sweeps = "k_vel_sweep", t_vel, 77 * ncell, 77 * ncell
    else:
        kname = {1: "k_pres_sweep", 2: "k_pres_sweep2", 3: "k_pres_sweep3", 4: "k_pres_sweep4"}[pres_depth]
        kms, kbytes = t_pd, (29 if pres_depth == 1 else 28) * ncell
        kbytes_sweeps = pres_depth * 29 * ncell
    achieved = kbytes / (kms * 1e-3) / 1e9
    tr = MEASURED_TRAFFIC_B_PER_CELL.get(args.workload, {}).get(kname)
    traffic = int(tr * ncell) if tr is not None else None
    u = s.velocity(3)
    finite = bool(torch.isfinite(torch.from_numpy(u)).all())
    s.close()
    # what a plain streaming kernel reaches on the same working sets (6 arrays read + 1 written for the
    # pressure launches' 28 B/cell, 16 + 1 for the momentum pairs' 73 B/cell) and on one far larger than
    # the last-level cache (HBM itself)
    ceil = {"pressure_like_28B_per_cell": round(stream_ceiling(ncell, 6), 1),
            "momentum_like_68B_per_cell": round(stream_ceiling(ncell, 16), 1),
            "hbm_1792MiB": round(stream_ceiling(64 << 20, 6, 10), 1)}
    kceil = ceil["momentum_like_68B_per_cell"] if kname.startswith("k_vel") else ceil["pressure_like_28B_per_cell"]
    b_comp = compulsory_bytes_per_cell_step(V, P, C, les, pres_depth, vel_depth)
    comp_gbps = b_comp * ncell / (ms_step * 1e-3) / 1e9
    b_meas = measured_bytes_per_cell_step(args.workload, V, P, C, les, pres_depth, vel_depth)
    if args.calibrate_copy:
        src = torch.ones(64 * 1024 * 1024, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(2):
            dst.copy_(src)
        torch.cuda.synchronize()

    out = {
        "metric": METRIC,
        "value": round(mlups, 2),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": w.name, "grid_xyz": [X, Y, Z], "cells": ncell, "vel_sweeps": V,
                   "pres_sweeps": P, "piso_correctors": C, "les": "smagorinsky" if les else "none",
                   "wall": ("rough log law" if int(cf.wallModel) == 1 else "smooth log law") if les else "none",
                   "bytes_per_cell_step": b_step},
        "hbm_frac": round(hbm_frac, 4),
        "hbm_frac_basis": "B_step: every Jacobi sweep counted as a separate pass over its arrays; the two-sweep "
                          "launches move less, so this can exceed 1",
        "algorithmic_GBps": round(b_step * ncell / (ms_step * 1e-3) / 1e9, 1),
        "step_bytes": {"compulsory_B_per_cell": b_comp, "compulsory_GBps": round(comp_gbps, 1),
                       "compulsory_frac_of_peak": round(comp_gbps / HBM_PEAK_GBS, 4),
                       "compulsory_frac_of_stream_ceiling": round(comp_gbps / ceil["momentum_like_68B_per_cell"], 4),
                       "basis": "each array once per launch of the kernels the step runs (multi-sweep launches "
                                "counted as one pass, per the handle's launch plan); the stream ceiling is the "
                                "plain streaming kernel on a working set of the momentum pair's size"},
        "step_bytes_measured": None if b_meas is None else {
            "B_per_cell": b_meas, "GBps": round(b_meas * ncell / (ms_step * 1e-3) / 1e9, 1),
            "frac_of_peak": round(b_meas * ncell / (ms_step * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
            "basis": "per-launch HBM bytes from the FETCH_SIZE and WRITE_SIZE counters of a profiled urban run, "
                     "summed over the step's launch plan"},
        "stream_ceiling_GBps": ceil,
        "roofline": {"bound": "hbm", "kernel": kname,
                     "selected_by": "the Jacobi sweep kernel with the larger share of the step (step_ms_share; "
                                    "where the two shares are close either can be named); frac_by_kernel gives "
                                    "every sweep kernel's fraction",
                     "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "frac_of_stream_ceiling": round(achieved / kceil, 4),
                     "algorithmic_bytes": kbytes, "single_sweeps_bytes": kbytes_sweeps,
                     "kernel_ms": {"k_vel_sweep": round(t_vel, 5), "k_vel_sweep2": round(t_vel2, 5),
                                   "k_pres_sweep": round(t_pres, 5), "k_pres_sweep2": round(t_pres2, 5),
                                   "k_pres_sweep3": round(t_pres3, 5), "k_pres_sweep4": round(t_pres4, 5)},
                     # the same fraction for every sweep kernel (algorithmic bytes of one launch over its
                     # duration), and each one's share of the step (launches per step x duration)
                     "frac_by_kernel": {k: round(b * ncell / (t * 1e-3) / 1e9 / HBM_PEAK_GBS, 4) for k, b, t in (
                         ("k_vel_sweep", 77, t_vel), ("k_vel_sweep2", 73, t_vel2), ("k_pres_sweep", 29, t_pres),
                         ("k_pres_sweep2", 28, t_pres2), ("k_pres_sweep3", 28, t_pres3),
                         ("k_pres_sweep4", 28, t_pres4))},
                     "step_ms_share": {"momentum_sweeps": round(share_vel, 4), "pressure_sweeps_main_kernel": round(share_pres, 4)}},
        "launch_plan": {"pres_sweeps_per_launch": pres_depth, "vel_sweeps_per_launch": vel_depth},
        "finite": finite,
    }
    if not args.no_ensemble:
        out["ensemble"] = ensemble_report(args.ensemble_steps, args.ensemble_check_steps, args.ensemble_launch)
    if not args.no_stretched:
        out["stretched"] = stretched_report()
    if not args.no_precision:
        only = [m for m in args.precision_modes.split(",") if m] or None
        out["precision"] = precision_report(args.precision_steps, args.precision_steps // 2, only)
    if not args.no_adjoint:
        out["adjoint"] = adjoint_report(args.adjoint_dirs)
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(args.cpu_budget)
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
