"""GPU precision modes, FP64 norms and long-run statistics against the oracle."""
from __future__ import annotations

import numpy as np
import pytest

from unitywindsimulation_amd import _native as N
from unitywindsimulation_amd import workloads

from .helpers import oracle_for, rel_l2, solver_for

pytestmark = pytest.mark.gpu


# (100, 21, 37): two x tiles, the second partial, Y and Z divided by no tile
# height or z chunk
@pytest.mark.parametrize("grid", [(16, 8, 12), (100, 21, 37)])
def test_fp64_gpu_run_matches_fp64_oracle(gpu, grid):
    w = workloads.tiny(grid=grid)
    w.cf.precision = N.PREC_FP64
    o = oracle_for(w, precision=1)
    with solver_for(w) as s:
        s.step(3)
        o.step(3)
        eu = rel_l2(s.field64(N.FIELD_VEL), o.velocity())
        ep = rel_l2(s.field64(N.FIELD_PRES), o.pressure())
        print(f"FP64 GPU vs FP64 oracle, 3 steps: u {eu:.2e} p {ep:.2e}")
        assert eu < 1e-12 and ep < 1e-12


def test_norms_match_oracle(gpu):
    w = workloads.tiny()
    w.cf.computeNorms = True
    # the oracle records the residual of every sweep with tolerance 0 (never breaks)
    ocf = workloads.tiny().cf
    ocf.calResidual, ocf.velResidualCheckInterval, ocf.presResidualCheckInterval = True, 1, 1
    ocf.velTolerance = ocf.presTolerance = 0.0
    wo = workloads.tiny()
    wo.cf = ocf
    o = oracle_for(wo)
    with solver_for(w) as s:
        s.step(1)
        o.step(1)
        nrm = s.norms()
        ores = o.residuals()
        assert rel_l2(s.velocity(3), o.velocity()) < 1e-5
        print("norms", nrm[:4], "oracle residuals", ores)
        assert nrm[N.NORM_VEL_RESIDUAL] == pytest.approx(ores[0], rel=1e-4)
        assert nrm[N.NORM_PRES_RESIDUAL] == pytest.approx(ores[1], rel=1e-4)
        assert nrm[N.NORM_NONFINITE] == 0
        # divergence: rms of the Rhie-Chow outflow of the final state with this step's D
        D = o.field(N.FIELD_D)
        o.set_field(N.FIELD_VEL_WORK, o.velocity())
        o.run_stage(N.STAGE_PRES_PRECOMP)
        div = np.sqrt(np.sum(o.field(N.FIELD_BP) ** 2) / o.n)
        assert nrm[N.NORM_DIVERGENCE] == pytest.approx(div, rel=1e-4)
        assert D.shape == (o.n, 3)


@pytest.mark.parametrize("prec,tol", [(N.PREC_FP32_FP16, 2e-3), (N.PREC_FP32_BF16, 2e-2)])
def test_low_precision_defect_correction(gpu, prec, tol):
    """R rounds of K 2-byte-storage sweeps equal R*K plain FP32 sweeps up to the
    storage rounding; the defect shrinks from round to round."""
    w = workloads.tiny(pres_iters=60)
    ref = oracle_for(w)            # FP32, 60 plain sweeps per pressure solve
    ref.step(1)
    w.cf.precision, w.cf.refineRounds, w.cf.refineSweeps, w.cf.computeNorms = prec, 3, 20, True
    with solver_for(w) as s:
        s.step(1)
        eu = rel_l2(s.velocity(3), ref.velocity())
        ep = rel_l2(s.pressure(), ref.pressure())
        nrm = s.norms()
        rounds = nrm[N.NORM_REFINE0:N.NORM_REFINE0 + 3]
        print(f"prec {prec}: u {eu:.2e} p {ep:.2e} defect by round {rounds}")
        assert eu < tol and ep < tol
        assert rounds[1] < rounds[0] and rounds[2] < rounds[1]


@pytest.mark.parametrize("prec", [N.PREC_FP32_FP16, N.PREC_FP32_BF16])
@pytest.mark.parametrize("sweeps,grid", [(20, (20, 14, 17)), (7, (20, 14, 17)), (7, (100, 21, 37))])
def test_low_precision_triples_equal_single_sweeps(gpu, monkeypatch, prec, sweeps, grid):
    """The three- and four-sweep low-precision kernels (their intermediate
    sweeps rounded to 2 bytes as the stored iterate would be, the last of a
    round writing p') give the fields and defect norms of single
    low-precision sweeps bit for bit, with the grid's z chunk and with 5-
    and 1-plane chunks."""
    w = workloads.tiny(grid=grid)
    w.cf.precision, w.cf.refineRounds, w.cf.refineSweeps, w.cf.computeNorms = prec, 3, sweeps, True
    runs = [{"UWS_SWEEP_FUSION": "0"}, {"UWS_SWEEP_FUSION": "2"}, {"UWS_SWEEP_FUSION": "2", "UWS_ZCHUNK3": "5"},
            {"UWS_SWEEP_FUSION": "2", "UWS_ZCHUNK3": "1"}, {"UWS_SWEEP_FUSION": "2", "UWS_SWEEP_DEPTH": "4"},
            {"UWS_SWEEP_FUSION": "2", "UWS_SWEEP_DEPTH": "4", "UWS_ZCHUNK4": "5"},
            {"UWS_SWEEP_FUSION": "2", "UWS_SWEEP_DEPTH": "4", "UWS_ZCHUNK4": "1"}]
    res = []
    for env in runs:
        for k in ("UWS_SWEEP_FUSION", "UWS_ZCHUNK3", "UWS_SWEEP_DEPTH", "UWS_ZCHUNK4"):
            monkeypatch.delenv(k, raising=False)
        for k, v in env.items():
            monkeypatch.setenv(k, v)
        with solver_for(w) as s:
            s.step(2)
            res.append((s.velocity(3), s.pressure(), s.norms()))
    for env, r in zip(runs[1:], res[1:]):
        for a, b in zip(r, res[0]):
            assert np.array_equal(a, b), env


def test_time_mean_matches_oracle(gpu):
    """200 steps of a small LES case; the mean velocity over the last 100
    steps agrees with the oracle's within 1e-2 relative L2."""
    w = workloads.tiny(grid=(32, 16, 16), vel_iters=20, pres_iters=50)
    o = oracle_for(w)
    with solver_for(w) as s:
        mg = np.zeros((w.ncells, 3))
        mo = np.zeros((w.ncells, 3))
        for k in range(200):
            s.step(1)
            o.step(1)
            if k >= 100:
                mg += s.velocity(3).reshape(-1, 3)
                mo += o.velocity()
        e_last = rel_l2(s.velocity(3), o.velocity())
        e_mean = rel_l2(mg / 100, mo / 100)
        print(f"200 steps: instantaneous rel-L2 {e_last:.2e}, time-mean rel-L2 {e_mean:.2e}")
        assert e_mean <= 1e-2


def test_time_mean_accumulator(gpu):
    """uws_mean_accumulate / uws_get_mean_velocity: the FP64 mean of the
    velocities after each step, equal to the mean of the downloaded states."""
    w = workloads.tiny()
    with solver_for(w) as s:
        s.step(1)
        s.mean_reset()
        frames = []
        for _ in range(4):
            s.step(1)
            s.mean_accumulate()
            frames.append(s.field64(N.FIELD_VEL))
        um, cnt = s.mean_velocity()
    assert cnt == 4
    ref = np.mean(frames, axis=0)
    assert np.max(np.abs(um - ref)) <= 1e-12 * max(1.0, np.max(np.abs(ref)))


def test_nonfinite_detection(gpu):
    """One NaN planted in u^n of a fluid cell: after a step with norms the
    FP64 count of non-finite cells is at least one and uws_get_norms returns
    UWS_ERR_NONFINITE (with the norms filled); a clean step returns 0 cells
    and UWS_OK."""
    w = workloads.tiny()
    w.cf.computeNorms = True
    w.cf.Init()
    from unitywindsimulation_amd.geometry import noise_velocity

    u0 = noise_velocity(w.flags, w.noise_seed, w.noise_amp).reshape(-1, 3).copy()
    for plant in (False, True):
        u = u0.copy()
        if plant:
            cell = int(np.flatnonzero(w.flags == 0)[w.ncells // 3])
            u[cell, 1] = np.nan
        with solver_for(w, u0=u) as s:
            s.step(1)
            out = np.zeros(N.NORM_COUNT)
            rc = s._lib.uws_get_norms(s._h, N.ptr(out), N.NORM_COUNT)
            nrm = s.norms()
            print(f"planted NaN {plant}: status {rc}, non-finite cells {nrm[N.NORM_NONFINITE]:.0f}")
            if plant:
                assert rc == N.UWS_ERR_NONFINITE and nrm[N.NORM_NONFINITE] >= 1
                assert out[N.NORM_NONFINITE] == nrm[N.NORM_NONFINITE]
                with pytest.raises(N.UwsError):
                    s.norms(strict=True)
            else:
                assert rc == N.UWS_OK and nrm[N.NORM_NONFINITE] == 0
