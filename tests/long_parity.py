"""Long-run parity of the GPU step against the CPU oracle (not collected by
pytest: it takes minutes).  Runs a workload for `--steps` steps on both, with
the same seeded initial state, and prints the relative L2 errors of u and p at
the end and of the time mean of u over the last `--mean-last` steps.

    python tests/long_parity.py --workload isolated_building --steps 200 --mean-last 100
"""
from __future__ import annotations

import argparse
import sys
import time
from pathlib import Path

import numpy as np

sys.path.insert(0, str(Path(__file__).resolve().parent.parent))
sys.path.insert(0, str(Path(__file__).resolve().parent))

from helpers import oracle_for, rel_l2, solver_for  # noqa: E402

from unitywindsimulation_amd import workloads  # noqa: E402


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default="isolated_building", choices=sorted(workloads.WORKLOADS))
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--mean-last", type=int, default=100)
    args = ap.parse_args()
    w = workloads.WORKLOADS[args.workload]()
    first = args.steps - args.mean_last
    t0 = time.perf_counter()
    o = oracle_for(w)
    acc = None
    for k in range(args.steps):
        o.step(1)
        if k >= first:
            u = o.velocity().astype(np.float64)
            acc = u if acc is None else acc + u
        if (k + 1) % 50 == 0:
            print(f"  oracle step {k + 1} ({time.perf_counter() - t0:.0f} s)", flush=True)
    o_mean, o_u, o_p = acc / args.mean_last, o.velocity(), o.pressure()
    t1 = time.perf_counter()
    with solver_for(w) as s:
        s.step(first)
        s.mean_reset()
        for _ in range(args.mean_last):
            s.step(1)
            s.mean_accumulate()
        g_mean, _ = s.mean_velocity()
        g_u, g_p = s.velocity(3), s.pressure()
    print(f"{w.name}: {args.steps} steps (oracle {t1 - t0:.0f} s)")
    print(f"  u   rel-L2 {rel_l2(g_u, o_u):.3e}")
    print(f"  p   rel-L2 {rel_l2(g_p, o_p):.3e}")
    print(f"  time mean of u over the last {args.mean_last} steps rel-L2 {rel_l2(g_mean, o_mean):.3e}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
