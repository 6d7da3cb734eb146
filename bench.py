"""Benchmark of the MI355X level-set update (one JSON line on stdout).

Headline workload: 256^3 FP64 grid, sphere r 0.15 at (0.35, 0.35, 0.35),
LeVeque deformation velocity with its cos(pi t/3) factor folded into dt
(velocity evaluated at t_n + dt/2), dt = h, RK2 backtrace + Catmull-Rom
tricubic SL followed by the 19-point curvature each step.  One step = one pass
over the grid; one grid-point update = one phi value written by the SL step
and one curvature value at the same node, counted over the (N-4)^3 interior
nodes the SL step advects.  Inputs are resident in HBM when timing starts.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 256] [--no-cpu]
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12           # MI355X HBM3E, bytes/s
B_SL_F64 = 8 + 24 + 8       # phi0 read, velocity read, phi write (per update)
B_CURV_F64 = 8 + 8          # phi read, kappa write
B_STEP_F64 = 48             # phi0 + velocity + phi + kappa


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=150)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-steps", type=int, default=2,
                    help="full-grid steps timed in the CPU baseline leg")
    return ap.parse_args()


def headline(args):
    import torch
    import dcell_amd as D
    from dcell_amd import configs

    n = args.n
    win = D.GridWindow.unit_box(n)
    dev = torch.device("cuda:0")
    phi = torch.from_numpy(configs.sphere_sdf(n)).to(dev)
    vel = torch.from_numpy(configs.deformation_velocity(n)).to(dev)
    bufs = [phi, torch.empty_like(phi)]
    kappa = torch.empty_like(phi)
    dts = [configs.deformation_dt(n, s) for s in range(args.warmup + args.steps)]

    def step(s, ev=None):
        src, dst = bufs[s % 2], bufs[(s + 1) % 2]
        if ev is not None:
            ev[0].record()
        D.advect_sl(src, vel, win, dts[s], "cubic", "rk2", out=dst)
        if ev is not None:
            ev[1].record()
        D.curvature(dst, win, out=kappa)
        if ev is not None:
            ev[2].record()

    for s in range(args.warmup):
        step(s)
    torch.cuda.synchronize()
    events = [[torch.cuda.Event(enable_timing=True) for _ in range(3)] for _ in range(args.steps)]
    t0 = time.perf_counter()
    for i in range(args.steps):
        step(args.warmup + i, events[i])
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    sl_ms = np.array([e[0].elapsed_time(e[1]) for e in events])
    cv_ms = np.array([e[1].elapsed_time(e[2]) for e in events])
    interior = (n - 4) ** 3
    ups = interior * args.steps / wall
    sl_avg = float(sl_ms.mean()) * 1e-3
    achieved = interior * B_SL_F64 / sl_avg
    return dict(
        n=n, wall=wall, ups=ups, ms_per_step=wall / args.steps * 1e3,
        sl_ms=float(sl_ms.mean()), curv_ms=float(cv_ms.mean()),
        roofline=dict(bound="hbm", kernel="sl_dense<double,3,cubic,rk2>",
                      achieved=achieved / 1e9, peak=HBM_PEAK / 1e9, unit="GB/s",
                      frac=achieved / HBM_PEAK, traffic=None,
                      bytes_per_launch=interior * B_SL_F64),
        step_frac=ups * B_STEP_F64 / HBM_PEAK,
        curv_gbs=interior * B_CURV_F64 / (float(cv_ms.mean()) * 1e-3) / 1e9,
    )


def cpu_baseline(args):
    """The oracle (a C restatement of the reference loops) on the host: the same
    workload (RK2 + tricubic SL, then the 19-point curvature, FP64, dense over
    the full grid) for a bounded number of steps, single-threaded like the
    reference's one-rank-per-cell execution (DWorld/DCellsArray.c:57-60), and
    with all host threads for comparison."""
    import oracle as O
    from dcell_amd import configs

    n = args.n
    h = 1.0 / (n - 1)
    w = O.Win((0, 0, 0), (n, n, n), (h, h, h))
    phi = configs.sphere_sdf(n)
    vel = configs.deformation_velocity(n)
    out = {}
    nthreads = len(os.sched_getaffinity(0))
    for threads in (1, nthreads):
        O.set_threads(threads)
        p = phi
        t0 = time.perf_counter()
        for s in range(args.cpu_steps):
            p = O.sl_advect(p, vel, w, configs.deformation_dt(n, s), O.CUBIC, O.RK2)
            O.curvature(p, w)
        out[threads] = (n - 4) ** 3 * args.cpu_steps / (time.perf_counter() - t0)
    O.set_threads(1)
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return dict(value=out[1], unit="grid-point updates/s", cores=1, kind="port",
                sample=f"{args.cpu_steps} full steps of the {n}^3 workload (RK2+tricubic SL + "
                       "19-point curvature, FP64, dense) by the oracle, 1 thread",
                all_cores=dict(value=out[nthreads], cores=nthreads), cpu=model)


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU: --gpus {args.gpus} is not supported "
              "(the workload is one GPU's share; see DESIGN.md for how it would shard)",
              file=sys.stderr)
        sys.exit(2)
    r = headline(args)
    line = {
        "metric": "grid-point updates/s (SL advect+curvature), % HBM roofline, FP32 vs FP64",
        "value": r["ups"],
        "unit": "grid-point updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"3D deformation, {r['n']}^3 FP64 sphere r=0.15, RK2 backtrace + "
                               "tricubic SL + 19-point curvature, dt=h",
                   "grid": [r["n"]] * 3, "interior_updates_per_step": (r["n"] - 4) ** 3,
                   "parallelism": "single GPU"},
        "roofline": r["roofline"],
        "step_roofline_frac": r["step_frac"],
        "kernels_ms": {"sl_advect": r["sl_ms"], "curvature": r["curv_ms"]},
    }
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(args)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
