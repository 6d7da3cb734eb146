"""Benchmark of the MI355X level-set update (one JSON line on stdout).

Headline workload: 256^3 FP64 grid, sphere r 0.15 at (0.35, 0.35, 0.35),
LeVeque deformation velocity with its cos(pi t/3) factor folded into dt
(velocity evaluated at t_n + dt/2), dt = h, RK2 backtrace + Catmull-Rom
tricubic SL followed by the 19-point curvature each step.  One step = one pass
over the grid; one grid-point update = one phi value written by the SL step
and one curvature value at the same node, counted over the (N-4)^3 interior
nodes the SL step advects.  Inputs are resident in HBM when timing starts;
the velocity is held as component planes (one n^3 plane per component, as
dcl_get_velocity_planes writes it from the fluid), the layout the FP64 SL
kernel gathers fastest from; the form fp64_config_scheme_256_per_node_velocity
times the same step with the reference's per-node (u, v, w) layout.

Beside the headline, "forms" reports the same metric for the reference's own
scheme (Euler + trilinear), the FP32 form at 384^3 and the other workloads,
each over its own timed steps.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 256] [--no-cpu] [--no-forms]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12  # MI355X HBM3E, bytes/s
COPY_BYTES = 256 << 20  # bytes read (and written) by one launch of the probe's copy kernel
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "sl_f64_cubic_rk2_256.traffic.txt")
SL_KERNELS = ("sl_dense", "sl_ring3d")  # the dense 3D SL kernels (name prefixes)


def b_alg(itemsize):
    """Algorithmic bytes per update: phi0 read, 3 velocity components read, phi
    and kappa written (48 B in FP64, 24 B in FP32); per kernel: the SL pass
    reads phi0 + velocity and writes phi, the curvature pass reads phi and
    writes kappa."""
    return dict(step=6 * itemsize, sl=5 * itemsize, curv=2 * itemsize)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=150)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-forms", action="store_true", help="only the headline workload")
    ap.add_argument("--cpu-steps", type=int, default=3, help="full-grid steps of the CPU baseline leg")
    ap.add_argument("--form", choices=sorted(FORMS),
                    help="run one form only (for profiling) and print its JSON")
    ap.add_argument("--probe", nargs="?", const="sl64", choices=sorted(PROBES),
                    help="a short program for rocprofv3 --pmc: sl64 (default) 5 launches of the headline SL "
                         "kernel and 5 of a copy kernel; sl32, curv, batch, adjoint: the other hot kernels")
    ap.add_argument("--counter-summary", nargs="+", metavar=("NAME", "PROBE DIR"),
                    help="NAME PROBE DIR...: write profiles/NAME.counters.txt from the rocprofv3 --pmc output "
                         "folders of --probe PROBE")
    ap.add_argument("--traffic-summary", nargs=2, metavar=("FETCH_DIR", "WRITE_DIR"),
                    help="write the HBM traffic summary from two rocprofv3 --pmc output folders")
    return ap.parse_args()


def deformation_velocity_dev(n, dtype, planes):
    """The deformation velocity on the device, per node (n, n, n, 3) or as
    component planes (3, n, n, n), the layout dcl_get_velocity_planes writes."""
    import torch
    from dcell_amd import configs
    v = configs.deformation_velocity(n)
    if planes:
        v = np.ascontiguousarray(np.moveaxis(v, -1, 0))
    return torch.from_numpy(v).to("cuda:0", dtype)


def dense_run(n, dtype, interp, bt, steps, warmup, deform=True, planes=False):
    """Time `steps` dense updates (SL kernel then curvature kernel) of an n^3
    sphere under the deformation field; per-kernel times from HIP events on
    the stream the kernels run on.  ``planes``: velocity as component planes."""
    import torch
    import dcell_amd as D
    from dcell_amd import configs

    win = D.GridWindow.unit_box(n)
    dev = torch.device("cuda:0")
    phi = torch.from_numpy(configs.sphere_sdf(n)).to(dev, dtype)
    vel = deformation_velocity_dev(n, dtype, planes)
    bufs = [phi, torch.empty_like(phi)]
    kappa = torch.empty_like(phi)
    dts = [configs.deformation_dt(n, s) for s in range(warmup + steps)]

    def step(s, ev=None):
        src, dst = bufs[s % 2], bufs[(s + 1) % 2]
        if ev is not None:
            ev[0].record()
        D.advect_sl(src, vel, win, dts[s], interp, bt, out=dst, planes=planes)
        if ev is not None:
            ev[1].record()
        D.curvature(dst, win, out=kappa)
        if ev is not None:
            ev[2].record()

    for s in range(warmup):
        step(s)
    torch.cuda.synchronize()
    events = [[torch.cuda.Event(enable_timing=True) for _ in range(3)] for _ in range(steps)]
    t0 = time.perf_counter()
    for i in range(steps):
        step(warmup + i, events[i])
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    sl_s = float(np.mean([e[0].elapsed_time(e[1]) for e in events])) * 1e-3
    cv_s = float(np.mean([e[1].elapsed_time(e[2]) for e in events])) * 1e-3
    interior = (n - 4) ** 3
    B = b_alg(torch.tensor([], dtype=dtype).element_size())
    ups = interior * steps / wall
    out = dict(n=n, velocity_layout="component planes" if planes else "per node", steps=steps, wall_s=wall, updates_per_s=ups, ms_per_step=wall / steps * 1e3,
               step_roofline_frac=ups * B["step"] / HBM_PEAK,
               sl_ms=sl_s * 1e3, curv_ms=cv_s * 1e3,
               sl_gbs=interior * B["sl"] / sl_s / 1e9, curv_gbs=interior * B["curv"] / cv_s / 1e9,
               bytes_per_update=B["step"])
    del phi, vel, bufs, kappa
    torch.cuda.empty_cache()
    return out


def probe(args):
    """Short programs for the counter runs (rocprofv3 --pmc), one per probe:
    sl64: 5 launches of the headline SL kernel on the headline workload
    (velocity as component planes), then
    5 launches of a plain copy kernel (torch's elementwise neg over COPY_BYTES
    of FP32: reads and writes COPY_BYTES), whose counter reading per byte
    scales the SL kernel's; the others run the kernels named in PROBES."""
    if args.probe != "sl64":
        PROBES[args.probe](args)
        return
    import torch
    import dcell_amd as D
    from dcell_amd import configs

    n = args.n
    win = D.GridWindow.unit_box(n)
    dev = torch.device("cuda:0")
    phi = torch.from_numpy(configs.sphere_sdf(n)).to(dev, torch.float64)
    vel = deformation_velocity_dev(n, torch.float64, True)
    out = torch.empty_like(phi)
    for s in range(5):
        D.advect_sl(phi, vel, win, configs.deformation_dt(n, s), "cubic", "rk2", out=out, planes=True)
    src = torch.ones(COPY_BYTES // 4, dtype=torch.float32, device=dev)
    dst = torch.empty_like(src)
    for _ in range(5):
        torch.neg(src, out=dst)
    torch.cuda.synchronize()
    print(f"probe done: 5 x SL (RK2 + tricubic, FP64) at {n}^3, 5 x neg over {COPY_BYTES} B", flush=True)


def probe_sl32(args):
    """5 launches of the FP32 dense SL kernel (RK2 + tricubic) at 384^3."""
    _probe_dense(384, "float32", curv=False)


def probe_curv(args):
    """5 launches of the dense curvature kernel on the 256^3 FP64 sphere and 5
    on the 384^3 FP32 one."""
    _probe_dense(256, "float64", sl=False)
    _probe_dense(384, "float32", sl=False)


def _probe_dense(n, dtype, sl=True, curv=True):
    import torch
    import dcell_amd as D
    from dcell_amd import configs
    dt = getattr(torch, dtype)
    win = D.GridWindow.unit_box(n)
    phi = torch.from_numpy(configs.sphere_sdf(n)).to("cuda:0", dt)
    vel = torch.from_numpy(configs.deformation_velocity(n)).to("cuda:0", dt)
    out, kap = torch.empty_like(phi), torch.empty_like(phi)
    for s in range(5):
        if sl:
            D.advect_sl(phi, vel, win, configs.deformation_dt(n, s), "cubic", "rk2", out=out)
        if curv:
            D.curvature(phi, win, out=kap)
    torch.cuda.synchronize()
    print(f"probe done: 5 x {'SL ' if sl else ''}{'curvature ' if curv else ''}at {n}^3 {dtype}", flush=True)


def probe_batch(args):
    """3 steps of the 512-cell FP32 batch (RK2 + tricubic) in the ring
    configuration the autotuner picks for it (shape 7, occupancy index 2:
    8 blocks per CU, static stride, the default z-chunk per size class)."""
    import torch
    import dcell_amd as D
    from dcell_amd import configs
    from dcell_amd.batch import Batch
    edges = configs.batch_edges(512, 4)
    g = configs.philox(4)
    b = Batch([D.GridWindow.unit_box(int(e)) for e in edges], torch.float32, "cubic", "rk2")
    phi0, phi1, kap, vel = b.new_field(), b.new_field(), b.new_field(), b.new_field(3)
    dts = []
    for c, e in enumerate(edges):
        phi, v, dt = configs.cell_fields(int(e), g, "cuda:0", torch.float32)
        b.cell(phi0, c).copy_(phi)
        b.cell(vel, c, 3).copy_(v)
        dts.append(dt)
    b.set_dt(dts)
    b.set_config(7, 2)
    b.set_claim(0)
    for _ in range(3):
        b.step(phi0, phi1, kap, vel)
    torch.cuda.synchronize()
    print("probe done: 3 batch steps, 512 cells", flush=True)
    b.close()


def probe_adjoint(args):
    """One checkpointed backward of the 128^3 FP32 differentiation form (50
    adjoint steps, 40 recomputed forward steps)."""
    r = adjoint_run(128, 50, 10)
    print(f"probe done: adjoint backward {r['backward_ms']:.1f} ms", flush=True)


PROBES = {"sl64": None, "sl32": probe_sl32, "curv": probe_curv, "batch": probe_batch, "adjoint": probe_adjoint}

# the counters of the summaries, one rocprofv3 --pmc pass per line
XCDS = 8  # MI355X: accelerator complex dies, each with its own GRBM_GUI_ACTIVE
COUNTER_PASSES = (
    "SQ_WAVES SQ_INSTS_VALU SQ_INSTS_VMEM_RD SQ_INSTS_VMEM_WR SQ_INSTS_LDS SQ_INSTS_SALU SQ_INSTS_BRANCH SQ_WAIT_INST_ANY",
    "VALUBusy VALUUtilization MeanOccupancyPerCU TA_BUSY_avr TD_TD_BUSY_sum SQ_LDS_BANK_CONFLICT SQ_ACTIVE_INST_LDS "
    "GRBM_GUI_ACTIVE",
    "FETCH_SIZE",
    "WRITE_SIZE",
)


def counter_summary(args):
    """profiles/NAME.counters.txt: per kernel, the mean per dispatch of every
    counter found in the rocprofv3 --pmc folders given, and per wave the
    instruction counts (one node per lane in these kernels)."""
    import collections
    import csv
    import glob
    name, probe_name, dirs = args.counter_summary[0], args.counter_summary[1], args.counter_summary[2:]
    vals = collections.defaultdict(lambda: collections.defaultdict(list))
    for d in dirs:
        for f in glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True):
            per = collections.defaultdict(float)
            kname = {}
            with open(f) as fh:
                for row in csv.DictReader(fh):
                    key = (row["Dispatch_Id"], row["Counter_Name"])
                    per[key] += float(row["Counter_Value"])
                    kname[row["Dispatch_Id"]] = row["Kernel_Name"]
            for (disp, cn), v in per.items():
                vals[kname[disp]][cn].append(v)
    out = os.path.join(ROOT, "profiles", f"{name}.counters.txt")
    with open(out, "w") as fo:
        fo.write(f"# rocprofv3 counters: {name}\n# one pass per line of counters (from the repository root):\n")
        for q, c in enumerate(COUNTER_PASSES):
            d = dirs[q] if q < len(dirs) else f"DIR{q}"
            fo.write(f"#   rocprofv3 --pmc {c} -d {d} --output-format csv -- python3 bench.py --probe {probe_name}\n")
        fo.write(f"#   python3 bench.py --counter-summary {name} {probe_name} " + " ".join(dirs) + "\n")
        fo.write("# values: mean per dispatch; '/wave' = per wavefront of 64 lanes;\n"
                 "# FETCH_SIZE / WRITE_SIZE in KiB (raw: the copy kernel reads 2x its FETCH_SIZE, see the\n"
                 f"# traffic summary); GRBM_GUI_ACTIVE sums the {XCDS} XCDs' busy cycles, so the kernel's\n"
                 f"# cycles are GRBM_GUI_ACTIVE / {XCDS} (about 2.2 GHz x its duration under the profiler);\n"
                 "# busy %: TA_BUSY_avr, TD_TD_BUSY_sum / 256 CUs and SQ_LDS_BANK_CONFLICT / 256 CUs over\n"
                 "# those cycles; VALUBusy as rocprofv3 derives it\n")
        for k in sorted(vals, key=lambda k: -len(vals[k])):
            cs = vals[k]
            if not any(c in cs for c in ("SQ_WAVES", "FETCH_SIZE", "GRBM_GUI_ACTIVE")):
                continue
            mean = {cn: sum(v) / len(v) for cn, v in cs.items()}
            fo.write(f"\nkernel: {k[:160]}\n  dispatches: {max(len(v) for v in cs.values())}\n")
            waves = mean.get("SQ_WAVES")
            for cn in sorted(mean):
                line = f"  {cn}: {mean[cn]:.6g}"
                if waves and cn.startswith("SQ_INSTS"):
                    line += f"  ({mean[cn] / waves:.2f} /wave)"
                fo.write(line + "\n")
            gui = mean.get("GRBM_GUI_ACTIVE")
            if gui:
                gui /= XCDS
                fo.write(f"  kernel cycles (GRBM_GUI_ACTIVE / {XCDS}): {gui:.6g}\n")
                if "TA_BUSY_avr" in mean:
                    fo.write(f"  TA busy: {100 * mean['TA_BUSY_avr'] / gui:.1f} %\n")
                if "TD_TD_BUSY_sum" in mean:
                    fo.write(f"  TD busy: {100 * mean['TD_TD_BUSY_sum'] / 256 / gui:.1f} %\n")
                if "SQ_LDS_BANK_CONFLICT" in mean:
                    fo.write(f"  LDS bank-conflict cycles: {100 * mean['SQ_LDS_BANK_CONFLICT'] / 256 / gui:.1f} %\n")
    print(open(out).read())


def _pmc_per_kernel(folder, counter):
    """Mean per-dispatch value of `counter` by kernel name, from the
    counter_collection CSV files rocprofv3 wrote under `folder`."""
    import csv
    import glob
    import collections
    per = collections.defaultdict(float)
    name = {}
    for f in glob.glob(os.path.join(folder, "**", "*counter_collection.csv"), recursive=True):
        with open(f) as fh:
            for row in csv.DictReader(fh):
                if row["Counter_Name"] != counter:
                    continue
                key = (f, row["Dispatch_Id"])
                per[key] += float(row["Counter_Value"])
                name[key] = row["Kernel_Name"]
    acc = collections.defaultdict(list)
    for key, v in per.items():
        acc[name[key]].append(v)
    return {k: sum(v) / len(v) for k, v in acc.items()}


def traffic_summary(args):
    """HBM bytes per launch of the headline SL kernel from a FETCH_SIZE run and
    a WRITE_SIZE run of `--probe` (values in KiB per dispatch), each scaled by
    COPY_BYTES / (the same counter for the copy kernel)."""
    fetch_dir, write_dir = args.traffic_summary
    lines = []
    res = {}
    kname = None
    for counter, folder in (("FETCH_SIZE", fetch_dir), ("WRITE_SIZE", write_dir)):
        k = _pmc_per_kernel(folder, counter)
        sl = [(name, v) for name, v in k.items() if any(p in name for p in SL_KERNELS)]
        cp = [v for name, v in k.items() if "neg" in name]
        if len(sl) != 1 or len(cp) != 1:
            sys.exit(f"{folder}: expected one SL kernel and one neg kernel, found {sorted(k)}")
        kname = sl[0][0]
        scale = COPY_BYTES / (cp[0] * 1024.0)
        res[counter] = sl[0][1] * 1024.0 * scale
        lines.append(f"{counter}: SL kernel {sl[0][1]:.1f} KiB/dispatch, copy kernel {cp[0]:.1f} KiB/dispatch "
                     f"for {COPY_BYTES} B, scale {scale:.4f}")
    n = args.n
    alg = (n - 4) ** 3 * b_alg(8)["sl"]
    with open(TRAFFIC_FILE, "w") as f:
        f.write(f"# HBM traffic of the dense SL kernel, {n}^3 FP64 deformation workload (RK2 + tricubic)\n")
        f.write("# commands (from the repository root):\n")
        f.write(f"#   rocprofv3 --pmc FETCH_SIZE -d {fetch_dir} --output-format csv -- python3 bench.py --probe sl64 --n {n}\n")
        f.write(f"#   rocprofv3 --pmc WRITE_SIZE -d {write_dir} --output-format csv -- python3 bench.py --probe sl64 --n {n}\n")
        f.write(f"#   python3 bench.py --n {n} --traffic-summary {fetch_dir} {write_dir}\n")
        for ln in lines:
            f.write(f"# {ln}\n")
        f.write(f"kernel: {kname}\n")
        f.write(f"n: {n}\n")
        f.write(f"fetch_bytes_per_launch: {res['FETCH_SIZE']:.0f}\n")
        f.write(f"write_bytes_per_launch: {res['WRITE_SIZE']:.0f}\n")
        f.write(f"algorithmic_bytes_per_launch: {alg}\n")
    print(open(TRAFFIC_FILE).read())


def read_traffic(n):
    """(kernel name, HBM bytes per launch) of the headline SL kernel from the
    committed counter summary, or (None, None) when there is none for this
    grid."""
    try:
        with open(TRAFFIC_FILE) as f:
            kv = dict((a.strip(), b.strip()) for a, b in
                      (ln.split(":", 1) for ln in f if ln.strip() and not ln.startswith("#")))
    except OSError:
        return None, None
    if int(kv["n"]) != n:
        return None, None
    return kv["kernel"], float(kv["fetch_bytes_per_launch"]) + float(kv["write_bytes_per_launch"])


COUNTERS_FILE = os.path.join(ROOT, "profiles", "sl_f64_cubic_rk2_256.counters.txt")


def read_limiter(kernel_prefix="dcl::sl_ring3d<double, 1, 1,"):
    """The unit that limits the headline SL kernel, from the committed counter
    summary: the busiest of the vector ALU, the texture address unit and the
    texture data unit (percent of the kernel's cycles), or None without a
    summary."""
    try:
        with open(COUNTERS_FILE) as f:
            lines = f.read().splitlines()
    except OSError:
        return None
    vals, inside = {}, False
    for ln in lines:
        if ln.startswith("kernel:"):
            inside = kernel_prefix in ln
        elif inside and ":" in ln:
            k, v = ln.strip().split(":", 1)
            vals[k] = v.strip()
    busy = {}
    for unit, key in (("VALU", "VALUBusy"), ("TA (texture address)", "TA busy"),
                      ("TD (texture data)", "TD busy")):
        if key in vals:
            busy[unit] = float(vals[key].split()[0])
    if not busy:
        return None
    top = max(busy, key=busy.get)
    return {"unit": top, "busy_pct": busy,
            "source": os.path.relpath(COUNTERS_FILE, ROOT)}


def host_cpu():
    """lscpu's model name of the host, with its family / model / stepping and
    socket / core counts (the model name alone can be a bare vendor string on
    virtual machines)."""
    import subprocess
    keys = ("Model name", "CPU family", "Model", "Stepping", "Socket(s)", "Core(s) per socket",
            "Thread(s) per core", "CPU(s)")
    info = {}
    try:
        txt = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
        for line in txt.splitlines():
            k, _, v = line.partition(":")
            if k.strip() in keys and k.strip() not in info:
                info[k.strip()] = v.strip()
    except (OSError, subprocess.SubprocessError):
        pass
    if "Model name" not in info:
        try:
            with open("/proc/cpuinfo") as f:
                for line in f:
                    if line.startswith("model name"):
                        info["Model name"] = line.split(":", 1)[1].strip()
                        break
        except OSError:
            pass
    return info


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
        p = O.sl_advect(phi, vel, w, configs.deformation_dt(n, 0), O.CUBIC, O.RK2)  # warm-up step
        O.curvature(p, w)
        times = []
        for s in range(1, 1 + args.cpu_steps):
            t0 = time.perf_counter()
            p = O.sl_advect(p, vel, w, configs.deformation_dt(n, s), O.CUBIC, O.RK2)
            O.curvature(p, w)
            times.append(time.perf_counter() - t0)
        out[threads] = (n - 4) ** 3 / float(np.median(times))
    O.set_threads(1)
    # the irregular-node list of the same sphere over its 6-cell band
    # (irregular_run's workload), by the oracle on one thread
    band = O.band_compact(phi, w, 6.0)
    O.irregular_nodes(phi, w, band[:1000])
    t0 = time.perf_counter()
    recs = O.irregular_nodes(phi, w, band)
    t_irr = time.perf_counter() - t0
    # the drop-in form's band step (Euler + trilinear over the 6-cell band,
    # the drop-in's default scheme) by the oracle on one thread
    vel_full = configs.deformation_velocity(n)
    dt0 = configs.deformation_dt(n, 0)
    O.sl_advect(phi, vel_full, w, dt0, O.LINEAR, O.EULER, band=band[:1000])
    t0 = time.perf_counter()
    O.sl_advect(phi, vel_full, w, dt0, O.LINEAR, O.EULER, band=band)
    t_band = time.perf_counter() - t0
    # the reinitialisation form's inputs, by the oracle: the 256^3 sphere on
    # one thread, and the 64-cell batch one cell per thread
    import concurrent.futures as cf
    wr = O.Win((0, 0, 0), (n, n, n), (1.0, 1.0, 1.0))
    phir = 1.5 * configs.sphere_sdf(n)
    recr = O.irregular_nodes(phir, wr, O.band_compact(phir, wr, 8.0))
    t0 = time.perf_counter()
    _, bandr = O.reinitialize(phir, wr, recr)
    t_re = time.perf_counter() - t0
    jobs = []
    for e, c in reinit_cells():
        we = O.Win((0, 0, 0), (e, e, e), (1.0, 1.0, 1.0))
        pe = 1.5 * configs.sphere_sdf(e, c, 0.25)
        jobs.append((pe, we, O.irregular_nodes(pe, we, O.band_compact(pe, we, 8.0))))

    def one(job):
        return len(O.reinitialize(*job)[1])

    t0 = time.perf_counter()
    nbatch = sum(one(j) for j in jobs)
    t_b1 = time.perf_counter() - t0
    with cf.ThreadPoolExecutor(nthreads) as ex:
        t0 = time.perf_counter()
        list(ex.map(one, jobs))
        t_bn = time.perf_counter() - t0
    model = host_cpu()
    return dict(value=out[1], unit="grid-point updates/s", cores=1, kind="port",
                sample=f"median of {args.cpu_steps} timed full steps (after 1 warm-up step) of the {n}^3 "
                       "workload (RK2+tricubic SL + 19-point curvature, FP64, dense) by the oracle, 1 thread",
                all_cores=dict(value=out[nthreads], cores=nthreads), cpu=model,
                irregular_nodes=dict(band_nodes_per_s=len(band) / t_irr, records=len(recs), cores=1,
                                     sample=f"the full {len(band)}-node band of the {n}^3 sphere, one call"),
                band_step=dict(ms=t_band * 1e3, band_nodes=len(band), cores=1,
                               sample=f"one Euler + trilinear SL step over the {len(band)}-node 6-cell band of "
                                      f"the {n}^3 sphere"),
                reinitialize=dict(single_ms=t_re * 1e3, single_band_nodes=len(bandr), cores=1,
                                  batch_ms=t_b1 * 1e3, batch_band_nodes=nbatch,
                                  batch_all_cores_ms=t_bn * 1e3, all_cores=nthreads,
                                  sample=f"the reinitialisation form's inputs: the {n}^3 sphere (one call), "
                                         "and the 64-cell batch, cells one after another and one per thread"))


def batch_run(ncells, seed, steps, warmup, interp="cubic", bt="rk2", autotune=True, dim=3, lo=32, hi=128,
              dtype=None):
    """The batching form: ncells cells (edges log-uniform in [lo, hi], seeded),
    FP32 by default, each a perturbed sphere (2D: circle) in random Fourier-mode
    velocity with its own dt; one step = one hipGraph launch (SL + curvature
    over the packed tile list).  Reports the autotuned configuration, with
    the ring z-chunk chosen for each cell-size class."""
    import torch
    import dcell_amd as D
    from dcell_amd import configs
    from dcell_amd.batch import CLAIMS, ZCHUNKS, Batch

    dtype = torch.float32 if dtype is None else dtype
    edges = configs.batch_edges(ncells, seed, lo, hi)
    g = configs.philox(seed)
    wins = [D.GridWindow.unit_box(int(n), dim) for n in edges]
    b = Batch(wins, dtype, interp, bt)
    phi0, phi1, kap, vel = b.new_field(), b.new_field(), b.new_field(), b.new_field(dim)
    dts = []
    fields = configs.cell_fields if dim == 3 else configs.cell_fields_2d
    for c, n in enumerate(edges):
        phi, v, dt = fields(int(n), g, "cuda:0", dtype)
        b.cell(phi0, c).copy_(phi)
        b.cell(vel, c, dim).copy_(v)
        dts.append(dt)
    b.set_dt(dts)
    tune = None
    if autotune:
        best, table, ztable, ctable = b.autotune(phi0, phi1, kap, vel, reps=3)
        tune = dict(best_ms=best, table_ms=table.round(4).tolist(), zchunks=list(ZCHUNKS),
                    class_zchunk_ms=ztable.round(4).tolist(), tiles_per_claim=list(CLAIMS),
                    claim_ms=ctable.round(4).tolist(), chosen=b.config(), classes=b.classes())
    graphs = [b.graph(phi0, phi1, kap, vel), b.graph(phi1, phi0, kap, vel)]
    for s in range(warmup):
        Batch.launch(graphs[s % 2])
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record()
    for s in range(steps):
        Batch.launch(graphs[s % 2])
    e1.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    gpu_s = e0.elapsed_time(e1) * 1e-3
    interior = int(sum((int(n) - 4) ** dim for n in edges))
    ups = interior * steps / wall
    isz = torch.tensor([], dtype=dtype).element_size()
    bpu = (3 + dim) * isz     # phi0 + velocity read, phi + kappa written
    out = dict(cells=ncells, dim=dim, edges=[lo, hi], nodes=int(b.nodes), interior_updates_per_step=interior,
               steps=steps, interp=interp, backtrace=bt, dtype=str(dtype).replace("torch.", ""),
               updates_per_s=ups, ms_per_step=wall / steps * 1e3, gpu_ms_per_step=gpu_s / steps * 1e3,
               bytes_per_update=bpu, step_roofline_frac=ups * bpu / HBM_PEAK, autotune=tune)
    b.close()
    del phi0, phi1, kap, vel
    torch.cuda.empty_cache()
    return out


def pls_oracle(phi_np, vel_np, P_np, n, dts):
    """The FP64 oracle's run of the precision form on the host with every host
    thread: per step the particle RK2, the SL step, the escaped-particle
    correction and the radius update, in LevelSetAdvectPLS's order
    (LevelSetMethod/ParticleLS.c:299-329).  Returns phi after the last step."""
    import oracle as O
    h = 1.0 / (n - 1)
    w = O.Win((0, 0, 0), (n, n, n), (h, h, h))
    O.set_threads(len(os.sched_getaffinity(0)))
    phi, P = phi_np, P_np
    for dt in dts:
        P = O.pls_advect(P, vel_np, w, w.d, dt)
        phi = O.sl_advect(phi, vel_np, w, dt, O.CUBIC, O.RK2)
        phi = O.pls_correct(P, phi, w, None, 0, 3)
        P = O.pls_adjust_radii(P, phi, w)
    O.set_threads(1)
    return phi


def pls_errors(phi, ref, n):
    """Errors of phi against the reference run ref (both float64 numpy): max
    and RMS |dphi| over the grid, max |dphi| where |ref| < 6, and the relative
    errors of the interface area and the delta-weighted mean curvature (the
    oracle's metric functions on both, eps = 1.5 cells)."""
    import oracle as O
    h = 1.0 / (n - 1)
    w = O.Win((0, 0, 0), (n, n, n), (h, h, h))
    d = phi - ref
    near = np.abs(ref) < 6
    O.set_threads(len(os.sched_getaffinity(0)))
    a, _, km, _ = O.interface_metrics(phi, w, O.curvature(phi, w))
    ar, _, kmr, _ = O.interface_metrics(ref, w, O.curvature(ref, w))
    O.set_threads(1)
    return dict(max_abs_phi=float(np.abs(d).max()), rms_phi=float(np.sqrt(np.mean(d * d))),
                max_abs_phi_band=float(np.abs(d[near]).max()),
                area=float(a), area_rel=float(abs(a - ar) / ar),
                mean_curvature=float(km), mean_curvature_rel=float(abs(km - kmr) / abs(kmr)))


def pls_run(n, steps, per_cell=32, seed=3, oracle=True):
    """The precision form: n^3 deformation workload with phi and velocity in
    FP32 (departure coordinates FP64), a particle level set of `per_cell`
    particles per cell in |phi| < 3 (seeded), per step in LevelSetAdvectPLS's
    order (LevelSetMethod/ParticleLS.c:299-329): particle RK2, SL, escaped-
    particle correction, then the curvature of the corrected phi and the
    radius update.  The same run in FP64 on the GPU, and the FP64 oracle's run
    on the host, give the error figures (max / RMS phi difference, interface
    area, mean curvature)."""
    import torch
    import dcell_amd as D
    from dcell_amd import configs
    from dcell_amd import metrics as M
    from dcell_amd import pls as PL

    win = D.GridWindow.unit_box(n)
    dev = torch.device("cuda:0")
    phi_np = configs.sphere_sdf(n)
    vel_np = configs.deformation_velocity(n)
    P_np = configs.seed_particles(phi_np, per_cell, 3.0, seed)
    dts = [configs.deformation_dt(n, s) for s in range(steps)]
    out = {}
    finals = {}
    for dtype in (torch.float32, torch.float64):
        phi = torch.from_numpy(phi_np).to(dev, dtype)
        vel = torch.from_numpy(vel_np).to(dev, dtype)
        P = torch.from_numpy(P_np).to(dev, dtype)
        other, kappa = torch.empty_like(phi), torch.empty_like(phi)
        ws = PL.CorrectionWorkspace(win.nodes, 8 * len(P), dtype)
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(6)]
        t_pa = t_sl = t_pc = t_cv = t_rad = 0.0
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for s in range(steps):
            ev[0].record()
            PL.advect_particles(P, vel, win, win.d, dts[s])
            ev[1].record()
            D.advect_sl(phi, vel, win, dts[s], "cubic", "rk2", out=other)
            ev[2].record()
            PL.correct(P, other, win, ws)
            ev[3].record()
            D.curvature(other, win, out=kappa)
            ev[4].record()
            PL.adjust_radii(P, other, win)
            ev[5].record()
            phi, other = other, phi
            torch.cuda.synchronize()
            t_pa += ev[0].elapsed_time(ev[1])
            t_sl += ev[1].elapsed_time(ev[2])
            t_pc += ev[2].elapsed_time(ev[3])
            t_cv += ev[3].elapsed_time(ev[4])
            t_rad += ev[4].elapsed_time(ev[5])
        wall = time.perf_counter() - t0
        area, kint, kmean, cnt = M.interface_metrics(phi, win, kappa)
        key = "fp32" if dtype == torch.float32 else "fp64"
        interior = (n - 4) ** 3
        out[key] = dict(ms_per_step=wall / steps * 1e3, sl_curv_ms=(t_sl + t_cv) / steps, sl_ms=t_sl / steps,
                        curv_ms=t_cv / steps, pls_advect_ms=t_pa / steps,
                        pls_correct_ms=t_pc / steps, pls_radii_ms=t_rad / steps,
                        sl_curv_updates_per_s=interior / ((t_sl + t_cv) / steps * 1e-3),
                        particle_updates_per_s=len(P) / ((t_pa + t_pc) / steps * 1e-3),
                        area=area, mean_curvature=kmean)
        finals[key] = phi
        ws.close()
        del vel, P, other, kappa
    a, b = finals["fp32"], finals["fp64"]
    mx, ss, mb, cnt = M.error_norms(a, b)
    near = (b.abs() < 6)
    out["errors_vs_fp64"] = dict(
        max_abs_phi=mx, rms_phi=float(np.sqrt(ss / cnt)),
        max_abs_phi_band=float((a.double() - b)[near].abs().max()),
        area_rel=abs(out["fp32"]["area"] - out["fp64"]["area"]) / out["fp64"]["area"],
        mean_curvature_rel=abs(out["fp32"]["mean_curvature"] - out["fp64"]["mean_curvature"])
        / abs(out["fp64"]["mean_curvature"]))
    out.update(n=n, steps=steps, particles=int(len(P_np)), per_cell=per_cell)
    host = {k: v.double().cpu().numpy() for k, v in finals.items()}
    del a, b, finals, near
    torch.cuda.empty_cache()
    if oracle:
        t0 = time.perf_counter()
        ref = pls_oracle(phi_np, vel_np, P_np, n, dts)
        t_orc = time.perf_counter() - t0
        out["errors_vs_oracle"] = {k: pls_errors(v, ref, n) for k, v in host.items()}
        out["errors_vs_oracle"].update(
            oracle_s=t_orc, oracle_threads=len(os.sched_getaffinity(0)),
            sample=f"the FP64 oracle's full {steps}-step run of the same inputs on the host")
    return out


def adjoint_run(n=128, steps=50, every=10):
    """The differentiation form: steady random Fourier velocity, sphere r 0.2,
    FP32 SL (RK2 + tricubic) for `steps` steps with a checkpoint every `every`
    steps, loss = sum (H(phi_T) - H(phi_target))^2, gradient w.r.t. velocity
    by the adjoint kernel.  Reports forward and backward times."""
    import torch
    import dcell_amd as D
    from dcell_amd import adjoint as A
    from dcell_amd import configs

    win = D.GridWindow.unit_box(n)
    h = 1.0 / (n - 1)
    dev = torch.device("cuda:0")
    phi0 = torch.from_numpy(configs.sphere_sdf(n, (0.5, 0.5, 0.5), 0.2)).to(dev, torch.float32)
    vel_np = configs.random_fourier_velocity(n, 6, 3, 0.5, seed=5)
    vt_np = configs.random_fourier_velocity(n, 6, 3, 0.5, seed=6)
    dt = 0.8 * h / max(np.abs(vel_np).max(), np.abs(vt_np).max())
    vt = torch.from_numpy(vt_np).to(dev, torch.float32)
    tgt = phi0
    for _ in range(steps):
        tgt = D.advect_sl(tgt, vt, win, dt, "cubic", "rk2")
    v = torch.from_numpy(vel_np).to(dev, torch.float32).requires_grad_(True)
    eps = 1.5
    # warm-up: one short forward + backward, so that first-use loading of the
    # kernels (ours and torch's) is outside the timed region
    w = A.advect_n(phi0, v, win, dt, 2, checkpoint_every=1)
    ((A.heaviside(w, eps) - A.heaviside(tgt, eps)) ** 2).sum().backward()
    v.grad = None
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    phiT = A.advect_n(phi0, v, win, dt, steps, checkpoint_every=every)
    loss = ((A.heaviside(phiT, eps) - A.heaviside(tgt, eps)) ** 2).sum()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    loss.backward()
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    interior = (n - 4) ** 3
    return dict(n=n, steps=steps, checkpoint_every=every, loss=float(loss), forward_ms=(t1 - t0) * 1e3,
                backward_ms=(t2 - t1) * 1e3,
                adjoint_updates_per_s=interior * steps / (t2 - t1),
                grad_norm=float(v.grad.norm()))


def zalesak_run(n=128, steps=628, interp="cubic", bt="rk2"):
    """C1: the Zalesak slotted disk (128^2 FP64) under rigid rotation for one
    revolution (628 steps, dt = 1/628), SL + 9-point curvature every step.
    Reports updates/s launched step by step and as one replayed HIP graph,
    and, after the revolution, the area of {phi < 0} and the max / mean
    |phi - phi_initial| within 3 cells of the interface."""
    import torch
    import dcell_amd as D
    from dcell_amd import configs

    h = 1.0 / (n - 1)
    win = D.GridWindow((0, 0, 0), (n, n, 1), (h, h, h))
    dev = torch.device("cuda:0")
    phi_np = configs.zalesak_sdf(n)
    phi = torch.from_numpy(phi_np).to(dev, torch.float64)
    vel = torch.from_numpy(configs.rotation_velocity(n)).to(dev, torch.float64)
    bufs = [phi, torch.empty_like(phi)]
    kappa = torch.empty_like(phi)
    for s in range(2):  # warm-up (no state carried over)
        D.advect_sl(bufs[0], vel, win, 1.0 / steps, interp, bt, out=bufs[1])
        D.curvature(bufs[1], win, out=kappa)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for s in range(steps):
        D.advect_sl(bufs[s % 2], vel, win, 1.0 / steps, interp, bt, out=bufs[(s + 1) % 2])
        D.curvature(bufs[(s + 1) % 2], win, out=kappa)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    final = bufs[steps % 2].cpu().numpy()
    # the same revolution as one HIP graph (1256 kernel nodes), replayed: the
    # two 16k-node kernels of a step are launch-bound one by one
    start = torch.from_numpy(phi_np).to(dev, torch.float64)
    g = torch.cuda.CUDAGraph()
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        D.advect_sl(bufs[0], vel, win, 1.0 / steps, interp, bt, out=bufs[1])
        D.curvature(bufs[1], win, out=kappa)
    torch.cuda.current_stream().wait_stream(side)
    with torch.cuda.graph(g):
        for s in range(steps):
            D.advect_sl(bufs[s % 2], vel, win, 1.0 / steps, interp, bt, out=bufs[(s + 1) % 2])
            D.curvature(bufs[(s + 1) % 2], win, out=kappa)
    bufs[0].copy_(start)
    g.replay()  # warm-up replay
    bufs[0].copy_(start)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    g.replay()
    torch.cuda.synchronize()
    gwall = time.perf_counter() - t0
    gfinal = bufs[steps % 2].cpu().numpy()
    band = np.abs(phi_np) < 3
    err = np.abs(final - phi_np)[band]
    interior = (n - 4) ** 2
    return dict(n=n, steps=steps, interp=interp, backtrace=bt, updates_per_s=interior * steps / wall,
                ms_per_step=wall / steps * 1e3, graph_updates_per_s=interior * steps / gwall,
                graph_ms_per_step=gwall / steps * 1e3, graph_equals_eager=bool(np.array_equal(gfinal, final)),
                area_initial=float((phi_np < 0).sum() * h * h),
                area_final=float((final < 0).sum() * h * h), max_abs_err_band=float(err.max()),
                mean_abs_err_band=float(err.mean()))


def irregular_run(n=256, steps=20, width=6.0):
    """The irregular-node list (LevelSetUpdateIrregularNodeList) of the
    headline sphere (n^3 FP64, r = 0.15) over its band |phi| < width cells
    (the reference's default band width, LevelSetMethod/LevelSet.c:8): band
    nodes/s and records/s of one dcl_irregular_nodes call (count, scan and
    write kernels and the count's return to the host), device-resident; then
    the surface quantities (dcl_surface_quantities) at every record."""
    import torch
    import dcell_amd as D
    from dcell_amd import configs
    from dcell_amd import metrics as M

    win = D.GridWindow.unit_box(n)
    phi = torch.from_numpy(configs.sphere_sdf(n)).to("cuda:0", torch.float64)
    band, nb = M.band_compact(phi, win, width)  # phi is in grid units
    band = band.contiguous()
    out = torch.empty(10 * nb * M.IRREGULAR_DTYPE.itemsize, dtype=torch.uint8, device="cuda:0")
    count = M.irregular_nodes_into(phi, win, band, out)  # warm-up
    torch.cuda.synchronize()
    times = []
    for _ in range(steps):
        t0 = time.perf_counter()
        M.irregular_nodes_into(phi, win, band, out)
        times.append(time.perf_counter() - t0)
    t = float(np.median(times))
    # the surface quantities (IIMUpdateSurfaceQuantities_3D) at every record
    surf = torch.empty(max(count, 1) * M.SURFACE_DTYPE.itemsize, dtype=torch.uint8, device="cuda:0")
    M.surface_quantities_into(phi, win, out, count, surf)
    torch.cuda.synchronize()
    stimes = []
    for _ in range(steps):
        t0 = time.perf_counter()
        M.surface_quantities_into(phi, win, out, count, surf)
        torch.cuda.synchronize()
        stimes.append(time.perf_counter() - t0)
    ts = float(np.median(stimes))
    return dict(n=n, band_nodes=nb, records=count, ms_per_call=t * 1e3, band_nodes_per_s=nb / t,
                records_per_s=count / t, band_width_cells=width,
                surface_quantities=dict(ms_per_call=ts * 1e3, records_per_s=count / ts))


def get_velocity_run(n=256, steps=20):
    """LevelSetGetVelocity from a device-resident fluid field: the (u, v, w)
    window of an n^3 level set out of an n^3 FP64 fluid array of (p, u, v, w)
    records, per node and as component planes.  Kernel time from HIP events on
    the current stream; 48 B per node (24 read, 24 written) against the HBM
    peak."""
    import torch
    import dcell_amd as D

    fluid = torch.rand((n, n, n, 4), dtype=torch.float64, device="cuda:0")
    win = D.GridWindow((0, 0, 0), (n, n, n), (1.0 / (n - 1),) * 3)
    res = dict(n=n, bytes_per_node=48)
    for planes in (False, True):
        out = torch.empty((3, n, n, n) if planes else (n, n, n, 3), dtype=torch.float64, device="cuda:0")
        D.get_velocity(fluid, win, out=out, planes=planes)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(steps):
            D.get_velocity(fluid, win, out=out, planes=planes)
        e1.record()
        torch.cuda.synchronize()
        t = e0.elapsed_time(e1) * 1e-3 / steps
        gbs = n ** 3 * 48 / t / 1e9
        r = dict(ms_per_call=t * 1e3, nodes_per_s=n ** 3 / t, achieved_gbs=gbs, roofline_frac=gbs * 1e9 / HBM_PEAK)
        if planes:
            res["planes"] = r
        else:
            res.update(r)
    return res


def band_mode_run(n=256, steps=30, warmup=3):
    """LevelSetAdvectAndReinit in band mode with the level set resident in HBM
    (dcell_amd.advect.DeviceLevelSet): the n^3 sphere in the deformation
    field, dt = h (cos(pi t / 3) folded in), Euler + trilinear SL over the
    6-cell band, the CFL count and the irregular-node list every step, and
    the parallel reinitialisation whenever the CFL count passes 3 (the
    reference's threshold).  Wall time per step over the timed steps."""
    import torch
    import dcell_amd as D
    from dcell_amd import configs
    from dcell_amd.advect import DeviceLevelSet

    win = D.GridWindow.unit_box(n)
    vel = torch.from_numpy(configs.deformation_velocity(n)).to("cuda:0")
    phi = torch.from_numpy(configs.sphere_sdf(n)).to("cuda:0")
    ls = DeviceLevelSet(phi, win, reinit="parallel")
    ls.update_irregular_nodes()
    for k in range(warmup):
        ls.advect_and_reinit(vel, configs.deformation_dt(n, k))
    torch.cuda.synchronize()
    r0, band_nodes = ls.reinits, 0
    t0 = time.perf_counter()
    for k in range(warmup, warmup + steps):
        band_nodes += ls.band.shape[0]
        ls.advect_and_reinit(vel, configs.deformation_dt(n, k))
    torch.cuda.synchronize()
    t = (time.perf_counter() - t0) / steps
    return dict(n=n, steps=steps, ms_per_step=t * 1e3, reinits=ls.reinits - r0,
                band_nodes_per_step=band_nodes / steps, band_updates_per_s=band_nodes / steps / t)


def formats_run(n=256, steps=20):
    """The output formats from device-resident data: the .Real64 bytes of an
    n^3 FP64 grid (16 B per value: 8 read, 8 written) with HIP events on the
    current stream against the HBM peak, the whole grid_write (pack, one copy
    to the host, the file write), and the node-list rows of the n^3 sphere's
    6-cell band (23 doubles per record; the rows call synchronises to return
    its row count, so its event time holds that wait as well as the kernel)."""
    import tempfile
    import torch
    import dcell_amd as D
    from dcell_amd import configs
    from dcell_amd import formats as F
    from dcell_amd import metrics as M

    win = D.GridWindow.unit_box(n)
    phi = torch.from_numpy(configs.sphere_sdf(n)).to("cuda:0")
    out = torch.empty(8 + 8 * n ** 3, dtype=torch.uint8, device="cuda:0")
    F.real64_pack(phi, out)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(steps):
        F.real64_pack(phi, out)
    e1.record()
    torch.cuda.synchronize()
    t = e0.elapsed_time(e1) * 1e-3 / steps
    gbs = n ** 3 * 16 / t / 1e9
    with tempfile.TemporaryDirectory() as d:
        F.grid_write(d, "phi", 0, phi, win)
        t0 = time.perf_counter()
        F.grid_write(d, "phi", 1, phi, win)
        t_write = time.perf_counter() - t0
    band, nb = M.band_compact(phi, win, 6.0)
    recs = torch.empty(10 * nb * M.IRREGULAR_DTYPE.itemsize, dtype=torch.uint8, device="cuda:0")
    count = M.irregular_nodes_into(phi, win, band.contiguous(), recs)
    surf = torch.empty(count * M.SURFACE_DTYPE.itemsize, dtype=torch.uint8, device="cuda:0")
    M.surface_quantities_into(phi, win, recs, count, surf)
    F.irregular_node_rows(recs, count, surf, 3)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    e0.record()
    for _ in range(steps):
        F.irregular_node_rows(recs, count, surf, 3)
    e1.record()
    torch.cuda.synchronize()
    t_rows = (time.perf_counter() - t0) / steps
    t_rows_dev = e0.elapsed_time(e1) * 1e-3 / steps
    row_bytes = M.IRREGULAR_DTYPE.itemsize + M.SURFACE_DTYPE.itemsize + 23 * 8  # read 88 + 104, write 184
    return dict(real64=dict(n=n, ms_per_call=t * 1e3, values_per_s=n ** 3 / t, achieved_gbs=gbs,
                            roofline_frac=gbs * 1e9 / HBM_PEAK, bytes_per_value=16),
                grid_write_ms=t_write * 1e3,
                irregular_rows=dict(records=count, ms_per_call=t_rows * 1e3, records_per_s=count / t_rows,
                                    device_ms_per_call=t_rows_dev * 1e3, bytes_per_record=row_bytes,
                                    achieved_gbs=count * row_bytes / t_rows_dev / 1e9))


def _reinit_inputs(n, scale, center, radius, torch, D, M, dev="cuda:0"):
    """phi = scale x the distance (grid units) to a sphere in an n^3 window,
    its 8-cell band and irregular-node records, all on the device."""
    from dcell_amd import configs
    win = D.GridWindow((0, 0, 0), (n, n, n), (1.0, 1.0, 1.0))
    phi = torch.from_numpy(scale * configs.sphere_sdf(n, center, radius)).to(dev, torch.float64)
    band, nb = M.band_compact(phi, win, 8.0)
    recs = torch.empty(max(10 * nb, 1) * M.IRREGULAR_DTYPE.itemsize, dtype=torch.uint8, device=dev)
    count = M.irregular_nodes_into(phi, win, band.contiguous(), recs)
    return win, phi, recs, count


def reinit_cells(ncells=64, seed=4):
    """The batch of the reinitialisation form: ncells spheres with edges
    log-uniform in [32, 96] (seed 4), radius n/4, centres jittered, phi a
    1.5x scaled distance."""
    rng = np.random.default_rng(seed)
    edges = np.round(np.exp(rng.uniform(np.log(32), np.log(96), ncells))).astype(int)
    jit = rng.uniform(-0.02, 0.02, (ncells, 3))
    return [(int(e), (0.5 + j[0], 0.5 + j[1], 0.5 + j[2])) for e, j in zip(edges, jit)]


def reinit_run(n=256, runs=5, exact=True):
    """LevelSetReinitialize on the GPU: the 256^3 sphere of the headline (r
    0.15, phi a 1.5x scaled distance) as one level set (two waves, one per
    sign), and a batch of 64 level sets (128 waves).  Median wall time of one
    dcl_reinitialize call; inputs (phi, records) resident on the device."""
    import torch
    import dcell_amd as D
    from dcell_amd import metrics as M

    def timed(cells_in, parallel=False):
        saved = [c[1].clone() for c in cells_in]
        ts, bands, sweeps = [], None, None
        for _ in range(runs + 1):
            for c, s in zip(cells_in, saved):
                c[1].copy_(s)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            args = [(c[1], c[0], c[2], c[3]) for c in cells_in]
            if parallel:
                bands, sweeps = M.reinitialize_parallel(args)
            else:
                bands = M.reinitialize(args)
            ts.append(time.perf_counter() - t0)
        phis = [c[1].clone() for c in cells_in]
        for c, s in zip(cells_in, saved):
            c[1].copy_(s)
        return float(np.median(ts[1:])), sum(int(b.shape[0]) for b in bands), phis, sweeps

    def agree(exact, par):
        """The parallel result against the exact one over the band nodes of both."""
        nb = eq = nx = 0
        dmax = 0.0
        for a, b in zip(exact, par):
            ma, mb = a.abs() < 7.5, b.abs() < 7.5
            m = ma & mb
            nb += int(m.sum())
            eq += int((a[m] == b[m]).sum())
            nx += int((ma ^ mb).sum())
            if m.any():
                dmax = max(dmax, float((a[m] - b[m]).abs().max()))
        return dict(bit_equal_frac=eq / max(nb, 1), max_abs_diff=dmax, band_nodes_in_one_only=nx)

    single = [_reinit_inputs(n, 1.5, (0.35, 0.35, 0.35), 0.15, torch, D, M)]
    batch = [_reinit_inputs(e, 1.5, c, 0.25, torch, D, M) for e, c in reinit_cells()]
    p1, pb1, q1, sw1 = timed(single, True)
    p2, pb2, q2, sw2 = timed(batch, True)
    res = dict(parallel_single=dict(n=n, band_nodes=pb1, ms_per_call=p1 * 1e3, band_nodes_per_s=pb1 / p1,
                                    sweeps=sw1[0]),
               parallel_batch=dict(cells=len(batch), band_nodes=pb2, ms_per_call=p2 * 1e3,
                                   band_nodes_per_s=pb2 / p2, sweeps_max=max(sw2)))
    if exact:
        t1, nb1, e1, _ = timed(single)
        t2, nb2, e2, _ = timed(batch)
        res["single"] = dict(n=n, band_nodes=nb1, ms_per_call=t1 * 1e3, band_nodes_per_s=nb1 / t1)
        res["batch"] = dict(cells=len(batch), band_nodes=nb2, ms_per_call=t2 * 1e3, band_nodes_per_s=nb2 / t2)
        res["parallel_single"]["vs_exact"] = agree(e1, q1)
        res["parallel_batch"]["vs_exact"] = agree(e2, q2)
    return res


def dropin_run(n=256, calls=10):
    """The C drop-in as DCell calls it: LevelSetAdvectSL on an n^3 FP64 level
    set (the headline sphere) over its 6-cell band, with a full n^3 velocity
    grid in host memory.  Each call uploads the boxes of phi0 and of the
    velocity grid that the band's departures reach, advects the band nodes on
    the GPU and writes them back into the host grid, so this is the
    PCIe-inclusive rate (bytes moved per call from the drop-in's own
    counters); the same band update through the C ABI with everything
    resident in HBM is timed beside it."""
    import ctypes as C
    import torch
    import dcell_amd as D
    from dcell_amd import configs
    from dcell_amd import dropin as DI
    from dcell_amd import metrics as M

    L = DI.lib()
    h = 1.0 / (n - 1)
    ls = DI.create_levelset((h, h, h), (0, 0, 0), (n, n, n))
    phi = DI.grid_array(ls.contents.phi)
    phi[:] = configs.sphere_sdf(n)
    win = D.GridWindow.unit_box(n)
    dphi = torch.from_numpy(phi.copy()).to("cuda:0")
    band, nb = M.band_compact(dphi, win, 6.0)
    band_np = band.cpu().numpy()
    DI.set_array(ls.contents.band, band_np, 12)
    vg = DI.create_grid((h, h, h), (0, 0, 0), (n, n, n), 3)
    vel = DI.grid_array(vg)
    vel[:] = configs.deformation_velocity(n)
    dt = configs.deformation_dt(n, 0)
    DI.check(L.LevelSetAdvectSL(ls, vg, C.c_double(dt)), "LevelSetAdvectSL")  # warm-up
    up0, down0 = C.c_ulonglong(), C.c_ulonglong()
    up1, down1 = C.c_ulonglong(), C.c_ulonglong()
    DI.check(L.LevelSetGPUTransferStats(C.byref(up0), C.byref(down0)), "LevelSetGPUTransferStats")
    ts = []
    for _ in range(calls):
        t0 = time.perf_counter()
        DI.check(L.LevelSetAdvectSL(ls, vg, C.c_double(dt)), "LevelSetAdvectSL")
        ts.append(time.perf_counter() - t0)
    DI.check(L.LevelSetGPUTransferStats(C.byref(up1), C.byref(down1)), "LevelSetGPUTransferStats")
    t_drop = float(np.median(ts))
    # device-resident: the same band update through the C ABI
    dvel = torch.from_numpy(vel.copy()).to("cuda:0")
    out = dphi.clone()
    D.advect_sl(dphi, dvel, win, dt, band=band, out=out)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(calls):
        D.advect_sl(dphi, dvel, win, dt, band=band, out=out)
    e1.record()
    torch.cuda.synchronize()
    t_dev = e0.elapsed_time(e1) * 1e-3 / calls
    L.GridDestroy(vg)
    L.LevelSetDestroy(ls)
    return dict(n=n, band_nodes=nb, dropin_ms_per_call=t_drop * 1e3, dropin_band_updates_per_s=nb / t_drop,
                host_bytes_uploaded_per_call=(up1.value - up0.value) // calls,
                host_bytes_downloaded_per_call=(down1.value - down0.value) // calls,
                whole_grids_bytes=(n ** 3) * 8 * 4, host_threads=int(os.environ.get("DCELL_GPU_HOST_THREADS", 8)),
                device_resident_ms_per_call=t_dev * 1e3, device_resident_band_updates_per_s=nb / t_dev)


def copy_ceiling():
    """The measured HBM ceiling: dcl_stream_copy (the library's streaming copy
    kernel) of 1 GiB, best of 10, read + write bytes over its time, with the
    same for torch's device copy beside it; and the streaming copy's rate
    over working sets of 64 MiB to 1 GiB copied again and again (a copy whose
    source and destination fit the 256 MB last-level cache between HBM and
    the XCDs runs faster than HBM allows)."""
    import ctypes as C
    import torch
    from dcell_amd import _lib as L
    from dcell_amd.levelset import stream_handle
    src = torch.ones(1 << 28, dtype=torch.float32, device="cuda:0")
    dst = torch.empty_like(src)

    def best_of(fn, nbytes, reps=10):
        fn()
        best = 1e30
        for _ in range(reps):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            fn()
            e1.record()
            torch.cuda.synchronize()
            best = min(best, e0.elapsed_time(e1) * 1e-3)
        return 2 * nbytes / best

    def own_copy(nbytes):
        return lambda: L.call("dcl_stream_copy", C.c_void_p(dst.data_ptr()), C.c_void_p(src.data_ptr()),
                              nbytes, stream_handle())

    own = best_of(own_copy(1 << 30), 1 << 30)
    torch_copy = best_of(lambda: dst.copy_(src), 1 << 30)
    by_size = {f"{m} MiB": round(best_of(own_copy(m << 20), m << 20) / 1e9, 1) for m in (64, 128, 256, 512, 1024)}
    del src, dst
    torch.cuda.empty_cache()
    return own, torch_copy, by_size


def _f64():
    import torch
    return torch.float64


def _f32():
    import torch
    return torch.float32


FORMS = {
    "batched_fp32_config_scheme": lambda: batch_run(512, 4, 100, 3, "cubic", "rk2"),
    "batched_fp32_reference_scheme": lambda: batch_run(512, 4, 100, 3, "linear", "euler"),
    "batched_2d_fp32_config_scheme": lambda: batch_run(512, 4, 100, 3, "cubic", "rk2", dim=2, lo=64, hi=512),
    "batched_2d_fp64_config_scheme": lambda: batch_run(512, 4, 100, 3, "cubic", "rk2", dim=2, lo=64, hi=512,
                                                       dtype=_f64()),
    "precision_pls_384": lambda: pls_run(384, 150),
    "adjoint_128": lambda: adjoint_run(128, 50, 10),
    "fp64_reference_scheme_256": lambda: dense_run(256, _f64(), "linear", "euler", 50, 2),
    "fp64_config_scheme_256_per_node_velocity": lambda: dense_run(256, _f64(), "cubic", "rk2", 50, 2),
    "fp32_config_scheme_384": lambda: dense_run(384, _f32(), "cubic", "rk2", 50, 2),
    "fp32_reference_scheme_384": lambda: dense_run(384, _f32(), "linear", "euler", 50, 2),
    "zalesak_2d_128_fp64_config_scheme": lambda: zalesak_run(128, 628, "cubic", "rk2"),
    "zalesak_2d_128_fp64_reference_scheme": lambda: zalesak_run(128, 628, "linear", "euler"),
    "irregular_nodes_256_fp64": lambda: irregular_run(256, 20),
    "get_velocity_256_fp64": lambda: get_velocity_run(256, 20),
    "reinitialize_fp64": lambda: reinit_run(256, 5),
    "reinitialize_parallel_fp64": lambda: reinit_run(256, 5, exact=False),
    "formats_256_fp64": lambda: formats_run(256, 20),
    "band_mode_256_fp64": lambda: band_mode_run(256, 30),
    "dropin_band_256_fp64": lambda: dropin_run(256, 10),
}


# forms run on their own only (the full line already holds their figures)
FORM_ONLY = {"reinitialize_parallel_fp64"}


def forms(args):
    out = {}
    for name, run in FORMS.items():
        if name in FORM_ONLY:
            continue
        t0 = time.perf_counter()
        try:
            out[name] = run()
        except Exception as e:  # a failing form is reported, the headline line still prints
            out[name] = {"error": f"{type(e).__name__}: {e}"}
        print(f"form {name}: {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU: --gpus {args.gpus} is not supported "
              "(the workload is one GPU's share; a multi-GPU run would give each GPU whole cells "
              "or a z-slab of one grid, which this build does not implement)",
              file=sys.stderr)
        sys.exit(2)
    if args.traffic_summary:
        traffic_summary(args)
        return
    if args.counter_summary:
        counter_summary(args)
        return
    if args.probe:
        probe(args)
        return
    if args.form:
        print(json.dumps({args.form: FORMS[args.form]()}), flush=True)
        return
    import torch
    r = dense_run(args.n, torch.float64, "cubic", "rk2", args.steps, args.warmup, planes=True)
    interior = (args.n - 4) ** 3
    B = b_alg(8)
    achieved = interior * B["sl"] / (r["sl_ms"] * 1e-3)
    kname, traffic = read_traffic(args.n)
    line = {
        "metric": "grid-point updates/s (SL advect+curvature), % HBM roofline, FP32 vs FP64",
        "value": r["updates_per_s"],
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
        "config": {"workload": f"3D deformation, {args.n}^3 FP64 sphere r=0.15, RK2 backtrace + "
                               "Catmull-Rom tricubic SL + 19-point curvature, dt=h",
                   "grid": [args.n] * 3, "interior_updates_per_step": interior,
                   "velocity_layout": "component planes [c][z][y][x] (dcl_get_velocity_planes)",
                   "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "kernel": kname or "dense SL (RK2 + tricubic, FP64)",
                     "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK, "traffic": traffic,
                     "bytes_per_launch": interior * B["sl"],
                     "avg_launch_ms": r["sl_ms"]},
        "step_roofline_frac": r["step_roofline_frac"],
        "measured_copy_gbs": None,
        "kernels_ms": {"sl_advect": r["sl_ms"], "curvature": r["curv_ms"]},
        "curvature_gbs": r["curv_gbs"],
    }
    own, torch_copy, by_size = copy_ceiling()
    line["measured_copy_gbs"] = own / 1e9
    line["torch_copy_gbs"] = torch_copy / 1e9
    line["copy_gbs_by_working_set"] = by_size
    line["roofline"]["frac_of_measured_copy"] = achieved / own
    line["roofline"]["limiter"] = read_limiter()
    if not args.no_forms:
        line["forms"] = forms(args)
    if not args.no_cpu:
        try:
            line["cpu_baseline"] = cpu_baseline(args)
        except Exception as e:  # reported in the line; the GPU figures still print
            line["cpu_baseline"] = {"error": f"{type(e).__name__}: {e}"}
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
