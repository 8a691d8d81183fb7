"""Benchmark of the heat-operator hot path on one MI355X.

Workload (one "step"): one FP64 application of the heat tangent
J = rho_cp/dt M + K_k on the unit cube, 64^3 hex cells, FE_Q(4) on
Gauss-Lobatto points with QGauss(5) (N = 257^3 = 16,974,593 DoFs), homogeneous
Dirichlet on all faces, k(x) = 1 + 0.5 sin(2 pi x) sin(2 pi y) sin(2 pi z) at
the quadrature points, rho_cp = 1, dt = 1e-3; input uniform[-1,1] on the free
DoFs from SplitMix64 seed 0. Inputs are resident in HBM before timing.

Printed: one JSON line with value = GDoF/s (N / time per vmult), the
roofline of the vmult kernel, the CPU baseline (the oracle's C++ restatement
timed on this host), and the CG + Jacobi time-to-1e-8 of the FP64 and the
FP32-mixed solver on the same operator.
"""
import argparse
import json
import statistics
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

N_CELLS = 64
DEGREE = 4
DT = 1e-3
K_AMP = 0.5
INNER_REL_TOLERANCE = 1e-4  # mixed CG: reduction asked of each inner FP32 solve
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6  # MI355X FP64 vector (VALU) peak; its FP64 matrix peak is the same
CLOCK_SETTLE_S = 0.5


def algorithmic_counts(n_cells, degree, bytes_per=8):
    """Bytes and flops of one vmult (DoF vectors read + written once; analytic
    coefficients and the implicit lattice index map cost no bytes; flops by the
    even-odd sum-factorisation count 12 n_q^4 + 36 n_q^3 per cell)."""
    n1 = n_cells * degree + 1
    N = n1**3
    nq = degree + 1
    cells = n_cells**3
    return {"N": N, "N_free": (n1 - 2) ** 3, "cells": cells, "bytes": 2 * bytes_per * N,
            "flops": cells * (12 * nq**4 + 36 * nq**3)}


def host_cpu_info():
    """What the CPU leg ran on: logical CPUs, the ones this process may use
    (affinity, cgroup quota, OMP_NUM_THREADS), model name and sockets (lscpu)."""
    import subprocess
    info = {"nproc": os.cpu_count()}
    try:
        info["affinity"] = len(os.sched_getaffinity(0))
    except AttributeError:
        info["affinity"] = os.cpu_count()
    quota = None
    try:
        with open("/sys/fs/cgroup/cpu.max") as f:
            q, per = f.read().split()[:2]
            if q != "max":
                quota = max(1, int(int(q) / int(per)))
    except (OSError, ValueError):
        pass
    info["cgroup_cpu_quota"] = quota
    info["omp_num_threads"] = os.environ.get("OMP_NUM_THREADS")
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            k, _, v = line.partition(":")
            if k.strip() == "Model name":
                info["model"] = v.strip()
            elif k.strip() == "Socket(s)":
                info["sockets"] = v.strip()
            elif k.strip() == "Core(s) per socket":
                info["cores_per_socket"] = v.strip()
    except (OSError, subprocess.SubprocessError):
        pass
    usable = [info["affinity"] or 1]
    if quota:
        usable.append(quota)
    if info["omp_num_threads"] and info["omp_num_threads"].isdigit():
        usable.append(int(info["omp_num_threads"]))
    info["threads_used"] = min(usable)
    return info


def cpu_baseline(threads=None, budget_s=10.0, min_samples=20):
    """Times the oracle's cell-wise sum-factorised FP64 vmult on the host cores
    (the reference cannot be built here: it needs deal.II + p4est + Trilinos +
    MPI): median of at least `min_samples` vmults after 3 warm-ups, sampling
    for about `budget_s` seconds, on every CPU this process may use (the
    smallest of its affinity set, its cgroup quota and OMP_NUM_THREADS); also
    one core, 3 vmults."""
    from oracle import cpu
    host = host_cpu_info()
    threads = threads or host["threads_used"]
    pb = cpu.Problem(N_CELLS, DEGREE, dt=DT, k_amplitude=K_AMP)
    u = pb.random(0)

    def sample(nthreads, warm, least, budget):
        cpu.set_threads(nthreads)
        for _ in range(warm):
            pb.vmult(u)
        ts = []
        t_start = time.perf_counter()
        while len(ts) < least or time.perf_counter() - t_start < budget:
            t0 = time.perf_counter()
            pb.vmult(u)
            ts.append(time.perf_counter() - t0)
            if len(ts) >= 10 * least:
                break
        return float(np.median(ts)), len(ts)

    t, k = sample(threads, 3, min_samples, budget_s)
    t1, k1 = sample(1, 0, 3, 0.0)
    cpu.set_threads(threads)
    return {"value": round(pb.N / t / 1e9, 4), "unit": "GDoF/s", "cores": threads, "kind": "port",
            "label": "CPU restatement, not deal.II",
            "sample": f"median of {k} FP64 vmults of the 64^3 Q4 operator after 3 warm-ups "
                      f"(C++ restatement of the reference cell loop, not deal.II; OpenMP {threads} "
                      f"threads), {t*1e3:.1f} ms each; 1 core: {t1*1e3:.0f} ms per vmult "
                      f"({pb.N / t1 / 1e9:.4f} GDoF/s, {k1} vmults)",
            "one_core_value": round(pb.N / t1 / 1e9, 5),
            "host": host}


def line_kernel_name(P, fp32, coef, vel, mf=False):
    """rocprof name of the line-form vmult kernel instantiated for degree P:
    the launch shape of the library's VShape (spread form from degree 4 without
    MFMA: the CPB cells of a 4-wavefront workgroup, one line per thread; else
    CW cells per wavefront in R lane slots) and its register budget."""
    spread = not mf and P >= 4
    nl = (P + 1) ** 2
    if mf:
        cww = {4: 2, 5: 3, 6: 1}.get(P, 4)
    else:
        cww = {1: 16, 2: 7, 3: 4}.get(P, 1)
    cw = (64 * 4) // nl if spread else cww
    r = 0 if spread else (cww * nl + 63) // 64
    lean = coef != 2 and not vel
    if lean and fp32 and not mf and P == 4:
        minw = 6
    elif (P >= 5 and not fp32) or (mf and P == 5):
        minw = 2
    else:
        minw = 4
    b = lambda v: "true" if v else "false"
    return (f"vmult_kernel<{P}, {'float' if fp32 else 'double'}, {coef}, {b(vel)}, {cw}, {r}, 4, "
            f"{minw}, {b(mf)}>")


def kernel_label(tuning):
    """The line kernel of the 64^3 Q4 FP64 workload (separable-sine k, no
    velocity) and its launches per vmult under the installed schedule."""
    k = line_kernel_name(DEGREE, False, 1, False)
    if tuning["variant"] != "line":
        k = f"vmult ({tuning['variant']} form)"
    sched = tuning["schedule"]
    if sched == "slabs":
        return k + " x 16 (8 colour passes x 2 z-slabs on two streams)", 16
    if sched == "ordered":
        return k.replace("vmult_kernel", "vmult_ordered_kernel") + " x 1 (ordered)", 1
    return k + " x 8 colour passes", 8


def measured_traffic(kernel_name):
    """HBM bytes per colour pass of the vmult kernel from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE summary, if it was taken on this kernel and
    schedule."""
    path = os.path.join(ROOT, "profiles", "vmult_q4_64_fp64_traffic.json")
    try:
        with open(path) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    if t.get("kernel") != kernel_name:
        return None
    return t.get("bytes_per_pass")


def copy_calibration(args, torch, dev):
    """Plain copy of one FP64 DoF vector of the workload (read N*8 B, write N*8 B):
    the known-size kernel that scales the FETCH_SIZE / WRITE_SIZE counters."""
    N = algorithmic_counts(N_CELLS, DEGREE)["N"]
    src = torch.rand(N, dtype=torch.float64, device=dev)
    dst = torch.empty_like(src)
    for _ in range(args.warmup):
        dst.copy_(src)
    torch.cuda.synchronize()
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    ev0.record()
    for _ in range(args.steps):
        dst.copy_(src)
    ev1.record()
    torch.cuda.synchronize()
    ms = ev0.elapsed_time(ev1) / args.steps
    print(json.dumps({"mode": "copy", "bytes_read": 8 * N, "bytes_written": 8 * N, "ms": ms,
                      "GBs": 16 * N / (ms * 1e-3) / 1e9}))


def time_solves(op, b, torch, LinearSolver, LinearSolverData, repeats=5):
    """CG + Jacobi to 1e-8 in FP64 and FP32-mixed: one untimed warm-up solve of
    each (workspace allocation), then `repeats` timed solves of each,
    alternating FP64 and mixed so that both see the same clocks; the median
    of each is reported (the samples beside it). The mixed solve asks each
    inner FP32 CG for a 1e-4 reduction: two outer FP64 corrections reach 1e-8
    (277 inner iterations, against four corrections and 286 iterations at the
    1e-2 default); the stopping test is the FP64 residual either way."""
    runs = (("fp64", False), ("fp32_mixed", True))
    data = {name: LinearSolverData(solver_type="CG", preconditioner_type="Diagonal",
                                   rel_tolerance=1e-8, mixed_precision=mixed,
                                   inner_rel_tolerance=INNER_REL_TOLERANCE)
            for name, mixed in runs}
    x = torch.zeros_like(b)
    for name, _ in runs:
        x.zero_()
        LinearSolver.solve(op, x, b, data[name])
    samples = {name: [] for name, _ in runs}
    solve = {}
    for rep in range(repeats):
        for name, _ in runs:
            x.zero_()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            LinearSolver.solve(op, x, b, data[name])
            torch.cuda.synchronize()
            samples[name].append(time.perf_counter() - t0)
            r = LinearSolver.last_report
            solve[name] = {"iterations": r["iterations"], "outer_iterations": r["outer_iterations"],
                           "rel_residual": r["final_residual"] / r["initial_residual"]}
    for name, _ in runs:
        solve[name]["seconds"] = round(statistics.median(samples[name]), 5)
        solve[name]["samples_s"] = [round(t, 5) for t in samples[name]]
    solve["fp32_mixed"]["speedup_vs_fp64"] = round(solve["fp64"]["seconds"] /
                                                   solve["fp32_mixed"]["seconds"], 3)
    solve["fp32_mixed"]["inner_rel_tolerance"] = INNER_REL_TOLERANCE
    return solve


def solve_legs(args, torch, dev):
    from meltpooldg_amd import HeatDiffuseOperator, LinearSolver, LinearSolverData
    from meltpooldg_amd.inputs import random_vector
    op = HeatDiffuseOperator(N_CELLS, DEGREE, conductivity_amplitude=K_AMP)
    op.reset_time_increment(DT)
    b = torch.tensor(random_vector(N_CELLS, DEGREE, 1), dtype=torch.float64, device=dev)
    print(json.dumps({"mode": "solve", "tuning": {"fp64": op.get_tuning("fp64"),
                                                  "fp32": op.get_tuning("fp32")},
                      "solve_time_to_1e-8": time_solves(op, b, torch, LinearSolver,
                                                        LinearSolverData)}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    # a vmult is 0.25 ms: 100 untimed ones (25 ms) bring the GPU to its
    # steady clocks; with 5 the first timed vmults ran about 10 % slower
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-solve", action="store_true", help="skip the CG time-to-solution legs")
    ap.add_argument("--tune", action="store_true",
                    help="autotune the operator (FP64 and FP32) before timing instead of using the "
                         "per-degree default tuning (itself the autotuner's pick on this workload)")
    ap.add_argument("--tuning", default=None,
                    help="fixed FP64 tuning z_layers,blocks_per_cu[,variant,tile_log2,schedule]")
    ap.add_argument("--mode", choices=["vmult", "copy", "solve", "robin", "solidification",
                                       "timestep", "q2"],
                    default="vmult",
                    help="copy: time a plain device copy of one DoF vector (HBM counter calibration); "
                         "solve: only the FP64 and FP32-mixed CG solves (solver profiling); "
                         "robin: the headline operator with a convecting and radiating face; "
                         "solidification: the tangent of a temperature-dependent material; "
                         "timestep: the reference's regression inputs, graphs vs eager; "
                         "q2: the degree-2 operators against the HBM and vector-ALU bounds")
    ap.add_argument("--cells", default="96,192",
                    help="q2 mode: comma-separated cells per direction (96,192 by default)")
    ap.add_argument("--config", type=int, default=1,
                    help="1: the headline line (default); 0, 2, 3, 4: the other workloads")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}); the multi-GPU design is "
              "described in DESIGN.md and is not built", file=sys.stderr)
        sys.exit(2)

    import torch
    from meltpooldg_amd import HeatDiffuseOperator, LinearSolver, LinearSolverData
    from meltpooldg_amd.inputs import random_vector

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    if args.mode == "copy":
        return copy_calibration(args, torch, dev)
    if args.mode == "solve":
        return solve_legs(args, torch, dev)
    if args.mode == "robin":
        import bench_configs
        return print(json.dumps(bench_configs.robin(args, torch)))
    if args.mode == "solidification":
        import bench_configs
        return print(json.dumps(bench_configs.solidification(args, torch)))
    if args.mode == "timestep":
        import bench_configs
        return print(json.dumps(bench_configs.timestep(args, torch)))
    if args.mode == "q2":
        import bench_configs
        return print(json.dumps(bench_configs.q2(args, torch)))
    if args.config != 1:
        import bench_configs
        return bench_configs.run(args.config, args)
    op = HeatDiffuseOperator(N_CELLS, DEGREE, conductivity_amplitude=K_AMP)
    op.reset_time_increment(DT)
    if args.tuning:
        f = args.tuning.split(",")
        op.set_tuning(z_layers=int(f[0]), blocks_per_cu=int(f[1]),
                      variant=f[2] if len(f) > 2 else None,
                      tile_log2=int(f[3]) if len(f) > 3 else None,
                      schedule=f[4] if len(f) > 4 else None, precision="fp64")
        tuning = op.get_tuning("fp64")
    elif args.tune:
        tuning = op.autotune("fp64", repeats=3)
    else:
        tuning = op.get_tuning("fp64")
    cnt = algorithmic_counts(N_CELLS, DEGREE)
    N = op.n_dofs()
    assert N == cnt["N"]
    src = torch.tensor(random_vector(N_CELLS, DEGREE, 0), dtype=torch.float64, device=dev)
    dst = torch.empty_like(src)

    # the library launches on the handle's stream: torch's current stream here
    stream = torch.cuda.current_stream()
    # clock settling, untimed and independent of --warmup: the GPU reaches its
    # steady clocks after some 100 ms of work (with only 5 vmults of warm-up
    # the first timed ones ran about 10 % slower)
    t_settle = time.perf_counter() + CLOCK_SETTLE_S
    while time.perf_counter() < t_settle:
        for _ in range(20):
            op.vmult(dst, src)
        torch.cuda.synchronize()
    for _ in range(args.warmup):
        op.vmult(dst, src)
    torch.cuda.synchronize()
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    ev0.record(stream)
    for _ in range(args.steps):
        op.vmult(dst, src)
    ev1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_event = ev0.elapsed_time(ev1) / args.steps
    ms_step = wall * 1e3 / args.steps
    gdofs = N / (ms_step * 1e-3) / 1e9

    # one vmult = 8 colour passes of the line kernel. With the passes schedule
    # a pass is one launch; with the slab schedule it is two launches, one per
    # z-slab on its own stream, running side by side. Per pass, algorithmic
    # flops and bytes are 1/8 of the vmult's and the time 1/8 of the event
    # time per vmult.
    t_k = ms_event * 1e-3
    ach_flops = cnt["flops"] / t_k / 1e12
    ach_bytes = cnt["bytes"] / t_k / 1e9
    kernel, launches = kernel_label(tuning)
    roofline = {"bound": "valu_fp64", "achieved": round(ach_flops, 3), "peak": FP64_PEAK_TFLOPS,
                "unit": "TFLOP/s", "frac": round(ach_flops / FP64_PEAK_TFLOPS, 4),
                "traffic": measured_traffic(kernel),
                "kernel": kernel, "launches_per_vmult": launches,
                "pass_ms": round(ms_event / 8, 5),
                "algorithmic_bytes_per_pass": cnt["bytes"] // 8,
                "algorithmic_flops_per_pass": cnt["flops"] // 8,
                "hbm_achieved_GBs": round(ach_bytes, 1),
                "hbm_frac": round(ach_bytes / HBM_PEAK_GBS, 4),
                "tuning": tuning,
                "note": "FP64 vector-FMA (VALU) kernel: peak is MI355X's FP64 vector rate. "
                        "achieved = flops of one vmult / its event time (per pass: 1/8 of both). "
                        "With the slab schedule a pass is two launches side by side, one per "
                        "z-half; a rocprof launch duration is one of them, so 16 launch "
                        "durations add up to more than the vmult's time"}

    solve = None
    if not args.no_solve:
        b = torch.tensor(random_vector(N_CELLS, DEGREE, 1), dtype=torch.float64, device=dev)
        fp32_tuning = op.autotune("fp32", repeats=3) if args.tune else op.get_tuning("fp32")
        solve = time_solves(op, b, torch, LinearSolver, LinearSolverData)
        solve["fp32_tuning"] = fp32_tuning

    # the degree-2 FP32 operator of the low-degree HBM target, 96^3 cells
    # (bench.py --mode q2 has the FP64 rows and 192^3)
    q2 = None
    if not args.no_solve:
        import bench_configs
        del op
        q2 = [{k: r[k] for k in ("k", "GDoF/s", "algorithmic_GB/s", "hbm_frac", "valu_frac")}
              for r in bench_configs.q2(argparse.Namespace(cells="96"), torch)["rows"]
              if r["precision"] == "fp32"]

    cpu = None if args.no_cpu else cpu_baseline()

    line = {
        "metric": "GDoF/s heat vmult (Q4, 64^3); CG time-to-1e-8: FP32-mixed vs FP64, 1 GPU",
        "value": round(gdofs, 3),
        "unit": "GDoF/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (SplitMix64 seed 0 uniform[-1,1] on free DoFs)",
        "config": {"workload": "heat vmult, unit cube 64^3 hex cells, FE_Q(4) GLL / QGauss(5), "
                               "k(x)=1+0.5 sin sin sin, rho_cp=1, dt=1e-3, Dirichlet all faces",
                   "n_dofs": N, "n_free_dofs": cnt["N_free"], "cells": cnt["cells"],
                   "parallelism": "single GPU"},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "solve_time_to_1e-8": solve,
        "q2_fp32_96": q2,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
