"""Batched heterogeneous heat solves (bench.py --config 3).

64 independent solves on unit cubes of n^3 cells, n ~ U{8..48}, degree
p ~ U{1..4}, conductivity k = 10^U(-1, 1) (numpy default_rng, master seed 0),
rho_cp = 1, dt = 1e-3, Dirichlet on all faces, RHS uniform[-1,1] from
SplitMix64 seed 1000 + s. FP32 cell loop inside FP64 defect-correction CG +
Jacobi to a relative residual of 1e-8, all members solved together; the same
members solved one after another with the single-operator solver are timed
beside it.
"""
import time

import numpy as np


def members(count=64, seed=0):
    from meltpooldg_amd import MaterialData, PhaseData
    from meltpooldg_amd.inputs import batch_members
    return [{"n_cells": m["n_cells"], "degree": m["degree"],
             "material": MaterialData(gas=PhaseData(1.0, 1.0, m["k"]))}
            for m in batch_members(count, seed)]


def run(args, torch):
    from meltpooldg_amd import HeatBatch, HeatDiffuseOperator, LinearSolver, LinearSolverData
    from meltpooldg_amd.inputs import random_vector
    ms = members()
    batch = HeatBatch(ms)
    batch.reset_time_increment(1e-3)
    bs = [torch.tensor(random_vector(m["n_cells"], m["degree"], 1000 + i), dtype=torch.float64,
                       device="cuda") for i, m in enumerate(ms)]
    N = [b.numel() for b in bs]
    data = LinearSolverData(solver_type="CG", preconditioner_type="Diagonal", rel_tolerance=1e-8,
                            mixed_precision=True)
    xs = [torch.zeros_like(b) for b in bs]
    batch.solve(xs, bs, data)  # warm-up (graph capture, workspace)
    xs = [torch.zeros_like(b) for b in bs]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    reps = batch.solve(xs, bs, data)
    torch.cuda.synchronize()
    t_batch = time.perf_counter() - t0
    done_t = np.array([r["seconds"] for r in reps])
    dof_it = sum(n * r["iterations"] for n, r in zip(N, reps))

    # the same members one after another (single-operator mixed solver)
    t_seq = 0.0
    its_seq = []
    for i, m in enumerate(ms):
        op = HeatDiffuseOperator(m["n_cells"], m["degree"], material=m["material"])
        op.reset_time_increment(1e-3)
        x = torch.zeros_like(bs[i])
        if i == 0:
            LinearSolver.solve(op, x, bs[i], data)
            x.zero_()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        LinearSolver.solve(op, x, bs[i], data)
        torch.cuda.synchronize()
        t_seq += time.perf_counter() - t1
        its_seq.append(LinearSolver.last_report["iterations"])
        # agreement with the batched solution
        if i < 8:
            rel = float((x - xs[i]).abs().max() / x.abs().max())
            assert rel < 1e-6, (i, rel)
    return {"config": 3, "workload": "64 solves, n~U{8..48}, p~U{1..4}, k~10^U(-1,1), FP32-mixed CG to 1e-8",
            "total_dofs": int(sum(N)), "batch_s": t_batch,
            "aggregate_GDoF_it_per_s": dof_it / t_batch / 1e9,
            "p50_completion_s": float(np.percentile(done_t, 50)),
            "p99_completion_s": float(np.percentile(done_t, 99)),
            "inner_iterations_total": int(sum(r["iterations"] for r in reps)),
            "all_converged": all(r["status"] == 0 for r in reps),
            "sequential_s": t_seq, "speedup_vs_sequential": t_seq / t_batch,
            "degrees": [m["degree"] for m in ms], "n_cells": [m["n_cells"] for m in ms]}
