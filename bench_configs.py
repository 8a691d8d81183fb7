"""Workload runners for the heat-operator configurations other than the
headline one (bench.py --config N). Each prints one JSON line with the GPU
measurements, the error against the oracle and, where bounded, the CPU
oracle's time on this host.

config 0: 16^3 Q2, gas (rho_cp = k = 1), dt = 1e-3, Dirichlet all faces: one
          vmult and CG + Jacobi to 1e-10 relative residual, FP64.
config 2: 96^3 Q3, two-phase sphere (radius 0.3, centre 0.5^3) from a nodal
          smoothed heaviside of width 2 cells, k 1:100 and rho*cp 1:10 (equal
          cp, liquid inside): FP32 cell loop in FP64 defect-correction CG to
          1e-8, errors against the oracle's FP64 solution and residual.
config 3: batched solves (see bench_batch.py).
config 4: 64^3 Q5 convection-diffusion, vortex velocity, dt = 5e-4,
          k = 1e-2: GMRES(30) + Jacobi, FP64 and FP32-mixed.
robin (bench.py --mode robin): the physics of the reference's
          unidirectional_heat_transfer Robin cases (steel: rho 7850, cp 460,
          k 55.563; dt = 0.01; T = 1000; convection alpha = 25 and radiation
          emissivity 0.98 to T_inf = 300) on the unit cube with 64^3 Q4 cells,
          the Robin face being the top face z = 1 and every other face
          insulated: the vmult with and without the face launches, the
          full-size vmult against the oracle, and one device Newton time step
          (GMRES + Jacobi, the heat defaults; Newton residual / correction
          tolerances 1e-7 / 1e-9 absolute).
"""
import json
import os
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))


def _t(a, torch, dtype=None):
    return torch.tensor(a, dtype=dtype or torch.float64, device="cuda")


def _oracle_vmult_check(pb, op, torch, threads=16):
    """One full-size vmult against the oracle (relative max-norm error) and the
    oracle's own time for it."""
    from oracle import cpu
    cpu.set_threads(threads)
    u = pb.random(0)
    t0 = time.perf_counter()
    ref = pb.vmult(u)
    t_cpu = time.perf_counter() - t0
    src = _t(u, torch)
    dst = torch.empty_like(src)
    op.vmult(dst, src)
    torch.cuda.synchronize()
    err = float(np.abs(dst.cpu().numpy() - ref).max() / np.abs(ref).max())
    return err, t_cpu


def _vmult_time(op, torch, dtype, n_rep=20):
    N = op.n_dofs()
    src = torch.rand(N, dtype=dtype, device="cuda")
    dst = torch.empty_like(src)
    op.vmult(dst, src)
    ms = op.time_vmult(dst, src, n_rep)
    return ms


def _solve(op, b, torch, **kw):
    """Solve from x = 0 after one untimed warm-up solve (workspace allocation)."""
    from meltpooldg_amd import LinearSolver, LinearSolverData
    x = torch.zeros_like(b)
    data = LinearSolverData(preconditioner_type="Diagonal", **kw)
    LinearSolver.solve(op, x, b, data)
    x.zero_()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    LinearSolver.solve(op, x, b, data)
    torch.cuda.synchronize()
    t = time.perf_counter() - t0
    rep = dict(LinearSolver.last_report)
    rep["wall_s"] = t
    return x, rep


def _oracle_reference(pb, b, solver, inv_diag, rel_tol=1e-10):
    """The oracle's FP64 solution of the same system (Jacobi with inv_diag)."""
    t0 = time.perf_counter()
    bn = b.cpu().numpy()
    ref = pb.solve(bn, solver=solver, restart=30, rel_tol=rel_tol, inv_diag=inv_diag)
    ref["s"] = time.perf_counter() - t0
    ref["b"] = bn
    return ref


def _oracle_errors(pb, x, b, ref):
    """Error of a GPU solution against the oracle's FP64 solution (max norm,
    relative) and its relative residual in the oracle's FP64 operator."""
    xn = x.cpu().numpy()
    r = ref["b"] - pb.vmult(xn)
    return {"err_vs_oracle_fp64": float(np.abs(xn - ref["x"]).max() / np.abs(ref["x"]).max()),
            "oracle_rel_residual": float(np.linalg.norm(r) / np.linalg.norm(ref["b"]))}


def config0(args, torch):
    from meltpooldg_amd import HeatDiffuseOperator
    from oracle import cpu
    n, p = 16, 2
    op = HeatDiffuseOperator(n, p)
    op.reset_time_increment(1e-3)
    pb = cpu.Problem(n, p, dt=1e-3)
    err, t_cpu_vmult = _oracle_vmult_check(pb, op, torch)
    b = pb.random(1)
    t0 = time.perf_counter()
    ref = pb.solve(b, rel_tol=1e-10, inv_diag=pb.inverse_diagonal())
    t_cpu_solve = time.perf_counter() - t0
    x, rep = _solve(op, _t(b, torch), torch, solver_type="CG", rel_tolerance=1e-10)
    xerr = float(np.abs(x.cpu().numpy() - ref["x"]).max() / np.abs(ref["x"]).max())
    ms = _vmult_time(op, torch, torch.float64)
    return {"config": 0, "workload": "16^3 Q2 gas, CG+Jacobi to 1e-10, FP64", "n_dofs": op.n_dofs(),
            "vmult_rel_err_vs_oracle": err, "gpu_vmult_ms": ms, "cpu_vmult_ms": t_cpu_vmult * 1e3,
            "gpu_cg_iterations": rep["iterations"], "oracle_cg_iterations": ref["iterations"],
            "gpu_cg_s": rep["wall_s"], "cpu_cg_s": t_cpu_solve, "solution_rel_err": xerr}


def config2(args, torch):
    from meltpooldg_amd import HeatDiffuseOperator, MaterialData, PhaseData
    from oracle import cpu
    n, p = 96, 3
    from meltpooldg_amd.inputs import sphere_heaviside
    H = sphere_heaviside(n, p)
    mat = MaterialData(type="gas_liquid", gas=PhaseData(1.0, 1.0, 1.0),
                       liquid=PhaseData(10.0, 1.0, 100.0))
    results = {"config": 2, "workload": "96^3 Q3 two-phase sphere, k 1:100, rho*cp 1:10, CG+Jacobi to 1e-8"}
    op = HeatDiffuseOperator(n, p, material=mat)
    op.reset_time_increment(1e-3)
    op.set_heaviside(_t(H, torch))
    op32 = HeatDiffuseOperator(n, p, material=mat, precision="fp32")
    op32.reset_time_increment(1e-3)
    op32.set_heaviside(_t(H, torch))
    results["tuning"] = {"fp64": op.autotune("fp64", repeats=3), "fp32": op.autotune("fp32", repeats=3),
                         "fp32_operator": op32.autotune("fp32", repeats=3)}
    pb = cpu.Problem(n, p, dt=1e-3, material={"type": "gas_liquid", "gas_rho": 1.0, "gas_cp": 1.0,
                                              "gas_k": 1.0, "liq_rho": 10.0, "liq_cp": 1.0,
                                              "liq_k": 100.0}, heaviside=H)
    err, t_cpu_vmult = _oracle_vmult_check(pb, op, torch)
    b = _t(pb.random(1), torch)
    x64, rep64 = _solve(op, b, torch, solver_type="CG", rel_tolerance=1e-8)
    # inner FP32 solves asked for a 1e-3 reduction: three outer corrections and
    # 687 inner iterations, against four and 776 at the 1e-2 default (0.322 vs
    # 0.365 s); 1e-4 took 720
    xmx, repmx = _solve(op, b, torch, solver_type="CG", rel_tolerance=1e-8, mixed_precision=True,
                        inner_rel_tolerance=1e-3)
    # reference: the oracle's FP64 CG + Jacobi solution to 1e-10 (the discrete
    # solution the 1e-8 solves approximate) and the oracle's own residuals
    ref = _oracle_reference(pb, b, "CG", pb.inverse_diagonal())
    e = lambda x: _oracle_errors(pb, x, b, ref)
    results.update({"n_dofs": op.n_dofs(), "vmult_rel_err_vs_oracle": err,
                    "cpu_vmult_ms": t_cpu_vmult * 1e3,
                    "gpu_vmult_ms_fp64": _vmult_time(op, torch, torch.float64),
                    "gpu_vmult_ms_fp32": _vmult_time(op32, torch, torch.float32),
                    "oracle_fp64_cg_1e-10": {"iterations": ref["iterations"], "s": ref["s"]},
                    "fp64": {"s": rep64["wall_s"], "iterations": rep64["iterations"],
                             **e(x64)},
                    "fp32_mixed": {"s": repmx["wall_s"], "inner_iterations": repmx["iterations"],
                                   "outer_iterations": repmx["outer_iterations"], **e(xmx)},
                    "speedup_mixed_vs_fp64": rep64["wall_s"] / repmx["wall_s"]})
    return results


def config4(args, torch):
    from meltpooldg_amd import HeatDiffuseOperator, MaterialData, PhaseData
    from oracle import cpu
    n, p = 64, 5
    mat = MaterialData(gas=PhaseData(1.0, 1.0, 1e-2))
    op = HeatDiffuseOperator(n, p, material=mat, velocity="vortex", velocity_w0=0.2)
    op.reset_time_increment(5e-4)
    op32 = HeatDiffuseOperator(n, p, material=mat, velocity="vortex", velocity_w0=0.2, precision="fp32")
    op32.reset_time_increment(5e-4)
    tuning = {"fp64": op.autotune("fp64", repeats=3), "fp32": op.autotune("fp32", repeats=3),
              "fp32_operator": op32.autotune("fp32", repeats=3)}
    pb = cpu.Problem(n, p, dt=5e-4, material={"type": "gas", "gas_rho": 1.0, "gas_cp": 1.0,
                                              "gas_k": 1e-2}, velocity=True, w0=0.2)
    err, t_cpu_vmult = _oracle_vmult_check(pb, op, torch)
    b = _t(pb.random(1), torch)
    x64, rep64 = _solve(op, b, torch, solver_type="GMRES", rel_tolerance=1e-8)
    xmx, repmx = _solve(op, b, torch, solver_type="GMRES", rel_tolerance=1e-8, mixed_precision=True)
    from oracle.kron import KronOperator
    inv = KronOperator(n, p, dt=5e-4, k=1e-2, velocity=True, w0=0.2).inverse_diagonal()
    ref = _oracle_reference(pb, b, "GMRES", inv)
    e = lambda x: _oracle_errors(pb, x, b, ref)
    return {"config": 4, "workload": "64^3 Q5 convection-diffusion (vortex), GMRES(30)+Jacobi to 1e-8",
            "n_dofs": op.n_dofs(), "tuning": tuning, "vmult_rel_err_vs_oracle": err, "cpu_vmult_ms": t_cpu_vmult * 1e3,
            "gpu_vmult_ms_fp64": _vmult_time(op, torch, torch.float64),
            "gpu_vmult_ms_fp32": _vmult_time(op32, torch, torch.float32),
            "oracle_fp64_gmres_1e-10": {"iterations": ref["iterations"], "s": ref["s"]},
            "fp64": {"s": rep64["wall_s"], "iterations": rep64["iterations"], **e(x64)},
            "fp32_mixed": {"s": repmx["wall_s"], "inner_iterations": repmx["iterations"],
                           "outer_iterations": repmx["outer_iterations"], **e(xmx)},
            "speedup_mixed_vs_fp64": rep64["wall_s"] / repmx["wall_s"]}


def robin(args, torch):
    from meltpooldg_amd import (ConvectionBC, HeatDiffuseOperation, HeatDiffuseOperator,
                                LinearSolverData, MaterialData, NonlinearSolverData, PhaseData,
                                RadiationBC)
    from oracle import cpu
    n, p, dt = 64, 4, 1e-2
    steel = MaterialData(gas=PhaseData(7850.0, 460.0, 55.563))
    conv, rad = ConvectionBC(25.0, 300.0), RadiationBC(0.98, 300.0)
    top = 5
    op = HeatDiffuseOperator(n, p, material=steel, dirichlet_faces=0, convection=conv,
                             radiation=rad, convection_boundary_ids=[top],
                             radiation_boundary_ids=[top])
    op.reset_time_increment(dt)
    plain = HeatDiffuseOperator(n, p, material=steel, dirichlet_faces=0)
    plain.reset_time_increment(dt)
    N = op.n_dofs()
    T = torch.full((N,), 1000.0, dtype=torch.float64, device="cuda")
    op.set_temperature(T)
    ms_faces = _vmult_time(op, torch, torch.float64, 50)
    ms_plain = _vmult_time(plain, torch, torch.float64, 50)
    pb = cpu.Problem(n, p, dt=dt, material={"type": "gas", "gas_rho": 7850.0, "gas_cp": 460.0,
                                            "gas_k": 55.563}, dirichlet_faces=0,
                     robin=dict(convection_faces=1 << top, convection_coefficient=25.0,
                                convection_t_inf=300.0, radiation_faces=1 << top,
                                emissivity=0.98, radiation_t_inf=300.0),
                     temperature=np.full(N, 1000.0))
    err, t_cpu_vmult = _oracle_vmult_check(pb, op, torch)
    # absolute Newton tolerances sized for 17 M DoFs: the residual's round-off
    # floor (entries ~ rho_cp/dt |cell| T, summed over the lattice) is ~1e-9,
    # the NonlinearSolverData default
    nl = NonlinearSolverData(residual_tolerance=1e-7, field_correction_tolerance=1e-9)
    heat = HeatDiffuseOperation(op, T, nlsolve=nl, linear_solver=LinearSolverData(
        solver_type="GMRES", preconditioner_type="Diagonal", rel_tolerance=1e-12))
    heat.advance(dt)  # warm-up step (workspace allocation)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    rep = heat.advance(dt)
    torch.cuda.synchronize()
    t_step = time.perf_counter() - t0
    face_nodes = (n * p + 1) ** 2
    return {"mode": "robin", "workload": "64^3 Q4 steel (7850, 460, 55.563), dt 0.01, top face "
                                         "convection 25 / radiation 0.98 (T_inf 300), T=1000, "
                                         "insulated elsewhere",
            "n_dofs": N, "face_nodes": face_nodes,
            "gpu_vmult_ms_with_faces": ms_faces, "gpu_vmult_ms_without_faces": ms_plain,
            "face_share_of_vmult": (ms_faces - ms_plain) / ms_faces,
            "vmult_rel_err_vs_oracle": err, "cpu_vmult_ms": t_cpu_vmult * 1e3,
            "newton_step_s": t_step, "newton": rep}


def solidification(args, torch):
    """Tangent of a temperature-dependent (gas_liquid_solid) material at full size:
    64^3 Q4 steel with a gas phase, temperature crossing the mushy zone, a
    heaviside of a plane level set and a constant velocity."""
    from meltpooldg_amd import (HeatDiffuseOperation, HeatDiffuseOperator, LinearSolverData,
                                MaterialData, NonlinearSolverData, PhaseData)
    from oracle import cpu
    n, p, dt = 64, 4, 1e-2
    gas, liq, sol = (3925.0, 230.0, 27.7815), (7850.0, 460.0, 55.563), (7850.0, 700.0, 17.0)
    mat = MaterialData(type="gas_liquid_solid", gas=PhaseData(*gas), liquid=PhaseData(*liq),
                       solid=PhaseData(*sol), two_phase_fluid_properties_transition_type="sharp",
                       solidus_temperature=1966.0, liquidus_temperature=1974.0)
    vel = (-0.005, 0.0, 0.0)
    op = HeatDiffuseOperator(n, p, material=mat, dirichlet_faces=0, velocity="constant",
                             velocity_value=vel)
    op.reset_time_increment(dt)
    plain = HeatDiffuseOperator(n, p, material=MaterialData(type="gas_liquid", gas=PhaseData(*gas),
                                                            liquid=PhaseData(*liq)),
                                dirichlet_faces=0, velocity="constant", velocity_value=vel)
    plain.reset_time_increment(dt)
    N = op.n_dofs()
    n1 = n * p + 1
    x = np.linspace(0.0, 1.0, n1)
    Z, Y, X = np.meshgrid(x, x, x, indexing="ij")
    T = (1960.0 + 20.0 * X + 2.0 * np.sin(2 * np.pi * Y) * np.sin(2 * np.pi * Z)).reshape(-1)
    To = T - 0.1
    H = np.clip(0.5 + (0.7 - Z.reshape(-1)) / 0.04, 0.0, 1.0)
    Tg, Tog, Hg = _t(T, torch), _t(To, torch), _t(H, torch)
    op.set_heaviside(Hg)
    plain.set_heaviside(Hg)
    op.set_temperature(Tg)
    op.set_temperature_old(Tog)
    ms_tmat = _vmult_time(op, torch, torch.float64, 30)
    ms_plain = _vmult_time(plain, torch, torch.float64, 30)
    inv = torch.empty_like(Tg)
    op.compute_inverse_diagonal_from_matrixfree(inv)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(10):
        op.compute_inverse_diagonal_from_matrixfree(inv)
    torch.cuda.synchronize()
    ms_diag = (time.perf_counter() - t0) / 10 * 1e3
    mdict = {"type": "gas_liquid_solid", "gas_rho": gas[0], "gas_cp": gas[1], "gas_k": gas[2],
             "liq_rho": liq[0], "liq_cp": liq[1], "liq_k": liq[2], "sol_rho": sol[0],
             "sol_cp": sol[1], "sol_k": sol[2], "solidus": 1966.0, "liquidus": 1974.0, "sharp": 1}
    pb = cpu.Problem(n, p, dt=dt, material=mdict, heaviside=H, dirichlet_faces=0,
                     velocity="constant", velocity_value=vel, temperature=T, temperature_old=To)
    err, t_cpu_vmult = _oracle_vmult_check(pb, op, torch)
    # a time step from T (old = T): absolute tolerances sized for 17 M DoFs
    nl = NonlinearSolverData(residual_tolerance=1e-6, field_correction_tolerance=1e-8)
    heat = HeatDiffuseOperation(op, Tg, nlsolve=nl, linear_solver=LinearSolverData(
        solver_type="GMRES", preconditioner_type="Diagonal", rel_tolerance=1e-10))
    heat.advance(dt)  # warm-up step (workspace allocation)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    rep = heat.advance(dt)
    torch.cuda.synchronize()
    t_step = time.perf_counter() - t0
    bytes_alg = 5 * 8 * N  # src, dst, T, T_old, H
    return {"mode": "solidification",
            "workload": "64^3 Q4 gas_liquid_solid steel (solidus 1966, liquidus 1974), T "
                        "crossing the mushy zone, plane heaviside, velocity (-0.005, 0, 0), dt 0.01",
            "n_dofs": N, "gpu_vmult_ms": ms_tmat, "gpu_vmult_gdofs": N / ms_tmat / 1e6,
            "algorithmic_bytes_per_vmult": bytes_alg,
            "algorithmic_gbs": bytes_alg / ms_tmat / 1e6,
            "gpu_vmult_ms_gas_liquid_same_fields": ms_plain, "gpu_diagonal_ms": ms_diag,
            "vmult_rel_err_vs_oracle": err, "cpu_vmult_ms": t_cpu_vmult * 1e3,
            "newton_step_s": t_step, "newton": rep}


def timestep(args, torch):
    """The reference's own unidirectional_heat_transfer regression inputs that the
    engine runs (tests/golden/unidirectional_cases.json) through the JSON front
    end, each whole run timed with the Krylov restart cycles replayed as
    hipGraphs and with eager launches (after one untimed warm-up run)."""
    from meltpooldg_amd import frontend as F
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tests", "golden",
                        "unidirectional_cases.json")
    cases = json.load(open(path))["cases"]
    out = []
    for name, c in sorted(cases.items()):
        try:
            prm = F.parse_parameters(c["input"])
            F.setup(prm)
        except (F.UnsupportedFeature, F.UnknownParameter):
            continue
        row = {"input": name}
        for graphs in (True, False):
            F.run(prm, log=lambda s: None, graphs=graphs)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            rec = F.run(prm, log=lambda s: None, graphs=graphs)
            torch.cuda.synchronize()
            key = "graphs" if graphs else "eager"
            row[key + "_s"] = time.perf_counter() - t0
            row["steps"] = len(rec)
            row["linear_iterations"] = sum(r["newton"]["linear_iterations"] for r in rec)
        row["speedup"] = row["eager_s"] / row["graphs_s"]
        out.append(row)
    return {"mode": "timestep", "workload": "the reference's unidirectional_heat_transfer "
                                            "regression inputs (2D / 1D Q1, extruded)",
            "runs": out}


def q2(args, torch):
    """Low-degree roofline leg (bench.py --mode q2): the FP32 operator of the
    mixed-precision solve (and the FP64 one) at degree 2 on 96^3 and 192^3
    cells, unit cube, separable-sine k and a constant k, Dirichlet on all
    faces, the default tuning. Each vmult's algorithmic bytes (src read + dst
    written, b N each) and flops (cells x (12 n_q^4 + 36 n_q^3), n_q = 3) are
    set against MI355X's 8 TB/s and against the vector ALU's non-packed rate
    (78.6 TFLOP/s for FP32 as for FP64: packed FP32 is disabled in these
    kernels). At 96^3 the FP32 vectors (2 x 28.8 MB) stay in the 256 MB
    last-level cache; at 192^3 (2 x 227 MB) they do not."""
    from meltpooldg_amd import HeatDiffuseOperator
    peak_tb, peak_tf = 8.0, 78.6
    rows = []
    for n in [int(c) for c in args.cells.split(",")]:
        for coef in ("sine", "constant"):
            kw = {"conductivity_amplitude": 0.5} if coef == "sine" else {}
            op = HeatDiffuseOperator(n, 2, **kw)
            op.reset_time_increment(1e-3)
            N = op.n_dofs()
            cells = n ** 3
            flops = cells * (12 * 3 ** 4 + 36 * 3 ** 3)
            for prec, dtype, b in (("fp32", torch.float32, 4), ("fp64", torch.float64, 8)):
                src = torch.rand(N, dtype=dtype, device="cuda")
                dst = torch.empty_like(src)
                t_end = time.perf_counter() + 0.3
                while time.perf_counter() < t_end:
                    for _ in range(10):
                        op.vmult(dst, src)
                    torch.cuda.synchronize()
                ms = op.time_vmult(dst, src, 100)
                gbs = 2 * b * N / (ms * 1e-3) / 1e9
                tfs = flops / (ms * 1e-3) / 1e12
                rows.append({"cells": n, "k": coef, "precision": prec, "n_dofs": N,
                             "ms_per_vmult": round(ms, 5), "GDoF/s": round(N / ms / 1e6, 2),
                             "algorithmic_GB/s": round(gbs, 1), "hbm_frac": round(gbs / 1e3 / peak_tb, 4),
                             "TFLOP/s": round(tfs, 2), "valu_frac": round(tfs / peak_tf, 4),
                             "tuning": op.get_tuning(prec)})
            del op
    return {"mode": "q2", "workload": f"Q2 vmult, unit cube, {args.cells} cells per direction, sine and constant k",
            "peaks": {"hbm_TB/s": peak_tb, "valu_TFLOP/s_nonpacked": peak_tf}, "rows": rows}


def run(config, args):
    import torch
    fn = {0: config0, 2: config2, 4: config4}
    if config == 3:
        import bench_batch
        out = bench_batch.run(args, torch)
    elif config in fn:
        out = fn[config](args, torch)
    else:
        raise SystemExit(f"unknown config {config}")
    print(json.dumps(out))
