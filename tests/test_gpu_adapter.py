"""GPU checks of the drop-in boundary beyond the operator itself:

* the header-only C++ adapter (include/mpdg_heat_operator.hpp) driving a device
  operator from a host program (tests/adapter_gpu.cpp): vmult, a
  JacobiPreconditioner updated by the caller, and both LinearSolver::solve
  overloads, against the oracle;
* JacobiPreconditioner::vmult (mpdg_heat_jacobi_vmult): dst = inv_diag .* src,
  bitwise;
* LinearSolver::solve with a caller-owned preconditioner
  (mpdg_heat_solve_preconditioned): the caller's inverse diagonal is applied as
  it stands, so a diagonal computed once serves several solves, and a
  different diagonal gives the oracle's iterations with that diagonal.

Tolerances: vmult 1e-12 relative (max norm); CG iteration counts equal to the
oracle's; solutions within 1e-9 relative of the oracle's.
"""
import os
import subprocess

import numpy as np
import pytest

from tests.cases import engine_operator, make_case, oracle_problem, rel_err

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


def _t(a):
    import torch
    return torch.tensor(a, dtype=torch.float64, device="cuda")


def _np(t):
    import torch
    torch.cuda.synchronize()
    return t.cpu().numpy()


def _adapter_binary():
    exe = os.path.join(HERE, "_build", "adapter_gpu")
    if not os.path.exists(exe):
        subprocess.run(["make", "-s", "-C", HERE], check=True)
    return exe


def test_cpp_adapter_vmult_and_solves_on_device(tmp_path):
    n, p = 6, 3
    case = make_case(n, p, k_amplitude=0.5)
    pb = oracle_problem(case)
    src, rhs = pb.random(0), pb.random(1)
    (tmp_path / "src.bin").write_bytes(src.tobytes())
    (tmp_path / "rhs.bin").write_bytes(rhs.tobytes())
    out = tmp_path / "out.bin"
    print("adapter_gpu", n, p, flush=True)
    r = subprocess.run([_adapter_binary(), str(n), str(p), str(tmp_path / "src.bin"),
                        str(tmp_path / "rhs.bin"), str(out)], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, (r.returncode, r.stdout, r.stderr)
    res = np.frombuffer(out.read_bytes(), dtype=np.float64).reshape(3, pb.N)
    assert rel_err(res[0], pb.vmult(src)) < 1e-12
    ref = pb.solve(rhs, rel_tol=1e-10, inv_diag=pb.inverse_diagonal())
    it1, it2 = (int(v) for v in r.stdout.split()[1:3])
    assert it1 == it2 == ref["iterations"], (it1, it2, ref["iterations"])
    assert rel_err(res[1], ref["x"]) < 1e-9
    assert rel_err(res[2], ref["x"]) < 1e-9


def test_jacobi_vmult_scales_by_inverse_diagonal():
    import torch
    from meltpooldg_amd import JacobiPreconditioner
    case = make_case((5, 4, 3), 2, k_amplitude=0.5)
    pb = oracle_problem(case)
    op = engine_operator(case)
    prec = JacobiPreconditioner(op)
    prec.update()
    inv = _np(prec.inverse_diag)
    assert rel_err(inv, pb.inverse_diagonal()) < 1e-12
    src = pb.random(3) + 0.25  # nonzero on constrained rows too
    dst = torch.full((pb.N,), float("nan"), dtype=torch.float64, device="cuda")
    prec.vmult(dst, _t(src))
    assert np.array_equal(_np(dst), inv * src)


def test_solve_applies_the_callers_preconditioner():
    import torch
    from meltpooldg_amd import JacobiPreconditioner, LinearSolver, LinearSolverData
    case = make_case(8, 3, k_amplitude=0.5)
    pb = oracle_problem(case)
    op = engine_operator(case)
    inv_ref = pb.inverse_diagonal()
    prec = JacobiPreconditioner(op)
    prec.update()
    data = LinearSolverData(solver_type="CG", preconditioner_type="Diagonal", rel_tolerance=1e-10)
    # one update, two solves (a time step's Newton increments reuse it)
    for seed in (1, 2):
        b = pb.random(seed)
        ref = pb.solve(b, rel_tol=1e-10, inv_diag=inv_ref)
        x = torch.zeros(pb.N, dtype=torch.float64, device="cuda")
        it = LinearSolver.solve(op, x, _t(b), data, prec)
        assert it == ref["iterations"], (seed, it, ref["iterations"])
        assert rel_err(_np(x), ref["x"]) < 1e-9
    # the caller's vector is applied as it stands: a diagonal of ones is
    # PreconditionIdentity, not the recomputed Jacobi diagonal
    prec.inverse_diag.fill_(1.0)
    b = pb.random(1)
    ref_id = pb.solve(b, rel_tol=1e-10, inv_diag=np.ones(pb.N))
    assert ref_id["iterations"] != pb.solve(b, rel_tol=1e-10, inv_diag=inv_ref)["iterations"]
    x = torch.zeros(pb.N, dtype=torch.float64, device="cuda")
    assert LinearSolver.solve(op, x, _t(b), data, prec) == ref_id["iterations"]
    assert rel_err(_np(x), ref_id["x"]) < 1e-9
    x.zero_()
    assert LinearSolver.solve(op, x, _t(b), data, "Identity") == ref_id["iterations"]
    # the mixed-precision solve takes the caller's diagonal as well
    prec.update()
    data_m = LinearSolverData(solver_type="CG", preconditioner_type="Diagonal",
                              rel_tolerance=1e-10, mixed_precision=True)
    x.zero_()
    LinearSolver.solve(op, x, _t(b), data_m, prec)
    ref = pb.solve(b, rel_tol=1e-12, inv_diag=inv_ref)
    rep = LinearSolver.last_report
    assert rep["final_residual"] <= 1e-10 * rep["initial_residual"]
    assert rel_err(_np(x), ref["x"]) < 1e-8


def test_preconditioner_of_another_operator_is_refused():
    from meltpooldg_amd import JacobiPreconditioner, LinearSolver, LinearSolverData
    import torch
    case = make_case(3, 2)
    op1, op2 = engine_operator(case), engine_operator(case)
    prec = JacobiPreconditioner(op2)
    x = torch.zeros(op1.n_dofs(), dtype=torch.float64, device="cuda")
    with pytest.raises(ValueError, match="another operator"):
        LinearSolver.solve(op1, x, x.clone(), LinearSolverData(solver_type="CG"), prec)
    with pytest.raises(ValueError, match="not built"):
        LinearSolver.solve(op1, x, x.clone(), LinearSolverData(solver_type="CG"), "ILU")
