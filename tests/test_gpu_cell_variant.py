"""GPU parity of the cell form of the vmult (one cell per lane, degree <= 2)
against the oracle, through the C ABI.

Same bar as the line form: FP64 within 1e-12 relative (max-norm scaled by the
max entry) of the FP64 oracle, FP32 within 2e-5, CG iteration counts equal to
the oracle's. Every tile shape is run on meshes that are not multiples of the
tile (partial tiles, single cells, partial Dirichlet face sets) and with
z-chunks that do not divide the mesh.
"""
import numpy as np
import pytest

from tests.cases import engine_operator, make_case, oracle_problem, rel_err

pytestmark = pytest.mark.gpu

TWO = {"type": "gas_liquid", "gas_rho": 1.0, "gas_cp": 1.0, "gas_k": 1.0,
       "liq_rho": 10.0, "liq_cp": 1.0, "liq_k": 100.0}


def _t(a, dtype=None):
    import torch
    return torch.tensor(a, dtype=dtype or torch.float64, device="cuda")


def _np(t):
    import torch
    torch.cuda.synchronize()
    return t.double().cpu().numpy()


def _vmult(op, src_np, fp32=False):
    import torch
    dt = torch.float32 if fp32 else torch.float64
    src = _t(src_np, dt)
    dst = torch.full_like(src, float("nan"))  # every entry must be written
    op.vmult(dst, src)
    return _np(dst)


MESHES = [
    dict(n=(37, 3, 2)),                               # partial tiles along x
    dict(n=(5, 11, 7), k_amplitude=0.5),              # partial tiles along y, sine k
    dict(n=(9, 4, 5), velocity=True, dt=5e-4),        # vortex convection
    dict(n=(1, 1, 1)),
    dict(n=(2, 1, 1), faces=0),
    dict(n=(6, 7, 3), faces=0b100101, velocity=True, k_amplitude=0.25),
    dict(n=(17, 9, 4), faces=0b011010),
]


@pytest.mark.parametrize("p", [1, 2])
@pytest.mark.parametrize("tile", [3, 4, 5, 6])
def test_cell_vmult_vs_oracle(p, tile):
    for m in MESHES:
        kw = dict(m)
        n = kw.pop("n")
        case = make_case(n, p, **kw)
        pb = oracle_problem(case)
        u = pb.random(0)
        ref = pb.vmult(u)
        op = engine_operator(case)
        for zl in (1, 3):
            print("cell", p, tile, n, kw, "z_layers", zl, flush=True)
            op.set_tuning(variant="cell", tile_log2=tile, z_layers=zl)
            assert op.get_tuning()["variant"] == "cell"
            assert rel_err(_vmult(op, u), ref) < 1e-12


@pytest.mark.parametrize("p", [1, 2])
@pytest.mark.parametrize("mat", ["smooth", "sharp", "volspec"])
def test_cell_two_phase_vs_oracle(p, mat):
    m = dict(TWO)
    if mat == "sharp":
        m["sharp"] = 1
    if mat == "volspec":
        m.update(volume_specific=1, liq_cp=3.0)
    n = (9, 5, 4)
    from oracle.rng import random_vector
    H = 0.5 * (random_vector(n, p, 11, dirichlet_faces=0) + 1.0)
    for vel in (False, True):
        case = make_case(n, p, material=m, heaviside=H, velocity=vel)
        print("cell two-phase", p, mat, vel, flush=True)
        pb = oracle_problem(case)
        op = engine_operator(case)
        op.set_tuning(variant="cell", tile_log2=3, z_layers=2)
        u = pb.random(0)
        assert rel_err(_vmult(op, u), pb.vmult(u)) < 1e-12


@pytest.mark.parametrize("p", [1, 2])
def test_cell_fp32(p):
    case = make_case((13, 6, 5), p, k_amplitude=0.5, velocity=True, dt=5e-4)
    print("cell fp32", p, flush=True)
    pb = oracle_problem(case)
    op = engine_operator(case, precision="fp32")
    op.set_tuning(variant="cell", tile_log2=4, z_layers=2)
    u = pb.random(0)
    out = _vmult(op, u.astype(np.float32), fp32=True)
    assert rel_err(out, pb.vmult(u)) < 2e-5


def test_cell_deterministic_bitwise():
    import torch
    case = make_case((40, 33, 9), 2, k_amplitude=0.5)
    op = engine_operator(case)
    op.set_tuning(variant="cell", tile_log2=4, z_layers=4)
    pb = oracle_problem(case)
    src = _t(pb.random(0))
    a = torch.empty_like(src)
    b = torch.empty_like(src)
    op.vmult(a, src)
    for _ in range(3):
        op.vmult(b, src)
        torch.cuda.synchronize()
        assert torch.equal(a, b)


@pytest.mark.parametrize("p", [1, 2])
def test_cell_cg_iterations_match_oracle(p):
    """CG + Jacobi with the cell form's fused p . Ap (the PCG path)."""
    import torch
    from meltpooldg_amd import LinearSolver, LinearSolverData
    case = make_case((16, 12, 10), p, k_amplitude=0.5)
    pb = oracle_problem(case)
    b = pb.random(1)
    ref = pb.solve(b, rel_tol=1e-10, inv_diag=pb.inverse_diagonal())
    op = engine_operator(case)
    op.set_tuning(variant="cell", tile_log2=4, z_layers=3)
    x = torch.zeros(op.n_dofs(), dtype=torch.float64, device="cuda")
    data = LinearSolverData(solver_type="CG", preconditioner_type="Diagonal", rel_tolerance=1e-10)
    it = LinearSolver.solve(op, x, _t(b), data)
    print("cell cg", p, "iterations", it, "oracle", ref["iterations"], flush=True)
    assert it == ref["iterations"]
    assert rel_err(_np(x), ref["x"]) < 1e-9


@pytest.mark.parametrize("p", [2, 4])
def test_autotune_installs_a_valid_tuning(p):
    case = make_case((20, 18, 16), p, k_amplitude=0.5)
    pb = oracle_problem(case)
    u = pb.random(0)
    op = engine_operator(case)
    for prec in ("fp64", "fp32"):
        best = op.autotune(precision=prec, repeats=2)
        print("autotune", p, prec, best, flush=True)
        assert best == op.get_tuning(prec)
        allowed = {"line"} | ({"cell"} if p <= 2 else set()) | \
            ({"mfma"} if p >= 3 and prec == "fp32" else set())
        assert best["variant"] in allowed
    assert rel_err(_vmult(op, u), pb.vmult(u)) < 1e-12


def test_cell_variant_rejected_above_degree_2():
    case = make_case((4, 4, 4), 3)
    op = engine_operator(case)
    with pytest.raises(Exception):
        op.set_tuning(variant="cell", tile_log2=4)
    with pytest.raises(Exception):
        make_op = engine_operator(make_case((4, 4, 4), 2))
        make_op.set_tuning(variant="cell", tile_log2=7)
