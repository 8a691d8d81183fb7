"""Numbering helpers on the device: a caller whose DoFs are numbered otherwise
(a DoFHandler's numbering, here a random permutation of the support points)
maps its vectors onto the engine's lattice and back. The permutation moves
values bitwise, and the operator applied in the caller's numbering
(from_lattice(vmult(to_lattice(x)))) is the oracle's operator in that
numbering, to the same 1e-12 bar as the lattice-numbered vmult."""
import numpy as np
import pytest

from tests.cases import engine_operator, make_case, oracle_problem, rel_err, support_points

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("index_dtype", ["int32", "uint32"])
@pytest.mark.parametrize("n,p", [((5, 4, 3), 3), ((2, 3, 2), 6)])
def test_vectors_in_a_caller_numbering(n, p, index_dtype):
    import torch
    from meltpooldg_amd import lattice_index_of_points
    lo, hi = (0.0, -0.5, 1.0), (2.0, 0.5, 1.6)
    case = make_case(n, p, dt=1e-2, k_amplitude=0.3, velocity="constant",
                     velocity_value=(0.7, -0.2, 0.4), faces=0b100101, lo=lo, hi=hi)
    op = engine_operator(case)
    pb = oracle_problem(case)
    N = op.n_dofs()
    order = np.random.default_rng(3).permutation(N)
    index = lattice_index_of_points(n, p, support_points(n, p, lo, hi)[order], box=(lo, hi))
    assert np.array_equal(index, order.astype(np.uint32))
    print("lattice", n, p, index_dtype, N, flush=True)
    idx = torch.from_numpy(index.astype(index_dtype)).to("cuda")
    x_d = np.random.default_rng(4).uniform(-1.0, 1.0, N)
    xd = torch.tensor(x_d, dtype=torch.float64, device="cuda")
    xl = torch.empty_like(xd)
    op.to_lattice(xl, xd, idx)
    back = torch.full_like(xd, np.nan)
    op.from_lattice(back, xl, idx)
    torch.cuda.synchronize()
    x_l = xl.cpu().numpy()
    assert np.array_equal(x_l[order], x_d)
    assert np.array_equal(back.cpu().numpy(), x_d)
    # the operator in the caller's numbering
    yl = torch.empty_like(xd)
    op.vmult(yl, xl)
    yd = torch.empty_like(xd)
    op.from_lattice(yd, yl, idx)
    torch.cuda.synchronize()
    ref = pb.vmult(x_l)
    assert rel_err(yd.cpu().numpy(), ref[order]) <= 1e-12


def test_index_tensor_is_checked():
    import torch
    case = make_case((2, 2, 2), 2)
    op = engine_operator(case)
    v = torch.zeros(op.n_dofs(), dtype=torch.float64, device="cuda")
    w = torch.zeros_like(v)
    ok = torch.arange(op.n_dofs(), dtype=torch.int32, device="cuda")
    for bad in (ok.to(torch.int64), ok[:-1], ok.cpu()):
        with pytest.raises(ValueError, match="lattice_index"):
            op.to_lattice(v, w, bad)
    with pytest.raises(RuntimeError, match="in place"):
        op.from_lattice(v, v, ok)
    op.to_lattice(v, w, ok)
