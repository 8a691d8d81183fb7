"""The allocation-free FE L2 norm (mpdg_heat_l2_norm_noalloc) through the C
ABI against mpdg_heat_l2_norm, the entry point it is a variant of: two calls
in a row on one GPU, the second on a smaller mesh (the shared static scratch
keeps the first call's contents), each bitwise equal to the allocating
entry point's norm; then the first mesh once more. A mesh that needs more
reduction workgroups than the scratch holds is refused before any launch."""
import ctypes

import numpy as np
import pytest

from meltpooldg_amd.inputs import random_vector

pytestmark = pytest.mark.gpu


def _norms(n, p, seed):
    import torch
    from meltpooldg_amd import HeatDiffuseOperator
    from meltpooldg_amd import _lib as L
    op = HeatDiffuseOperator(n, p, conductivity_amplitude=0.5)
    op.reset_time_increment(1e-3)
    v = torch.tensor(random_vector(n, p, seed, dirichlet_faces=0) + 0.5, dtype=torch.float64,
                     device="cuda")
    out = torch.full((1,), float("nan"), dtype=torch.float64, device="cuda")
    print("no-alloc L2 norm", n, p, seed, flush=True)
    L.check(op._lib.mpdg_heat_l2_norm_noalloc(op._h, ctypes.c_void_p(v.data_ptr()),
                                               ctypes.c_void_p(out.data_ptr())))
    torch.cuda.synchronize()
    got = float(out.cpu()[0])
    ref = ctypes.c_double()
    L.check(op._lib.mpdg_heat_l2_norm(op._h, ctypes.c_void_p(v.data_ptr()), ctypes.byref(ref)))
    return got, ref.value


def test_noalloc_l2_norm_matches_the_allocating_entry_point():
    big = _norms((24, 20, 16), 3, 1)
    small = _norms((5, 4, 3), 2, 2)
    again = _norms((24, 20, 16), 3, 1)
    for got, ref in (big, small, again):
        assert np.isfinite(got) and got > 0.0
        assert got == ref, (got, ref)
    assert small[0] != big[0] and again[0] == big[0]


def test_noalloc_l2_norm_refuses_meshes_beyond_its_scratch():
    import torch
    from meltpooldg_amd import HeatDiffuseOperator
    from meltpooldg_amd import _lib as L
    # degree 1: 32-cell row segments (at most 2^19 segments x rows x layers)
    op = HeatDiffuseOperator((32, 1024, 520), 1, dirichlet_faces=0)
    op.reset_time_increment(1e-3)
    out = torch.zeros(1, dtype=torch.float64, device="cuda")
    st = op._lib.mpdg_heat_l2_norm_noalloc(op._h, ctypes.c_void_p(out.data_ptr()),
                                            ctypes.c_void_p(out.data_ptr()))
    assert st == L.MPDG_ERR_UNSUPPORTED
    assert "static scratch" in op._lib.mpdg_last_error().decode()
