"""CPU checks of the C-ABI library: it loads, exports every entry point the
header declares, and rejects bad arguments before touching the GPU. Also
checks that the product package never reaches the oracle."""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    text = open(os.path.join(ROOT, "include", "mpdg_heat.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(mpdg_[a-z0-9_]+)\s*\(", text)))


@pytest.fixture(scope="module")
def lib():
    from meltpooldg_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        from meltpooldg_amd import build
        build.build()
    return _lib.load()


def test_exports_every_declared_symbol(lib):
    from meltpooldg_amd import _lib
    names = header_functions()
    assert len(names) >= 14
    for name in names:
        assert hasattr(lib, name), name
        assert name in _lib.SIGNATURES, f"{name} missing from the ctypes binding"
    assert set(_lib.SIGNATURES) == set(names)


def test_argument_errors_without_gpu(lib):
    from meltpooldg_amd import _lib as L
    h = ctypes.c_void_p()
    assert lib.mpdg_heat_create(None, ctypes.byref(h)) == L.MPDG_ERR_ARGUMENT
    assert lib.mpdg_last_error()
    d = L.HeatDesc()
    d.n_cells[:] = [4, 4, 4]
    d.box_hi[:] = [1.0, 1.0, 1.0]
    d.degree = 9
    assert lib.mpdg_heat_create(ctypes.byref(d), ctypes.byref(h)) == L.MPDG_ERR_UNSUPPORTED
    d.degree = 2
    d.n_cells[1] = 0
    assert lib.mpdg_heat_create(ctypes.byref(d), ctypes.byref(h)) == L.MPDG_ERR_ARGUMENT
    d.n_cells[1] = 4
    d.box_hi[2] = 0.0
    assert lib.mpdg_heat_create(ctypes.byref(d), ctypes.byref(h)) == L.MPDG_ERR_ARGUMENT
    d.box_hi[2] = 1.0
    d.material = 7
    assert lib.mpdg_heat_create(ctypes.byref(d), ctypes.byref(h)) == L.MPDG_ERR_UNSUPPORTED
    d.material = 0
    d.convection_faces = 0x40  # no seventh face
    assert lib.mpdg_heat_create(ctypes.byref(d), ctypes.byref(h)) == L.MPDG_ERR_ARGUMENT
    d.convection_faces = 0x02
    d.emissivity = float("nan")
    assert lib.mpdg_heat_create(ctypes.byref(d), ctypes.byref(h)) == L.MPDG_ERR_ARGUMENT
    d.emissivity = 0.0
    assert lib.mpdg_heat_set_temperature(None, None) == L.MPDG_ERR_ARGUMENT
    assert lib.mpdg_heat_set_dt(None, 1.0) == L.MPDG_ERR_ARGUMENT
    assert lib.mpdg_heat_vmult(None, None, None) == L.MPDG_ERR_ARGUMENT
    assert lib.mpdg_heat_n_dofs(None) == -1


def test_vector_size_limit_is_enforced(lib):
    """The cell kernels address each vector through a raw buffer with 32-bit
    byte sizes and offsets: a mesh whose FP64 vector reaches 2^31 bytes is
    refused before anything is allocated (no GPU needed), also for an FP32 cell
    loop (its FP64 outer vectors go through the same kernels), and also as a
    batch member."""
    from meltpooldg_amd import _lib as L
    h = ctypes.c_void_p()
    d = L.HeatDesc()
    d.box_hi[:] = [1.0, 1.0, 1.0]
    d.gas.density = d.gas.specific_heat_capacity = d.gas.thermal_conductivity = 1.0
    d.dirichlet_faces = 0x3F
    # 128^3 Q6: 769^3 = 454,756,609 DoFs, 3.6 GB per FP64 vector
    d.n_cells[:] = [128, 128, 128]
    d.degree = 6
    for prec in (L.MPDG_FP64, L.MPDG_FP32):
        d.precision = prec
        assert lib.mpdg_heat_create(ctypes.byref(d), ctypes.byref(h)) == L.MPDG_ERR_UNSUPPORTED
        msg = lib.mpdg_last_error().decode()
        assert "2^31 bytes" in msg and "454756609" in msg, msg
        assert not h.value
    # the first size past the limit: N1 = 646 gives 269,586,136 DoFs (> 2^28)
    d.precision = L.MPDG_FP64
    d.n_cells[:] = [645, 645, 645]
    d.degree = 1
    assert 646 ** 3 * 8 >= 2 ** 31 > 645 ** 3 * 8
    assert lib.mpdg_heat_create(ctypes.byref(d), ctypes.byref(h)) == L.MPDG_ERR_UNSUPPORTED
    descs = (L.HeatDesc * 2)(d, d)
    descs[0].n_cells[:] = [4, 4, 4]
    b = ctypes.c_void_p()
    assert lib.mpdg_heat_batch_create(descs, 2, ctypes.byref(b)) != L.MPDG_OK
    assert not b.value


def test_solver_defaults_follow_linear_solver_data(lib):
    from meltpooldg_amd import _lib as L
    p = L.SolverParams()
    lib.mpdg_solver_params_default(ctypes.byref(p))
    # LinearSolverData defaults (linear_solver_data.hpp:37-51), CG + Diagonal
    assert p.max_iterations == 10000
    assert p.rel_tolerance == 1e-12
    assert p.abs_tolerance == 1e-20
    assert p.solver_type == 0 and p.preconditioner_type == 3


def test_robin_boundary_ids_to_face_masks():
    from meltpooldg_amd.heat import _face_mask
    assert _face_mask(()) == 0
    assert _face_mask([1]) == 0b10
    assert _face_mask((0, 5, 1)) == 0b100011
    with pytest.raises(ValueError):
        _face_mask([6])


def test_python_mirror_defaults():
    from meltpooldg_amd import LinearSolverData
    d = LinearSolverData()
    assert d.solver_type == "GMRES" and d.preconditioner_type == "Identity"
    assert d.max_iterations == 10000 and d.rel_tolerance == 1e-12 and d.abs_tolerance == 1e-20


def test_product_never_imports_oracle():
    pkg = os.path.join(ROOT, "meltpooldg_amd")
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".cpp", ".hip", ".hpp", ".h")):
                text = open(os.path.join(dirpath, f)).read()
                assert "oracle" not in text.lower(), os.path.join(dirpath, f)


def test_cpp_adapter_compiles_and_links(lib, tmp_path):
    """The header-only C++ adapter builds with g++ against the C ABI library and
    turns status codes into exceptions (error paths need no GPU)."""
    import subprocess
    exe = tmp_path / "adapter_check"
    libdir = os.path.join(ROOT, "meltpooldg_amd")
    r = subprocess.run(["g++", "-std=c++17", "-O1", "-I", os.path.join(ROOT, "include"),
                        os.path.join(ROOT, "tests", "adapter_check.cpp"), "-L", libdir,
                        "-lmpdg_heat", f"-Wl,-rpath,{libdir}", "-o", str(exe)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([str(exe)], capture_output=True, text=True)
    assert r.returncode == 0 and "adapter ok" in r.stdout, (r.returncode, r.stdout, r.stderr)


def test_lattice_index_of_points_inverts_any_numbering():
    """mpdg_lattice_index_of_points (host, no GPU): a shuffled numbering of the
    box's support points (as a DoFHandler would give) maps back to the lattice;
    a point off the lattice or a repeated one is an error."""
    from meltpooldg_amd import heat
    from tests.cases import support_points
    n, p, lo, hi = (3, 2, 4), 3, (0.0, -1.0, 0.5), (1.5, 1.0, 2.0)
    pts = support_points(n, p, lo, hi)
    order = np.random.default_rng(1).permutation(len(pts))
    shuffled = pts[order] + np.random.default_rng(2).uniform(-1e-12, 1e-12, pts.shape)
    idx = heat.lattice_index_of_points(n, p, shuffled, box=(lo, hi))
    assert idx.dtype == np.uint32 and np.array_equal(idx, order.astype(np.uint32))
    bad = shuffled.copy()
    bad[5, 2] += 0.03  # z nodes are >= 0.10 apart here; a match allows a quarter of that
    with pytest.raises(RuntimeError, match="not a support point"):
        heat.lattice_index_of_points(n, p, bad, box=(lo, hi))
    dup = shuffled.copy()
    dup[7] = dup[8]
    with pytest.raises(RuntimeError, match="given twice"):
        heat.lattice_index_of_points(n, p, dup, box=(lo, hi))
    with pytest.raises(RuntimeError, match="count"):
        heat.lattice_index_of_points(n, p, shuffled[:-1], box=(lo, hi))
