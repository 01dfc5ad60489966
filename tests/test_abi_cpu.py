"""Host-side checks that need no GPU: the C ABI library loads and exports every
symbol that include/mesh_bem_amd.h declares (nothing is called that touches a
device), the ctypes table covers the header, and the host-side tables of the
Laplace path equal the oracle's."""

import ctypes
import os
import re

import numpy as np
import pytest

import oracle as O
from mesh_bem_amd import _lib
from mesh_bem_amd._build import LIB_PATH
from mesh_bem_amd.laplace import relation_tables

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "mesh_bem_amd.h")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"//[^\n]*", "", text)
    names = re.findall(r"\b(mba_[a-z0-9_]+)\s*\(", text)
    return sorted(set(names))


def test_header_declares_entry_points():
    names = declared_symbols()
    assert len(names) >= 30
    for must in ("mba_mom_zmat_faceedge_ref", "mba_mom_zmat_faceedge_ns", "mba_lu_factor", "mba_lu_solve",
                 "mba_laplace_layer_ref", "mba_laplace_p0_assemble"):
        assert must in names


@pytest.mark.skipif(not os.path.exists(LIB_PATH), reason="library not built (run __graft_entry__.build())")
def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(LIB_PATH)
    missing = [n for n in declared_symbols() if not hasattr(lib, n)]
    assert not missing, missing


def test_ctypes_table_matches_header():
    assert sorted(_lib.SIGNATURES) == declared_symbols()


@pytest.mark.skipif(not os.path.exists(LIB_PATH), reason="library not built")
def test_status_strings_and_sizes_without_device():
    lib = ctypes.CDLL(LIB_PATH)
    lib.mba_status_string.restype = ctypes.c_char_p
    lib.mba_status_string.argtypes = [ctypes.c_int]
    assert lib.mba_status_string(0) == b"ok"
    lib.mba_lu_workspace_bytes.restype = ctypes.c_size_t
    lib.mba_lu_workspace_bytes.argtypes = [ctypes.c_int, ctypes.c_int64]
    assert lib.mba_lu_workspace_bytes(0, 1000) > 0


@pytest.mark.parametrize("Q", [1, 2, 3, 4, 5])
def test_relation_tables_equal_oracle(Q):
    gp, gw = O.gl_ref(Q)
    tab, off = relation_tables(gp, gw)
    ot = O.laplace_tables(gp, gw)
    assert list(np.diff(off)) == [Q ** 4, 2 * Q ** 4, 5 * Q ** 4, 3 * Q ** 4]
    for r in range(4):
        assert np.array_equal(tab[off[r]:off[r + 1]], ot[r])


def test_product_has_no_cpu_fallback(monkeypatch):
    """Without a device, product entry points raise instead of computing."""
    import torch

    from mesh_bem_amd.dense import gemm

    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    A = torch.zeros((4, 4), dtype=torch.complex64)
    with pytest.raises(Exception):
        gemm(A, A)


def test_reference_ops_check_shapes_before_launch():
    """The reference-API ops reject mis-shaped operands with ValueError before
    any device check or launch (CPU tensors suffice)."""
    import torch

    from mesh_bem_amd.kernels import MeshMOMInterpolateJsKernel, MeshMOMRhsKernel, MeshMOMZmatKernel

    v = torch.zeros(4, 3, dtype=torch.float64)
    f = torch.tensor([[0, 1, 2], [0, 2, 3]], dtype=torch.int32)
    a = torch.ones(2, 1, dtype=torch.float64)
    n = torch.zeros(2, 3, dtype=torch.float64)
    g = torch.tensor([0.5], dtype=torch.float64)
    w = torch.tensor([1.0], dtype=torch.float64)
    with pytest.raises(ValueError):
        MeshMOMZmatKernel.apply(v.reshape(-1), f, a, n, g, w, g, w, 1.0)
    with pytest.raises(ValueError):
        MeshMOMZmatKernel.apply(v, f, a[:1], n, g, w, g, w, 1.0)
    with pytest.raises(ValueError):
        MeshMOMRhsKernel.apply(v, f, a, n[:, :2], g, w, 1.0)
    with pytest.raises(ValueError):
        MeshMOMRhsKernel.apply(v, f, a, n, g, torch.ones(2, dtype=torch.float64), 1.0)
    with pytest.raises(ValueError):
        MeshMOMInterpolateJsKernel.apply(v, f, torch.zeros(5, dtype=torch.int32), a,
                                         torch.zeros(1, 5, dtype=torch.complex128))


def test_mom_mesh_build_checks_indices():
    import numpy as np

    from mesh_bem_amd.mom import MomMesh

    v = np.zeros((4, 3))
    with pytest.raises(ValueError):
        MomMesh.build(v, np.array([[0, 1, 4]]), device="cpu")
    with pytest.raises(ValueError):
        MomMesh.build(v, np.array([[0, 1]]), device="cpu")


def test_laplace_entry_points_check_shapes():
    import torch

    from mesh_bem_amd import laplace

    v = torch.zeros(4, 3, dtype=torch.float64)
    f = torch.tensor([[0, 1, 2], [0, 2, 3]], dtype=torch.int32)
    with pytest.raises(ValueError):
        laplace.layer_ref(v[:, :2], f, 2, 1, 1)
    with pytest.raises(ValueError):
        laplace.pack(torch.zeros(2, 3, 3, 3, dtype=torch.float64), f, 4)
    with pytest.raises(ValueError):
        laplace.dirichlet_rhs(torch.zeros(4, 4, dtype=torch.float64), torch.zeros(3), torch.zeros(4))
    with pytest.raises(ValueError):
        laplace.p0_operators(v, f, 1, out_v=torch.zeros(3, 3))
