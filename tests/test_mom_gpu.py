"""GPU parity of the MoM kernels against the CPU FP64 oracle.

Bars: FP64 kernels agree with the oracle to 1e-11 (relative, Frobenius and
entrywise over entries >= 1e-3 max|Z|); FP32 kernels to 1e-5 (the Z-entry
tolerance of the mixed-precision path).
"""

import numpy as np
import pytest
import torch

import oracle as O
from mesh_bem_amd import _lib
from mesh_bem_amd.geometry import (compute_face_areas, compute_face_normals, get_edge_unique,
                                   icosphere, star_mesh)
from mesh_bem_amd.kernels import MeshMOMInterpolateJsKernel, MeshMOMRhsKernel, MeshMOMZmatKernel
from mesh_bem_amd.mom import MomMesh, assemble_ns, rhs_ns, zmat_faceedge_ns
from mesh_bem_amd.quadrature import get_gaussion_integration_points_and_weights

pytestmark = pytest.mark.gpu


def relerr(a, b, floor=1e-3):
    a = np.asarray(a)
    b = np.asarray(b)
    fro = np.linalg.norm(a - b) / np.linalg.norm(b)
    mask = np.abs(b) >= floor * np.abs(b).max()
    ent = (np.abs(a - b)[mask] / np.abs(b)[mask]).max()
    return fro, ent


def _mesh(dtype, dev, sub=1):
    V, F = icosphere(sub)
    return (torch.tensor(V, dtype=dtype, device=dev), torch.tensor(F, dtype=torch.int32, device=dev),
            V, F)


@pytest.mark.parametrize("dtype,tol", [(torch.float64, 1e-11), (torch.float32, 1e-5)])
@pytest.mark.parametrize("Q,k", [(4, 1.0), (3, 2.5)])
def test_ref_zmat_matches_oracle(dev, dtype, tol, Q, k):
    print(f"ref zmat dtype={dtype} Q={Q} k={k}", flush=True)
    v, f, V, F = _mesh(dtype, dev)
    fa = compute_face_areas(v, f, keepdim=True)
    fn = compute_face_normals(v, f)
    gpx, gwx = get_gaussion_integration_points_and_weights(Q, device=dev, dtype=dtype)
    gpy, gwy = get_gaussion_integration_points_and_weights(Q - 1, device=dev, dtype=dtype)
    Z = MeshMOMZmatKernel.apply(v, f, fa, fn, gpx, gwx, gpy, gwy, k)
    assert Z.shape == (1, 3 * len(F), 3 * len(F))
    gx, wx = O.gl_ref(Q)
    gy, wy = O.gl_ref(Q - 1)
    Zo = O.ref_zmat_faceedge(V if dtype == torch.float64 else v.double().cpu().numpy(), F, gx, wx, gy, wy, k)
    fro, ent = relerr(Z[0].cpu().numpy(), Zo)
    print(f"  fro={fro:.3e} ent={ent:.3e}")
    assert fro < tol and ent < 10 * tol


@pytest.mark.parametrize("dtype,tol", [(torch.float64, 1e-12), (torch.float32, 1e-5)])
def test_ref_rhs_matches_oracle(dev, dtype, tol):
    v, f, V, F = _mesh(dtype, dev)
    fa = compute_face_areas(v, f, keepdim=True)
    fn = compute_face_normals(v, f)
    gp, gw = get_gaussion_integration_points_and_weights(4, device=dev, dtype=dtype)
    R = MeshMOMRhsKernel.apply(v, f, fa, fn, gp, gw, 3.0)
    assert R.shape == (1, len(F), 3, 1)
    g, w = O.gl_ref(4)
    Ro = O.ref_rhs_faceedge(v.double().cpu().numpy(), F, g, w, 3.0)
    fro, ent = relerr(R.reshape(-1).cpu().numpy(), Ro.reshape(-1))
    assert fro < tol and ent < 10 * tol


@pytest.mark.parametrize("dtype,tol", [(torch.float64, 1e-13), (torch.float32, 1e-6)])
def test_interp_matches_oracle(dev, dtype, tol):
    v, f, V, F = _mesh(dtype, dev)
    edges, fe, _ = get_edge_unique(v, f)
    E = edges.shape[0]
    rng = np.random.default_rng(0)
    I = rng.standard_normal((2, E)) + 1j * rng.standard_normal((2, E))
    fa = compute_face_areas(v, f, keepdim=True)
    J = MeshMOMInterpolateJsKernel.apply(v, f, fe, fa, torch.tensor(I, device=dev))
    Jo = O.interp_faceedge(v.double().cpu().numpy(), F, fe.cpu().numpy(), I)
    fro, _ = relerr(J.cpu().numpy(), Jo)
    assert fro < tol


@pytest.mark.parametrize("dtype,tol", [(torch.float64, 1e-11), (torch.float32, 1e-5)])
@pytest.mark.parametrize("k", [1.0, 6.0])
def test_ns_faceedge_matches_oracle(dev, dtype, tol, k):
    print(f"ns faceedge dtype={dtype} k={k}", flush=True)
    V, F = icosphere(2)
    mesh = MomMesh.build(torch.tensor(V, dtype=dtype), F, device=dev)
    zf = zmat_faceedge_ns(mesh, k, _lib.prec_of(dtype))
    ptr = mesh.near_ptr.cpu().numpy()
    idx = mesh.near_idx.cpu().numpy()
    Vd = mesh.verts.double().cpu().numpy()
    Zo = O.ns_zmat_faceedge(Vd, F, k, ptr, idx)
    fro, ent = relerr(zf.cpu().numpy(), Zo)
    print(f"  fro={fro:.3e} ent={ent:.3e}")
    assert fro < tol and ent < 20 * tol


@pytest.mark.parametrize("dtype,tol", [(torch.float64, 1e-11), (torch.float32, 1e-5)])
def test_ns_rwg_matrix_and_windows(dev, dtype, tol):
    """RWG gather over several source windows equals the one-shot oracle system."""
    V, F, _ = star_mesh(3, 4)
    mesh = MomMesh.build(torch.tensor(V, dtype=dtype), F, device=dev)
    prec = _lib.prec_of(dtype)
    Z = assemble_ns(mesh, 3.0, prec, window_faces=37)
    Z1 = assemble_ns(mesh, 3.0, prec)
    # FP64 pairs with both faces in one window are evaluated once and
    # mirrored, so the windowing changes roundings only
    assert (Z - Z1).abs().max().item() <= 4 * tol * Z1.abs().max().item()
    if dtype == torch.float32:
        assert torch.equal(Z, Z1)
    dirs = np.array([[0, 0, 1.0], [0.6, 0, 0.8]])
    pols = np.array([[1.0, 0, 0], [0, 1.0, 0]])
    Vr = rhs_ns(mesh, 3.0, prec, dirs, pols)
    Zo, Vo, _, _ = O.ns_system(mesh.verts.double().cpu().numpy(), F, 3.0, dirs, pols)
    fro, ent = relerr(Z.cpu().numpy(), Zo)
    print(f"  Z fro={fro:.3e} ent={ent:.3e}")
    assert fro < tol and ent < 20 * tol
    fro, _ = relerr(Vr.cpu().numpy(), Vo)
    assert fro < tol


def test_ns_rwg_gather_large_windows(dev):
    """E >= 8192 (eight rows per gather workgroup): several windows give the
    single-window matrix to rounding, and sampled entries equal the four
    face-edge terms summed in the stated order from the face-edge matrix."""
    V, F, _ = star_mesh(2, 17)  # 5780 faces, 8670 RWG
    mesh = MomMesh.build(torch.tensor(V, dtype=torch.float64), F, device=dev)
    E, nF = mesh.n_edges, mesh.n_faces
    assert E >= 8192
    prec = _lib.MBA_F64
    Z1 = assemble_ns(mesh, 4.0, prec)
    Zw = assemble_ns(mesh, 4.0, prec, window_faces=1000)
    assert (Z1 - Zw).abs().max().item() <= 1e-13 * Z1.abs().max().item()
    zf = zmat_faceedge_ns(mesh, 4.0, prec, 0, nF)
    fe = mesh.rwg_fe.long()
    rng = np.random.default_rng(3)
    m = torch.tensor(rng.integers(0, E, 4096), device=dev)
    n = torch.tensor(rng.integers(0, E, 4096), device=dev)
    a0, a1, b0, b1 = fe[m, 0], fe[m, 1], fe[n, 0], fe[n, 1]
    want = (zf[a0, b0] + zf[a1, b0]) + (zf[a0, b1] + zf[a1, b1])
    assert torch.equal(Z1[m, n], want)


@pytest.mark.parametrize("dtype,tol", [(torch.float64, 1e-11)])
def test_ns_faceedge_symmetric(dev, dtype, tol):
    """FP64, one window over all faces: each far pair is evaluated once and mirrored,
    so far blocks are exactly symmetric (zf[3a+i, 3b+j] == zf[3b+j, 3a+i]);
    near blocks carry their own analytic corrections.  Both orders agree with
    the oracle, which evaluates every ordered pair directly."""
    V, F, _ = star_mesh(4, 5)
    mesh = MomMesh.build(torch.tensor(V, dtype=dtype), F, device=dev)
    prec = _lib.prec_of(dtype)
    zf = zmat_faceedge_ns(mesh, 2.5, prec, 0, len(F)).cpu().numpy()
    nF = len(F)
    near = np.zeros((nF, nF), dtype=bool)
    ptr, idx = mesh.near_ptr.cpu().numpy(), mesh.near_idx.cpu().numpy()
    for b in range(nF):
        near[idx[ptr[b]:ptr[b + 1]], b] = True
    assert np.array_equal(near, near.T)
    far = np.repeat(np.repeat(~near, 3, axis=0), 3, axis=1)
    assert np.array_equal(zf[far], zf.T[far])
    Vd = mesh.verts.double().cpu().numpy()
    Zo = O.ns_zmat_faceedge(Vd, F, 2.5, ptr, idx)
    fro, ent = relerr(zf, Zo)
    assert fro < tol and ent < 20 * tol
