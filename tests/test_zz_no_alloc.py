"""mba_mom_zmat_faceedge_ref_noalloc (the allocation-free variant of the
reference-rule face-edge assembly, create_MOM_Zmat_3d_forward,
src/mesh_BEM/cuda/mesh_MOM_3d.cpp:52-77) against mba_mom_zmat_faceedge_ref,
both through the C ABI: bit-identical blocks on two calls in a row, the
second on a smaller mesh (the shared device array keeps the first call's
records), in FP32 and FP64; meshes beyond the array's capacity are rejected
on the host before any launch."""

import ctypes

import numpy as np
import pytest
import torch

from mesh_bem_amd import _lib
from mesh_bem_amd.geometry import icosphere
from mesh_bem_amd.quadrature import get_gaussion_integration_points_and_weights

RECORD_BYTES = {_lib.MBA_F32: 96, _lib.MBA_F64: 192}
SCRATCH_BYTES = 4 << 20


def operands(sub, dtype, dev, Q=3):
    V, F = icosphere(sub)
    verts = torch.tensor(V, dtype=dtype, device=dev)
    faces = torch.tensor(np.asarray(F), dtype=torch.int32, device=dev)
    P = verts[faces.long()]
    cr = torch.cross(P[:, 1] - P[:, 0], P[:, 2] - P[:, 1], dim=-1)
    areas = (0.5 * cr.norm(dim=-1)).contiguous()
    normals = (cr / cr.norm(dim=-1, keepdim=True)).contiguous()
    gpx, gwx = get_gaussion_integration_points_and_weights(Q, dtype=dtype)
    gpy, gwy = get_gaussion_integration_points_and_weights(Q - 1, dtype=dtype)
    rule = [t.to(dev).contiguous() for t in (gpx, gwx, gpy, gwy)]
    return verts, faces, areas, normals, rule


def call(name, prec, verts, faces, areas, normals, rule, k, zf):
    lib = _lib.load()
    F = faces.shape[0]
    gpx, gwx, gpy, gwy = rule
    return getattr(lib, name)(prec, _lib.ptr(verts), _lib.ptr(faces), _lib.ptr(areas), _lib.ptr(normals),
                              _lib.ptr(gpx), _lib.ptr(gwx), gpx.numel(), _lib.ptr(gpy), _lib.ptr(gwy), gpy.numel(),
                              float(k), verts.shape[0], F, _lib.ptr(zf), zf.shape[1], _lib.stream())


@pytest.mark.gpu
@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_noalloc_equals_allocating_entry_point(dev, dtype):
    prec = _lib.prec_of(dtype)
    cdt = torch.complex64 if dtype == torch.float32 else torch.complex128
    for sub, k in ((3, 2.0), (1, 0.7)):  # second call smaller than the first
        ops = operands(sub, dtype, dev)
        F = ops[1].shape[0]
        print(f"noalloc parity: dtype={dtype} F={F} k={k}", flush=True)
        assert F * RECORD_BYTES[prec] <= SCRATCH_BYTES
        za = torch.full((3 * F, 3 * F), float("nan"), dtype=cdt, device=dev)
        zb = torch.full((3 * F, 3 * F), float("nan"), dtype=cdt, device=dev)
        assert call("mba_mom_zmat_faceedge_ref", prec, *ops, k, za) == 0
        assert call("mba_mom_zmat_faceedge_ref_noalloc", prec, *ops, k, zb) == 0
        torch.cuda.synchronize()
        assert torch.isfinite(torch.view_as_real(zb)).all()
        assert torch.equal(torch.view_as_real(za), torch.view_as_real(zb))


@pytest.mark.gpu
def test_noalloc_rejects_meshes_beyond_capacity(dev):
    """Size check on the host: no launch happens (the buffers are small, the
    announced face count is not)."""
    verts, faces, areas, normals, rule = operands(0, torch.float64, dev)
    lib = _lib.load()
    zf = torch.zeros((1, 1), dtype=torch.complex128, device=dev)
    F_big = SCRATCH_BYTES // RECORD_BYTES[_lib.MBA_F64] + 1
    gpx, gwx, gpy, gwy = rule
    st = lib.mba_mom_zmat_faceedge_ref_noalloc(
        _lib.MBA_F64, _lib.ptr(verts), _lib.ptr(faces), _lib.ptr(areas), _lib.ptr(normals), _lib.ptr(gpx),
        _lib.ptr(gwx), gpx.numel(), _lib.ptr(gpy), _lib.ptr(gwy), gpy.numel(), 1.0, verts.shape[0], F_big,
        _lib.ptr(zf), 3 * F_big, _lib.stream())
    assert st == 2  # MBA_ERR_SIZE
